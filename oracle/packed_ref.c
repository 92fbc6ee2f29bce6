/*
 * ORACLE (test infrastructure only): packed per-lane machine-program simulator.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * this library, and only as the checker; the product never links it.
 *
 * Restates rrtl_sim_ir::PackedSimulator (crates/rrtl-sim-ir/src/lib.rs:9033-10255):
 *   - storage: signals word-major values[(offset+limb)*lanes + lane]
 *     (value_index, :10045), memories lane-major
 *     memories[lane*total_mem_words + offset + addr*limbs + limb] (:10049-10054);
 *   - eval_combinational = Comb then AsyncResetComb (:9278-9281);
 *   - tick = eval, tick_from_evaluated_no_post_eval, eval (:9283-9293);
 *   - tick_from_evaluated_no_post_eval = capture tick_next into a buffer
 *     (capture_register_value :9647, reset value when the reset is asserted),
 *     run tick_commit (memory writes, :9740-9777), commit (:9688);
 *   - tick_many(n) = eval once, then n x {tick_from_evaluated, eval} (:9295-9307);
 *   - instruction semantics eval_machine_instr_lane (:9796-9907) with the limb
 *     helpers of :12942-13196 (fit_limbs, add/sub/mul_limbs, eq_values at the
 *     wider operand width, lt_values at the lhs width, slice/sext/concat,
 *     decode_usize saturating high limbs, decode_bool = bit 0).
 * Instruction-outer, lane-inner, exactly as the reference evaluates a block.
 *
 * Pinning: tests/test_oracle_pins.py checks every op against Python big-int
 * closed forms (exhaustive at small widths, random at 31..128 bits) and checks
 * whole designs against oracle/scalar_sim.py and closed forms.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "blob.h"

typedef struct { uint32_t sig; uint32_t limbs; uint32_t *vals; } p_capture;

typedef struct pref_sim {
    p_program P;
    p_signal *sig; uint32_t nsig;
    p_memory *mem; uint32_t nmem;
    uint32_t tsw, tmw;
    p_block *blk;
    uint32_t lanes;
    uint32_t *values;
    uint32_t *memories;
    p_capture *caps; uint32_t ncaps;
} pref_sim;

/* ----------------------------------------------------- per-lane value access */
static void ws_load(const p_block *bk, uint32_t lanes, uint32_t v, uint32_t lane, uint32_t *out, uint32_t want) {
    uint32_t have = nlimbs(bk->vwidth[v]);
    for (uint32_t l = 0; l < want; l++) out[l] = l < have ? bk->ws[v][(size_t)l * lanes + lane] : 0;
}
static void load_layout(const pref_sim *s, const p_signal *g, uint32_t lane, uint32_t *out) {
    for (uint32_t l = 0; l < g->limbs; l++) out[l] = s->values[(size_t)(g->offset + l) * s->lanes + lane];
    out[g->limbs - 1] &= final_mask(g->width);
}
static void store_layout(pref_sim *s, const p_signal *g, uint32_t lane, const uint32_t *v, uint32_t have) {
    uint32_t tmp[MAXL];
    for (uint32_t l = 0; l < g->limbs; l++) tmp[l] = l < have ? v[l] : 0;
    tmp[g->limbs - 1] &= final_mask(g->width);
    for (uint32_t l = 0; l < g->limbs; l++) s->values[(size_t)(g->offset + l) * s->lanes + lane] = tmp[l];
}
static size_t mem_index(const pref_sim *s, const p_memory *m, uint32_t lane, uint64_t addr, uint32_t limb) {
    return (size_t)lane * s->tmw + m->offset + (size_t)addr * m->limbs + limb;
}
/* decode_usize (:12989): low 64 bits, saturating when a higher limb is set */
static uint64_t decode_usize(const uint32_t *v, uint32_t n) {
    uint64_t out = v[0] | (n > 1 ? (uint64_t)v[1] << 32 : 0);
    for (uint32_t l = 2; l < n; l++) if (v[l]) return UINT64_MAX;
    return out;
}
static int reset_asserted(const pref_sim *s, uint32_t rsig, uint32_t rpol, uint32_t lane) {
    uint32_t tmp[MAXL];
    load_layout(s, &s->sig[rsig], lane, tmp);
    int b = tmp[0] & 1;
    return rpol == 0 ? b : !b;
}

static void eval_lane(pref_sim *s, const p_block *bk, const p_instr *in, uint32_t lane, uint32_t *out) {
    uint32_t L = nlimbs(in->width), a[MAXL], b[MAXL], c[MAXL];
    uint32_t lanes = s->lanes;
    memset(out, 0, sizeof(uint32_t) * L);
    switch (in->op) {
    case OP_LIT:
        for (uint32_t l = 0; l < L && l < in->nargs; l++) out[l] = in->args[l];
        break;
    case OP_SIGNAL: {
        const p_signal *g = &s->sig[in->args[0]];
        load_layout(s, g, lane, a);
        for (uint32_t l = 0; l < L; l++) out[l] = l < g->limbs ? a[l] : 0;
        break; }
    case OP_NOT:
        ws_load(bk, lanes, in->args[0], lane, a, L);
        for (uint32_t l = 0; l < L; l++) out[l] = ~a[l];
        break;
    case OP_AND: case OP_OR: case OP_XOR:
        ws_load(bk, lanes, in->args[0], lane, a, L);
        ws_load(bk, lanes, in->args[1], lane, b, L);
        for (uint32_t l = 0; l < L; l++)
            out[l] = in->op == OP_AND ? (a[l] & b[l]) : in->op == OP_OR ? (a[l] | b[l]) : (a[l] ^ b[l]);
        break;
    case OP_ADD: {
        ws_load(bk, lanes, in->args[0], lane, a, L);
        ws_load(bk, lanes, in->args[1], lane, b, L);
        uint64_t carry = 0;
        for (uint32_t l = 0; l < L; l++) { uint64_t t = (uint64_t)a[l] + b[l] + carry; out[l] = (uint32_t)t; carry = t >> 32; }
        break; }
    case OP_SUB: {
        ws_load(bk, lanes, in->args[0], lane, a, L);
        ws_load(bk, lanes, in->args[1], lane, b, L);
        uint64_t borrow = 0;
        for (uint32_t l = 0; l < L; l++) {
            uint64_t sub = (uint64_t)b[l] + borrow, lw = a[l];
            out[l] = (uint32_t)(lw - sub); borrow = lw < sub;
        }
        break; }
    case OP_MUL:
        ws_load(bk, lanes, in->args[0], lane, a, L);
        ws_load(bk, lanes, in->args[1], lane, b, L);
        /* truncated schoolbook, equal to mul_limbs' u128 column sums mod 2^(32L) */
        for (uint32_t j = 0; j < L; j++) {
            uint64_t carry = 0;
            for (uint32_t k = 0; j + k < L; k++) {
                uint64_t t = (uint64_t)a[j] * b[k] + out[j + k] + carry;
                out[j + k] = (uint32_t)t; carry = t >> 32;
            }
        }
        break;
    case OP_EQ: case OP_NE: {
        uint32_t wa = bk->vwidth[in->args[0]], wb = bk->vwidth[in->args[1]];
        uint32_t wd = wa > wb ? wa : wb, n = nlimbs(wd);
        ws_load(bk, lanes, in->args[0], lane, a, n);
        ws_load(bk, lanes, in->args[1], lane, b, n);
        fit(a, n, wd); fit(b, n, wd);
        int eq = 1;
        for (uint32_t l = 0; l < n; l++) if (a[l] != b[l]) eq = 0;
        out[0] = in->op == OP_EQ ? eq : !eq;
        break; }
    case OP_LT: {
        uint32_t wd = bk->vwidth[in->args[0]], n = nlimbs(wd);
        ws_load(bk, lanes, in->args[0], lane, a, n);
        ws_load(bk, lanes, in->args[1], lane, b, n);
        fit(a, n, wd); fit(b, n, wd);
        int res = 0, decided = 0;
        if (in->args[2]) {
            int sa = get_bit(a, n, wd - 1), sb = get_bit(b, n, wd - 1);
            if (sa != sb) { res = sa; decided = 1; }
        }
        for (int l = (int)n - 1; l >= 0 && !decided; l--)
            if (a[l] != b[l]) { res = a[l] < b[l]; decided = 1; }
        out[0] = res;
        break; }
    case OP_MUX:
        ws_load(bk, lanes, in->args[0], lane, c, 1);
        ws_load(bk, lanes, (c[0] & 1) ? in->args[1] : in->args[2], lane, out, L);
        break;
    case OP_SLICE: {
        uint32_t v = in->args[0], lsb = in->args[1], n = nlimbs(bk->vwidth[v]);
        ws_load(bk, lanes, v, lane, a, n);
        for (uint32_t bit = 0; bit < in->width; bit++)
            if (get_bit(a, n, lsb + bit)) out[bit / 32] |= 1u << (bit % 32);
        break; }
    case OP_ZEXT: case OP_TRUNC: case OP_CAST:
        ws_load(bk, lanes, in->args[0], lane, out, L);
        break;
    case OP_SEXT: {
        uint32_t iw = bk->vwidth[in->args[0]], n = nlimbs(iw);
        ws_load(bk, lanes, in->args[0], lane, a, n);
        int sign = get_bit(a, n, iw - 1);
        fit(a, n, iw);
        for (uint32_t l = 0; l < L; l++) out[l] = l < n ? a[l] : 0;
        if (sign) for (uint32_t bit = iw; bit < in->width; bit++) out[bit / 32] |= 1u << (bit % 32);
        break; }
    case OP_CONCAT: {
        uint32_t off = 0;
        for (int k = (int)in->nargs - 1; k >= 0; k--) {
            uint32_t v = in->args[k], pw = bk->vwidth[v], n = nlimbs(pw);
            ws_load(bk, lanes, v, lane, a, n);
            for (uint32_t bit = 0; bit < pw; bit++)
                if (get_bit(a, n, bit) && off + bit < 32 * L) out[(off + bit) / 32] |= 1u << ((off + bit) % 32);
            off += pw;
        }
        break; }
    case OP_MEMREAD: {
        const p_memory *m = &s->mem[in->args[0]];
        uint32_t v = in->args[1], n = nlimbs(bk->vwidth[v]);
        ws_load(bk, lanes, v, lane, a, n);
        uint64_t addr = decode_usize(a, n);
        if (addr < m->depth) { /* load_memory masks to the data width (:10240) */
            for (uint32_t l = 0; l < m->limbs; l++) b[l] = s->memories[mem_index(s, m, lane, addr, l)];
            b[m->limbs - 1] &= final_mask(m->width);
            for (uint32_t l = 0; l < m->limbs && l < L; l++) out[l] = b[l];
        }
        break; }
    }
    fit(out, L, in->width);
}

static void eval_instr(pref_sim *s, p_block *bk, const p_instr *in) {
    uint32_t L = nlimbs(in->width), out[MAXL];
    uint32_t *slot = bk->ws[in->dst];
    if (!slot) slot = bk->ws[in->dst] = calloc((size_t)L * s->lanes, sizeof(uint32_t));
    for (uint32_t lane = 0; lane < s->lanes; lane++) {
        eval_lane(s, bk, in, lane, out);
        for (uint32_t l = 0; l < L; l++) slot[(size_t)l * s->lanes + lane] = out[l];
    }
}

/* execute_machine_effect (:9740-9777) */
static void exec_effect(pref_sim *s, p_block *bk, const p_effect *ef) {
    uint32_t a[MAXL], b[MAXL];
    if (ef->kind == EF_STORE) {
        const p_signal *g = &s->sig[ef->dst];
        uint32_t n = nlimbs(bk->vwidth[ef->value]);
        for (uint32_t lane = 0; lane < s->lanes; lane++) { ws_load(bk, s->lanes, ef->value, lane, a, n); store_layout(s, g, lane, a, n); }
    } else if (ef->kind == EF_CAPTURE) {
        if (!ef->has_reset || ef->rkind != 1) return; /* only async resets act here */
        const p_signal *g = &s->sig[ef->dst];
        uint32_t n = nlimbs(bk->vwidth[ef->value]);
        for (uint32_t lane = 0; lane < s->lanes; lane++)
            if (reset_asserted(s, ef->rsig, ef->rpol, lane)) { ws_load(bk, s->lanes, ef->value, lane, a, n); store_layout(s, g, lane, a, n); }
    } else {
        const p_memory *m = &s->mem[ef->dst];
        for (uint32_t lane = 0; lane < s->lanes; lane++) {
            ws_load(bk, s->lanes, ef->en, lane, a, 1);
            if (!(a[0] & 1)) continue;
            uint32_t n = nlimbs(bk->vwidth[ef->addr]);
            ws_load(bk, s->lanes, ef->addr, lane, a, n);
            uint64_t addr = decode_usize(a, n);
            if (addr >= m->depth) continue;
            ws_load(bk, s->lanes, ef->data, lane, b, m->limbs);
            fit(b, m->limbs, m->width);
            for (uint32_t l = 0; l < m->limbs; l++) s->memories[mem_index(s, m, lane, addr, l)] = b[l];
        }
    }
}

static void exec_block(pref_sim *s, int bi) {
    p_block *bk = &s->blk[bi];
    for (uint32_t k = 0; k < bk->npackets; k++) {
        p_packet *pk = &bk->packets[k];
        for (uint32_t i = 0; i < pk->ninstr; i++) eval_instr(s, bk, &pk->instrs[i]);
        for (uint32_t e = 0; e < pk->neff; e++) exec_effect(s, bk, &pk->effects[e]);
    }
}

/* capture_register_stream + capture_register_value (:9613-9686) */
static void capture_registers(pref_sim *s) {
    p_block *bk = &s->blk[2];
    uint32_t a[MAXL];
    s->ncaps = 0;
    for (uint32_t k = 0; k < bk->npackets; k++) {
        p_packet *pk = &bk->packets[k];
        for (uint32_t i = 0; i < pk->ninstr; i++) eval_instr(s, bk, &pk->instrs[i]);
        for (uint32_t e = 0; e < pk->neff; e++) {
            const p_effect *ef = &pk->effects[e];
            if (ef->kind != EF_CAPTURE) continue;
            const p_signal *g = &s->sig[ef->dst];
            p_capture *cp = &s->caps[s->ncaps++];
            cp->sig = ef->dst;
            if (!cp->vals || cp->limbs < g->limbs) { free(cp->vals); cp->vals = calloc((size_t)g->limbs * s->lanes, sizeof(uint32_t)); }
            cp->limbs = g->limbs;
            for (uint32_t lane = 0; lane < s->lanes; lane++) {
                if (ef->has_reset && reset_asserted(s, ef->rsig, ef->rpol, lane)) {
                    for (uint32_t l = 0; l < g->limbs; l++) a[l] = l < ef->nlimbs ? ef->rlimbs[l] : 0;
                } else {
                    ws_load(bk, s->lanes, ef->value, lane, a, g->limbs);
                }
                fit(a, g->limbs, g->width);
                for (uint32_t l = 0; l < g->limbs; l++) cp->vals[(size_t)l * s->lanes + lane] = a[l];
            }
        }
    }
}

static void commit_captures(pref_sim *s) {
    for (uint32_t c = 0; c < s->ncaps; c++) {
        p_capture *cp = &s->caps[c];
        const p_signal *g = &s->sig[cp->sig];
        for (uint32_t l = 0; l < g->limbs; l++)
            memcpy(&s->values[(size_t)(g->offset + l) * s->lanes], &cp->vals[(size_t)l * s->lanes], sizeof(uint32_t) * s->lanes);
    }
}

/* ------------------------------------------------------------- public API */
void pref_eval_combinational(pref_sim *s) { exec_block(s, 1); exec_block(s, 0); }
void pref_tick_from_evaluated(pref_sim *s) { capture_registers(s); exec_block(s, 3); commit_captures(s); }
void pref_tick(pref_sim *s) { pref_eval_combinational(s); pref_tick_from_evaluated(s); pref_eval_combinational(s); }
void pref_tick_many(pref_sim *s, uint64_t steps) {
    if (!steps) return;
    pref_eval_combinational(s);
    for (uint64_t i = 0; i < steps; i++) { pref_tick_from_evaluated(s); pref_eval_combinational(s); }
}

pref_sim *pref_create(const void *blob, size_t nbytes, uint32_t lanes, char *err, size_t errlen) {
    if (lanes == 0) { snprintf(err, errlen, "E_SIM_IR_LANES: packed simulator requires at least one lane"); return NULL; }
    pref_sim *s = calloc(1, sizeof(pref_sim));
    if (load_program(&s->P, blob, nbytes, err, errlen)) { free_program(&s->P); free(s); return NULL; }
    s->sig = s->P.sig; s->nsig = s->P.nsig; s->mem = s->P.mem; s->nmem = s->P.nmem;
    s->tsw = s->P.tsw; s->tmw = s->P.tmw; s->blk = s->P.blk;
    s->lanes = lanes;
    s->values = calloc((size_t)s->tsw * lanes + 1, sizeof(uint32_t));
    s->memories = calloc((size_t)s->tmw * lanes + 1, sizeof(uint32_t));
    uint32_t ncap = 0;
    for (uint32_t k = 0; k < s->blk[2].npackets; k++) ncap += s->blk[2].packets[k].neff;
    s->caps = calloc(ncap + 1, sizeof(p_capture));
    pref_eval_combinational(s);
    return s;
}

void pref_destroy(pref_sim *s) {
    if (!s) return;
    uint32_t ncap = 0;
    for (uint32_t k = 0; k < s->blk[2].npackets; k++) ncap += s->blk[2].packets[k].neff;
    for (uint32_t c = 0; c < ncap; c++) free(s->caps[c].vals);
    free(s->caps); free(s->values); free(s->memories);
    free_program(&s->P);
    free(s);
}

/* set_packed_signal_limbs (:9909-9943): lane_limbs is [limbs][lanes]; re-settles when eval != 0 */
int pref_set_signal_limbs(pref_sim *s, uint32_t sig, const uint32_t *vals, int eval) {
    if (sig >= s->nsig) return -1;
    const p_signal *g = &s->sig[sig];
    uint32_t a[MAXL];
    for (uint32_t lane = 0; lane < s->lanes; lane++) {
        for (uint32_t l = 0; l < g->limbs; l++) a[l] = vals[(size_t)l * s->lanes + lane];
        store_layout(s, g, lane, a, g->limbs);
    }
    if (eval) pref_eval_combinational(s);
    return 0;
}
int pref_get_signal_limbs(const pref_sim *s, uint32_t sig, uint32_t *out) {
    if (sig >= s->nsig) return -1;
    const p_signal *g = &s->sig[sig];
    uint32_t a[MAXL];
    for (uint32_t lane = 0; lane < s->lanes; lane++) {
        load_layout(s, g, lane, a);
        for (uint32_t l = 0; l < g->limbs; l++) out[(size_t)l * s->lanes + lane] = a[l];
    }
    return 0;
}
/* memory values are [lanes][depth][limbs] (the lane-major storage order) */
int pref_set_memory_limbs(pref_sim *s, uint32_t mi, const uint32_t *vals, int eval) {
    if (mi >= s->nmem) return -1;
    const p_memory *m = &s->mem[mi];
    uint32_t a[MAXL];
    for (uint32_t lane = 0; lane < s->lanes; lane++)
        for (uint32_t d = 0; d < m->depth; d++) {
            for (uint32_t l = 0; l < m->limbs; l++) a[l] = vals[((size_t)lane * m->depth + d) * m->limbs + l];
            fit(a, m->limbs, m->width);
            for (uint32_t l = 0; l < m->limbs; l++) s->memories[mem_index(s, m, lane, d, l)] = a[l];
        }
    if (eval) pref_eval_combinational(s);
    return 0;
}
int pref_get_memory_limbs(const pref_sim *s, uint32_t mi, uint32_t *out) {
    if (mi >= s->nmem) return -1;
    const p_memory *m = &s->mem[mi];
    for (uint32_t lane = 0; lane < s->lanes; lane++)
        for (uint32_t d = 0; d < m->depth; d++) {
            for (uint32_t l = 0; l < m->limbs; l++) {
                uint32_t v = s->memories[mem_index(s, m, lane, d, l)];
                if (l == m->limbs - 1) v &= final_mask(m->width);
                out[((size_t)lane * m->depth + d) * m->limbs + l] = v;
            }
        }
    return 0;
}
void pref_snapshot(const pref_sim *s, uint32_t *values, uint32_t *memories) {
    memcpy(values, s->values, sizeof(uint32_t) * (size_t)s->tsw * s->lanes);
    memcpy(memories, s->memories, sizeof(uint32_t) * (size_t)s->tmw * s->lanes);
}
void pref_restore(pref_sim *s, const uint32_t *values, const uint32_t *memories) {
    memcpy(s->values, values, sizeof(uint32_t) * (size_t)s->tsw * s->lanes);
    memcpy(s->memories, memories, sizeof(uint32_t) * (size_t)s->tmw * s->lanes);
}
uint32_t pref_lanes(const pref_sim *s) { return s->lanes; }
