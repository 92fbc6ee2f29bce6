/*
 * ORACLE (test infrastructure only): host batch-SIMD engine, the timed CPU
 * baseline for word-level workloads.
 *
 * Restates rrtl_sim_ir::SimdCpuSimulator (crates/rrtl-sim-ir/src/lib.rs:5375-7093,
 * simd_words :7459-8842): every block is executed instruction-outer, lane-inner
 * over a contiguous limbs x lanes u32 slot per value (execute_machine_block
 * :5953, eval_simd_instr_into :6109); one-limb instructions run as straight
 * loops over lanes that the compiler vectorises (AVX2 here, the reference's
 * simd_words AVX2 path :8799-8841), wider values fall back to per-lane limb code
 * (the reference's scalar fallback :6886-7032, shared with oracle/packed_ref.c's
 * semantics).  tick_many is comb-fused (:5784-5803).  Lane shards run on
 * separate threads, one engine per thread (ThreadedReplayRunner, lib.rs:4517).
 *
 * The run harness matches the device runs: per cycle drive inputs, eval,
 * observe (PRE), edge, [eval, observe (POST)]; a final eval after PRE runs.
 * Validated lane for lane against oracle/packed_ref.c (tests).
 */
#include "blob.h"
#include "stim.h"
#include <omp.h>

typedef struct { uint32_t sig, limbs; uint32_t *vals; } cap_t;

typedef struct simd_shard {
    const p_program *P;
    uint32_t lanes, lane_base;
    uint32_t *values;     /* [tsw][lanes] */
    uint32_t *memories;   /* [lanes][tmw] */
    uint32_t **ws[4];     /* per block per value: [limbs][lanes] */
    cap_t *caps; uint32_t ncaps;
    uint64_t *digest;     /* [lanes] */
} simd_shard;

typedef struct simd_engine {
    p_program P;
    uint32_t lanes, nshards;
    simd_shard *shards;
    int threads;
} simd_engine;

static inline uint32_t *slot(simd_shard *s, int b, uint32_t v) { return s->ws[b][v]; }

static uint32_t vlimbs(const p_block *bk, uint32_t v) { return nlimbs(bk->vwidth[v]); }

/* generic per-lane fallback with the packed oracle's semantics */
static void lane_load(simd_shard *s, const p_block *bk, int b, uint32_t v, uint32_t lane, uint32_t *out, uint32_t want) {
    uint32_t have = vlimbs(bk, v);
    for (uint32_t l = 0; l < want; l++) out[l] = l < have ? s->ws[b][v][(size_t)l * s->lanes + lane] : 0;
}
static uint64_t decode_usize(const uint32_t *v, uint32_t n) {
    uint64_t out = v[0] | (n > 1 ? (uint64_t)v[1] << 32 : 0);
    for (uint32_t l = 2; l < n; l++) if (v[l]) return UINT64_MAX;
    return out;
}

static void eval_generic(simd_shard *s, const p_block *bk, int b, const p_instr *in, uint32_t *dst) {
    const p_program *P = s->P;
    uint32_t L = nlimbs(in->width);
    for (uint32_t lane = 0; lane < s->lanes; lane++) {
        uint32_t a[MAXL], c[MAXL], r[MAXL];
        memset(r, 0, sizeof(uint32_t) * L);
        switch (in->op) {
        case OP_LIT: for (uint32_t l = 0; l < L && l < in->nargs; l++) r[l] = in->args[l]; break;
        case OP_SIGNAL: {
            const p_signal *g = &P->sig[in->args[0]];
            for (uint32_t l = 0; l < g->limbs && l < L; l++) r[l] = s->values[(size_t)(g->offset + l) * s->lanes + lane];
            if (g->limbs <= L) r[g->limbs - 1] &= final_mask(g->width);
            break; }
        case OP_NOT: lane_load(s, bk, b, in->args[0], lane, a, L); for (uint32_t l = 0; l < L; l++) r[l] = ~a[l]; break;
        case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: {
            lane_load(s, bk, b, in->args[0], lane, a, L); lane_load(s, bk, b, in->args[1], lane, c, L);
            if (in->op == OP_AND) for (uint32_t l = 0; l < L; l++) r[l] = a[l] & c[l];
            else if (in->op == OP_OR) for (uint32_t l = 0; l < L; l++) r[l] = a[l] | c[l];
            else if (in->op == OP_XOR) for (uint32_t l = 0; l < L; l++) r[l] = a[l] ^ c[l];
            else if (in->op == OP_ADD) { uint64_t cy = 0; for (uint32_t l = 0; l < L; l++) { uint64_t t = (uint64_t)a[l] + c[l] + cy; r[l] = (uint32_t)t; cy = t >> 32; } }
            else if (in->op == OP_SUB) { uint64_t br = 0; for (uint32_t l = 0; l < L; l++) { uint64_t sb = (uint64_t)c[l] + br, x = a[l]; r[l] = (uint32_t)(x - sb); br = x < sb; } }
            else for (uint32_t j = 0; j < L; j++) { uint64_t cy = 0; for (uint32_t k = 0; j + k < L; k++) { uint64_t t = (uint64_t)a[j] * c[k] + r[j + k] + cy; r[j + k] = (uint32_t)t; cy = t >> 32; } }
            break; }
        case OP_EQ: case OP_NE: {
            uint32_t wa = bk->vwidth[in->args[0]], wb = bk->vwidth[in->args[1]], wd = wa > wb ? wa : wb, n = nlimbs(wd);
            lane_load(s, bk, b, in->args[0], lane, a, n); lane_load(s, bk, b, in->args[1], lane, c, n);
            int eq = 1; for (uint32_t l = 0; l < n; l++) if (a[l] != c[l]) eq = 0;
            r[0] = in->op == OP_EQ ? eq : !eq; break; }
        case OP_LT: {
            uint32_t wd = bk->vwidth[in->args[0]], n = nlimbs(wd);
            lane_load(s, bk, b, in->args[0], lane, a, n); lane_load(s, bk, b, in->args[1], lane, c, n);
            fit(a, n, wd); fit(c, n, wd);
            int res = 0, dec = 0;
            if (in->args[2]) { int sa = get_bit(a, n, wd - 1), sb = get_bit(c, n, wd - 1); if (sa != sb) { res = sa; dec = 1; } }
            for (int l = (int)n - 1; l >= 0 && !dec; l--) if (a[l] != c[l]) { res = a[l] < c[l]; dec = 1; }
            r[0] = res; break; }
        case OP_MUX: lane_load(s, bk, b, in->args[0], lane, c, 1); lane_load(s, bk, b, (c[0] & 1) ? in->args[1] : in->args[2], lane, r, L); break;
        case OP_SLICE: {
            uint32_t v = in->args[0], n = vlimbs(bk, v); lane_load(s, bk, b, v, lane, a, n);
            for (uint32_t bit = 0; bit < in->width; bit++) if (get_bit(a, n, in->args[1] + bit)) r[bit / 32] |= 1u << (bit % 32);
            break; }
        case OP_ZEXT: case OP_TRUNC: case OP_CAST: lane_load(s, bk, b, in->args[0], lane, r, L); break;
        case OP_SEXT: {
            uint32_t iw = bk->vwidth[in->args[0]], n = nlimbs(iw); lane_load(s, bk, b, in->args[0], lane, a, n);
            int sign = get_bit(a, n, iw - 1); fit(a, n, iw);
            for (uint32_t l = 0; l < L; l++) r[l] = l < n ? a[l] : 0;
            if (sign) for (uint32_t bit = iw; bit < in->width; bit++) r[bit / 32] |= 1u << (bit % 32);
            break; }
        case OP_CONCAT: {
            uint32_t off = 0;
            for (int k = (int)in->nargs - 1; k >= 0; k--) {
                uint32_t v = in->args[k], pw = bk->vwidth[v], n = nlimbs(pw); lane_load(s, bk, b, v, lane, a, n);
                for (uint32_t bit = 0; bit < pw; bit++) if (get_bit(a, n, bit) && off + bit < 32 * L) r[(off + bit) / 32] |= 1u << ((off + bit) % 32);
                off += pw;
            }
            break; }
        case OP_MEMREAD: {
            const p_memory *m = &P->mem[in->args[0]];
            uint32_t v = in->args[1], n = vlimbs(bk, v); lane_load(s, bk, b, v, lane, a, n);
            uint64_t addr = decode_usize(a, n);
            if (addr < m->depth) {
                for (uint32_t l = 0; l < m->limbs; l++) c[l] = s->memories[(size_t)lane * P->tmw + m->offset + (size_t)addr * m->limbs + l];
                c[m->limbs - 1] &= final_mask(m->width);
                for (uint32_t l = 0; l < m->limbs && l < L; l++) r[l] = c[l];
            }
            break; }
        }
        fit(r, L, in->width);
        for (uint32_t l = 0; l < L; l++) dst[(size_t)l * s->lanes + lane] = r[l];
    }
}

/* one-limb fast paths: whole-lane loops over contiguous u32 slots */
static int eval_fast(simd_shard *s, const p_block *bk, int b, const p_instr *in, uint32_t *restrict d) {
    const p_program *P = s->P;
    const uint32_t n = s->lanes, m = final_mask(in->width);
    if (nlimbs(in->width) != 1) return 0;
#define SRC(k) (vlimbs(bk, in->args[k]) == 1 ? slot(s, b, in->args[k]) : NULL)
    switch (in->op) {
    case OP_LIT: { uint32_t v = (in->nargs ? in->args[0] : 0) & m; for (uint32_t i = 0; i < n; i++) d[i] = v; return 1; }
    case OP_SIGNAL: {
        const p_signal *g = &P->sig[in->args[0]];
        const uint32_t *src = &s->values[(size_t)g->offset * n];
        uint32_t mm = m & (g->limbs == 1 ? final_mask(g->width) : 0xFFFFFFFFu);
        for (uint32_t i = 0; i < n; i++) d[i] = src[i] & mm;
        return 1; }
    case OP_NOT: { const uint32_t *a = slot(s, b, in->args[0]); for (uint32_t i = 0; i < n; i++) d[i] = ~a[i] & m; return 1; }
    case OP_ZEXT: case OP_TRUNC: case OP_CAST: { const uint32_t *a = slot(s, b, in->args[0]); for (uint32_t i = 0; i < n; i++) d[i] = a[i] & m; return 1; }
    case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: {
        const uint32_t *a = slot(s, b, in->args[0]), *c = slot(s, b, in->args[1]);
        switch (in->op) {
        case OP_AND: for (uint32_t i = 0; i < n; i++) d[i] = (a[i] & c[i]) & m; break;
        case OP_OR: for (uint32_t i = 0; i < n; i++) d[i] = (a[i] | c[i]) & m; break;
        case OP_XOR: for (uint32_t i = 0; i < n; i++) d[i] = (a[i] ^ c[i]) & m; break;
        case OP_ADD: for (uint32_t i = 0; i < n; i++) d[i] = (a[i] + c[i]) & m; break;
        case OP_SUB: for (uint32_t i = 0; i < n; i++) d[i] = (a[i] - c[i]) & m; break;
        default: for (uint32_t i = 0; i < n; i++) d[i] = (a[i] * c[i]) & m; break;
        }
        return 1; }
    case OP_MUX: {
        const uint32_t *c = slot(s, b, in->args[0]), *t = SRC(1), *e = SRC(2);
        if (!t || !e) return 0;
        for (uint32_t i = 0; i < n; i++) d[i] = ((c[i] & 1) ? t[i] : e[i]) & m;
        return 1; }
    case OP_EQ: case OP_NE: {
        const uint32_t *a = SRC(0), *c = SRC(1);
        if (!a || !c) return 0;
        if (in->op == OP_EQ) for (uint32_t i = 0; i < n; i++) d[i] = a[i] == c[i];
        else for (uint32_t i = 0; i < n; i++) d[i] = a[i] != c[i];
        return 1; }
    case OP_LT: {
        const uint32_t *a = SRC(0), *c = SRC(1);
        uint32_t wl = bk->vwidth[in->args[0]];
        if (!a || !c || wl > 32) return 0;
        uint32_t lm = final_mask(wl);
        if (in->args[2]) {
            uint32_t sh = 32 - wl;
            for (uint32_t i = 0; i < n; i++) d[i] = (int32_t)((a[i] & lm) << sh) < (int32_t)((c[i] & lm) << sh);
        } else {
            for (uint32_t i = 0; i < n; i++) d[i] = (a[i] & lm) < (c[i] & lm);
        }
        return 1; }
    case OP_SLICE: {
        const uint32_t *a = SRC(0);
        if (!a) return 0;
        uint32_t lsb = in->args[1];
        for (uint32_t i = 0; i < n; i++) d[i] = (lsb < 32 ? a[i] >> lsb : 0) & m;
        return 1; }
    }
#undef SRC
    return 0;
}

static void exec_block(simd_shard *s, int b) {
    const p_program *P = s->P;
    const p_block *bk = &P->blk[b];
    const uint32_t n = s->lanes;
    uint32_t a[MAXL];
    for (uint32_t k = 0; k < bk->npackets; k++) {
        const p_packet *pk = &bk->packets[k];
        for (uint32_t i = 0; i < pk->ninstr; i++) {
            const p_instr *in = &pk->instrs[i];
            uint32_t *d = s->ws[b][in->dst];
            if (!eval_fast(s, bk, b, in, d)) eval_generic(s, bk, b, in, d);
        }
        for (uint32_t e = 0; e < pk->neff; e++) {
            const p_effect *ef = &pk->effects[e];
            if (ef->kind == EF_STORE) {
                const p_signal *g = &P->sig[ef->dst];
                uint32_t have = vlimbs(bk, ef->value);
                for (uint32_t l = 0; l < g->limbs; l++) {
                    uint32_t mm = l + 1 == g->limbs ? final_mask(g->width) : 0xFFFFFFFFu;
                    uint32_t *dst = &s->values[(size_t)(g->offset + l) * n];
                    if (l < have) { const uint32_t *src = &s->ws[b][ef->value][(size_t)l * n]; for (uint32_t i = 0; i < n; i++) dst[i] = src[i] & mm; }
                    else memset(dst, 0, 4 * (size_t)n);
                }
            } else if (ef->kind == EF_CAPTURE) {
                if (b == 2 || !ef->has_reset || ef->rkind != 1) continue;
                const p_signal *g = &P->sig[ef->dst], *rs = &P->sig[ef->rsig];
                const uint32_t *rv = &s->values[(size_t)rs->offset * n];
                for (uint32_t lane = 0; lane < n; lane++) {
                    int as = (int)(rv[lane] & 1); if (ef->rpol) as = !as;
                    if (!as) continue;
                    lane_load(s, bk, b, ef->value, lane, a, g->limbs);
                    fit(a, g->limbs, g->width);
                    for (uint32_t l = 0; l < g->limbs; l++) s->values[(size_t)(g->offset + l) * n + lane] = a[l];
                }
            } else if (ef->kind == EF_MEMWRITE) {
                const p_memory *m = &P->mem[ef->dst];
                for (uint32_t lane = 0; lane < n; lane++) {
                    lane_load(s, bk, b, ef->en, lane, a, 1);
                    if (!(a[0] & 1)) continue;
                    uint32_t na = vlimbs(bk, ef->addr);
                    lane_load(s, bk, b, ef->addr, lane, a, na);
                    uint64_t addr = decode_usize(a, na);
                    if (addr >= m->depth) continue;
                    lane_load(s, bk, b, ef->data, lane, a, m->limbs);
                    fit(a, m->limbs, m->width);
                    for (uint32_t l = 0; l < m->limbs; l++) s->memories[(size_t)lane * P->tmw + m->offset + (size_t)addr * m->limbs + l] = a[l];
                }
            }
        }
    }
}

static void capture(simd_shard *s) {
    const p_program *P = s->P;
    const p_block *bk = &P->blk[2];
    const uint32_t n = s->lanes;
    exec_block(s, 2);
    s->ncaps = 0;
    uint32_t a[MAXL];
    for (uint32_t k = 0; k < bk->npackets; k++) {
        const p_packet *pk = &bk->packets[k];
        for (uint32_t e = 0; e < pk->neff; e++) {
            const p_effect *ef = &pk->effects[e];
            if (ef->kind != EF_CAPTURE) continue;
            const p_signal *g = &P->sig[ef->dst];
            cap_t *cp = &s->caps[s->ncaps++];
            cp->sig = ef->dst; cp->limbs = g->limbs;
            uint32_t have = vlimbs(bk, ef->value);
            for (uint32_t l = 0; l < g->limbs; l++) {
                uint32_t mm = l + 1 == g->limbs ? final_mask(g->width) : 0xFFFFFFFFu;
                uint32_t *dst = &cp->vals[(size_t)l * n];
                if (l < have) { const uint32_t *src = &s->ws[2][ef->value][(size_t)l * n]; for (uint32_t i = 0; i < n; i++) dst[i] = src[i] & mm; }
                else memset(dst, 0, 4 * (size_t)n);
            }
            if (ef->has_reset) {
                const p_signal *rs = &P->sig[ef->rsig];
                const uint32_t *rv = &s->values[(size_t)rs->offset * n];
                for (uint32_t l = 0; l < g->limbs; l++) a[l] = l < ef->nlimbs ? ef->rlimbs[l] : 0;
                fit(a, g->limbs, g->width);
                for (uint32_t l = 0; l < g->limbs; l++) {
                    uint32_t *dst = &cp->vals[(size_t)l * n];
                    const uint32_t pol = ef->rpol, v = a[l];
                    for (uint32_t i = 0; i < n; i++) dst[i] = (((rv[i] & 1) ^ pol) ? v : dst[i]);
                }
            }
        }
    }
}

static void eval_comb(simd_shard *s) { exec_block(s, 1); exec_block(s, 0); }
static void edge(simd_shard *s) {
    capture(s);
    exec_block(s, 3);
    for (uint32_t c = 0; c < s->ncaps; c++) {
        const p_signal *g = &s->P->sig[s->caps[c].sig];
        memcpy(&s->values[(size_t)g->offset * s->lanes], s->caps[c].vals, 4 * (size_t)g->limbs * s->lanes);
    }
}

simd_engine *simd_create(const void *blob, size_t nbytes, uint32_t lanes, int threads, uint32_t shard_lanes, char *err, size_t errlen) {
    if (!lanes) { snprintf(err, errlen, "E_SIM_IR_LANES: packed simulator requires at least one lane"); return NULL; }
    simd_engine *e = calloc(1, sizeof(simd_engine));
    if (load_program(&e->P, blob, nbytes, err, errlen)) { free_program(&e->P); free(e); return NULL; }
    e->lanes = lanes;
    e->threads = threads > 0 ? threads : omp_get_max_threads();
    if (!shard_lanes) shard_lanes = (lanes + e->threads - 1) / e->threads;
    e->nshards = (lanes + shard_lanes - 1) / shard_lanes;
    e->shards = calloc(e->nshards, sizeof(simd_shard));
    uint32_t ncap = 0;
    for (uint32_t k = 0; k < e->P.blk[2].npackets; k++) ncap += e->P.blk[2].packets[k].neff;
    for (uint32_t i = 0; i < e->nshards; i++) {
        simd_shard *s = &e->shards[i];
        s->P = &e->P; s->lane_base = i * shard_lanes;
        s->lanes = (s->lane_base + shard_lanes <= lanes) ? shard_lanes : lanes - s->lane_base;
        s->values = calloc((size_t)e->P.tsw * s->lanes + 1, 4);
        s->memories = calloc((size_t)e->P.tmw * s->lanes + 1, 4);
        for (int b = 0; b < 4; b++) {
            const p_block *bk = &e->P.blk[b];
            s->ws[b] = calloc(bk->nvalues + 1, sizeof(uint32_t *));
            for (uint32_t v = 0; v < bk->nvalues; v++) s->ws[b][v] = calloc((size_t)vlimbs(bk, v) * s->lanes + 1, 4);
        }
        s->caps = calloc(ncap + 1, sizeof(cap_t));
        uint32_t c = 0;
        for (uint32_t k = 0; k < e->P.blk[2].npackets; k++)
            for (uint32_t x = 0; x < e->P.blk[2].packets[k].neff; x++) {
                const p_effect *ef = &e->P.blk[2].packets[k].effects[x];
                if (ef->kind == EF_CAPTURE) s->caps[c++].vals = calloc((size_t)e->P.sig[ef->dst].limbs * s->lanes + 1, 4);
            }
        s->digest = calloc(s->lanes, 8);
        eval_comb(s);
    }
    return e;
}

void simd_destroy(simd_engine *e) {
    if (!e) return;
    uint32_t ncap = 0;
    for (uint32_t k = 0; k < e->P.blk[2].npackets; k++) ncap += e->P.blk[2].packets[k].neff;
    for (uint32_t i = 0; i < e->nshards; i++) {
        simd_shard *s = &e->shards[i];
        for (int b = 0; b < 4; b++) { for (uint32_t v = 0; v < e->P.blk[b].nvalues; v++) free(s->ws[b][v]); free(s->ws[b]); }
        for (uint32_t c = 0; c < ncap; c++) free(s->caps[c].vals);
        free(s->caps); free(s->values); free(s->memories); free(s->digest);
    }
    free(e->shards); free_program(&e->P); free(e);
}

enum { IN_HOLD = 0, IN_RANDOM = 1, IN_CONST = 2, IN_PULSE = 3 };

static void observe_shard(simd_shard *s, uint32_t n_obs, const uint32_t *obs_sig) {
    const uint32_t n = s->lanes;
    for (uint32_t j = 0; j < n_obs; j++) {
        const p_signal *g = &s->P->sig[obs_sig[j]];
        for (uint32_t l = 0; l < g->limbs; l++) {
            const uint32_t *v = &s->values[(size_t)(g->offset + l) * n];
            for (uint32_t lane = 0; lane < n; lane++) s->digest[lane] = lane_digest_update(s->digest[lane], v[lane], j, l);
        }
    }
}

/* Runs `cycles` cycles with stimulus and observation on every shard in
 * parallel; returns elapsed seconds.  digest_out: [lanes] u64 chains (or NULL). */
double simd_run(simd_engine *e, uint64_t seed, uint32_t cycles, uint32_t n_in, const uint32_t *in_sig, const uint32_t *in_mode,
                const uint64_t *in_arg, uint32_t n_obs, const uint32_t *obs_sig, int observe_post, uint64_t *digest_out) {
    double t0 = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
    for (uint32_t i = 0; i < e->nshards; i++) {
        simd_shard *s = &e->shards[i];
        const p_program *P = s->P;
        const uint32_t n = s->lanes;
        for (uint32_t c = 0; c < cycles; c++) {
            for (uint32_t k = 0; k < n_in; k++) {
                const p_signal *g = &P->sig[in_sig[k]];
                if (in_mode[k] == IN_HOLD) continue;
                for (uint32_t l = 0; l < g->limbs; l++) {
                    uint32_t *dst = &s->values[(size_t)(g->offset + l) * n];
                    uint32_t mm = l + 1 == g->limbs ? final_mask(g->width) : 0xFFFFFFFFu;
                    if (in_mode[k] == IN_RANDOM) {
                        memset(dst, 0, 4 * (size_t)n);
                        uint32_t nb = g->width - 32 * l < 32 ? g->width - 32 * l : 32;
                        for (uint32_t bit = 0; bit < nb; bit++) {
                            uint64_t key = stim_stream_key(seed, ((uint64_t)in_sig[k] << 8) | (32 * l + bit), c);
                            uint64_t wcur = UINT64_MAX, w = 0;
                            for (uint32_t lane = 0; lane < n; lane++) {
                                uint64_t gl = (uint64_t)s->lane_base + lane;
                                if ((gl >> 6) != wcur) { wcur = gl >> 6; w = stim_word_from_key(key, wcur); }
                                dst[lane] |= (uint32_t)((w >> (gl & 63)) & 1) << bit;
                            }
                        }
                    } else {
                        uint64_t a = in_arg[k];
                        uint32_t v = in_mode[k] == IN_CONST ? (l == 0 ? (uint32_t)a : l == 1 ? (uint32_t)(a >> 32) : 0)
                                   : (l == 0 ? (uint32_t)(c < (uint32_t)a ? (a >> 32) & 1 : (a >> 33) & 1) : 0);
                        for (uint32_t lane = 0; lane < n; lane++) dst[lane] = v & mm;
                    }
                }
            }
            eval_comb(s);
            if (!observe_post) observe_shard(s, n_obs, obs_sig);
            edge(s);
            if (observe_post) { eval_comb(s); observe_shard(s, n_obs, obs_sig); }
        }
        if (!observe_post) eval_comb(s);
        if (digest_out) memcpy(&digest_out[s->lane_base], s->digest, 8 * (size_t)n);
    }
    return omp_get_wtime() - t0;
}

void simd_tick_many(simd_engine *e, uint64_t steps) {
    if (!steps) return;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
    for (uint32_t i = 0; i < e->nshards; i++) {
        simd_shard *s = &e->shards[i];
        eval_comb(s);
        for (uint64_t k = 0; k < steps; k++) { edge(s); eval_comb(s); }
    }
}

/* signal access by lane slice: [limbs][lanes] across shards */
void simd_get_signal_limbs(const simd_engine *e, uint32_t sig, uint32_t *out) {
    const p_signal *g = &e->P.sig[sig];
    for (uint32_t i = 0; i < e->nshards; i++) {
        const simd_shard *s = &e->shards[i];
        for (uint32_t l = 0; l < g->limbs; l++)
            memcpy(&out[(size_t)l * e->lanes + s->lane_base], &s->values[(size_t)(g->offset + l) * s->lanes], 4 * (size_t)s->lanes);
    }
}
void simd_set_signal_limbs(simd_engine *e, uint32_t sig, const uint32_t *vals) {
    const p_signal *g = &e->P.sig[sig];
    for (uint32_t i = 0; i < e->nshards; i++) {
        simd_shard *s = &e->shards[i];
        for (uint32_t l = 0; l < g->limbs; l++) {
            uint32_t mm = l + 1 == g->limbs ? final_mask(g->width) : 0xFFFFFFFFu;
            for (uint32_t lane = 0; lane < s->lanes; lane++)
                s->values[(size_t)(g->offset + l) * s->lanes + lane] = vals[(size_t)l * e->lanes + s->lane_base + lane] & mm;
        }
        eval_comb(s);
    }
}

/* memory access by lane: [lanes][depth][limbs] (set_memory_limbs,
 * crates/rrtl-sim-ir/src/lib.rs:10136); a write re-settles comb like
 * set_packed_signal_limbs */
void simd_set_memory_limbs(simd_engine *e, uint32_t mem, const uint32_t *vals) {
    const p_memory *m = &e->P.mem[mem];
    const size_t per_lane = (size_t)m->depth * m->limbs;
    for (uint32_t i = 0; i < e->nshards; i++) {
        simd_shard *s = &e->shards[i];
        for (uint32_t lane = 0; lane < s->lanes; lane++)
            for (size_t k = 0; k < per_lane; k++) {
                uint32_t v = vals[(size_t)(s->lane_base + lane) * per_lane + k];
                if (k % m->limbs == m->limbs - 1) v &= final_mask(m->width);
                s->memories[(size_t)lane * e->P.tmw + m->offset + k] = v;
            }
        eval_comb(s);
    }
}
void simd_get_memory_limbs(const simd_engine *e, uint32_t mem, uint32_t *out) {
    const p_memory *m = &e->P.mem[mem];
    const size_t per_lane = (size_t)m->depth * m->limbs;
    for (uint32_t i = 0; i < e->nshards; i++) {
        const simd_shard *s = &e->shards[i];
        for (uint32_t lane = 0; lane < s->lanes; lane++)
            memcpy(&out[(size_t)(s->lane_base + lane) * per_lane], &s->memories[(size_t)lane * e->P.tmw + m->offset], 4 * per_lane);
    }
}
