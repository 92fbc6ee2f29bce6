/*
 * ORACLE (test infrastructure only): host bit-parallel gate engines.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * this library, and only as the checker / the timed CPU baseline.
 *
 * 1. bp_*  -- restates rrtl_sim_ir::bitparallel::BitParallelSimulator
 *    (crates/rrtl-sim-ir/src/bitparallel.rs:31-355): 64 lanes per u64, state
 *    state[s*words + w]; gate ops And/Or/Xor/Not/Mux/Signal/Lit on 1-bit
 *    signals only (:47-85, errors E_BITPARALLEL); settle = async-reset-comb then
 *    comb, each packet's stores applied before the next packet (:256-300);
 *    tick = settle, capture tick_next with the reset mux (a&rv)|(~a&v), commit,
 *    apply_forces, settle (:303-336); set_force/clamp_signal re-clamp a forced
 *    (signal, lane) after every comb store of that signal and after commit
 *    (:96-131).  One deliberate difference: in settle the reference writes an
 *    async-reset register as (asserted & rv) | (~asserted & literal) (:281-287),
 *    i.e. the reset literal also on lanes whose reset is NOT asserted, which
 *    disagrees with PackedSimulator and the scalar Simulator; this restatement
 *    keeps the register on non-asserted lanes, as those two do.
 *
 * 2. bpg_* -- the fast host engine timed as the CPU baseline: the
 *    BitParallelAot structure (bitparallel.rs:414-525): the program is
 *    flattened to a gate list once, state is group-major (a group is G u64
 *    words = 512 lanes, every slot's G words contiguous), the cycle loop runs
 *    per group with compiler-vectorised G-word gate ops, and groups are spread
 *    over threads like tick_many_parallel (:513).  Validated against bp_*.
 *
 * Both engines share the run harness below: per cycle, drive stimulus into the
 * inputs, settle, [observe], capture+commit+forces, [settle, observe]; a final
 * settle after the last cycle.  Stimulus and digests follow oracle/stim.h.
 */
#include "blob.h"
#include "stim.h"
#include <omp.h>

#define G 8 /* u64 words per group in the fast engine */

/* ------------------------------------------------------------ run spec */
enum { IN_HOLD = 0, IN_RANDOM = 1, IN_CONST = 2, IN_PULSE = 3, IN_GROUP_RANDOM = 5 };
typedef struct {
    uint64_t seed;
    uint32_t cycles;
    uint32_t n_inputs;
    const uint32_t *input_sig;   /* signal index */
    const uint32_t *input_mode;  /* IN_* */
    const uint64_t *input_arg;   /* CONST: bit0; PULSE: low32 = n, bit32 = value while c<n, bit33 = after */
    uint32_t n_obs;
    const uint32_t *obs_sig;
    uint32_t observe_post;       /* 0: sample after the pre-edge settle; 1: after the post-tick settle */
    uint64_t word_base;          /* global index of this engine's word 0 (stimulus is per global word) */
    uint64_t *digest;            /* [64][words] or NULL */
    uint64_t *trace;             /* [cycles][n_obs][words] or NULL */
    const uint64_t *group_key;   /* IN_GROUP_RANDOM: key per local u64 word (one 64-lane group), or NULL = global word */
} bp_run_spec;

/* gword: global u64 word (64-lane group) index; lword: its index in this engine */
static uint64_t input_word(const bp_run_spec *r, uint32_t i, uint32_t cyc, uint64_t key, uint64_t gword, uint64_t lword) {
    uint64_t arg = r->input_arg ? r->input_arg[i] : 0;
    switch (r->input_mode[i]) {
    case IN_RANDOM: return stim_word_from_key(key, gword);
    case IN_GROUP_RANDOM:
        /* one bit per group, shared by its 64 lanes */
        return (stim_word_from_key(key, r->group_key ? r->group_key[lword] : gword) & 1) ? ~0ull : 0ull;
    case IN_CONST: return (arg & 1) ? ~0ull : 0ull;
    case IN_PULSE: {
        uint32_t n = (uint32_t)arg;
        int v = cyc < n ? (int)((arg >> 32) & 1) : (int)((arg >> 33) & 1);
        return v ? ~0ull : 0ull; }
    }
    return 0;
}

/* =================================================== 1. faithful engine */
typedef struct { uint32_t sig, lane, value; } force_t;

typedef struct bp_sim {
    p_program P;
    uint32_t lanes, words, nsig;
    uint64_t *state;
    uint64_t *work[4];
    uint64_t *next;
    force_t *forces; uint32_t nforces, capforces;
} bp_sim;

static int bp_validate(const p_program *P, char *err, size_t errlen) {
    for (uint32_t i = 0; i < P->nsig; i++)
        if (P->sig[i].width != 1) { snprintf(err, errlen, "E_BITPARALLEL: signal %u is %u bits", i, P->sig[i].width); return -1; }
    if (P->nmem) { snprintf(err, errlen, "E_BITPARALLEL: memories are not supported"); return -1; }
    for (int b = 0; b < 4; b++)
        for (uint32_t k = 0; k < P->blk[b].npackets; k++) {
            const p_packet *pk = &P->blk[b].packets[k];
            for (uint32_t i = 0; i < pk->ninstr; i++) {
                uint32_t op = pk->instrs[i].op;
                if (!(op == OP_AND || op == OP_OR || op == OP_XOR || op == OP_NOT || op == OP_MUX || op == OP_SIGNAL || op == OP_LIT)) {
                    snprintf(err, errlen, "E_BITPARALLEL: op %u is not a gate-level bitwise op", op); return -1;
                }
            }
            for (uint32_t e = 0; e < pk->neff; e++)
                if (pk->effects[e].kind == EF_MEMWRITE) { snprintf(err, errlen, "E_BITPARALLEL: memory writes are not supported"); return -1; }
        }
    return 0;
}

bp_sim *bp_create(const void *blob, size_t nbytes, uint32_t lanes, char *err, size_t errlen) {
    bp_sim *s = calloc(1, sizeof(bp_sim));
    if (load_program(&s->P, blob, nbytes, err, errlen) || bp_validate(&s->P, err, errlen)) { free_program(&s->P); free(s); return NULL; }
    s->lanes = lanes; s->words = lanes ? (lanes + 63) / 64 : 1; s->nsig = s->P.nsig;
    s->state = calloc((size_t)s->nsig * s->words + 1, 8);
    for (int b = 0; b < 4; b++) s->work[b] = calloc((size_t)(s->P.blk[b].nvalues + 1) * s->words, 8);
    s->next = calloc((size_t)(s->nsig + 1) * s->words, 8);
    return s;
}
void bp_destroy(bp_sim *s) {
    if (!s) return;
    for (int b = 0; b < 4; b++) free(s->work[b]);
    free(s->state); free(s->next); free(s->forces); free_program(&s->P); free(s);
}
void bp_set_force(bp_sim *s, uint32_t sig, uint32_t lane, int value) {
    if (s->nforces == s->capforces) { s->capforces = s->capforces ? 2 * s->capforces : 64; s->forces = realloc(s->forces, s->capforces * sizeof(force_t)); }
    s->forces[s->nforces++] = (force_t){sig, lane, (uint32_t)(value != 0)};
}
void bp_clear_forces(bp_sim *s) { s->nforces = 0; }
static void set_bit_word(bp_sim *s, uint32_t sig, uint32_t lane, int v) {
    uint64_t *w = &s->state[(size_t)sig * s->words + lane / 64];
    if (v) *w |= 1ull << (lane % 64); else *w &= ~(1ull << (lane % 64));
}
static void apply_forces(bp_sim *s) { for (uint32_t i = 0; i < s->nforces; i++) set_bit_word(s, s->forces[i].sig, s->forces[i].lane, s->forces[i].value); }
static void clamp_signal(bp_sim *s, uint32_t sig) {
    for (uint32_t i = 0; i < s->nforces; i++) if (s->forces[i].sig == sig) set_bit_word(s, sig, s->forces[i].lane, s->forces[i].value);
}
void bp_set_signal_words(bp_sim *s, uint32_t sig, const uint64_t *w) { memcpy(&s->state[(size_t)sig * s->words], w, 8 * (size_t)s->words); }
void bp_get_signal_words(const bp_sim *s, uint32_t sig, uint64_t *w) { memcpy(w, &s->state[(size_t)sig * s->words], 8 * (size_t)s->words); }

/* eval_packet (:171-240) */
static void eval_packet(bp_sim *s, const p_packet *pk, uint64_t *work) {
    size_t W = s->words;
    for (uint32_t i = 0; i < pk->ninstr; i++) {
        const p_instr *in = &pk->instrs[i];
        uint64_t *d = &work[(size_t)in->dst * W];
        switch (in->op) {
        case OP_SIGNAL: memcpy(d, &s->state[(size_t)in->args[0] * W], 8 * W); break;
        case OP_LIT: { uint64_t f = (in->nargs && (in->args[0] & 1)) ? ~0ull : 0; for (size_t k = 0; k < W; k++) d[k] = f; break; }
        case OP_NOT: { const uint64_t *a = &work[(size_t)in->args[0] * W]; for (size_t k = 0; k < W; k++) d[k] = ~a[k]; break; }
        case OP_AND: case OP_OR: case OP_XOR: {
            const uint64_t *a = &work[(size_t)in->args[0] * W], *b = &work[(size_t)in->args[1] * W];
            if (in->op == OP_AND) for (size_t k = 0; k < W; k++) d[k] = a[k] & b[k];
            else if (in->op == OP_OR) for (size_t k = 0; k < W; k++) d[k] = a[k] | b[k];
            else for (size_t k = 0; k < W; k++) d[k] = a[k] ^ b[k];
            break; }
        case OP_MUX: {
            const uint64_t *c = &work[(size_t)in->args[0] * W], *t = &work[(size_t)in->args[1] * W], *e = &work[(size_t)in->args[2] * W];
            for (size_t k = 0; k < W; k++) d[k] = (c[k] & t[k]) | (~c[k] & e[k]);
            break; }
        }
    }
}
static uint64_t asserted_word(const bp_sim *s, uint32_t rsig, uint32_t rpol, size_t k) {
    uint64_t v = s->state[(size_t)rsig * s->words + k];
    return rpol == 0 ? v : ~v;
}
/* settle (:256-300) */
void bp_settle(bp_sim *s) {
    size_t W = s->words;
    for (int b = 0; b < 2; b++) {
        const p_block *bk = &s->P.blk[b];
        uint64_t *work = s->work[b];
        for (uint32_t k = 0; k < bk->npackets; k++) {
            const p_packet *pk = &bk->packets[k];
            eval_packet(s, pk, work);
            for (uint32_t e = 0; e < pk->neff; e++) {
                const p_effect *ef = &pk->effects[e];
                if (ef->kind == EF_STORE) {
                    memcpy(&s->state[(size_t)ef->dst * W], &work[(size_t)ef->value * W], 8 * W);
                    if (s->nforces) clamp_signal(s, ef->dst);
                } else if (ef->kind == EF_CAPTURE && ef->has_reset) {
                    uint64_t rv = (ef->nlimbs && (ef->rlimbs[0] & 1)) ? ~0ull : 0;
                    uint64_t *d = &s->state[(size_t)ef->dst * W];
                    for (size_t w = 0; w < W; w++) {
                        uint64_t a = asserted_word(s, ef->rsig, ef->rpol, w);
                        d[w] = (a & rv) | (~a & d[w]);
                    }
                    if (s->nforces) clamp_signal(s, ef->dst);
                }
            }
        }
    }
}
/* the edge of tick (:303-336) without its two settles */
static void bp_edge(bp_sim *s) {
    size_t W = s->words;
    const p_block *bk = &s->P.blk[2];
    uint64_t *work = s->work[2];
    for (uint32_t k = 0; k < bk->npackets; k++) eval_packet(s, &bk->packets[k], work);
    uint32_t nc = 0;
    uint32_t *dsts = malloc(sizeof(uint32_t) * (s->nsig + 1));
    for (uint32_t k = 0; k < bk->npackets; k++) {
        const p_packet *pk = &bk->packets[k];
        for (uint32_t e = 0; e < pk->neff; e++) {
            const p_effect *ef = &pk->effects[e];
            if (ef->kind != EF_CAPTURE) continue;
            uint64_t *nv = &s->next[(size_t)nc * W];
            memcpy(nv, &work[(size_t)ef->value * W], 8 * W);
            if (ef->has_reset) {
                uint64_t rv = (ef->nlimbs && (ef->rlimbs[0] & 1)) ? ~0ull : 0;
                for (size_t w = 0; w < W; w++) { uint64_t a = asserted_word(s, ef->rsig, ef->rpol, w); nv[w] = (a & rv) | (~a & nv[w]); }
            }
            dsts[nc++] = ef->dst;
        }
    }
    for (uint32_t c = 0; c < nc; c++) memcpy(&s->state[(size_t)dsts[c] * W], &s->next[(size_t)c * W], 8 * W);
    free(dsts);
    apply_forces(s);
}
void bp_tick(bp_sim *s) { bp_settle(s); bp_edge(s); bp_settle(s); }
void bp_tick_many(bp_sim *s, uint64_t n) { for (uint64_t i = 0; i < n; i++) bp_tick(s); }

static void bp_observe(bp_sim *s, const bp_run_spec *r, uint32_t cyc) {
    size_t W = s->words;
    for (uint32_t j = 0; j < r->n_obs; j++)
        if (r->trace) memcpy(&r->trace[((size_t)cyc * r->n_obs + j) * W], &s->state[(size_t)r->obs_sig[j] * W], 8 * W);
    if (!r->digest) return;
    uint64_t *v = malloc(8 * (size_t)r->n_obs + 8);
    for (size_t w = 0; w < W; w++) {
        for (uint32_t j = 0; j < r->n_obs; j++) v[j] = s->state[(size_t)r->obs_sig[j] * W + w];
        sig_digest_cycle(&r->digest[w], W, cyc, v, r->n_obs, 1);
    }
    free(v);
}
int bp_run(bp_sim *s, const bp_run_spec *r) {
    size_t W = s->words;
    for (uint32_t c = 0; c < r->cycles; c++) {
        for (uint32_t i = 0; i < r->n_inputs; i++) {
            if (r->input_mode[i] == IN_HOLD) continue;
            uint64_t key = stim_stream_key(r->seed, (uint64_t)r->input_sig[i] << 8, c);
            uint64_t *d = &s->state[(size_t)r->input_sig[i] * W];
            for (size_t w = 0; w < W; w++) d[w] = input_word(r, i, c, key, r->word_base + w, w);
        }
        bp_settle(s);
        if (!r->observe_post) bp_observe(s, r, c);
        bp_edge(s);
        if (r->observe_post) { bp_settle(s); bp_observe(s, r, c); }
    }
    bp_settle(s);
    return 0;
}

/* ================================================ 2. fast group engine */
enum { G_COPY, G_NOT, G_AND, G_OR, G_XOR, G_NAND, G_NOR, G_XNOR, G_MUX };
typedef struct { uint32_t op, a, b, c, dst; } gate_t;
typedef struct { uint32_t dst, val, rsig, rpol, has_reset, rbit; } commit_t;

typedef struct bpg_engine {
    p_program P;
    uint32_t lanes, words, groups, nsig, nslots;
    uint32_t c0, c1;              /* constant slots */
    uint32_t stage;               /* first capture staging slot */
    gate_t *async_g; uint32_t n_async; commit_t *async_c; uint32_t n_asyncc;
    gate_t *comb_g; uint32_t n_comb;
    gate_t *next_g; uint32_t n_next; commit_t *commits; uint32_t n_commits;
    uint64_t *st;                 /* [groups][nslots][G] */
    uint8_t *forced;              /* per slot: any force */
    force_t *forces; uint32_t nforces, capforces;
    int threads;
} bpg_engine;

typedef struct { gate_t *g; uint32_t n, cap; } gvec;
static void gpush(gvec *v, gate_t g) { if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 256; v->g = realloc(v->g, v->cap * sizeof(gate_t)); } v->g[v->n++] = g; }

/* Flatten one block.  Values are local to an op-packet group (the run of
 * machine packets ending in one carrying effects). */
static uint32_t compile_block(bpg_engine *e, const p_block *bk, gvec *out, commit_t **commits, uint32_t *ncommit,
                              int mode /*0 comb/async stores, 1 tick_next captures*/, uint32_t temp_base) {
    uint32_t nv = bk->nvalues + 1;
    uint32_t *slot = malloc(sizeof(uint32_t) * nv);
    int32_t *gate_of = malloc(sizeof(int32_t) * nv);
    uint32_t *uses = calloc(nv, sizeof(uint32_t));
    uint8_t *read_sig = calloc(e->nsig + 2, 1);
    uint32_t temp = temp_base, max_temp = temp_base;
    uint32_t gstart = 0;
    *ncommit = 0;
    if (commits) *commits = malloc(sizeof(commit_t) * (e->nsig + 1));
    uint32_t k0 = 0;
    for (uint32_t k = 0; k < bk->npackets; k++) {
        if (bk->packets[k].neff == 0 && k + 1 < bk->npackets) continue;
        /* group = packets k0..k */
        for (uint32_t q = k0; q <= k; q++) {
            const p_packet *pk = &bk->packets[q];
            for (uint32_t i = 0; i < pk->ninstr; i++) {
                const p_instr *in = &pk->instrs[i];
                if (in->op == OP_SIGNAL) read_sig[in->args[0]] = 1;
                else if (in->op == OP_NOT) uses[in->args[0]]++;
                else if (in->op == OP_AND || in->op == OP_OR || in->op == OP_XOR) { uses[in->args[0]]++; uses[in->args[1]]++; }
                else if (in->op == OP_MUX) { uses[in->args[0]]++; uses[in->args[1]]++; uses[in->args[2]]++; }
            }
            for (uint32_t f = 0; f < pk->neff; f++) uses[pk->effects[f].value]++;
        }
        temp = temp_base; gstart = out->n;
        for (uint32_t q = k0; q <= k; q++) {
            const p_packet *pk = &bk->packets[q];
            for (uint32_t i = 0; i < pk->ninstr; i++) {
                const p_instr *in = &pk->instrs[i];
                gate_of[in->dst] = -1;
                switch (in->op) {
                case OP_SIGNAL: slot[in->dst] = in->args[0]; break;
                case OP_LIT: slot[in->dst] = (in->nargs && (in->args[0] & 1)) ? e->c1 : e->c0; break;
                case OP_NOT: {
                    int32_t gi = gate_of[in->args[0]];
                    if (gi >= 0 && uses[in->args[0]] == 1 && out->g[gi].op >= G_AND && out->g[gi].op <= G_XOR) {
                        out->g[gi].op += 3; /* AND->NAND, OR->NOR, XOR->XNOR */
                        slot[in->dst] = out->g[gi].dst; gate_of[in->dst] = gi;
                    } else {
                        slot[in->dst] = temp++; gate_of[in->dst] = (int32_t)out->n;
                        gpush(out, (gate_t){G_NOT, slot[in->args[0]], 0, 0, slot[in->dst]});
                    }
                    break; }
                case OP_AND: case OP_OR: case OP_XOR:
                    slot[in->dst] = temp++; gate_of[in->dst] = (int32_t)out->n;
                    gpush(out, (gate_t){in->op == OP_AND ? G_AND : in->op == OP_OR ? G_OR : G_XOR,
                                        slot[in->args[0]], slot[in->args[1]], 0, slot[in->dst]});
                    break;
                case OP_MUX:
                    slot[in->dst] = temp++; gate_of[in->dst] = (int32_t)out->n;
                    gpush(out, (gate_t){G_MUX, slot[in->args[1]], slot[in->args[2]], slot[in->args[0]], slot[in->dst]});
                    break;
                }
            }
            if (temp > max_temp) max_temp = temp;
        }
        const p_packet *last = &bk->packets[k];
        for (uint32_t f = 0; f < last->neff; f++) {
            const p_effect *ef = &last->effects[f];
            if (mode == 0 && ef->kind == EF_STORE) {
                int32_t gi = gate_of[ef->value];
                if (gi >= 0 && (uint32_t)gi >= gstart && uses[ef->value] == 1 && !read_sig[ef->dst]) out->g[gi].dst = ef->dst;
                else gpush(out, (gate_t){G_COPY, slot[ef->value], 0, 0, ef->dst});
            } else if (ef->kind == EF_CAPTURE) {
                if (mode == 0 && !ef->has_reset) continue;
                commit_t c = {ef->dst, slot[ef->value], ef->rsig, ef->rpol, ef->has_reset,
                              (uint32_t)(ef->nlimbs && (ef->rlimbs[0] & 1))};
                if (mode == 0) c.val = 0; /* async capture writes the literal only on asserted lanes */
                (*commits)[(*ncommit)++] = c;
            }
        }
        for (uint32_t q = k0; q <= k; q++) {
            const p_packet *pk = &bk->packets[q];
            for (uint32_t i = 0; i < pk->ninstr; i++) if (pk->instrs[i].op == OP_SIGNAL) read_sig[pk->instrs[i].args[0]] = 0;
            for (uint32_t i = 0; i < pk->ninstr; i++) uses[pk->instrs[i].dst] = 0;
        }
        k0 = k + 1;
    }
    free(slot); free(gate_of); free(uses); free(read_sig);
    return max_temp;
}

bpg_engine *bpg_create(const void *blob, size_t nbytes, uint32_t lanes, int threads, char *err, size_t errlen) {
    bpg_engine *e = calloc(1, sizeof(bpg_engine));
    if (load_program(&e->P, blob, nbytes, err, errlen) || bp_validate(&e->P, err, errlen)) { free_program(&e->P); free(e); return NULL; }
    e->lanes = lanes; e->words = (lanes + 63) / 64; e->groups = (e->words + G - 1) / G;
    e->nsig = e->P.nsig; e->c0 = e->nsig; e->c1 = e->nsig + 1;
    uint32_t base = e->nsig + 2, mx = base;
    gvec ga = {0}, gc = {0}, gn = {0};
    commit_t *dummy = NULL; uint32_t nd = 0;
    uint32_t t;
    t = compile_block(e, &e->P.blk[0], &ga, &e->async_c, &e->n_asyncc, 0, base); if (t > mx) mx = t;
    t = compile_block(e, &e->P.blk[1], &gc, &dummy, &nd, 0, base); if (t > mx) mx = t;
    free(dummy);
    t = compile_block(e, &e->P.blk[2], &gn, &e->commits, &e->n_commits, 1, base); if (t > mx) mx = t;
    e->async_g = ga.g; e->n_async = ga.n; e->comb_g = gc.g; e->n_comb = gc.n; e->next_g = gn.g; e->n_next = gn.n;
    e->stage = mx;                 /* one staging slot per capture, after the temps */
    e->nslots = mx + e->n_commits;
    e->st = calloc((size_t)e->groups * e->nslots * G + 1, 8);
    for (uint32_t g = 0; g < e->groups; g++) for (int w = 0; w < G; w++) e->st[((size_t)g * e->nslots + e->c1) * G + w] = ~0ull;
    e->forced = calloc(e->nslots + 1, 1);
    e->threads = threads > 0 ? threads : omp_get_max_threads();
    return e;
}
void bpg_destroy(bpg_engine *e) {
    if (!e) return;
    free(e->async_g); free(e->async_c); free(e->comb_g); free(e->next_g); free(e->commits);
    free(e->st); free(e->forced); free(e->forces); free_program(&e->P); free(e);
}
uint32_t bpg_gate_count(const bpg_engine *e) { return e->n_comb; }
void bpg_set_force(bpg_engine *e, uint32_t sig, uint32_t lane, int value) {
    if (e->nforces == e->capforces) { e->capforces = e->capforces ? 2 * e->capforces : 64; e->forces = realloc(e->forces, e->capforces * sizeof(force_t)); }
    e->forces[e->nforces++] = (force_t){sig, lane, (uint32_t)(value != 0)};
    e->forced[sig] = 1;
}
void bpg_set_signal_words(bpg_engine *e, uint32_t sig, const uint64_t *w) {
    for (uint32_t k = 0; k < e->words; k++) e->st[((size_t)(k / G) * e->nslots + sig) * G + k % G] = w[k];
}
void bpg_get_signal_words(const bpg_engine *e, uint32_t sig, uint64_t *w) {
    for (uint32_t k = 0; k < e->words; k++) w[k] = e->st[((size_t)(k / G) * e->nslots + sig) * G + k % G];
}

static inline void clamp_group(const bpg_engine *e, uint64_t *st, uint32_t g, uint32_t slot) {
    for (uint32_t i = 0; i < e->nforces; i++) {
        const force_t *f = &e->forces[i];
        if (f->sig != slot || f->lane / (64 * G) != g) continue;
        uint64_t *w = &st[(size_t)slot * G + (f->lane / 64) % G];
        if (f->value) *w |= 1ull << (f->lane % 64); else *w &= ~(1ull << (f->lane % 64));
    }
}
static inline void run_gates(const bpg_engine *e, uint64_t *st, const gate_t *gs, uint32_t n, uint32_t g) {
    for (uint32_t i = 0; i < n; i++) {
        const gate_t q = gs[i];
        uint64_t *restrict d = &st[(size_t)q.dst * G];
        const uint64_t *a = &st[(size_t)q.a * G], *b = &st[(size_t)q.b * G], *c = &st[(size_t)q.c * G];
        uint64_t r[G];
        switch (q.op) {
        case G_COPY: for (int w = 0; w < G; w++) r[w] = a[w]; break;
        case G_NOT: for (int w = 0; w < G; w++) r[w] = ~a[w]; break;
        case G_AND: for (int w = 0; w < G; w++) r[w] = a[w] & b[w]; break;
        case G_OR: for (int w = 0; w < G; w++) r[w] = a[w] | b[w]; break;
        case G_XOR: for (int w = 0; w < G; w++) r[w] = a[w] ^ b[w]; break;
        case G_NAND: for (int w = 0; w < G; w++) r[w] = ~(a[w] & b[w]); break;
        case G_NOR: for (int w = 0; w < G; w++) r[w] = ~(a[w] | b[w]); break;
        case G_XNOR: for (int w = 0; w < G; w++) r[w] = ~(a[w] ^ b[w]); break;
        default: for (int w = 0; w < G; w++) r[w] = (c[w] & a[w]) | (~c[w] & b[w]); break;
        }
        for (int w = 0; w < G; w++) d[w] = r[w];
        if (e->forced[q.dst]) clamp_group(e, st, g, q.dst);
    }
}
static inline void group_settle(const bpg_engine *e, uint64_t *st, uint32_t g) {
    run_gates(e, st, e->async_g, e->n_async, g);
    for (uint32_t i = 0; i < e->n_asyncc; i++) {
        const commit_t *c = &e->async_c[i];
        uint64_t rv = c->rbit ? ~0ull : 0;
        for (int w = 0; w < G; w++) {
            uint64_t a = st[(size_t)c->rsig * G + w]; if (c->rpol) a = ~a;
            st[(size_t)c->dst * G + w] = (a & rv) | (~a & st[(size_t)c->dst * G + w]);
        }
        if (e->forced[c->dst]) clamp_group(e, st, g, c->dst);
    }
    run_gates(e, st, e->comb_g, e->n_comb, g);
}
static inline void group_edge(const bpg_engine *e, uint64_t *st, uint32_t g) {
    run_gates(e, st, e->next_g, e->n_next, g);
    /* stage every capture from the pre-edge state, then commit */
    for (uint32_t i = 0; i < e->n_commits; i++) {
        const commit_t *c = &e->commits[i];
        uint64_t *s = &st[(size_t)(e->stage + i) * G];
        const uint64_t *v = &st[(size_t)c->val * G];
        if (c->has_reset) {
            uint64_t rv = c->rbit ? ~0ull : 0;
            for (int w = 0; w < G; w++) { uint64_t a = st[(size_t)c->rsig * G + w]; if (c->rpol) a = ~a; s[w] = (a & rv) | (~a & v[w]); }
        } else {
            for (int w = 0; w < G; w++) s[w] = v[w];
        }
    }
    for (uint32_t i = 0; i < e->n_commits; i++) {
        const commit_t *c = &e->commits[i];
        for (int w = 0; w < G; w++) st[(size_t)c->dst * G + w] = st[(size_t)(e->stage + i) * G + w];
    }
    for (uint32_t i = 0; i < e->nforces; i++) {
        const force_t *f = &e->forces[i];
        if (f->lane / (64 * G) != g) continue;
        uint64_t *w = &st[(size_t)f->sig * G + (f->lane / 64) % G];
        if (f->value) *w |= 1ull << (f->lane % 64); else *w &= ~(1ull << (f->lane % 64));
    }
}

static void bpg_digest_word(const bpg_engine *e, const bp_run_spec *r, const uint64_t *st, uint32_t c, uint32_t w, uint32_t wl) {
    uint64_t v[4096];
    uint32_t n = r->n_obs < 4096 ? r->n_obs : 4096;
    for (uint32_t j = 0; j < n; j++) v[j] = st[(size_t)r->obs_sig[j] * G + wl];
    sig_digest_cycle(&r->digest[w], e->words, c, v, n, 1);
}

/* Runs every group for r->cycles cycles; group-parallel over threads.  The
 * digest/trace arrays use the same [64][words] / [cycles][n_obs][words] layout
 * as bp_run.  Returns elapsed seconds of the cycle loop. */
double bpg_run(bpg_engine *e, const bp_run_spec *r) {
    double t0 = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic, 1) num_threads(e->threads)
    for (uint32_t g = 0; g < e->groups; g++) {
        uint64_t *st = &e->st[(size_t)g * e->nslots * G];
        uint32_t wlo = g * G, whi = wlo + G < e->words ? wlo + G : e->words;
        for (uint32_t c = 0; c < r->cycles; c++) {
            for (uint32_t i = 0; i < r->n_inputs; i++) {
                if (r->input_mode[i] == IN_HOLD) continue;
                uint64_t key = stim_stream_key(r->seed, (uint64_t)r->input_sig[i] << 8, c);
                for (uint32_t w = 0; w < G; w++)
                    st[(size_t)r->input_sig[i] * G + w] = input_word(r, i, c, key, r->word_base + wlo + w, wlo + w < e->words ? wlo + w : 0);
            }
            group_settle(e, st, g);
            if (r->observe_post) group_edge(e, st, g), group_settle(e, st, g);
            for (uint32_t j = 0; j < r->n_obs; j++) {
                const uint64_t *v = &st[(size_t)r->obs_sig[j] * G];
                if (r->trace) for (uint32_t w = wlo; w < whi; w++) r->trace[((size_t)c * r->n_obs + j) * e->words + w] = v[w - wlo];
            }
            if (r->digest)
                for (uint32_t w = wlo; w < whi; w++)
                    bpg_digest_word(e, r, st, c, w, w - wlo);
            if (!r->observe_post) group_edge(e, st, g);
        }
        group_settle(e, st, g);
    }
    return omp_get_wtime() - t0;
}
