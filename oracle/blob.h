/*
 * ORACLE (test infrastructure only): parser for the packed machine-program blob
 * shared by the oracle engines.  Independent of the HIP library's parser.
 */
#ifndef RRTL_ORACLE_BLOB_H
#define RRTL_ORACLE_BLOB_H
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define MAXL 64 /* widest value handled: 2048 bits */

enum { OP_LIT, OP_SIGNAL, OP_NOT, OP_AND, OP_OR, OP_XOR, OP_ADD, OP_SUB, OP_MUL,
       OP_EQ, OP_NE, OP_LT, OP_MUX, OP_SLICE, OP_ZEXT, OP_SEXT, OP_TRUNC, OP_CAST,
       OP_CONCAT, OP_MEMREAD, OP_COUNT };
enum { EF_STORE, EF_CAPTURE, EF_MEMWRITE };

typedef struct { uint32_t op, dst, width, sgn, nargs; const uint32_t *args; } p_instr;
typedef struct {
    uint32_t kind, dst, value, has_reset, rsig, rkind, rpol, nlimbs;
    const uint32_t *rlimbs;
    uint32_t en, addr, data;
} p_effect;
typedef struct { uint32_t ninstr, neff; p_instr *instrs; p_effect *effects; } p_packet;
typedef struct { uint32_t nvalues, npackets; p_packet *packets; uint32_t *vwidth; uint32_t **ws; } p_block;
typedef struct { uint32_t width, sgn, offset, limbs, kind; } p_signal;
typedef struct { uint32_t width, sgn, depth, aw, offset, limbs; } p_memory;

typedef struct p_program {
    uint32_t *words; size_t nwords;
    p_signal *sig; uint32_t nsig;
    p_memory *mem; uint32_t nmem;
    uint32_t tsw, tmw;
    p_block blk[4];
} p_program;

static inline uint32_t nlimbs(uint32_t w) { uint32_t l = (w + 31) / 32; return l ? l : 1; }
static inline uint32_t final_mask(uint32_t w) { uint32_t r = w % 32; return r ? ((1u << r) - 1) : 0xFFFFFFFFu; }

/* fit_limbs: resize to limbs(width) (zero-extend / truncate) and mask the top limb */
static inline void fit(uint32_t *v, uint32_t have, uint32_t width) {
    uint32_t n = nlimbs(width);
    for (uint32_t i = have; i < n; i++) v[i] = 0;
    v[n - 1] &= final_mask(width);
}
static inline int get_bit(const uint32_t *v, uint32_t n, uint32_t bit) {
    uint32_t l = bit / 32;
    return l < n ? (v[l] >> (bit % 32)) & 1 : 0;
}

/* ---------------------------------------------------------------- parsing */
static inline int parse_program(p_program *s, char *err, size_t errlen) {
    const uint32_t *w = s->words; size_t n = s->nwords, p = 0;
#define NEED(k) do { if (p + (k) > n) { snprintf(err, errlen, "E_PROGRAM_TRUNCATED"); return -1; } } while (0)
    NEED(8);
    if (w[0] != 0x4C545252u || w[1] != 0x31475250u || w[2] != 1) { snprintf(err, errlen, "E_PROGRAM_MAGIC"); return -1; }
    s->nsig = w[3]; s->nmem = w[4]; s->tsw = w[5]; s->tmw = w[6]; p = 8;
    s->sig = calloc(s->nsig + 1, sizeof(p_signal));
    s->mem = calloc(s->nmem + 1, sizeof(p_memory));
    NEED((size_t)5 * s->nsig);
    for (uint32_t i = 0; i < s->nsig; i++, p += 5) {
        p_signal *g = &s->sig[i];
        g->width = w[p]; g->sgn = w[p + 1]; g->offset = w[p + 2]; g->limbs = w[p + 3]; g->kind = w[p + 4];
    }
    NEED((size_t)6 * s->nmem);
    for (uint32_t i = 0; i < s->nmem; i++, p += 6) {
        p_memory *m = &s->mem[i];
        m->width = w[p]; m->sgn = w[p + 1]; m->depth = w[p + 2]; m->aw = w[p + 3]; m->offset = w[p + 4]; m->limbs = w[p + 5];
    }
    for (int b = 0; b < 4; b++) {
        p_block *bk = &s->blk[b];
        NEED(2);
        bk->nvalues = w[p]; bk->npackets = w[p + 1]; p += 2;
        bk->packets = calloc(bk->npackets + 1, sizeof(p_packet));
        bk->vwidth = calloc(bk->nvalues + 1, sizeof(uint32_t));
        for (uint32_t k = 0; k < bk->npackets; k++) {
            p_packet *pk = &bk->packets[k];
            NEED(2);
            pk->ninstr = w[p]; pk->neff = w[p + 1]; p += 2;
            pk->instrs = calloc(pk->ninstr + 1, sizeof(p_instr));
            pk->effects = calloc(pk->neff + 1, sizeof(p_effect));
            for (uint32_t i = 0; i < pk->ninstr; i++) {
                NEED(5);
                p_instr *in = &pk->instrs[i];
                in->op = w[p]; in->dst = w[p + 1]; in->width = w[p + 2]; in->sgn = w[p + 3]; in->nargs = w[p + 4];
                p += 5; NEED(in->nargs); in->args = &w[p]; p += in->nargs;
                if (in->dst >= bk->nvalues || in->op >= OP_COUNT || in->width == 0) { snprintf(err, errlen, "E_PROGRAM_INSTR"); return -1; }
                if (nlimbs(in->width) > MAXL) { snprintf(err, errlen, "E_PROGRAM_WIDE: value wider than %d bits", MAXL * 32); return -1; }
                bk->vwidth[in->dst] = in->width;
            }
            for (uint32_t e = 0; e < pk->neff; e++) {
                p_effect *ef = &pk->effects[e];
                NEED(1); ef->kind = w[p++];
                if (ef->kind == EF_STORE) { NEED(2); ef->dst = w[p]; ef->value = w[p + 1]; p += 2; }
                else if (ef->kind == EF_CAPTURE) {
                    NEED(7);
                    ef->dst = w[p]; ef->value = w[p + 1]; ef->has_reset = w[p + 2]; ef->rsig = w[p + 3];
                    ef->rkind = w[p + 4]; ef->rpol = w[p + 5]; ef->nlimbs = w[p + 6]; p += 7;
                    NEED(ef->nlimbs); ef->rlimbs = &w[p]; p += ef->nlimbs;
                } else if (ef->kind == EF_MEMWRITE) {
                    NEED(4); ef->dst = w[p]; ef->en = w[p + 1]; ef->addr = w[p + 2]; ef->data = w[p + 3]; p += 4;
                } else { snprintf(err, errlen, "E_PROGRAM_EFFECT"); return -1; }
            }
        }
        bk->ws = calloc(bk->nvalues + 1, sizeof(uint32_t *));
    }
#undef NEED
    return 0;
}


static inline int load_program(p_program *s, const void *blob, size_t nbytes, char *err, size_t errlen) {
    memset(s, 0, sizeof(*s));
    s->nwords = nbytes / 4;
    s->words = malloc(s->nwords * 4 + 4);
    memcpy(s->words, blob, s->nwords * 4);
    return parse_program(s, err, errlen);
}

static inline void free_program(p_program *s) {
    for (int b = 0; b < 4; b++) {
        p_block *bk = &s->blk[b];
        for (uint32_t k = 0; k < bk->npackets; k++) { free(bk->packets[k].instrs); free(bk->packets[k].effects); }
        if (bk->ws) for (uint32_t v = 0; v < bk->nvalues; v++) free(bk->ws[v]);
        free(bk->packets); free(bk->vwidth); free(bk->ws);
    }
    free(s->sig); free(s->mem); free(s->words);
}
#endif
