// Device side of the word-level engine: record encoding, run arguments and the
// per-lane interpreter body.  Shared by the word engine's own launches and by
// the multi-design batch kernel.
#pragma once
#include "common.hpp"

namespace rrtl {
namespace word {

constexpr uint32_t MAXL = 8;           // widest value: 256 bits
constexpr uint32_t SP_SCRATCH = 0u << 30, SP_SIGNAL = 1u << 30, SP_CONST = 2u << 30, SP_MASK = 3u << 30;
constexpr uint32_t NONE = 0xFFFFFFFFu;

// record ops: machine ops keep their numbers, effects follow
enum : uint32_t { R_STORE = 32, R_ASYNC_CAPTURE, R_CAPTURE, R_COMMIT, R_MEMWRITE, R_SIGLOAD };

struct Rec { uint32_t w[8]; };  // op|limbs<<8, dst, a, b, c, width, aux0, aux1

// ------------------------------------------------------------------ kernel
enum : uint32_t { F_EVAL_FIRST = 1, F_EVAL_PRE = 2, F_EVAL_POST = 4, F_FINAL_EVAL = 8, F_OBS_POST = 16, F_EDGE = 32 };

struct WordArgs {
  uint32_t* values; uint32_t* mems; uint32_t* scratch;
  const Rec* blk[4]; uint32_t nrec[4];
  const uint32_t* cpool; const uint32_t* mem_table;
  uint32_t lanes, flags;
  uint64_t cycles, cycle_base, seed;
  // stimulus: per input (offset, limbs, width, mode), arg, stream base
  const uint32_t* in_desc; const uint64_t* in_arg; uint32_t n_in;
  const uint32_t* trace_in; uint32_t trace_in_limbs;      // [cycles][trace_in_limbs][lanes]
  const uint32_t* obs_desc; uint32_t n_obs, n_obs_limbs;  // per observed: (offset, limbs)
  uint64_t* digest; uint32_t* trace;                      // [lanes] / [cycles][n_obs_limbs][lanes]
  uint32_t halt_offset; uint64_t* halt_cycles;            // NONE = no halting
};

struct Ctx {
  const WordArgs* A;
  uint32_t lane;
  uint32_t* scr; uint32_t sstride;
  __device__ __forceinline__ uint32_t ld(uint32_t code, uint32_t limb) const {
    const uint32_t idx = code & ~SP_MASK;
    switch (code & SP_MASK) {
      case SP_SCRATCH: return scr[size_t(idx + limb) * sstride];
      case SP_SIGNAL: return A->values[size_t(idx + limb) * A->lanes + lane];
      default: return A->cpool[idx + limb];
    }
  }
  // operand of `have` limbs zero-extended to `n`
  __device__ __forceinline__ void ldv(uint32_t code, uint32_t have, uint32_t n, uint32_t* v) const {
    for (uint32_t l = 0; l < n; l++) v[l] = l < have ? ld(code, l) : 0u;
  }
  __device__ __forceinline__ void st(uint32_t slot, uint32_t limb, uint32_t v) const { scr[size_t(slot + limb) * sstride] = v; }
  __device__ __forceinline__ void stsig(uint32_t off, uint32_t limb, uint32_t v) const { A->values[size_t(off + limb) * A->lanes + lane] = v; }
};

__device__ __forceinline__ uint32_t fmask(uint32_t w) { uint32_t r = w & 31; return r ? ((1u << r) - 1) : 0xFFFFFFFFu; }
__device__ __forceinline__ uint32_t get_bit(const uint32_t* v, uint32_t n, uint32_t bit) { uint32_t l = bit >> 5; return l < n ? (v[l] >> (bit & 31)) & 1u : 0u; }
__device__ __forceinline__ uint64_t decode_usize(const uint32_t* v, uint32_t n) {
  uint64_t out = v[0] | (n > 1 ? uint64_t(v[1]) << 32 : 0);
  for (uint32_t l = 2; l < n; l++) if (v[l]) return ~0ull;
  return out;
}

__device__ void exec_rec(const Ctx& C, const Rec& R) {
  const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF, width = R.w[5];
  uint32_t a[MAXL], b[MAXL], r[MAXL];
  if (L == 1 && op <= OP_MEMREAD && op != OP_CONCAT && op != OP_MEMREAD && op != OP_SLICE && op != OP_EQ && op != OP_NE &&
      op != OP_LT && op != OP_SEXT) {
    // one-limb fast path
    const uint32_t x = C.ld(R.w[2], 0);
    uint32_t v;
    switch (op) {
      case OP_NOT: v = ~x; break;
      case OP_AND: v = x & C.ld(R.w[3], 0); break;
      case OP_OR: v = x | C.ld(R.w[3], 0); break;
      case OP_XOR: v = x ^ C.ld(R.w[3], 0); break;
      case OP_ADD: v = x + C.ld(R.w[3], 0); break;
      case OP_SUB: v = x - C.ld(R.w[3], 0); break;
      case OP_MUL: v = x * C.ld(R.w[3], 0); break;
      case OP_MUX: v = (C.ld(R.w[4], 0) & 1u) ? x : C.ld(R.w[3], 0); break;
      default: v = x; break;  // ZEXT / TRUNC / CAST
    }
    C.st(R.w[1], 0, v & fmask(width));
    return;
  }
  switch (op) {
    case R_SIGLOAD: {
      C.ldv(R.w[2], R.w[7], L, r);
      break;
    }
    case OP_NOT: C.ldv(R.w[2], R.w[7], L, a); for (uint32_t l = 0; l < L; l++) r[l] = ~a[l]; break;
    case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: {
      C.ldv(R.w[2], R.w[7] & 0xFF, L, a);
      C.ldv(R.w[3], (R.w[7] >> 8) & 0xFF, L, b);
      if (op == OP_AND) for (uint32_t l = 0; l < L; l++) r[l] = a[l] & b[l];
      else if (op == OP_OR) for (uint32_t l = 0; l < L; l++) r[l] = a[l] | b[l];
      else if (op == OP_XOR) for (uint32_t l = 0; l < L; l++) r[l] = a[l] ^ b[l];
      else if (op == OP_ADD) { uint64_t c = 0; for (uint32_t l = 0; l < L; l++) { uint64_t t = uint64_t(a[l]) + b[l] + c; r[l] = uint32_t(t); c = t >> 32; } }
      else if (op == OP_SUB) { uint64_t br = 0; for (uint32_t l = 0; l < L; l++) { uint64_t s = uint64_t(b[l]) + br, x = a[l]; r[l] = uint32_t(x - s); br = x < s; } }
      else {
        for (uint32_t l = 0; l < L; l++) r[l] = 0;
        for (uint32_t j = 0; j < L; j++) {
          uint64_t c = 0;
          for (uint32_t k = 0; j + k < L; k++) { uint64_t t = uint64_t(a[j]) * b[k] + r[j + k] + c; r[j + k] = uint32_t(t); c = t >> 32; }
        }
      }
      break;
    }
    case OP_EQ: case OP_NE: {
      uint32_t la = R.w[7] & 0xFF, lb = (R.w[7] >> 8) & 0xFF, n = la > lb ? la : lb;
      uint32_t eq = 1;
      for (uint32_t l = 0; l < n; l++) eq &= (l < la ? C.ld(R.w[2], l) : 0u) == (l < lb ? C.ld(R.w[3], l) : 0u);
      r[0] = op == OP_EQ ? eq : (eq ^ 1u);
      for (uint32_t l = 1; l < L; l++) r[l] = 0;
      break;
    }
    case OP_LT: {
      const uint32_t wl = R.w[6], n = (wl + 31) / 32;
      C.ldv(R.w[2], R.w[7] & 0xFF, n, a);
      C.ldv(R.w[3], (R.w[7] >> 8) & 0xFF, n, b);
      a[n - 1] &= fmask(wl); b[n - 1] &= fmask(wl);
      uint32_t res = 0, decided = 0;
      if (R.w[7] >> 24) {
        uint32_t sa = get_bit(a, n, wl - 1), sb = get_bit(b, n, wl - 1);
        if (sa != sb) { res = sa; decided = 1; }
      }
      for (int l = int(n) - 1; l >= 0 && !decided; l--) if (a[l] != b[l]) { res = a[l] < b[l]; decided = 1; }
      r[0] = res;
      for (uint32_t l = 1; l < L; l++) r[l] = 0;
      break;
    }
    case OP_MUX: {
      const bool c = C.ld(R.w[4], 0) & 1u;
      C.ldv(c ? R.w[2] : R.w[3], c ? (R.w[7] & 0xFF) : ((R.w[7] >> 8) & 0xFF), L, r);
      break;
    }
    case OP_SLICE: {
      const uint32_t n = R.w[7], lsb = R.w[6];
      C.ldv(R.w[2], n, n, a);
      for (uint32_t l = 0; l < L; l++) {
        const uint32_t bit = lsb + 32 * l, li = bit >> 5, sh = bit & 31;
        uint32_t lo = li < n ? a[li] : 0u, hi = li + 1 < n ? a[li + 1] : 0u;
        r[l] = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
      }
      break;
    }
    case OP_ZEXT: case OP_TRUNC: case OP_CAST: C.ldv(R.w[2], R.w[7], L, r); break;
    case OP_SEXT: {
      const uint32_t iw = R.w[6], n = R.w[7];
      C.ldv(R.w[2], n, L, r);
      const uint32_t sign = get_bit(r, L, iw - 1);
      // clear above iw, then fill with the sign
      for (uint32_t l = 0; l < L; l++) {
        uint32_t lo_bit = 32 * l;
        if (lo_bit >= iw) r[l] = sign ? ~0u : 0u;
        else if (lo_bit + 32 > iw) {
          uint32_t keep = (1u << (iw - lo_bit)) - 1;
          r[l] = (r[l] & keep) | (sign ? ~keep : 0u);
        }
      }
      break;
    }
    case OP_CONCAT: {
      for (uint32_t l = 0; l < L; l++) r[l] = 0;
      uint32_t off = 0;
      for (int k = int(R.w[3]) - 1; k >= 0; k--) {
        const uint32_t code = C.A->cpool[R.w[2] + 2 * k], wd = C.A->cpool[R.w[2] + 2 * k + 1] & 0xFFFF,
                       n = C.A->cpool[R.w[2] + 2 * k + 1] >> 16;
        for (uint32_t l = 0; l < n; l++) {
          uint32_t v = C.ld(code, l);
          if (l == n - 1) v &= fmask(wd);
          const uint32_t bit = off + 32 * l, li = bit >> 5, sh = bit & 31;
          if (li < L) r[li] |= v << sh;
          if (sh && li + 1 < L) r[li + 1] |= v >> (32 - sh);
        }
        off += wd;
      }
      break;
    }
    case OP_MEMREAD: {
      const uint32_t* mt = &C.A->mem_table[4 * R.w[6]];
      C.ldv(R.w[2], R.w[7], R.w[7], a);
      const uint64_t addr = decode_usize(a, R.w[7]);
      for (uint32_t l = 0; l < L; l++) r[l] = 0;
      if (addr < mt[1]) {
        for (uint32_t l = 0; l < mt[2] && l < L; l++)
          r[l] = C.A->mems[(size_t(mt[0]) + addr * mt[2] + l) * C.A->lanes + C.lane];
        if (mt[2] <= L) r[mt[2] - 1] &= fmask(mt[3]);
      }
      break;
    }
  }
  r[L - 1] &= fmask(width);
  for (uint32_t l = 0; l < L; l++) C.st(R.w[1], l, r[l]);
}

__device__ __forceinline__ bool reset_asserted(const Ctx& C, uint32_t rsig_code, uint32_t pol) {
  const uint32_t v = C.ld(rsig_code, 0) & 1u;
  return pol == 0 ? v != 0 : v == 0;
}

__device__ void exec_effect(const Ctx& C, const Rec& R) {
  const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF;
  uint32_t v[MAXL];
  switch (op) {
    case R_STORE:
      C.ldv(R.w[2], R.w[7], L, v);
      v[L - 1] &= fmask(R.w[5]);
      for (uint32_t l = 0; l < L; l++) C.stsig(R.w[1], l, v[l]);
      break;
    case R_ASYNC_CAPTURE:
      if (reset_asserted(C, R.w[3], R.w[6])) {
        C.ldv(R.w[2], R.w[7], L, v);
        v[L - 1] &= fmask(R.w[5]);
        for (uint32_t l = 0; l < L; l++) C.stsig(R.w[1], l, v[l]);
      }
      break;
    case R_CAPTURE:
      if ((R.w[6] & 2u) && reset_asserted(C, R.w[3], R.w[6] & 1u)) for (uint32_t l = 0; l < L; l++) v[l] = C.ld(R.w[4], l);
      else C.ldv(R.w[2], R.w[7], L, v);
      v[L - 1] &= fmask(R.w[5]);
      for (uint32_t l = 0; l < L; l++) C.st(R.w[1], l, v[l]);
      break;
    case R_COMMIT:
      for (uint32_t l = 0; l < L; l++) C.stsig(R.w[1], l, C.ld(R.w[2], l));
      break;
    case R_MEMWRITE: {
      if (!(C.ld(R.w[2], 0) & 1u)) break;
      const uint32_t* mt = &C.A->mem_table[4 * R.w[1]];
      uint32_t la = R.w[7] & 0xFF, ld = (R.w[7] >> 8) & 0xFF;
      C.ldv(R.w[3], la, la, v);
      const uint64_t addr = decode_usize(v, la);
      if (addr >= mt[1]) break;
      C.ldv(R.w[4], ld, L, v);
      v[L - 1] &= fmask(mt[3]);
      for (uint32_t l = 0; l < L; l++) C.A->mems[(size_t(mt[0]) + addr * mt[2] + l) * C.A->lanes + C.lane] = v[l];
      break;
    }
  }
}

__device__ void exec_block(const Ctx& C, int b) {
  const Rec* recs = C.A->blk[b];
  const uint32_t n = C.A->nrec[b];
  for (uint32_t i = 0; i < n; i++) {
    const Rec R = recs[i];  // uniform: every lane runs the same record
    if ((R.w[0] & 0xFF) >= R_STORE && (R.w[0] & 0xFF) != R_SIGLOAD) exec_effect(C, R);
    else exec_rec(C, R);
  }
}

__device__ __forceinline__ void eval_comb(const Ctx& C) { exec_block(C, B_COMB); exec_block(C, B_ASYNC); }
__device__ __forceinline__ void edge(const Ctx& C) { exec_block(C, B_NEXT); exec_block(C, B_COMMIT); }

__device__ void apply_stimulus(const Ctx& C, uint64_t c, uint64_t cyc) {
  const WordArgs& A = *C.A;
  uint32_t trace_limb = 0;
  for (uint32_t i = 0; i < A.n_in; i++) {
    const uint32_t off = A.in_desc[4 * i], limbs = A.in_desc[4 * i + 1], width = A.in_desc[4 * i + 2], mode = A.in_desc[4 * i + 3];
    const uint64_t arg = A.in_arg[i];
    const uint32_t sig_index = uint32_t(arg >> 40);  // stream id for IN_RANDOM
    for (uint32_t l = 0; l < limbs; l++) {
      uint32_t v = 0;
      if (mode == IN_RANDOM) {
        const uint32_t nb = min(32u, width - 32 * l);
        for (uint32_t bit = 0; bit < nb; bit++) {
          const uint64_t stream = (uint64_t(sig_index) << 8) | (32 * l + bit);
          const uint64_t w = stim_from_key(stim_key(A.seed, stream, cyc), C.lane >> 6);
          v |= uint32_t((w >> (C.lane & 63)) & 1u) << bit;
        }
      } else if (mode == IN_CONST) {
        v = l == 0 ? uint32_t(arg) : (l == 1 ? uint32_t(arg >> 32) : 0u);
      } else if (mode == IN_PULSE) {
        v = l == 0 ? uint32_t(cyc < (arg & 0xFFFFFFFFull) ? (arg >> 32) & 1 : (arg >> 33) & 1) : 0u;
      } else if (mode == IN_TRACE) {
        v = A.trace_in[(size_t(c) * A.trace_in_limbs + trace_limb + l) * A.lanes + C.lane];
      } else {
        continue;
      }
      if (l == limbs - 1) v &= fmask(width);
      C.stsig(off, l, v);
    }
    if (mode == IN_TRACE) trace_limb += limbs;
  }
}

__device__ void observe(const Ctx& C, uint64_t c, uint64_t& d) {
  const WordArgs& A = *C.A;
  uint32_t k = 0;
  for (uint32_t j = 0; j < A.n_obs; j++) {
    const uint32_t off = A.obs_desc[2 * j], limbs = A.obs_desc[2 * j + 1];
    for (uint32_t l = 0; l < limbs; l++, k++) {
      const uint32_t v = A.values[size_t(off + l) * A.lanes + C.lane];
      if (A.trace) A.trace[(size_t(c) * A.n_obs_limbs + k) * A.lanes + C.lane] = v;
      if (A.digest) d = lane_digest(d, v, j, l);
    }
  }
}

// One block of lanes; lanes past the batch skip (no barriers inside).
template <bool kLds>
__device__ void word_body(const WordArgs& A, uint32_t block, uint32_t* lds) {
  const uint32_t lane = block * blockDim.x + threadIdx.x;
  if (lane >= A.lanes) return;
  Ctx C;
  C.A = &A; C.lane = lane;
  if (kLds) { C.scr = lds + threadIdx.x; C.sstride = blockDim.x; }
  else { C.scr = A.scratch + lane; C.sstride = A.lanes; }
  uint64_t d = A.digest ? A.digest[lane] : 0;
  uint64_t done = 0;
  bool halted = false;
  if (A.halt_cycles) { done = A.halt_cycles[lane] >> 1; halted = A.halt_cycles[lane] & 1; }
  if (A.flags & F_EVAL_FIRST) eval_comb(C);
  for (uint64_t c = 0; c < A.cycles && !halted; c++) {
    const uint64_t cyc = A.cycle_base + c;
    if (A.n_in) apply_stimulus(C, c, cyc);
    if (A.flags & F_EVAL_PRE) eval_comb(C);
    if (A.n_obs && !(A.flags & F_OBS_POST)) observe(C, c, d);
    if (A.flags & F_EDGE) edge(C);
    if (A.flags & F_EVAL_POST) eval_comb(C);
    if (A.n_obs && (A.flags & F_OBS_POST)) observe(C, c, d);
    done++;
    if (A.halt_offset != NONE && (A.values[size_t(A.halt_offset) * A.lanes + lane] & 1u)) halted = true;
  }
  if (A.flags & F_FINAL_EVAL) eval_comb(C);
  if (A.digest) A.digest[lane] = d;
  if (A.halt_cycles) A.halt_cycles[lane] = (done << 1) | (halted ? 1u : 0u);
}

template <bool kLds>
__global__ void __launch_bounds__(256) word_kernel(WordArgs A) {
  extern __shared__ uint32_t lds[];
  word_body<kLds>(A, blockIdx.x, lds);
}

}  // namespace word
}  // namespace rrtl
