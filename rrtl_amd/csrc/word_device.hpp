// Device side of the word-level engine: record encoding, run arguments and the
// per-lane interpreter body.  Shared by the word engine's own launches and by
// the multi-design batch kernel.
#pragma once
#include "device_defs.hpp"

namespace rrtl {
namespace word {

constexpr uint32_t MAXL = 8;           // widest value: 256 bits
constexpr uint32_t SP_SCRATCH = 0u << 30, SP_SIGNAL = 1u << 30, SP_CONST = 2u << 30, SP_MASK = 3u << 30;
constexpr uint32_t NONE = 0xFFFFFFFFu;

// record ops: machine ops keep their numbers, effects follow
enum : uint32_t { R_STORE = 40, R_ASYNC_CAPTURE, R_CAPTURE, R_COMMIT, R_MEMWRITE, R_SIGLOAD, R_MACRO };
// effect records (run by exec_effect); the others produce a value
__host__ __device__ inline bool is_effect_op(uint32_t op) { return op >= R_STORE && op < R_SIGLOAD; }

// Macro records (the interpreter's fused superoperator, the counterpart of the
// reference's fuse_macros, crates/rrtl-gpu-sim/src/interp.rs:696-826): a cone
// of one-limb operations whose intermediate values each have one reader runs
// as one record, the intermediates in the record's own registers instead of
// scratch slots.  w[2]: offset of the sub-operation records (8 words each) in
// the constant pool, w[3]: their count (at most MACRO_REGS, the root last),
// w[1]: the root's scratch slot.  A sub-operation writes register k (its
// index, in w[1]); operand code SP_LOCAL | k reads it.
constexpr uint32_t SP_LOCAL = 3u << 30;
constexpr uint32_t MACRO_REGS = 16;
// Fused one-limb records (the fusion menu of the reference's interpreter,
// crates/rrtl-gpu-sim/src/interp.rs:351-495): an inner operation whose value
// has no other reader folded into its reader, operands a, b, c in w[2..4] and
// for the four-input cells d in w[6].  Results are masked to the record's width
// like every one-limb record.
enum : uint32_t {
  FOP_NAND = OP_COUNT,  // ~(a & b)
  FOP_NOR,              // ~(a | b)
  FOP_XNOR,             // ~(a ^ b)
  FOP_ANDOR,            // (a & b) | c
  FOP_ORAND,            // (a | b) & c
  FOP_XOR3,             // a ^ b ^ c
  FOP_ADD3,             // a + b + c
  FOP_MULADD,           // a * b + c
  FOP_AOI21,            // ~((a & b) | c)
  FOP_OAI21,            // ~((a | b) & c)
  FOP_MULSUB,           // a * b - c
  FOP_AOI22,            // ~((a & b) | (c & d))
  FOP_OAI22,            // ~((a | b) & (c | d))
  FOP_END
};
static_assert(FOP_END <= R_STORE, "fused opcodes must stay below the effect records");

struct Rec { uint32_t w[8]; };  // op|limbs<<8, dst, a, b, c, width, aux0, aux1

// ------------------------------------------------------------------ kernel
enum : uint32_t { F_EVAL_FIRST = 1, F_EVAL_PRE = 2, F_EVAL_POST = 4, F_FINAL_EVAL = 8, F_OBS_POST = 16, F_EDGE = 32 };

struct WordArgs {
  uint32_t* values; uint32_t* mems; uint32_t* scratch;
  const Rec* blk[4]; uint32_t nrec[4];
  const uint32_t* cpool; const uint32_t* mem_table;
  uint32_t lanes, flags;
  uint64_t cycles, cycle_base, seed;
  // stimulus: per input (offset, limbs, width, mode | trace limb base << 8), arg;
  // the first n_in_once inputs hold one value for the whole run (IN_CONST) and
  // are written once per launch, the rest every cycle
  const uint32_t* in_desc; const uint64_t* in_arg; uint32_t n_in, n_in_once;
  const uint32_t* trace_in; uint32_t trace_in_limbs;      // [cycles][trace_in_limbs][lanes]
  const uint32_t* obs_desc; uint32_t n_obs, n_obs_limbs;  // per observed: (offset, limbs)
  uint64_t* digest; uint32_t* trace;                      // [lanes] / [cycles][n_obs_limbs][lanes]
  uint32_t halt_offset; uint64_t* halt_cycles;            // NONE = no halting
  uint32_t n_scratch, n_sig_words;                        // scratch slots; signal words (LDS signal mode)
  uint32_t block_lanes;                                   // lanes per workgroup (0: one per thread); a
                                                          // smaller count leaves the higher threads idle
  // Trace replay (SimBackend::replay_trace moved on chip): inputs shared by all
  // lanes, and checks of the observed signals against expected values.
  const uint32_t* trace_uni; uint32_t trace_uni_limbs;    // IN_TRACE_UNIFORM values [cycles][trace_uni_limbs]
  const uint32_t* expect; uint32_t expect_per_lane;       // [cycles][n_obs_limbs] or [cycles][n_obs_limbs][lanes]
  const uint32_t* check_active; uint32_t check_words;     // [cycles][check_words]: bit j checks observed j; null: all
  uint32_t check_lanes;                                   // lanes below this compare (UINT32_MAX: all)
  uint32_t* mm_count; uint32_t* mm_rec; uint32_t mm_max;  // per lane: mismatch count; first mm_max records
                                                          // [lanes][mm_max][6] = (cycle, check, actual limbs 0..3)
  // Lane compaction (null lane_map: thread slot t runs lane t).  Workgroups are
  // dealt round-robin to the XCDs, so the slots of workgroup b belong to
  // partition b % 8; slot j of partition x runs lane lane_map[x * map_cap + j]
  // for j < live[x].
  const uint32_t* lane_map; const uint32_t* live; uint32_t map_cap;
};

// Where a block keeps its per-lane scratch and signal values during a launch:
// WM_HBM both in HBM; WM_SCRATCH_LDS scratch in LDS; WM_ALL_LDS scratch and the
// block's signal values in LDS (loaded at the start of the launch, written back
// at the end; memories stay in HBM).  LDS columns are per thread, rows per
// word, so a block access is conflict-free.
enum : int { WM_HBM = 0, WM_SCRATCH_LDS = 1, WM_ALL_LDS = 2 };

// A read of launch-constant data (records, constant pool, memory table)
// through the constant address space, so uniform reads become scalar loads.
typedef const __attribute__((address_space(4))) uint32_t const_u32;
__device__ __forceinline__ uint32_t cload(const uint32_t* p, uint32_t i) {
  return ((const_u32*)p)[i];  // address-space cast
}

__device__ __forceinline__ uint32_t* lds_words() {
  extern __shared__ __align__(16) uint32_t word_lds[];
  return word_lds;
}

// Per-lane view of the run: operand spaces and this lane's index.  Held by
// value (never a pointer to the kernel arguments, which would force them into
// private memory); LDS areas are addressed from the block's LDS base so every
// access compiles to an LDS instruction.
template <int kMode>
struct Ctx {
  uint32_t* values; uint32_t* mems; const uint32_t* cpool; const uint32_t* mem_table;
  uint32_t lanes, lane;
  uint32_t* scr; uint32_t sstride;   // HBM scratch (WM_HBM)
  uint32_t col, stride, sig0, mem0;  // LDS: thread column, block stride, first signal / memory row offsets
  uint32_t wave;                     // which of the waves sharing these lanes (generated multi-wave kernels)
  __device__ __forceinline__ uint32_t ld(uint32_t code, uint32_t limb) const {
    const uint32_t idx = code & ~SP_MASK;
    switch (code & SP_MASK) {
      case SP_SCRATCH:
        if constexpr (kMode >= WM_SCRATCH_LDS) return lds_words()[(idx + limb) * stride + col];
        else return scr[size_t(idx + limb) * sstride];
      case SP_SIGNAL:
        if constexpr (kMode == WM_ALL_LDS) return lds_words()[sig0 + (idx + limb) * stride + col];
        else return values[size_t(idx + limb) * lanes + lane];
      default: return cload(cpool, idx + limb);
    }
  }
  __device__ __forceinline__ void st(uint32_t slot, uint32_t limb, uint32_t v) const {
    if constexpr (kMode >= WM_SCRATCH_LDS) lds_words()[(slot + limb) * stride + col] = v;
    else scr[size_t(slot + limb) * sstride] = v;
  }
  __device__ __forceinline__ void stsig(uint32_t off, uint32_t limb, uint32_t v) const {
    if constexpr (kMode == WM_ALL_LDS) lds_words()[sig0 + (off + limb) * stride + col] = v;
    else values[size_t(off + limb) * lanes + lane] = v;
  }
  __device__ __forceinline__ uint32_t& mem(uint32_t base, uint64_t word) const { return mems[(size_t(base) + word) * lanes + lane]; }
  // A memory staged in LDS for the launch (generated kernels, WM_ALL_LDS):
  // word `word` of the memory whose first LDS row is `row`.
  __device__ __forceinline__ uint32_t& lmem(uint32_t row, uint32_t word) const {
    return lds_words()[mem0 + (row + word) * stride + col];
  }
};

__device__ __forceinline__ uint32_t fmask(uint32_t w) { uint32_t r = w & 31; return r ? ((1u << r) - 1) : 0xFFFFFFFFu; }

// Multi-limb values live in fixed MAXL-limb register arrays: every loop runs
// to the constant MAXL with a guard and every variable index goes through
// pick(), so the arrays never need private memory.
using Limbs = uint32_t[MAXL];
__device__ __forceinline__ uint32_t pick(const Limbs& v, uint32_t i) {
  uint32_t x = 0;
#pragma unroll
  for (uint32_t m = 0; m < MAXL; m++) x = m == i ? v[m] : x;
  return x;
}
__device__ __forceinline__ uint32_t get_bit(const Limbs& v, uint32_t n, uint32_t bit) {
  const uint32_t l = bit >> 5;
  return l < n ? (pick(v, l) >> (bit & 31)) & 1u : 0u;
}
// operand of `have` limbs zero-extended to MAXL
template <class Cx>
__device__ __forceinline__ void ldv(const Cx& C, uint32_t code, uint32_t have, Limbs& v) {
#pragma unroll
  for (uint32_t l = 0; l < MAXL; l++) v[l] = l < have ? C.ld(code, l) : 0u;
}
__device__ __forceinline__ uint64_t decode_usize(const Limbs& v, uint32_t n) {
  uint64_t out = v[0] | (n > 1 ? uint64_t(v[1]) << 32 : 0);
#pragma unroll
  for (uint32_t l = 2; l < MAXL; l++) if (l < n && v[l]) out = ~0ull;
  return out;
}
// address operand of `n` limbs as usize: a high limb set saturates
template <class Cx>
__device__ __forceinline__ uint64_t ld_usize(const Cx& C, uint32_t code, uint32_t n) {
  uint64_t out = C.ld(code, 0) | (n > 1 ? uint64_t(C.ld(code, 1)) << 32 : 0);
  for (uint32_t l = 2; l < n; l++) if (C.ld(code, l)) return ~0ull;
  return out;
}

// Record i of an array.  Every lane runs the same record, and records never
// change during a launch, so they are read through the constant address
// space: scalar loads and scalar branches on the opcode.
__device__ __forceinline__ Rec load_rec(const Rec* recs, uint32_t i) {
  const uint32_t* p = reinterpret_cast<const uint32_t*>(recs + i);
  Rec R;
#pragma unroll
  for (int k = 0; k < 8; k++) R.w[k] = cload(p, k);
  return R;
}

// One-limb results (every value of at most 32 bits): straight scalar code.
template <class Cx>
__device__ __forceinline__ bool exec_rec1(const Cx& C, const Rec& R) {
  const uint32_t op = R.w[0] & 0xFF, width = R.w[5];
  uint32_t v;
  switch (op) {
    case OP_NOT: v = ~C.ld(R.w[2], 0); break;
    case FOP_NAND: v = ~(C.ld(R.w[2], 0) & C.ld(R.w[3], 0)); break;
    case FOP_NOR: v = ~(C.ld(R.w[2], 0) | C.ld(R.w[3], 0)); break;
    case FOP_XNOR: v = ~(C.ld(R.w[2], 0) ^ C.ld(R.w[3], 0)); break;
    case FOP_ANDOR: v = (C.ld(R.w[2], 0) & C.ld(R.w[3], 0)) | C.ld(R.w[4], 0); break;
    case FOP_ORAND: v = (C.ld(R.w[2], 0) | C.ld(R.w[3], 0)) & C.ld(R.w[4], 0); break;
    case FOP_XOR3: v = C.ld(R.w[2], 0) ^ C.ld(R.w[3], 0) ^ C.ld(R.w[4], 0); break;
    case FOP_ADD3: v = C.ld(R.w[2], 0) + C.ld(R.w[3], 0) + C.ld(R.w[4], 0); break;
    case FOP_MULADD: v = C.ld(R.w[2], 0) * C.ld(R.w[3], 0) + C.ld(R.w[4], 0); break;
    case FOP_AOI21: v = ~((C.ld(R.w[2], 0) & C.ld(R.w[3], 0)) | C.ld(R.w[4], 0)); break;
    case FOP_OAI21: v = ~((C.ld(R.w[2], 0) | C.ld(R.w[3], 0)) & C.ld(R.w[4], 0)); break;
    case FOP_MULSUB: v = C.ld(R.w[2], 0) * C.ld(R.w[3], 0) - C.ld(R.w[4], 0); break;
    case FOP_AOI22: v = ~((C.ld(R.w[2], 0) & C.ld(R.w[3], 0)) | (C.ld(R.w[4], 0) & C.ld(R.w[6], 0))); break;
    case FOP_OAI22: v = ~((C.ld(R.w[2], 0) | C.ld(R.w[3], 0)) & (C.ld(R.w[4], 0) | C.ld(R.w[6], 0))); break;
    case OP_AND: v = C.ld(R.w[2], 0) & C.ld(R.w[3], 0); break;
    case OP_OR: v = C.ld(R.w[2], 0) | C.ld(R.w[3], 0); break;
    case OP_XOR: v = C.ld(R.w[2], 0) ^ C.ld(R.w[3], 0); break;
    case OP_ADD: v = C.ld(R.w[2], 0) + C.ld(R.w[3], 0); break;
    case OP_SUB: v = C.ld(R.w[2], 0) - C.ld(R.w[3], 0); break;
    case OP_MUL: v = C.ld(R.w[2], 0) * C.ld(R.w[3], 0); break;
    case OP_MUX: v = (C.ld(R.w[4], 0) & 1u) ? C.ld(R.w[2], 0) : C.ld(R.w[3], 0); break;
    case OP_ZEXT: case OP_TRUNC: case OP_CAST: case R_SIGLOAD: v = C.ld(R.w[2], 0); break;
    case OP_EQ: case OP_NE: {
      const uint32_t la = R.w[7] & 0xFF, lb = (R.w[7] >> 8) & 0xFF, n = la > lb ? la : lb;
      uint32_t eq = 1;
      for (uint32_t l = 0; l < n; l++) eq &= (l < la ? C.ld(R.w[2], l) : 0u) == (l < lb ? C.ld(R.w[3], l) : 0u);
      v = op == OP_EQ ? eq : (eq ^ 1u);
      break;
    }
    case OP_LT: {
      const uint32_t wl = R.w[6];
      if (wl > 32) return false;
      const uint32_t a = C.ld(R.w[2], 0) & fmask(wl), b = C.ld(R.w[3], 0) & fmask(wl);
      if (R.w[7] >> 24) {
        const uint32_t sa = (a >> (wl - 1)) & 1u, sb = (b >> (wl - 1)) & 1u;
        v = sa != sb ? sa : uint32_t(a < b);
      } else {
        v = a < b;
      }
      break;
    }
    case OP_SLICE: {
      const uint32_t n = R.w[7], lsb = R.w[6], li = lsb >> 5, sh = lsb & 31;
      const uint32_t lo = li < n ? C.ld(R.w[2], li) : 0u, hi = (sh && li + 1 < n) ? C.ld(R.w[2], li + 1) : 0u;
      v = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
      break;
    }
    case OP_SEXT: {
      const uint32_t iw = R.w[6], x = C.ld(R.w[2], 0);
      const uint32_t keep = iw >= 32 ? ~0u : ((1u << iw) - 1), sign = (x >> (iw - 1)) & 1u;
      v = (x & keep) | (sign ? ~keep : 0u);
      break;
    }
    case OP_CONCAT: {
      v = 0;
      uint32_t off = 0;
      for (int k = int(R.w[3]) - 1; k >= 0; k--) {
        const uint32_t code = cload(C.cpool, R.w[2] + 2 * k), wd = cload(C.cpool, R.w[2] + 2 * k + 1) & 0xFFFF;
        v |= (C.ld(code, 0) & fmask(wd)) << off;
        off += wd;
      }
      break;
    }
    case OP_MEMREAD: {
      const uint32_t* mt = &C.mem_table[4 * R.w[6]];
      const uint64_t addr = ld_usize(C, R.w[2], R.w[7]);
      v = addr < mt[1] ? C.mem(mt[0], addr * mt[2]) & (mt[2] == 1 ? fmask(mt[3]) : ~0u) : 0u;
      break;
    }
    default: return false;
  }
  C.st(R.w[1], 0, v & fmask(width));
  return true;
}

// A macro record's view of the operands: SP_LOCAL codes name the record's
// registers (a 16-entry vector indexed by wave-uniform record fields, so it
// stays in VGPRs), everything else reads through the block's context; a
// sub-operation's result goes to its register.
typedef uint32_t MacroRegs __attribute__((ext_vector_type(MACRO_REGS)));
template <class Cx>
struct MacroCx {
  const Cx& C;
  MacroRegs& loc;
  const uint32_t* cpool; const uint32_t* mem_table;
  __device__ __forceinline__ uint32_t ld(uint32_t code, uint32_t limb) const {
    if ((code & SP_MASK) == SP_LOCAL) return limb ? 0u : loc[code & (MACRO_REGS - 1)];
    return C.ld(code, limb);
  }
  __device__ __forceinline__ void st(uint32_t reg, uint32_t, uint32_t v) const { loc[reg & (MACRO_REGS - 1)] = v; }
  __device__ __forceinline__ uint32_t& mem(uint32_t base, uint64_t word) const { return C.mem(base, word); }
};

template <class Cx>
__device__ __forceinline__ void exec_macro(const Cx& C, const Rec& R) {
  MacroRegs loc;
  const MacroCx<Cx> M{C, loc, C.cpool, C.mem_table};
  const Rec* sub = reinterpret_cast<const Rec*>(C.cpool + R.w[2]);
  const uint32_t n = R.w[3];
  for (uint32_t k = 0; k < n; k++) exec_rec1(M, load_rec(sub, k));
  C.st(R.w[1], 0, loc[(n - 1) & (MACRO_REGS - 1)]);
}

template <class Cx>
__device__ __forceinline__ void exec_rec(const Cx& C, const Rec& R) {
  const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF, width = R.w[5];
  if (op == R_MACRO) { exec_macro(C, R); return; }
  if (L == 1 && exec_rec1(C, R)) return;
  Limbs a, b, r;
#pragma unroll
  for (uint32_t l = 0; l < MAXL; l++) r[l] = 0;
  switch (op) {
    case R_SIGLOAD: ldv(C, R.w[2], R.w[7], r); break;
    case OP_NOT: {
      ldv(C, R.w[2], R.w[7], a);
#pragma unroll
      for (uint32_t l = 0; l < MAXL; l++) r[l] = ~a[l];
      break;
    }
    case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: {
      ldv(C, R.w[2], R.w[7] & 0xFF, a);
      ldv(C, R.w[3], (R.w[7] >> 8) & 0xFF, b);
      if (op == OP_AND) {
#pragma unroll
        for (uint32_t l = 0; l < MAXL; l++) r[l] = a[l] & b[l];
      } else if (op == OP_OR) {
#pragma unroll
        for (uint32_t l = 0; l < MAXL; l++) r[l] = a[l] | b[l];
      } else if (op == OP_XOR) {
#pragma unroll
        for (uint32_t l = 0; l < MAXL; l++) r[l] = a[l] ^ b[l];
      } else if (op == OP_ADD) {
        uint64_t c = 0;
#pragma unroll
        for (uint32_t l = 0; l < MAXL; l++) { const uint64_t t = uint64_t(a[l]) + b[l] + c; r[l] = uint32_t(t); c = t >> 32; }
      } else if (op == OP_SUB) {
        uint64_t br = 0;
#pragma unroll
        for (uint32_t l = 0; l < MAXL; l++) { const uint64_t s = uint64_t(b[l]) + br, x = a[l]; r[l] = uint32_t(x - s); br = x < s; }
      } else {
#pragma unroll
        for (uint32_t j = 0; j < MAXL; j++) {
          uint64_t c = 0;
#pragma unroll
          for (uint32_t k = 0; j + k < MAXL; k++) {
            const uint64_t t = uint64_t(a[j]) * b[k] + r[j + k] + c;
            r[j + k] = uint32_t(t); c = t >> 32;
          }
        }
      }
      break;
    }
    case OP_EQ: case OP_NE: {
      const uint32_t la = R.w[7] & 0xFF, lb = (R.w[7] >> 8) & 0xFF, n = la > lb ? la : lb;
      uint32_t eq = 1;
      for (uint32_t l = 0; l < n; l++) eq &= (l < la ? C.ld(R.w[2], l) : 0u) == (l < lb ? C.ld(R.w[3], l) : 0u);
      r[0] = op == OP_EQ ? eq : (eq ^ 1u);
      break;
    }
    case OP_LT: {
      const uint32_t wl = R.w[6], n = (wl + 31) / 32;
      ldv(C, R.w[2], min(R.w[7] & 0xFF, n), a);
      ldv(C, R.w[3], min((R.w[7] >> 8) & 0xFF, n), b);
#pragma unroll
      for (uint32_t l = 0; l < MAXL; l++)
        if (l + 1 == n) { a[l] &= fmask(wl); b[l] &= fmask(wl); }
      uint32_t res = 0, decided = 0;
      if (R.w[7] >> 24) {
        const uint32_t sa = get_bit(a, n, wl - 1), sb = get_bit(b, n, wl - 1);
        if (sa != sb) { res = sa; decided = 1; }
      }
#pragma unroll
      for (int l = int(MAXL) - 1; l >= 0; l--)
        if (uint32_t(l) < n && !decided && a[l] != b[l]) { res = a[l] < b[l]; decided = 1; }
      r[0] = res;
      break;
    }
    case OP_MUX: {
      const bool c = C.ld(R.w[4], 0) & 1u;
      ldv(C, c ? R.w[2] : R.w[3], c ? (R.w[7] & 0xFF) : ((R.w[7] >> 8) & 0xFF), r);
      break;
    }
    case OP_SLICE: {
      const uint32_t n = R.w[7], lsb = R.w[6];
      ldv(C, R.w[2], n, a);
#pragma unroll
      for (uint32_t l = 0; l < MAXL; l++) {
        const uint32_t bit = lsb + 32 * l, li = bit >> 5, sh = bit & 31;
        const uint32_t lo = li < n ? pick(a, li) : 0u, hi = li + 1 < n ? pick(a, li + 1) : 0u;
        r[l] = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
      }
      break;
    }
    case OP_ZEXT: case OP_TRUNC: case OP_CAST: ldv(C, R.w[2], R.w[7], r); break;
    case OP_SEXT: {
      const uint32_t iw = R.w[6];
      ldv(C, R.w[2], R.w[7], r);
      const uint32_t sign = get_bit(r, MAXL, iw - 1);
#pragma unroll
      for (uint32_t l = 0; l < MAXL; l++) {
        const uint32_t lo_bit = 32 * l;
        if (lo_bit >= iw) r[l] = sign ? ~0u : 0u;
        else if (lo_bit + 32 > iw) {
          const uint32_t keep = (1u << (iw - lo_bit)) - 1;
          r[l] = (r[l] & keep) | (sign ? ~keep : 0u);
        }
      }
      break;
    }
    case OP_CONCAT: {
      uint32_t off = 0;
      for (int k = int(R.w[3]) - 1; k >= 0; k--) {
        const uint32_t code = cload(C.cpool, R.w[2] + 2 * k), wd = cload(C.cpool, R.w[2] + 2 * k + 1) & 0xFFFF,
                       n = cload(C.cpool, R.w[2] + 2 * k + 1) >> 16;
        for (uint32_t l = 0; l < n; l++) {
          uint32_t v = C.ld(code, l);
          if (l == n - 1) v &= fmask(wd);
          const uint32_t bit = off + 32 * l, li = bit >> 5, sh = bit & 31;
#pragma unroll
          for (uint32_t j = 0; j < MAXL; j++) {
            if (j == li) r[j] |= v << sh;
            if (sh && j == li + 1) r[j] |= v >> (32 - sh);
          }
        }
        off += wd;
      }
      break;
    }
    case OP_MEMREAD: {
      const uint32_t* mt = &C.mem_table[4 * R.w[6]];
      ldv(C, R.w[2], R.w[7], a);
      const uint64_t addr = decode_usize(a, R.w[7]);
      if (addr < mt[1]) {
#pragma unroll
        for (uint32_t l = 0; l < MAXL; l++)
          if (l < mt[2] && l < L) r[l] = C.mem(mt[0], addr * mt[2] + l) & (l + 1 == mt[2] ? fmask(mt[3]) : ~0u);
      }
      break;
    }
  }
#pragma unroll
  for (uint32_t l = 0; l < MAXL; l++)
    if (l < L) C.st(R.w[1], l, l + 1 == L ? r[l] & fmask(width) : r[l]);
}

template <class Cx>
__device__ __forceinline__ bool reset_asserted(const Cx& C, uint32_t rsig_code, uint32_t pol) {
  const uint32_t v = C.ld(rsig_code, 0) & 1u;
  return pol == 0 ? v != 0 : v == 0;
}

template <class Cx>
__device__ __forceinline__ void exec_effect(const Cx& C, const Rec& R) {
  const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF;
  switch (op) {
    case R_STORE:
      for (uint32_t l = 0; l < L; l++) {
        const uint32_t v = l < R.w[7] ? C.ld(R.w[2], l) : 0u;
        C.stsig(R.w[1], l, l + 1 == L ? v & fmask(R.w[5]) : v);
      }
      break;
    case R_ASYNC_CAPTURE:
      if (reset_asserted(C, R.w[3], R.w[6]))
        for (uint32_t l = 0; l < L; l++) {
          const uint32_t v = l < R.w[7] ? C.ld(R.w[2], l) : 0u;
          C.stsig(R.w[1], l, l + 1 == L ? v & fmask(R.w[5]) : v);
        }
      break;
    case R_CAPTURE: {
      const bool rst = (R.w[6] & 2u) && reset_asserted(C, R.w[3], R.w[6] & 1u);
      for (uint32_t l = 0; l < L; l++) {
        const uint32_t v = rst ? C.ld(R.w[4], l) : (l < R.w[7] ? C.ld(R.w[2], l) : 0u);
        C.st(R.w[1], l, l + 1 == L ? v & fmask(R.w[5]) : v);
      }
      break;
    }
    case R_COMMIT:
      for (uint32_t l = 0; l < L; l++) C.stsig(R.w[1], l, C.ld(R.w[2], l));
      break;
    case R_MEMWRITE: {
      if (!(C.ld(R.w[2], 0) & 1u)) break;
      const uint32_t* mt = &C.mem_table[4 * R.w[1]];
      const uint32_t la = R.w[7] & 0xFF, ld = (R.w[7] >> 8) & 0xFF;
      const uint64_t addr = ld_usize(C, R.w[3], la);
      if (addr >= mt[1]) break;
      for (uint32_t l = 0; l < L; l++) {
        const uint32_t v = l < ld ? C.ld(R.w[4], l) : 0u;
        C.mem(mt[0], addr * mt[2] + l) = l + 1 == L ? v & fmask(mt[3]) : v;
      }
      break;
    }
  }
}

// The interpreter loop over one block's records.
template <class Cx>
__device__ __forceinline__ void run_records(const Rec* recs, uint32_t n, const Cx& C) {
  for (uint32_t i = 0; i < n; i++) {
    const Rec R = load_rec(recs, i);
    if (is_effect_op(R.w[0] & 0xFF)) exec_effect(C, R);
    else exec_rec(C, R);
  }
}

// How a block's records are executed: InterpExec walks the record arrays of
// the run; a design compiled at run time supplies its own executor with the
// same run(A, block id, C) entry.
struct InterpExec {
  template <class Cx>
  __device__ __forceinline__ static void run(const WordArgs& A, uint32_t id, const Cx& C) {
    run_records(A.blk[id], A.nrec[id], C);
  }
  // staging of LDS-resident memories at the start and end of a launch (none here)
  template <class Cx>
  __device__ __forceinline__ static void enter(const Cx&) {}
  template <class Cx>
  __device__ __forceinline__ static void leave(const Cx&) {}
};

template <class Cx>
__device__ __forceinline__ void apply_stimulus(const WordArgs& A, const Cx& C, uint64_t c, uint64_t cyc, uint32_t i0,
                                               uint32_t i1) {
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t off = A.in_desc[4 * i], limbs = A.in_desc[4 * i + 1], width = A.in_desc[4 * i + 2];
    const uint32_t mode = A.in_desc[4 * i + 3] & 0xFFu, tl = A.in_desc[4 * i + 3] >> 8;
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
        v = A.trace_in[(size_t(c) * A.trace_in_limbs + tl + l) * A.lanes + C.lane];
      } else if (mode == IN_TRACE_UNIFORM) {
        v = A.trace_uni[size_t(c) * A.trace_uni_limbs + tl + l];
      } else {
        continue;
      }
      if (l == limbs - 1) v &= fmask(width);
      C.stsig(off, l, v);
    }
  }
}

// Observation point: trace, digest and, in a replay, the checks: an active
// check whose value differs from its expected value in any limb counts one
// mismatch for the lane and, among the lane's first mm_max, is recorded with
// its cycle, check number and actual value.
template <class Cx>
__device__ __forceinline__ void observe(const WordArgs& A, const Cx& C, uint64_t c, uint64_t& d, uint32_t& mm) {
  const bool check = A.expect && C.lane < A.check_lanes;
  uint32_t k = 0;
  for (uint32_t j = 0; j < A.n_obs; j++) {
    const uint32_t off = A.obs_desc[2 * j], limbs = A.obs_desc[2 * j + 1];
    uint32_t diff = 0, act[4] = {0, 0, 0, 0};
    for (uint32_t l = 0; l < limbs; l++, k++) {
      const uint32_t v = C.ld(SP_SIGNAL | off, l);
      if (A.trace && C.wave == 0) A.trace[(size_t(c) * A.n_obs_limbs + k) * A.lanes + C.lane] = v;
      if (A.digest) d = lane_digest(d, v, j, l);
      if (check) {
        const uint32_t e = A.expect_per_lane ? A.expect[(size_t(c) * A.n_obs_limbs + k) * A.lanes + C.lane]
                                             : A.expect[size_t(c) * A.n_obs_limbs + k];
        diff |= v ^ e;
        if (l == 0) act[0] = v;
        else if (l == 1) act[1] = v;
        else if (l == 2) act[2] = v;
        else if (l == 3) act[3] = v;
      }
    }
    if (check && diff && (!A.check_active || ((A.check_active[size_t(c) * A.check_words + j / 32] >> (j % 32)) & 1u))) {
      if (mm < A.mm_max && C.wave == 0) {
        uint32_t* r = A.mm_rec + (size_t(C.lane) * A.mm_max + mm) * 6;
        r[0] = uint32_t(A.cycle_base + c); r[1] = j;
        r[2] = act[0]; r[3] = act[1]; r[4] = act[2]; r[5] = act[3];
      }
      mm++;
    }
  }
}

// One block of lanes; lanes past the batch skip (no barriers inside).  The
// launch is one loop over steps -- the first settle, every cycle, the final
// settle -- and each step runs a uniform sequence of record blocks and
// observation points through a single inlined copy of the interpreter.
constexpr uint32_t STEP_OBS = 15;  // sequence entry: observe (else a block id)

// kWaves > 1 (generated kernels only): a workgroup of kWaves waves runs one
// set of 64 lanes, every wave on the same LDS columns.  The generated code
// splits each evaluation between the waves (the wave index selects its part,
// barriers order the parts); the rest of the step -- stimulus, observation,
// halting -- runs in every wave alike, so its LDS writes are the same values
// and its decisions agree.  Only wave 0 writes results back to HBM.
template <int kMode, class Exec = InterpExec, int kWaves = 1>
__device__ __forceinline__ void word_body(const WordArgs& A, uint32_t block) {
  const uint32_t col = kWaves > 1 ? threadIdx.x & 63u : threadIdx.x;
  const uint32_t per_block = A.block_lanes ? A.block_lanes : (kWaves > 1 ? 64u : blockDim.x);
  const uint32_t slot = block * per_block + col;
  if (col >= per_block) return;
  uint32_t lane = slot;
  if (A.lane_map) {
    const uint32_t part = block % 8u, j = (block / 8u) * per_block + col;
    if (j >= A.live[part]) return;
    lane = A.lane_map[size_t(part) * A.map_cap + j];
  }
  if (lane >= A.lanes) return;
  Ctx<kMode> C;
  C.values = A.values; C.mems = A.mems; C.cpool = A.cpool; C.mem_table = A.mem_table;
  C.lanes = A.lanes; C.lane = lane;
  C.scr = A.scratch + lane; C.sstride = A.lanes;
  const uint32_t stride = kWaves > 1 ? per_block : blockDim.x;
  C.col = col; C.stride = stride; C.sig0 = A.n_scratch * stride;
  C.mem0 = C.sig0 + A.n_sig_words * stride;
  C.wave = kWaves > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : 0u;
  const bool primary = C.wave == 0;
  if constexpr (kMode == WM_ALL_LDS)
    for (uint32_t w = 0; w < A.n_sig_words; w++) lds_words()[C.sig0 + w * C.stride + C.col] = A.values[size_t(w) * A.lanes + lane];
  Exec::enter(C);
  // inputs that hold one value all run: the design never writes an input, so
  // once per launch is the same as every cycle
  if (A.n_in_once) apply_stimulus(A, C, 0, A.cycle_base, 0, A.n_in_once);
  // every wave's copies are in LDS before any part of an evaluation writes state
  if constexpr (kWaves > 1) __syncthreads();
  uint64_t d = A.digest ? A.digest[lane] : 0;
  uint32_t mm = A.mm_count ? A.mm_count[lane] : 0;
  uint64_t done = 0;
  bool halted = false;
  if (A.halt_cycles) { done = A.halt_cycles[lane] >> 1; halted = A.halt_cycles[lane] & 1; }
  // per-cycle sequence, 4 bits per entry
  uint32_t cyc_seq = 0, cyc_n = 0;
  auto push = [&](uint32_t x) { cyc_seq |= x << (4 * cyc_n); cyc_n++; };
  if (A.flags & F_EVAL_PRE) { push(B_COMB); push(B_ASYNC); }
  if (A.n_obs && !(A.flags & F_OBS_POST)) push(STEP_OBS);
  if (A.flags & F_EDGE) { push(B_NEXT); push(B_COMMIT); }
  if (A.flags & F_EVAL_POST) { push(B_COMB); push(B_ASYNC); }
  if (A.n_obs && (A.flags & F_OBS_POST)) push(STEP_OBS);
  const uint32_t settle_seq = B_COMB | (B_ASYNC << 4);
  const uint64_t steps = A.cycles + 2;
  for (uint64_t s = 0; s < steps; s++) {
    const bool first = s == 0, last = s == steps - 1;
    uint32_t seq, n;
    uint64_t c = 0;
    if (first) { seq = settle_seq; n = (A.flags & F_EVAL_FIRST) ? 2 : 0; }
    else if (last) { seq = settle_seq; n = (A.flags & F_FINAL_EVAL) ? 2 : 0; }
    else {
      // a wave whose lanes have all halted skips to the final step
      if (A.halt_offset != NONE && __builtin_amdgcn_ballot_w64(!halted) == 0) {
        s = steps - 2;
        continue;
      }
      if (halted) continue;
      c = s - 1;
      if (A.n_in > A.n_in_once) apply_stimulus(A, C, c, A.cycle_base + c, A.n_in_once, A.n_in);
      seq = cyc_seq; n = cyc_n;
    }
    // the sequence is the same on every active lane; make that visible
    const uint32_t seq_u = __builtin_amdgcn_readfirstlane(seq), n_u = __builtin_amdgcn_readfirstlane(n);
    for (uint32_t k = 0; k < n_u; k++) {
      const uint32_t id = (seq_u >> (4 * k)) & 15u;
      if (id == STEP_OBS) observe(A, C, c, d, mm);
      else Exec::run(A, id, C);
    }
    if (!first && !last) {
      done++;
      if (A.halt_offset != NONE && (C.ld(SP_SIGNAL | A.halt_offset, 0) & 1u)) halted = true;
    }
  }
  if (!primary) return;
  Exec::leave(C);
  if constexpr (kMode == WM_ALL_LDS)
    for (uint32_t w = 0; w < A.n_sig_words; w++) A.values[size_t(w) * A.lanes + lane] = lds_words()[C.sig0 + w * C.stride + C.col];
  if (A.digest) A.digest[lane] = d;
  if (A.mm_count) A.mm_count[lane] = mm;
  if (A.halt_cycles) A.halt_cycles[lane] = (done << 1) | (halted ? 1u : 0u);
}

template <int kMode>
__global__ void __launch_bounds__(256) word_kernel(WordArgs A) {
  word_body<kMode>(A, blockIdx.x);
}

}  // namespace word
}  // namespace rrtl
