// Device side of the bit-sliced gate engine: gate encoding, the tile program
// arguments and the body of the LDS-resident tile kernel.  Shared by the gate
// engine's own launches and by the multi-design batch kernel, which is
// generated per batch and compiled at run time (so no host-only headers here).
#pragma once
#include "device_defs.hpp"
#ifndef __HIPCC_RTC__
#include <type_traits>
#endif

namespace rrtl {
namespace gate {

// 4-bit gate op: bit0 invert a, bit1 invert b, bit2 invert result, bit3 xor.
// r = (x ? a ^ b : (a ^ ia) & (b ^ ib)) ^ io covers every 2-input gate kind.
enum : uint32_t { GI_A = 1, GI_B = 2, GI_O = 4, GI_X = 8 };
constexpr uint32_t G_AND = 0, G_OR = GI_A | GI_B | GI_O, G_XOR = GI_X, G_ANDN = GI_B, G_ORN = GI_A | GI_O;
constexpr uint32_t G_NOT = GI_A;  // b = ones
constexpr uint32_t G_COPY = 0;    // b = ones
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t CLAMP_COMMIT = 0xFFFFFFFEu;
constexpr uint32_t TILE_THREADS = 256;
// Waves per SIMD the tile kernels are register-budgeted for.  The one-pass
// kernel runs when every tile is resident at once, which for the 20k-gate
// netlist (about 40 KB of LDS per 64-lane tile) means four workgroups per CU, one
// wave of each on every SIMD: at most 128 VGPRs.  The work-queue kernel runs
// when tiles outnumber the resident slots, typically at three tiles per CU: at
// most 168 VGPRs.
constexpr int TILE_WAVES_PER_SIMD = 4;
constexpr int QUEUE_WAVES_PER_SIMD = 3;

struct CGate { uint32_t op, a, b, dst; };
struct CCommit { uint32_t reg, val, rsig, pol, rbit, has_reset; };

__host__ __device__ inline uint32_t gate_eval(uint32_t op, uint32_t a, uint32_t b) {
  uint32_t ia = 0u - (op & 1u), ib = 0u - ((op >> 1) & 1u), io = 0u - ((op >> 2) & 1u);
  uint32_t r = (op & GI_X) ? (a ^ b) : ((a ^ ia) & (b ^ ib));
  return r ^ io;
}

// ------------------------------------------------------------------ kernels
struct TileArgs {
  uint32_t* state; uint64_t nwords;             // HBM rows [node][nwords]
  const uint2* desc;                            // [stage][thread][position]; STAGE_AHEAD rows past the end repeat the first stages
  uint32_t n_cycle_stages, n_post_stages, any_fix;  // any_fix: some gate applies its output inversion
  const uint32_t* load_slot; const uint32_t* load_node; uint32_t n_load;
  const uint32_t* stim_slot; const uint32_t* stim_mode; const uint32_t* stim_stream; const uint64_t* stim_arg; uint32_t n_stim;
  const uint4* commit; uint32_t n_commit;
  // commits without a reset, at byte offsets: reg | (val | stored inverted) << 16;
  // at most FAST_COMMITS_PER_THREAD x TILE_THREADS, held in registers for the run
  const uint32_t* commit_fast; uint32_t n_commit_fast;
  // 1: the PRE observation, the commit and the next cycle's inputs touch
  // disjoint slots and share one barrier (no post program)
  uint32_t merge_tail;
  const uint32_t* obs_slot; uint32_t n_obs;     // slot | stored value inverted << 31
  // per tile: 5-word entries (pos, slot, word, and, or); tile t owns
  // [force_start[t], force_start[t+1]), of which [force_mid[t], ...) clamp at the commit
  const uint32_t* force; const uint32_t* force_start; const uint32_t* force_mid;
  const uint64_t* group_key;                           // IN_GROUP_RANDOM key per 64-lane group, or null (= group index)
  uint32_t* detect;                                    // [nwords] lanes whose observed values left lane 0 of their group, or null
  uint32_t* digest;                                    // [64][nwords] or null
  uint32_t* trace;                                     // [cycles][n_obs][nwords] or null
  uint64_t seed, cycle_base, cycles;
  uint32_t n_slots, c0, c1, observe_post;
};

// W consecutive u32 lane-words of one slot, moved with one LDS access
template <int W> struct Words { uint32_t v[W]; };

template <int W>
__device__ __forceinline__ Words<W> ld_words(const uint32_t* lds, uint32_t slot) {
  Words<W> r;
  const uint32_t* p = lds + size_t(slot) * W;
  if constexpr (W == 1) { r.v[0] = p[0]; }
  else if constexpr (W == 2) { uint2 x = *reinterpret_cast<const uint2*>(p); r.v[0] = x.x; r.v[1] = x.y; }
  else {
#pragma unroll
    for (int k = 0; k < W; k += 4) {
      uint4 x = *reinterpret_cast<const uint4*>(p + k);
      r.v[k] = x.x; r.v[k + 1] = x.y; r.v[k + 2] = x.z; r.v[k + 3] = x.w;
    }
  }
  return r;
}

template <int W>
__device__ __forceinline__ void st_words(uint32_t* lds, uint32_t slot, const Words<W>& r) {
  uint32_t* p = lds + size_t(slot) * W;
  if constexpr (W == 1) { p[0] = r.v[0]; }
  else if constexpr (W == 2) { *reinterpret_cast<uint2*>(p) = make_uint2(r.v[0], r.v[1]); }
  else {
#pragma unroll
    for (int k = 0; k < W; k += 4) *reinterpret_cast<uint4*>(p + k) = make_uint4(r.v[k], r.v[k + 1], r.v[k + 2], r.v[k + 3]);
  }
}

// Gate descriptor, one per thread, stage and position (8 bytes).  When the tile
// fits in 64 KiB (kBA) a, b and dst are LDS byte offsets and
//   x = a | b << 16 at position 0; at position 1 operand a comes from
//       registers, so its low half holds the mode bits of every position k
//       of the thread: xor << 2k | fix << (2k + 1); later positions: b << 16,
//   y = dst | (ia ? 0xFF : 0) << 16 | (ib ? 0xFF : 0) << 24,
// so the input inversions are sign-extended bytes that the XORs read directly
// (no mask is built for them).  Otherwise a, b and dst are slot indices:
// x = a | b << 16, y = dst | op << 16.
// XOR-family ops carry no input inversions (they cancel into the output
// inversion).  A
// position without a gate reads an operand some other lane of its half-wave
// reads (or constant 0) and writes one of 64 sink slots, so every thread runs
// the same straight-line code.
template <int W, bool kBA>
__device__ __forceinline__ uint32_t* lds_at(uint32_t* lds, uint32_t field) {
  return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(lds) + (kBA ? field : field * uint32_t(W * 4)));
}

template <int W>
__device__ __forceinline__ Words<W> ld_words_p(const uint32_t* p) {
  Words<W> r;
  if constexpr (W == 1) { r.v[0] = p[0]; }
  else if constexpr (W == 2) { uint2 x = *reinterpret_cast<const uint2*>(p); r.v[0] = x.x; r.v[1] = x.y; }
  else {
#pragma unroll
    for (int k = 0; k < W; k += 4) {
      uint4 x = *reinterpret_cast<const uint4*>(p + k);
      r.v[k] = x.x; r.v[k + 1] = x.y; r.v[k + 2] = x.z; r.v[k + 3] = x.w;
    }
  }
  return r;
}

// x ? A ^ B : A & B in one v_bitop3_b32.  The truth table indexes
// (S0 << 2 | S1 << 1 | S2) = (A, B, x): evaluated on A = 0xF0, B = 0xCC, x = 0xAA
// it is (0xAA & 0x3C) | (0x55 & 0xC0) = 0x68.  Written as the builtin because
// the compiler turns the bitwise form into per-word selects.
__device__ __forceinline__ uint32_t op3_xor_or_and(uint32_t A, uint32_t B, uint32_t x) {
  return __builtin_amdgcn_bitop3_b32(A, B, x, 0x68);
}

template <int M> struct StageDesc { uint2 d[M]; };

// A thread's M descriptors of a stage are contiguous (8 M bytes), so the fetch
// is one 16-byte load per two positions.
template <int M>
__device__ __forceinline__ void load_stage(StageDesc<M>& D, const uint2* row) {
  const uint2* p = row + threadIdx.x * uint32_t(M);
  if constexpr (M % 2 == 0) {
#pragma unroll
    for (int k = 0; k < M; k += 2) {
      const uint4 q = *reinterpret_cast<const uint4*>(p + k);
      D.d[k] = make_uint2(q.x, q.y); D.d[k + 1] = make_uint2(q.z, q.w);
    }
  } else {
#pragma unroll
    for (int k = 0; k < M; k++) D.d[k] = p[k];
  }
}

// One stage of this thread: M gate positions over every lane-word of the tile,
//   t = x ? A ^ B : A & B,  A = a ^ ia, B = b ^ ib  (per-thread masks),
// stored as is: a gate's output inversion is carried by its readers, except
// for the gates whose io bit is set (values that must hold their true
// polarity), which store t ^ io.  kFix: some gate of the program sets io (else
// the XOR is compiled out).
// Position 0 reads a and b; a later position takes a from the previous
// position's t (registers) and reads b.  All LDS reads of the stage are issued
// first (their operands come from earlier stages), then the positions run in
// order, each storing its result.
template <int W, bool kBA, int M, bool kFix>
__device__ __forceinline__ void eval_stage(uint32_t* lds, const StageDesc<M>& D) {
  Words<W> va, vb[M];
  static_assert(!kBA || M >= 2, "the mode bits live in position 1's descriptor");
  va = ld_words_p<W>(lds_at<W, kBA>(lds, D.d[0].x & 0xFFFFu));
#pragma unroll
  for (int k = 0; k < M; k++) vb[k] = ld_words_p<W>(lds_at<W, kBA>(lds, D.d[k].x >> 16));
  // Once its reads are issued, a wave runs at raised issue priority until the
  // stage barrier (stage_step drops it): waves holding their operands finish
  // their writes and reach the barrier first (20k-gate netlist: +4 % at
  // priority 3, +2.6 % at 2; raising it for the reads instead lost 3 %).
  __builtin_amdgcn_s_setprio(3);
  Words<W> prev;
#pragma unroll
  for (int k = 0; k < M; k++) {
    const uint32_t y = D.d[k].y;
    uint32_t ia, ib, x, io;
    if constexpr (kBA) {
      const uint32_t modes = D.d[M > 1 ? 1 : 0].x;
      ia = uint32_t(int32_t(int8_t(y >> 16))); ib = uint32_t(int32_t(int8_t(y >> 24)));
      x = 0u - ((modes >> (2 * k)) & 1u); io = 0u - ((modes >> (2 * k + 1)) & 1u);
    } else {
      ia = 0u - ((y >> 16) & 1u); ib = 0u - ((y >> 17) & 1u); x = 0u - ((y >> 19) & 1u); io = 0u - ((y >> 18) & 1u);
    }
    Words<W> r;
#pragma unroll
    for (int w = 0; w < W; w++) r.v[w] = op3_xor_or_and((k == 0 ? va.v[w] : prev.v[w]) ^ ia, vb[k].v[w] ^ ib, x);
    if constexpr (kFix) {
#pragma unroll
      for (int w = 0; w < W; w++) r.v[w] ^= io;
    }
    uint32_t* o = lds_at<W, kBA>(lds, y & 0xFFFFu);
    if constexpr (W == 1) { o[0] = r.v[0]; }
    else if constexpr (W == 2) { *reinterpret_cast<uint2*>(o) = make_uint2(r.v[0], r.v[1]); }
    else {
#pragma unroll
      for (int q = 0; q < W; q += 4) *reinterpret_cast<uint4*>(o + q) = make_uint4(r.v[q], r.v[q + 1], r.v[q + 2], r.v[q + 3]);
    }
    prev = r;
  }
}

template <int W>
__device__ __forceinline__ void apply_clamps(const TileArgs& A, uint32_t* lds, uint32_t pos, uint32_t fb, uint32_t fe) {
  for (uint32_t i = fb + threadIdx.x; i < fe; i += TILE_THREADS) {
    const uint32_t* f = &A.force[size_t(i) * 5];
    if (f[0] != pos) continue;
    uint32_t* v = &lds[f[1] * W + f[2]];
    *v = (*v & f[3]) | f[4];
  }
}

// The tile's commit clamps (forced registers, entries [fm, fe), all at
// position CLAMP_COMMIT) held one per thread in registers for the run when
// there are at most 256 of them, so the per-cycle clamp reads no table.
struct CommitClamp { uint32_t off, andm, orm; bool on, held; };

template <int W>
__device__ __forceinline__ CommitClamp load_commit_clamp(const TileArgs& A, uint32_t fm, uint32_t fe) {
  CommitClamp c{0u, ~0u, 0u, false, fe - fm <= TILE_THREADS};
  const uint32_t i = fm + threadIdx.x;
  if (c.held && i < fe) {
    const uint32_t* f = &A.force[size_t(i) * 5];
    c.off = f[1] * W + f[2]; c.andm = f[3]; c.orm = f[4]; c.on = true;
  }
  return c;
}

template <int W>
__device__ __forceinline__ void apply_commit_clamps(const TileArgs& A, uint32_t* lds, const CommitClamp& c,
                                                    uint32_t fm, uint32_t fe) {
  if (c.held) {
    if (c.on) lds[c.off] = (lds[c.off] & c.andm) | c.orm;
  } else {
    apply_clamps<W>(A, lds, CLAMP_COMMIT, fm, fe);
  }
}

// Signature of the observed outputs: per cycle c, S[(j + c) & 63] ^= v_j for every
// observed word v_j, then one nonlinear round S'[p] = S[p] ^ (S[p+17] & ~S[p+41]).
// Bitwise, so every lane keeps its own 64-bit signature.  Lane p of wave 0
// holds S[p] (all W words) in registers for the whole tile run; the round
// reads S[p+17] and S[p+41] from other lanes of the same wave with cross-lane
// moves, so it needs no barrier.
//
// Detection (fault simulation): lane 0 of every 64-lane group (bit 0 of the
// group's even u32 word) is the reference lane; a lane is flagged once any
// observed value differs from its group's lane 0.  An inverted stored value
// flips both sides of the comparison, so the inversion flag drops out.  The
// flags accumulate in the thread's registers (thread j owns observed signals
// j, j + 256, ...) and reach HBM once per tile run.
constexpr int OBS_REGS = 4;

// Observed slots a thread reads every cycle, held in registers for the run:
// for detection signals tid, tid + 256, ... when there are at most 1024 of
// them; else for the signature (wave 0) signals tid, tid + 64, ... when there
// are at most 256.  Otherwise the table is read every cycle.
struct ObsRegs { uint32_t o[OBS_REGS]; bool det_on, dig_on; };

__device__ __forceinline__ ObsRegs load_obs(const TileArgs& A) {
  ObsRegs R;
  const uint32_t tid = threadIdx.x;
  R.det_on = A.detect && A.n_obs <= TILE_THREADS * OBS_REGS;
  R.dig_on = !R.det_on && A.digest && A.n_obs <= 64 * OBS_REGS;
  const uint32_t stride = R.det_on ? TILE_THREADS : 64u;
#pragma unroll
  for (int k = 0; k < OBS_REGS; k++) {
    const uint32_t j = tid + stride * k;
    R.o[k] = (R.det_on || (R.dig_on && tid < 64)) && j < A.n_obs ? A.obs_slot[j] : NONE;
  }
  return R;
}

template <int W>
__device__ __forceinline__ void observe_values(const TileArgs& A, uint32_t* lds, const ObsRegs& R, Words<W>& sreg,
                                               uint64_t c, uint64_t cyc, uint64_t w0, Words<W>& det) {
  const uint32_t tid = threadIdx.x;
  if (R.det_on) {
    Words<W> v[OBS_REGS];
#pragma unroll
    for (int k = 0; k < OBS_REGS; k++)
      if (R.o[k] != NONE) v[k] = ld_words<W>(lds, R.o[k] & 0x7FFFFFFFu);
#pragma unroll
    for (int k = 0; k < OBS_REGS; k++)
      if (R.o[k] != NONE)
#pragma unroll
        for (int q = 0; q < W; q++) det.v[q] |= v[k].v[q] ^ (0u - (v[k].v[q & ~1] & 1u));
  } else if (A.detect)
    for (uint32_t j = tid; j < A.n_obs; j += TILE_THREADS) {
      const Words<W> v = ld_words<W>(lds, A.obs_slot[j] & 0x7FFFFFFFu);
#pragma unroll
      for (int k = 0; k < W; k++) det.v[k] |= v.v[k] ^ (0u - (v.v[k & ~1] & 1u));
    }
  if (A.trace)
    for (uint32_t i = tid; i < A.n_obs * W; i += TILE_THREADS) {
      uint32_t j = i / W, w = i % W;
      uint64_t gw = w0 + w;
      const uint32_t o = A.obs_slot[j];
      if (gw < A.nwords) A.trace[(c * A.n_obs + j) * A.nwords + gw] = lds[(o & 0x7FFFFFFFu) * W + w] ^ (0u - (o >> 31));
    }
  if (A.digest && tid < 64) {
    if (R.dig_on) {
      // lane q folds signals j = q (mod 64); S[p] takes them from lane (p - cyc) mod 64
      Words<W> v[OBS_REGS], x;
#pragma unroll
      for (int k = 0; k < OBS_REGS; k++)
        if (R.o[k] != NONE) v[k] = ld_words<W>(lds, R.o[k] & 0x7FFFFFFFu);
#pragma unroll
      for (int q = 0; q < W; q++) x.v[q] = 0u;
#pragma unroll
      for (int k = 0; k < OBS_REGS; k++)
        if (R.o[k] != NONE)
#pragma unroll
          for (int q = 0; q < W; q++) x.v[q] ^= v[k].v[q] ^ (0u - (R.o[k] >> 31));
      const int src = int((tid - uint32_t(cyc & 63)) & 63);
#pragma unroll
      for (int q = 0; q < W; q++) sreg.v[q] ^= uint32_t(__shfl(int(x.v[q]), src, 64));
    } else
      for (uint32_t j = (tid - uint32_t(cyc & 63)) & 63; j < A.n_obs; j += 64) {
        const uint32_t o = A.obs_slot[j];
        const Words<W> v = ld_words<W>(lds, o & 0x7FFFFFFFu);
        const uint32_t inv = 0u - (o >> 31);
#pragma unroll
        for (int k = 0; k < W; k++) sreg.v[k] ^= v.v[k] ^ inv;
      }
    const int pa = int((tid + 17) & 63), pb = int((tid + 41) & 63);
#pragma unroll
    for (int k = 0; k < W; k++) {
      const uint32_t a = uint32_t(__shfl(int(sreg.v[k]), pa, 64)), b = uint32_t(__shfl(int(sreg.v[k]), pb, 64));
      sreg.v[k] ^= a & ~b;
    }
  }
}

template <int W>
__device__ __forceinline__ void observe(const TileArgs& A, uint32_t* lds, const ObsRegs& R, Words<W>& sreg, uint64_t c,
                                        uint64_t cyc, uint64_t w0, Words<W>& det) {
  observe_values<W>(A, lds, R, sreg, c, cyc, w0, det);
  // the commit (or the next cycle's inputs) may overwrite an observed slot
  __syncthreads();
}

constexpr uint32_t STAGE_AHEAD = 2;  // stages of descriptor look-ahead (3 or 4 measured no faster)
constexpr uint32_t SINK_SLOTS = 64;     // per-tile slots written by lanes with no output to keep

constexpr int FAST_COMMITS_PER_THREAD = 8;
constexpr int FAST_COMMIT_BATCH = 8;

// This thread's reset-free commits (entries tid, tid + 256, ...), loaded once per run.
struct FastCommits { uint32_t e[FAST_COMMITS_PER_THREAD]; uint32_t n; };

__device__ __forceinline__ FastCommits load_fast_commits(const TileArgs& A) {
  FastCommits F;
  const uint32_t tid = threadIdx.x;
  F.n = A.n_commit_fast > tid ? (A.n_commit_fast - tid + TILE_THREADS - 1) / TILE_THREADS : 0u;
#pragma unroll
  for (int j = 0; j < FAST_COMMITS_PER_THREAD; j++) F.e[j] = uint32_t(j) < F.n ? A.commit_fast[tid + j * TILE_THREADS] : 0u;
  return F;
}

// Reset-free commits: reg = val (^ ~0 when the value's slot holds its
// complement).  No value slot is a committed register's slot (such values are
// staged through copies before the commit), so every read may go ahead of
// every write: the reads of a batch are issued together and the thread waits
// for LDS once per batch instead of once per register.
template <int W>
__device__ __forceinline__ void commit_fast(uint32_t* lds, const FastCommits& F) {
#pragma unroll
  for (int j0 = 0; j0 < FAST_COMMITS_PER_THREAD; j0 += FAST_COMMIT_BATCH) {
    if (uint32_t(j0) >= F.n) break;
    Words<W> v[FAST_COMMIT_BATCH];
#pragma unroll
    for (int j = 0; j < FAST_COMMIT_BATCH; j++)
      if (uint32_t(j0 + j) < F.n) v[j] = ld_words_p<W>(lds_at<W, true>(lds, (F.e[j0 + j] >> 16) & ~1u));
#pragma unroll
    for (int j = 0; j < FAST_COMMIT_BATCH; j++)
      if (uint32_t(j0 + j) < F.n) {
        const uint32_t e = F.e[j0 + j], inv = 0u - ((e >> 16) & 1u);
#pragma unroll
        for (int q = 0; q < W; q++) v[j].v[q] ^= inv;
        uint32_t* o = lds_at<W, true>(lds, e & 0xFFFFu);
        if constexpr (W == 1) { o[0] = v[j].v[0]; }
        else if constexpr (W == 2) { *reinterpret_cast<uint2*>(o) = make_uint2(v[j].v[0], v[j].v[1]); }
        else {
#pragma unroll
          for (int q = 0; q < W; q += 4) *reinterpret_cast<uint4*>(o + q) = make_uint4(v[j].v[q], v[j].v[q + 1], v[j].v[q + 2], v[j].v[q + 3]);
        }
      }
  }
}

// Commits with a reset (or any commit when the tile addresses slots by index):
// the general form, descriptors read from the table.
template <int W>
__device__ __forceinline__ void commit_general(const TileArgs& A, uint32_t* lds) {
  for (uint32_t k = threadIdx.x; k < A.n_commit; k += TILE_THREADS) {
    const uint4 cm = A.commit[k];
    Words<W> v = ld_words<W>(lds, cm.y);
    const uint32_t vinv = 0u - ((cm.w >> 3) & 1u);  // the value's slot holds its complement
#pragma unroll
    for (int q = 0; q < W; q++) v.v[q] ^= vinv;
    if (cm.w & 4) {
      const Words<W> a = ld_words<W>(lds, cm.z);
      const uint32_t pol = 0u - (cm.w & 1u), rv = 0u - ((cm.w >> 1) & 1u);
#pragma unroll
      for (int q = 0; q < W; q++) { const uint32_t as = a.v[q] ^ pol; v.v[q] = (as & rv) | (~as & v.v[q]); }
    }
    st_words<W>(lds, cm.x, v);
  }
}

// Input words of one tile for one cycle, for the thread that drives an input
// (mode, stream and argument of its table entry).
template <int W>
__device__ __forceinline__ Words<W> stimulus_of(const TileArgs& A, uint32_t mode, uint32_t stream, uint64_t arg,
                                                uint64_t cyc, uint64_t w0, const uint64_t* gkey) {
  Words<W> v;
  if (mode == IN_RANDOM) {
    const uint64_t key = stim_key(A.seed, stream, cyc);
#pragma unroll
    for (int k = 0; k < W; k += 2) {
      const uint64_t gw = w0 + k;
      const uint64_t word = stim_from_key(key, gw >> 1);
      v.v[k] = uint32_t(word >> (32 * (gw & 1)));
      if (k + 1 < W) v.v[k + 1] = uint32_t(word >> 32);
    }
  } else if (mode == IN_GROUP_RANDOM) {
    // one bit per 64-lane group, the same in all of its lanes
    const uint64_t key = stim_key(A.seed, stream, cyc);
#pragma unroll
    for (int k = 0; k < W; k += 2) {
      const uint32_t b = 0u - uint32_t(stim_from_key(key, gkey[k / 2]) & 1u);
      v.v[k] = b;
      if (k + 1 < W) v.v[k + 1] = b;
    }
  } else {
    const uint32_t bit = mode == IN_CONST ? uint32_t(arg & 1)
                                          : uint32_t((cyc < (arg & 0xFFFFFFFFull)) ? (arg >> 32) & 1 : (arg >> 33) & 1);
#pragma unroll
    for (int k = 0; k < W; k++) v.v[k] = bit ? ~0u : 0u;
  }
  return v;
}

// The input table entry of input tid (the first input a thread drives), held
// in registers for the run; inputs tid + 256, ... are read from the table.
// stream_mode: the stream (a multiple of 256) | the mode; slot NONE: no input.
struct StimReg { uint32_t slot, stream_mode; uint64_t arg; };

__device__ __forceinline__ StimReg load_stim(const TileArgs& A) {
  StimReg r{NONE, 0u, 0ull};
  if (threadIdx.x < A.n_stim) {
    r.slot = A.stim_slot[threadIdx.x];
    r.stream_mode = A.stim_stream[threadIdx.x] | A.stim_mode[threadIdx.x];
    r.arg = A.stim_arg[threadIdx.x];
  }
  return r;
}

// Writes every input of the tile for cycle cyc (the caller's barrier publishes them).
template <int W>
__device__ __forceinline__ void drive_inputs(const TileArgs& A, uint32_t* lds, const StimReg& R, uint64_t cyc,
                                             uint64_t w0, const uint64_t* gkey) {
  if (R.slot != NONE)
    st_words<W>(lds, R.slot, stimulus_of<W>(A, R.stream_mode & 0xFFu, R.stream_mode & ~0xFFu, R.arg, cyc, w0, gkey));
  for (uint32_t s = threadIdx.x + TILE_THREADS; s < A.n_stim; s += TILE_THREADS)
    st_words<W>(lds, A.stim_slot[s], stimulus_of<W>(A, A.stim_mode[s], A.stim_stream[s], A.stim_arg[s], cyc, w0, gkey));
}

// The merged per-cycle tail (byte-addressed tiles without an output trace):
// the PRE observation, the reset-free commits and the next cycle's inputs in
// one barrier phase, ordered so that every LDS read of the phase (committed
// values, observed values) is issued first and the input words of cycle
// cyc + 1 are hashed while they are in flight; then the commit writes, the
// observation's folds and the input writes.  The host enables the merged tail
// only when observed slots, committed registers and input slots are disjoint
// and no committed value is read from an input slot, so the reads may all go
// ahead of the writes.
template <int W>
__device__ __forceinline__ void merged_tail(const TileArgs& A, uint32_t* lds, const ObsRegs& R, const FastCommits& F,
                                            const StimReg& SR, Words<W>& sreg, Words<W>& det, uint64_t cyc, bool more,
                                            uint64_t w0, const uint64_t* gkey) {
  const uint32_t tid = threadIdx.x;
  Words<W> cv[FAST_COMMITS_PER_THREAD];
#pragma unroll
  for (int j = 0; j < FAST_COMMITS_PER_THREAD; j++)
    if (uint32_t(j) < F.n) cv[j] = ld_words_p<W>(lds_at<W, true>(lds, (F.e[j] >> 16) & ~1u));
  const bool dig = R.dig_on && A.digest && A.n_obs && tid < 64;
  Words<W> ov[OBS_REGS];
  if (R.det_on || dig) {
#pragma unroll
    for (int k = 0; k < OBS_REGS; k++)
      if (R.o[k] != NONE) ov[k] = ld_words<W>(lds, R.o[k] & 0x7FFFFFFFu);
  }
  const bool drive = more && SR.slot != NONE;
  Words<W> iv;
  if (drive) iv = stimulus_of<W>(A, SR.stream_mode & 0xFFu, SR.stream_mode & ~0xFFu, SR.arg, cyc + 1, w0, gkey);
#pragma unroll
  for (int j = 0; j < FAST_COMMITS_PER_THREAD; j++)
    if (uint32_t(j) < F.n) {
      const uint32_t e = F.e[j], inv = 0u - ((e >> 16) & 1u);
#pragma unroll
      for (int q = 0; q < W; q++) cv[j].v[q] ^= inv;
      uint32_t* o = lds_at<W, true>(lds, e & 0xFFFFu);
      if constexpr (W == 1) { o[0] = cv[j].v[0]; }
      else if constexpr (W == 2) { *reinterpret_cast<uint2*>(o) = make_uint2(cv[j].v[0], cv[j].v[1]); }
      else {
#pragma unroll
        for (int q = 0; q < W; q += 4) *reinterpret_cast<uint4*>(o + q) = make_uint4(cv[j].v[q], cv[j].v[q + 1], cv[j].v[q + 2], cv[j].v[q + 3]);
      }
    }
  if (R.det_on) {
#pragma unroll
    for (int k = 0; k < OBS_REGS; k++)
      if (R.o[k] != NONE)
#pragma unroll
        for (int q = 0; q < W; q++) det.v[q] |= ov[k].v[q] ^ (0u - (ov[k].v[q & ~1] & 1u));
  }
  if (dig) {
    // lane q folds signals j = q (mod 64); S[p] takes them from lane (p - cyc) mod 64
    Words<W> x;
#pragma unroll
    for (int q = 0; q < W; q++) x.v[q] = 0u;
#pragma unroll
    for (int k = 0; k < OBS_REGS; k++)
      if (R.o[k] != NONE)
#pragma unroll
        for (int q = 0; q < W; q++) x.v[q] ^= ov[k].v[q] ^ (0u - (R.o[k] >> 31));
    const int src = int((tid - uint32_t(cyc & 63)) & 63);
#pragma unroll
    for (int q = 0; q < W; q++) sreg.v[q] ^= uint32_t(__shfl(int(x.v[q]), src, 64));
    const int pa = int((tid + 17) & 63), pb = int((tid + 41) & 63);
#pragma unroll
    for (int k = 0; k < W; k++) {
      const uint32_t a = uint32_t(__shfl(int(sreg.v[k]), pa, 64)), b = uint32_t(__shfl(int(sreg.v[k]), pb, 64));
      sreg.v[k] ^= a & ~b;
    }
  }
  if (drive) st_words<W>(lds, SR.slot, iv);
  if (more)
    for (uint32_t s = tid + TILE_THREADS; s < A.n_stim; s += TILE_THREADS)
      st_words<W>(lds, A.stim_slot[s], stimulus_of<W>(A, A.stim_mode[s], A.stim_stream[s], A.stim_arg[s], cyc + 1, w0, gkey));
}

// One stage: this thread's positions (descriptors fetched STAGE_AHEAD stages
// earlier), the fetch for STAGE_AHEAD stages later into the same registers,
// the stage barrier and, when forces exist, their clamps.  No other global
// access happens in this loop, so the fetches stay in flight.  The fetch
// address is the uniform row of stage S + STAGE_AHEAD plus the thread's fixed
// offset.  The table's STAGE_AHEAD rows past the last stage repeat its first
// rows, so the fetches of a cycle's last stages bring in the next cycle's
// first descriptors (and need no bound check).  kFix: some stage has gates
// that apply their output inversion.
template <int W, bool kBA, int M, bool kClamp, bool kFix>
__device__ __forceinline__ void stage_step(const TileArgs& A, uint32_t* lds, uint32_t S, StageDesc<M>& D, uint32_t fb,
                                           uint32_t fe) {
  eval_stage<W, kBA, M, kFix>(lds, D);
  load_stage<M>(D, A.desc + size_t(S + STAGE_AHEAD) * (uint32_t(M) * TILE_THREADS));
  __syncthreads();
  __builtin_amdgcn_s_setprio(0);  // raised in eval_stage after the reads
  if (kClamp) { apply_clamps<W>(A, lds, S + 1, fb, fe); __syncthreads(); }
}

// Stages [S0, S1).  Stage S runs on descriptor register set D[S % 2] (two
// sets: STAGE_AHEAD = 2), which its fetch then fills with stage S + 2; the
// sets live across the run's cycles (tile_body).
static_assert(STAGE_AHEAD == 2, "descriptor sets are indexed by stage parity");
template <int W, bool kBA, int M, bool kClamp, bool kFix>
__device__ __forceinline__ void run_stages(const TileArgs& A, uint32_t* lds, uint32_t S0, uint32_t S1, uint32_t fb,
                                           uint32_t fe, StageDesc<M> (&D)[2]) {
  uint32_t S = S0;
  if ((S & 1u) && S < S1) { stage_step<W, kBA, M, kClamp, kFix>(A, lds, S, D[1], fb, fe); S++; }
  for (; S + 2 <= S1; S += 2) {
    stage_step<W, kBA, M, kClamp, kFix>(A, lds, S, D[0], fb, fe);
    stage_step<W, kBA, M, kClamp, kFix>(A, lds, S + 1, D[1], fb, fe);
  }
  if (S < S1) stage_step<W, kBA, M, kClamp, kFix>(A, lds, S, D[0], fb, fe);
}

template <int W, bool kBA, int M>
__device__ __forceinline__ void run_program(const TileArgs& A, uint32_t* lds, uint32_t S0, uint32_t S1, uint32_t fb,
                                            uint32_t fe, StageDesc<M> (&D)[2]) {
  if (fe > fb) {
    if (A.any_fix) run_stages<W, kBA, M, true, true>(A, lds, S0, S1, fb, fe, D);
    else run_stages<W, kBA, M, true, false>(A, lds, S0, S1, fb, fe, D);
  } else {
    if (A.any_fix) run_stages<W, kBA, M, false, true>(A, lds, S0, S1, fb, fe, D);
    else run_stages<W, kBA, M, false, false>(A, lds, S0, S1, fb, fe, D);
  }
}

template <int W, bool kBA, int M>
__device__ void tile_body(const TileArgs& A, uint32_t tile, uint32_t* lds) {
  uint32_t* sig = lds + size_t(A.n_slots) * W;          // 64 x W signature words
  const uint32_t tid = threadIdx.x;
  const uint64_t w0 = uint64_t(tile) * W;
  const uint32_t nst = A.n_cycle_stages + A.n_post_stages;
  // stage clamps [fb, fm), commit clamps [fm, fe)
  const uint32_t fb = A.force_start ? A.force_start[tile] : 0, fe = A.force_start ? A.force_start[tile + 1] : 0;
  const uint32_t fm = A.force_start ? A.force_mid[tile] : 0;
  uint64_t gkey[(W + 1) / 2];
#pragma unroll
  for (int k = 0; k < W; k += 2) {
    const uint64_t g = (w0 + k) >> 1;
    gkey[k / 2] = A.group_key && w0 + k < A.nwords ? A.group_key[g] : g;
  }
  Words<W> det;
#pragma unroll
  for (int k = 0; k < W; k++) det.v[k] = 0u;
  for (uint32_t i = tid; i < A.n_load * W; i += TILE_THREADS) {
    uint32_t s = i / W, w = i % W;
    uint64_t gw = w0 + w;
    lds[A.load_slot[s] * W + w] = gw < A.nwords ? A.state[A.load_node[s] * A.nwords + gw] : 0u;
  }
  for (uint32_t w = tid; w < W; w += TILE_THREADS) { lds[A.c0 * W + w] = 0u; lds[A.c1 * W + w] = ~0u; }
  for (uint32_t i = tid; i < 64 * W; i += TILE_THREADS) {
    uint32_t p = i / W, w = i % W;
    uint64_t gw = w0 + w;
    sig[i] = (A.digest && gw < A.nwords) ? A.digest[p * A.nwords + gw] : 0u;
  }
  __syncthreads();
  Words<W> sreg;  // wave 0: signature word S[tid] for the run
#pragma unroll
  for (int k = 0; k < W; k++) sreg.v[k] = 0u;
  if (A.digest && tid < 64) sreg = ld_words<W>(sig, tid);
  FastCommits F;
  if constexpr (kBA) F = load_fast_commits(A);
  else F.n = 0;
  const StimReg SR = load_stim(A);
  const ObsRegs OR = load_obs(A);
  // observations the merged tail covers: detection or the signature from
  // register-held slots, no output trace
  const bool fast_tail = !A.trace && (A.n_obs == 0 || ((!A.detect || OR.det_on) && (!A.digest || OR.dig_on)));
  // descriptor sets of stages 0 and 1; each cycle's last fetches refill them
  // with the repeated rows, in the order of the stage count's parity
  const CommitClamp CC = load_commit_clamp<W>(A, fm, fe);
  StageDesc<M> D[2];
  load_stage<M>(D[0], A.desc);
  load_stage<M>(D[1], A.desc + size_t(M) * TILE_THREADS);
  for (uint64_t c = 0; c < A.cycles; c++) {
    const uint64_t cyc = A.cycle_base + c;
    // primary inputs for this cycle (thread per input, all words of the tile);
    // with a merged tail they were written by the previous cycle's tail
    if (c == 0 || !A.merge_tail) {
      drive_inputs<W>(A, lds, SR, cyc, w0, gkey);
      __syncthreads();
    }
    // cycle program, observe before the edge, commit, post program
    run_program<W, kBA, M>(A, lds, 0, A.n_cycle_stages, fb, fm, D);
    if (A.merge_tail) {
      // observed slots, committed registers and input slots are disjoint, and
      // no committed value is read from an input slot: one barrier phase
      if (kBA && fast_tail) {
        merged_tail<W>(A, lds, OR, F, SR, sreg, det, cyc, c + 1 < A.cycles, w0, gkey);
        commit_general<W>(A, lds);
        __builtin_amdgcn_s_setprio(3);  // as in the stages: reads issued, on to the barrier
      } else {
        if (A.n_obs) observe_values<W>(A, lds, OR, sreg, c, cyc, w0, det);
        if constexpr (kBA) commit_fast<W>(lds, F);
        commit_general<W>(A, lds);
        __builtin_amdgcn_s_setprio(3);  // as in the stages: reads issued, on to the barrier
        if (c + 1 < A.cycles) drive_inputs<W>(A, lds, SR, cyc + 1, w0, gkey);
      }
      __syncthreads();
      __builtin_amdgcn_s_setprio(0);
      if (fe > fm) { apply_commit_clamps<W>(A, lds, CC, fm, fe); __syncthreads(); }
    } else {
      if (!A.observe_post && A.n_obs) observe<W>(A, lds, OR, sreg, c, cyc, w0, det);
      if constexpr (kBA) commit_fast<W>(lds, F);
      commit_general<W>(A, lds);
      __syncthreads();
      if (fe > fm) { apply_commit_clamps<W>(A, lds, CC, fm, fe); __syncthreads(); }
      if (A.n_post_stages) run_program<W, kBA, M>(A, lds, A.n_cycle_stages, nst, fb, fm, D);
      if (A.observe_post && A.n_obs) observe<W>(A, lds, OR, sreg, c, cyc, w0, det);
    }
  }
  for (uint32_t i = tid; i < A.n_load * W; i += TILE_THREADS) {
    uint32_t s = i / W, w = i % W;
    uint64_t gw = w0 + w;
    if (gw < A.nwords) A.state[A.load_node[s] * A.nwords + gw] = lds[A.load_slot[s] * W + w];
  }
  if (A.digest) {
    if (tid < 64) st_words<W>(sig, tid, sreg);
    __syncthreads();
    for (uint32_t i = tid; i < 64 * W; i += TILE_THREADS) {
      uint32_t p = i / W, w = i % W;
      uint64_t gw = w0 + w;
      if (gw < A.nwords) A.digest[p * A.nwords + gw] = sig[i];
    }
  }
  if (A.detect && tid < A.n_obs)
#pragma unroll
    for (int k = 0; k < W; k++)
      if (det.v[k] && w0 + k < A.nwords) atomicOr(&A.detect[w0 + k], det.v[k]);
}

template <int W, bool kBA, int M>
__global__ void __launch_bounds__(TILE_THREADS, TILE_WAVES_PER_SIMD) tile_kernel(TileArgs A) {
  extern __shared__ __align__(16) uint32_t lds[];
  tile_body<W, kBA, M>(A, blockIdx.x, lds);
}

#ifndef __HIPCC_RTC__
// Calls f(W, byte_addr, M) with each as a std::integral_constant, for the
// compiled variants: W in {1, 2, 4, 8}, M in {2, 3, 4}.
template <class F>
__host__ __device__ inline void with_tile_variant(uint32_t W, bool byte_addr, uint32_t M, F&& f) {
  auto m_ = [&](auto w, auto ba) {
    if (M == 4) f(w, ba, std::integral_constant<int, 4>{});
    else if (M == 3) f(w, ba, std::integral_constant<int, 3>{});
    else f(w, ba, std::integral_constant<int, 2>{});
  };
  auto b_ = [&](auto w) {
    if (byte_addr) m_(w, std::true_type{});
    else m_(w, std::false_type{});
  };
  switch (W) {
    case 1: b_(std::integral_constant<int, 1>{}); break;
    case 2: b_(std::integral_constant<int, 2>{}); break;
    case 4: b_(std::integral_constant<int, 4>{}); break;
    default: b_(std::integral_constant<int, 8>{}); break;
  }
}
#endif

// Work-queue schedule for runs with more tiles than the device holds at once.
// Item i is chunk k = i / n_tiles of tile i % n_tiles: cycles [k * chunk,
// (k + 1) * chunk) of that tile, continuing from the state and signature that
// chunk k - 1 left in HBM.  Items are taken in index order, so the chunk an item
// waits for was taken earlier by a running workgroup and always finishes; the
// wait is still bounded by a wall-clock limit which, if reached, raises the
// error flag and lets every remaining item drain without work.
struct QueueArgs {
  uint32_t* ctl;          // [0] next item, [1] error flag, [2 + tile] chunks finished
  uint32_t n_tiles, n_chunks;
  uint64_t chunk;         // cycles per chunk
  uint64_t wait_limit;    // wall-clock ticks
};

// Control flow: thread 0's work of an iteration (publish the item finished
// last, take the next one; later the dependency wait) sits in single regions,
// and the values every thread branches on are made wave-uniform with
// readfirstlane, so the loop exit and the skip are uniform branches.  (With a
// thread-0 region on each side of the back edge and an LDS-read exit test the
// compiler may build an inner exec-mask loop that runs the other lanes again on
// the old item while thread 0 waits at its exit.)
template <int W, bool kBA, int M>
__global__ void __launch_bounds__(TILE_THREADS, QUEUE_WAVES_PER_SIMD) tile_queue_kernel(TileArgs A, QueueArgs Q) {
  extern __shared__ __align__(16) uint32_t lds[];
  __shared__ uint32_t s_item, s_err;
  const uint32_t total = Q.n_tiles * Q.n_chunks;
  uint32_t last_tile = NONE, last_k = 0;  // item this workgroup finished last
  for (;;) {
    if (threadIdx.x == 0) {
      // every thread fenced its stores of that item before the loop-end barrier
      if (last_tile != NONE)
        __hip_atomic_store(&Q.ctl[2 + last_tile], last_k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      s_item = atomicAdd(&Q.ctl[0], 1u);
    }
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= total) break;
    const uint32_t k = it / Q.n_tiles, tile = it - k * Q.n_tiles;
    if (threadIdx.x == 0) {
      uint32_t err = __hip_atomic_load(&Q.ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (k > 0 && !err) {
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(&Q.ctl[2 + tile], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < k) {
          if (wall_clock64() - t0 > Q.wait_limit) {
            __hip_atomic_store(&Q.ctl[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            err = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(8);
        }
      }
      s_err = err;
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(s_err) == 0) {
      __threadfence();  // acquire: chunk k - 1's state and signature stores
      TileArgs B = A;
      B.cycle_base = A.cycle_base + k * Q.chunk;
      B.cycles = min(Q.chunk, A.cycles - k * Q.chunk);
      if (A.trace) B.trace = A.trace + k * Q.chunk * A.n_obs * A.nwords;
      tile_body<W, kBA, M>(B, tile, lds);
      __threadfence();  // release: this thread's stores before the chunk count moves
    }
    last_tile = tile; last_k = k;
    __syncthreads();  // LDS and s_err are reused by the next item
  }
}

}  // namespace gate
}  // namespace rrtl
