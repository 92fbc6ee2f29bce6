// Learned triage surrogate: a ReLU MLP with one logit output, trained and run
// on the GPU by hand-written gfx950 kernels (include/rrtl_mlp.h).
//
// Every matrix product is one NT GEMM kernel, C[M][N] = A[M][K] . B[N][K]^T
// with both operands K-contiguous bf16 and fp32 accumulation on the matrix
// cores (v_mfma_f32_32x32x16_bf16).  The three products of a layer map onto it
// by keeping transposed bf16 copies of what the backward pass contracts over
// the batch:
//   forward   H[l+1] = relu(H[l] . W[l]^T + b[l])     A = H[l],       B = W[l]
//   data grad dZ[l]  = (dZ[l+1] . W[l]) * [H[l] > 0]  A = dZ[l+1],    B = W[l]^T
//   weight    dW[l]  = dZ[l+1]^T . H[l]               A = dZ[l+1]^T,  B = H[l]^T
// The epilogues fuse bias + ReLU (with the transposed copy written from the
// same registers), the ReLU derivative, or the fp32 gradient store.  The
// 1-wide output layer is a per-row dot product (no matrix-core shape), and
// Adam updates fp32 master weights and writes their bf16 copies.
#include "../../include/rrtl_mlp.h"
#include "common.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

namespace rrtl {
namespace mlp {

typedef uint16_t u16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 pattern, round to nearest even (finite inputs)
__host__ __device__ inline u16 f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u16(u >> 16);
}
__host__ __device__ inline float bf2f(u16 h) { return __builtin_bit_cast(float, uint32_t(h) << 16); }

// ------------------------------------------------------------------ GEMM
// Four waves in a 2 x 2 grid; each wave owns TM x TN MFMA 32x32 tiles, so the
// workgroup tile is (64 TM) x (64 TN) with K steps of 64: 128 x 128 (TM = TN =
// 2) for short products, 256 x 128 (TM = 4) when M is large -- per 16-deep K
// slice a wave then reads 6 fragments for 8 MFMAs instead of 4 for 4.  LDS
// rows are padded to 72 bf16 (144 bytes): the eight lanes of a 128-byte LDS
// pass read rows 36 banks apart, which covers all 64 banks once.  The next K
// step's global loads are in flight while the current one computes.  With
// gridDim.y > 1 each y-slice takes k_chunk of K and writes its own fp32
// partial product (split K, summed by splitk_reduce_kernel).
constexpr uint32_t BK = 64, LDK = BK + 8, GT = 256;
enum : int { EPI_RELU = 0, EPI_DRELU = 1, EPI_F32 = 2 };

struct GemmArgs {
  const u16* A; const u16* B;  // [M][lda], [N][ldb]
  uint32_t M, N, K, lda, ldb;
  const float* bias;           // EPI_RELU: [N]
  const u16* act;              // EPI_DRELU: [M][ldo], the forward activation whose ReLU is undone
  u16* out; uint32_t ldo;      // bf16 [M][ldo] (EPI_RELU, EPI_DRELU)
  u16* out_t; uint32_t ldt;    // bf16 [N][ldt] transposed copy, or null
  float* out_f; uint32_t ldf;  // EPI_F32: [M][ldf]
  uint32_t k_chunk;            // K per y-slice
  size_t split_stride;         // EPI_F32: elements between the y-slices' partial products
};

// Epilogue: the wave's TM x TN accumulator tiles, rows from mb, columns from
// nb.  C tile layout: lane holds column (lane & 31) and rows 8 (r / 4) +
// 4 (lane / 32) + r % 4 of each 32 x 32 tile.
// The bias values of the lane's TN columns are read before any store: a load
// issued after stores makes its wait count the stores too (the vector-memory
// counter is in order), which stalled every group on the stores before it.
template <int EPI, int TM, int TN>
__device__ __forceinline__ void store_acc(const GemmArgs& g, const f32x16 (&acc)[TM][TN], uint32_t mb, uint32_t nb,
                                          uint32_t lane) {
  float* out_f = g.out_f + (EPI == EPI_F32 ? blockIdx.y * g.split_stride : 0);
  float bias[TN];
  if constexpr (EPI == EPI_RELU) {
#pragma unroll
    for (int j = 0; j < TN; j++) bias[j] = g.bias[nb + j * 32u + (lane & 31u)];
  }
  // ReLU derivative: which of the lane's 16 values per tile keep their
  // gradient (forward activation > 0), one bit each, also read before any store
  uint32_t keep[TM][TN];
  if constexpr (EPI == EPI_DRELU) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) {
        const uint32_t n = nb + j * 32u + (lane & 31u);
        uint32_t m = 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const u16 a = g.act[size_t(mb + i * 32u + 8u * q + 4u * (lane >> 5) + r) * g.ldo + n];
            m |= ((a & 0x8000u) || a == 0) ? 0u : (1u << (4 * q + r));
          }
        keep[i][j] = m;
      }
  }
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const uint32_t n = nb + j * 32u + (lane & 31u);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t mq = mb + i * 32u + 8u * q + 4u * (lane >> 5);
        if constexpr (EPI == EPI_F32) {
#pragma unroll
          for (int r = 0; r < 4; r++) out_f[size_t(mq + r) * g.ldf + n] = acc[i][j][4 * q + r];
        } else {
          u16 h[4];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            float v = acc[i][j][4 * q + r];
            if constexpr (EPI == EPI_RELU) {
              v = fmaxf(v + bias[j], 0.f);
            } else {
              if (!((keep[i][j] >> (4 * q + r)) & 1u)) v = 0.f;
            }
            h[r] = f2bf(v);
            g.out[size_t(mq + r) * g.ldo + n] = h[r];
          }
          if (g.out_t) {
            uint2 pk;
            pk.x = uint32_t(h[0]) | (uint32_t(h[1]) << 16);
            pk.y = uint32_t(h[2]) | (uint32_t(h[3]) << 16);
            *reinterpret_cast<uint2*>(&g.out_t[size_t(n) * g.ldt + mq]) = pk;
          }
        }
      }
    }
}

template <int EPI, int TM, int TN>
__global__ void __launch_bounds__(GT, 2) gemm_nt_kernel(GemmArgs g) {
  constexpr uint32_t BMt = 64 * TM, BNt = 64 * TN, NA = BMt / 32, NB = BNt / 32;
  __shared__ __attribute__((aligned(16))) u16 sA[BMt * LDK];
  __shared__ __attribute__((aligned(16))) u16 sB[BNt * LDK];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t tiles_n = g.N / BNt, nb = gridDim.x;
  uint32_t bid = blockIdx.x;
  // Workgroups are dealt round-robin over the 8 XCDs; renumber so each XCD
  // (its own L2) works through a contiguous run of tiles sharing A rows.
  if (nb % 8u == 0) bid = (bid % 8u) * (nb / 8u) + bid / 8u;
  const uint32_t m0 = (bid / tiles_n) * BMt, n0 = (bid % tiles_n) * BNt;
  const uint32_t kbeg = blockIdx.y * g.k_chunk, kend = min(g.K, kbeg + g.k_chunk);
  const uint32_t wm = wave >> 1, wn = wave & 1u;
  const uint32_t crow = tid >> 3, cch = (tid & 7u) * 8u;  // copy: 16-byte chunk cch of rows crow + 32 i
  const u16* pa = g.A + size_t(m0 + crow) * g.lda + cch;
  const u16* pb = g.B + size_t(n0 + crow) * g.ldb + cch;
  const size_t sa = size_t(32) * g.lda, sb = size_t(32) * g.ldb;
  u32x4 ra[NA], rb[NB];
#pragma unroll
  for (uint32_t i = 0; i < NA; i++) ra[i] = *reinterpret_cast<const u32x4*>(pa + i * sa + kbeg);
#pragma unroll
  for (uint32_t i = 0; i < NB; i++) rb[i] = *reinterpret_cast<const u32x4*>(pb + i * sb + kbeg);
  f32x16 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;
  const uint32_t frow = lane & 31u, fk = (lane >> 5) * 8u;
  u32x4* da = reinterpret_cast<u32x4*>(&sA[crow * LDK + cch]);
  u32x4* db = reinterpret_cast<u32x4*>(&sB[crow * LDK + cch]);
  constexpr uint32_t R32 = 32u * LDK / 8u;  // 32 LDS rows in uint4 units
  for (uint32_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();  // the previous step's fragment reads are done
#pragma unroll
    for (uint32_t i = 0; i < NA; i++) da[i * R32] = ra[i];
#pragma unroll
    for (uint32_t i = 0; i < NB; i++) db[i * R32] = rb[i];
    __syncthreads();
    if (k0 + BK < kend) {
#pragma unroll
      for (uint32_t i = 0; i < NA; i++) ra[i] = *reinterpret_cast<const u32x4*>(pa + i * sa + k0 + BK);
#pragma unroll
      for (uint32_t i = 0; i < NB; i++) rb[i] = *reinterpret_cast<const u32x4*>(pb + i * sb + k0 + BK);
    }
#pragma unroll
    for (uint32_t kk = 0; kk < BK; kk += 16) {
      bf16x8 a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++)
        a[i] = *reinterpret_cast<const bf16x8*>(&sA[(wm * (32u * TM) + i * 32u + frow) * LDK + kk + fk]);
#pragma unroll
      for (int j = 0; j < TN; j++)
        b[j] = *reinterpret_cast<const bf16x8*>(&sB[(wn * (32u * TN) + j * 32u + frow) * LDK + kk + fk]);
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
  store_acc<EPI, TM, TN>(g, acc, m0 + wm * (32u * TM), n0 + wn * (32u * TN), lane);
}

// 256 x 256 workgroup tiles for large products (the surrogate's scoring of a
// million candidates).  Four waves in a 2 x 2 grid, each owning 4 x 4 MFMA
// 32 x 32 tiles: per 16-deep K slice a wave reads 8 fragments for 16 MFMAs
// (half the LDS bytes per MFMA of the 256 x 128 tiles), and its 256 fp32
// accumulators live in the AccVGPRs (one wave per SIMD).  K steps of 64 go
// through two LDS buffers (dynamic LDS, 144 KiB: one workgroup per CU): at the
// top of step k the operands of step k + 1, already in registers, are stored
// into the other buffer and the global loads of step k + 2 are issued, then
// the MFMAs of step k run from the current buffer; one barrier per step.
constexpr uint32_t BIG = 256;
constexpr size_t BIG_LDS = size_t(4) * BIG * LDK * sizeof(u16);  // two buffers of A and B tiles: 147,456 B

// Stores the registers of one step's operand tiles into an LDS buffer: this
// thread's 16-byte chunk of rows r, r + 32, ..., r + 224 of A, then of B.
__device__ __forceinline__ void big_stash(u32x4* __restrict__ d, const u32x4 (&ra)[8], const u32x4 (&rb)[8]) {
  constexpr uint32_t R32 = 32u * LDK / 8u;  // 32 LDS rows in uint4 units
#pragma unroll
  for (uint32_t i = 0; i < 8; i++) d[i * R32] = ra[i];
#pragma unroll
  for (uint32_t i = 0; i < 8; i++) d[BIG * LDK / 8u + i * R32] = rb[i];
}

// One K step of gemm_nt_big_kernel: the MFMAs of the step in buffer fa / fb,
// the next step's operands (in registers) stored through `st` into the other
// buffer, and the global loads of step knext.  The restrict qualifiers tell
// the compiler the fragment reads and the stores touch different buffers, so
// it may interleave them.
__device__ __forceinline__ void big_step(f32x16 (&acc)[4][4], const u16* __restrict__ fa, const u16* __restrict__ fb,
                                         u32x4* __restrict__ st, u32x4 (&ra)[8], u32x4 (&rb)[8], const u16* pa,
                                         const u16* pb, size_t sa, size_t sb, uint32_t knext) {
  bf16x8 a[2][4], b[2][4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    a[0][i] = *reinterpret_cast<const bf16x8*>(fa + i * 32u * LDK);
    b[0][i] = *reinterpret_cast<const bf16x8*>(fb + i * 32u * LDK);
  }
  big_stash(st, ra, rb);
#pragma unroll
  for (uint32_t i = 0; i < 8; i++) ra[i] = *reinterpret_cast<const u32x4*>(pa + i * sa + knext);
#pragma unroll
  for (uint32_t i = 0; i < 8; i++) rb[i] = *reinterpret_cast<const u32x4*>(pb + i * sb + knext);
#pragma unroll
  for (uint32_t kk = 0; kk < BK / 16u; kk++) {
    const uint32_t c = kk & 1u;
    if (kk + 1 < BK / 16u) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[c ^ 1u][i] = *reinterpret_cast<const bf16x8*>(fa + i * 32u * LDK + (kk + 1) * 16u);
        b[c ^ 1u][i] = *reinterpret_cast<const bf16x8*>(fb + i * 32u * LDK + (kk + 1) * 16u);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[c][i], b[c][j], acc[i][j], 0, 0, 0);
  }
  // Issue order: the first slice's 8 fragment reads, then in every slice its
  // 16 MFMAs in pairs, each pair followed by one of the next slice's fragment
  // reads, and every second pair by one store into the other buffer and one
  // global load: the 16 stores and loads spread over the whole step.  (Left
  // to itself the scheduler reuses one fragment register and waits for each
  // read right before its MFMAs; with all the stores in the first slice, the
  // reads queued behind them held up the second slice's MFMAs: 748 against
  // 838 TFLOP/s at 65,536 x 1,024 x 1,024.)
  __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
#pragma unroll
  for (int sl = 0; sl < 4; sl++)
#pragma unroll
    for (int q = 0; q < 8; q++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
      if (sl < 3) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      if (q & 1) {
        __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      }
    }
}

template <int EPI>
__global__ void __launch_bounds__(GT, 1) gemm_nt_big_kernel(GemmArgs g) {
  extern __shared__ __attribute__((aligned(16))) u16 big_lds[];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t tiles_n = g.N / BIG, nb = gridDim.x;
  uint32_t bid = blockIdx.x;
  // dealt round-robin over the 8 XCDs: each XCD takes a contiguous run of tiles sharing A rows
  if (nb % 8u == 0) bid = (bid % 8u) * (nb / 8u) + bid / 8u;
  const uint32_t m0 = (bid / tiles_n) * BIG, n0 = (bid % tiles_n) * BIG;
  const uint32_t kbeg = blockIdx.y * g.k_chunk, kend = min(g.K, kbeg + g.k_chunk);
  const uint32_t wm = wave >> 1, wn = wave & 1u;
  // copy: 16-byte chunk (tid & 7) of rows (tid >> 3) + 32 i, i < 8, of each operand
  const uint32_t crow = tid >> 3, cch = (tid & 7u) * 8u;
  const u16* pa = g.A + size_t(m0 + crow) * g.lda + cch;
  const u16* pb = g.B + size_t(n0 + crow) * g.ldb + cch;
  const size_t sa = size_t(32) * g.lda, sb = size_t(32) * g.ldb;
  u32x4 ra[8], rb[8];
  auto load = [&](uint32_t k) {
#pragma unroll
    for (uint32_t i = 0; i < 8; i++) ra[i] = *reinterpret_cast<const u32x4*>(pa + i * sa + k);
#pragma unroll
    for (uint32_t i = 0; i < 8; i++) rb[i] = *reinterpret_cast<const u32x4*>(pb + i * sb + k);
  };
  f32x16 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;
  // Each step is one basic block (big_step): the stores and loads for later
  // steps run unconditionally (past the end they reload the last step into
  // the buffer nobody reads again), so its scheduling groups can interleave
  // them with the MFMAs.
  const uint32_t klast = kend - BK;
  load(kbeg);
  big_stash(reinterpret_cast<u32x4*>(big_lds + crow * LDK + cch), ra, rb);
  load(min(kbeg + BK, klast));
  __syncthreads();
  const uint32_t frow = lane & 31u, fk = (lane >> 5) * 8u;
  const uint32_t fa_off = (wm * 128u + frow) * LDK + fk, fb_off = BIG * LDK + (wn * 128u + frow) * LDK + fk;
  uint32_t cur = 0;
  for (uint32_t k0 = kbeg; k0 < kend; k0 += BK) {
    u16* const buf = big_lds + cur * (2 * BIG * LDK);
    u16* const other = big_lds + (cur ^ 1u) * (2 * BIG * LDK);
    big_step(acc, buf + fa_off, buf + fb_off, reinterpret_cast<u32x4*>(other + crow * LDK + cch), ra, rb, pa, pb, sa, sb,
             min(k0 + 2 * BK, klast));
    __syncthreads();  // this step's fragment reads and the next step's stores are done
    cur ^= 1u;
  }
  store_acc<EPI, 4, 4>(g, acc, m0 + wm * 128u, n0 + wn * 128u, lane);
}

// out[i] = sum over s < splits of part[s * stride + i], i < n (n, stride multiples of 4)
__global__ void __launch_bounds__(256) splitk_reduce_kernel(const float* part, size_t stride, uint32_t splits, float* out,
                                                            size_t n) {
  for (size_t i = (size_t(blockIdx.x) * 256u + threadIdx.x) * 4u; i < n; i += size_t(gridDim.x) * 1024u) {
    float4 a = *reinterpret_cast<const float4*>(part + i);
    for (uint32_t s = 1; s < splits; s++) {
      const float4 b = *reinterpret_cast<const float4*>(part + s * stride + i);
      a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w;
    }
    *reinterpret_cast<float4*>(out + i) = a;
  }
}

// ------------------------------------------------------------------ small kernels
// Output layer and loss: one wave per row, logit = H[row] . w + b; with labels,
// the mean binary cross-entropy and its logit gradient (rows >= n_valid are
// padding: zero gradient, zero loss).
struct HeadArgs {
  const u16* H; uint32_t ldh, K, rows, n_valid;
  const u16* w; const float* b;
  const float* y;      // labels [rows] or null (inference)
  float* logit;        // [rows]
  float* dlogit;       // [rows] when y
  float* loss_sum;     // accumulated sum of row losses when y
};
__global__ void __launch_bounds__(256) head_fwd_kernel(HeadArgs h) {
  // 64 rows per workgroup, 16 per wave; the loss is summed in the workgroup
  // and added to the total once
  __shared__ float part[4];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  float loss = 0.f;
  for (uint32_t q = 0; q < 16u; q++) {
    const uint32_t row = blockIdx.x * 64u + wave * 16u + q;
    if (row >= h.rows) break;
    float s = 0.f;
    for (uint32_t k = lane * 8u; k < h.K; k += 512u) {
      const uint4 hv = *reinterpret_cast<const uint4*>(h.H + size_t(row) * h.ldh + k);
      const uint4 wv = *reinterpret_cast<const uint4*>(h.w + k);
      const uint32_t hx[4] = {hv.x, hv.y, hv.z, hv.w}, wx[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
      for (int t = 0; t < 4; t++) {
        s = fmaf(bf2f(u16(hx[t])), bf2f(u16(wx[t])), s);
        s = fmaf(bf2f(u16(hx[t] >> 16)), bf2f(u16(wx[t] >> 16)), s);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    const float l = s + h.b[0];
    if (lane == 0) {
      h.logit[row] = l;
      if (h.y) {
        if (row < h.n_valid) {
          const float y = h.y[row];
          const float p = 1.f / (1.f + expf(-l));
          h.dlogit[row] = (p - y) / float(h.n_valid);
          // softplus(l) - y l, evaluated without overflow
          loss += fmaxf(l, 0.f) - y * l + log1pf(expf(-fabsf(l)));
        } else {
          h.dlogit[row] = 0.f;
        }
      }
    }
  }
  if (!h.y) return;
  if (lane == 0) part[wave] = loss;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(h.loss_sum, part[0] + part[1] + part[2] + part[3]);
}

// dZ[row][k] = dlogit[row] * w[k] * [H[row][k] > 0], bf16; 8 columns per thread
__global__ void __launch_bounds__(256) head_bwd_kernel(const u16* H, uint32_t ld, uint32_t K, uint32_t rows,
                                                       const float* dlogit, const u16* w, u16* dZ) {
  const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
  const uint32_t per_row = K / 8u;
  if (i >= uint64_t(rows) * per_row) return;
  const uint32_t row = uint32_t(i / per_row), k = uint32_t(i % per_row) * 8u;
  const uint4 hv = *reinterpret_cast<const uint4*>(H + size_t(row) * ld + k);
  const uint4 wv = *reinterpret_cast<const uint4*>(w + k);
  const uint32_t hx[4] = {hv.x, hv.y, hv.z, hv.w}, wx[4] = {wv.x, wv.y, wv.z, wv.w};
  const float d = dlogit[row];
  uint32_t o[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const u16 h0 = u16(hx[q]), h1 = u16(hx[q] >> 16);
    const float v0 = ((h0 & 0x8000u) || h0 == 0) ? 0.f : d * bf2f(u16(wx[q]));
    const float v1 = ((h1 & 0x8000u) || h1 == 0) ? 0.f : d * bf2f(u16(wx[q] >> 16));
    o[q] = uint32_t(f2bf(v0)) | (uint32_t(f2bf(v1)) << 16);
  }
  *reinterpret_cast<uint4*>(dZ + size_t(row) * ld + k) = make_uint4(o[0], o[1], o[2], o[3]);
}

// out[r] = sum_c M[r][c] * (v ? v[c] : 1); one wave per row
__global__ void __launch_bounds__(256) rowdot_kernel(const u16* M, uint32_t ld, uint32_t cols, uint32_t rows,
                                                     const float* v, float* out) {
  const uint32_t lane = threadIdx.x & 63u, row = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (row >= rows) return;
  float s = 0.f;
  for (uint32_t c = lane * 8u; c < cols; c += 512u) {
    const uint4 x = *reinterpret_cast<const uint4*>(M + size_t(row) * ld + c);
    const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float a = bf2f(u16(xs[q])), b = bf2f(u16(xs[q] >> 16));
      if (v) { s = fmaf(a, v[c + 2 * q], s); s = fmaf(b, v[c + 2 * q + 1], s); }
      else { s += a + b; }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) out[row] = s;
}

// out[0] = sum of v[0..n)
__global__ void __launch_bounds__(256) vsum_kernel(const float* v, uint32_t n, float* out) {
  __shared__ float part[4];
  float s = 0.f;
  for (uint32_t i = threadIdx.x; i < n; i += 256u) s += v[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = part[0] + part[1] + part[2] + part[3];
}

// bf16 [R][C] (ld_in) -> [C][R] (ld_out) through a 64 x 64 LDS tile
__global__ void __launch_bounds__(256) transpose_kernel(const u16* in, uint32_t ld_in, u16* out, uint32_t ld_out,
                                                        uint32_t R, uint32_t C) {
  __shared__ u16 t[64][66];
  const uint32_t tiles_c = C / 64u;
  const uint32_t r0 = (blockIdx.x / tiles_c) * 64u, c0 = (blockIdx.x % tiles_c) * 64u;
  const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
  for (uint32_t r = ty; r < 64u; r += 4u) t[r][tx] = in[size_t(r0 + r) * ld_in + c0 + tx];
  __syncthreads();
  for (uint32_t c = ty; c < 64u; c += 4u) out[size_t(c0 + c) * ld_out + r0 + tx] = t[tx][c];
  (void)R;
}

// Batch assembly: row r of X = concat(ta[ia[p]], tb[ib[p]], ex[p]) with
// p = perm[base + r] (or base + r), zeros past n_valid; labels gathered alongside.
struct GatherArgs {
  const u16* ta; const u16* tb; const u16* ex; uint32_t da, db, de;
  const uint32_t* ia; const uint32_t* ib; const float* y; const uint32_t* perm;
  uint64_t base; uint32_t n_valid, rows;
  u16* X; float* yb;
};
__global__ void __launch_bounds__(256) gather_kernel(GatherArgs a) {
  const uint32_t d = a.da + a.db + a.de, w = d / 8u;
  const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
  if (i >= uint64_t(a.rows) * w) return;
  const uint32_t r = uint32_t(i / w), c = uint32_t(i % w) * 8u;
  uint4 v = make_uint4(0, 0, 0, 0);
  if (r < a.n_valid) {
    const uint64_t p = a.perm ? a.perm[a.base + r] : a.base + r;
    if (c < a.da) v = *reinterpret_cast<const uint4*>(a.ta + size_t(a.ia[p]) * a.da + c);
    else if (c < a.da + a.db) v = *reinterpret_cast<const uint4*>(a.tb + size_t(a.ib[p]) * a.db + (c - a.da));
    else v = *reinterpret_cast<const uint4*>(a.ex + size_t(p) * a.de + (c - a.da - a.db));
    if (c == 0 && a.yb) a.yb[r] = a.y ? a.y[p] : 0.f;
  } else if (c == 0 && a.yb) {
    a.yb[r] = 0.f;
  }
  *reinterpret_cast<uint4*>(a.X + size_t(r) * d + c) = v;
}

// Adam with decoupled weight decay over the flat parameter vector; writes the
// bf16 copies.  c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t).
__global__ void __launch_bounds__(256) adam_kernel(float* w, const float* g, float* m, float* v, u16* w16, uint64_t n,
                                                   float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
  for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256u) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi; v[i] = vi;
    const float wi = w[i] - lr * ((mi * c1) / (sqrtf(vi * c2) + eps) + wd * w[i]);
    w[i] = wi;
    w16[i] = f2bf(wi);
  }
}

__global__ void __launch_bounds__(256) to_bf16_kernel(const float* w, u16* w16, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x; i < n; i += uint64_t(gridDim.x) * 256u) w16[i] = f2bf(w[i]);
}

__global__ void __launch_bounds__(256) sigmoid_kernel(const float* l, float* p, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = 1.f / (1.f + expf(-l[i]));
}

// ------------------------------------------------------------------ GEMM launch
// Workgroup tile and K split of one product.  256 x 256 tiles when there are
// enough of them for two rounds of one workgroup per CU, else 256 x 128 when
// those fill the device, else 128 x 128; fp32 products split K (powers of
// two) while the tiles alone cannot occupy the CUs and every slice keeps at
// least four K steps.  Shared by the network's own products and the
// stand-alone entry points, so both run the same kernels in the same
// summation order.
struct GemmPlan {
  bool wide, big;
  uint32_t tiles, splits;
  size_t partial_floats(uint32_t M, uint32_t N) const { return splits > 1 ? size_t(splits) * M * N : 0; }
};

GemmPlan plan_gemm(int epi, uint32_t M, uint32_t N, uint32_t K, uint32_t tile) {
  const bool big = tile ? tile == 3 : (M % 256 == 0 && N % 256 == 0 && uint64_t(M / 256) * (N / 256) >= 512);
  if (big && (M % 256 || N % 256)) throw Error("E_MLP_GEMM: 256 x 256 tiles need M and N to be multiples of 256");
  const bool wide = !big && (tile ? tile == 2 : (M % 256 == 0 && uint64_t(M / 256) * (N / 128) >= 1024));
  if (wide && M % 256) throw Error("E_MLP_GEMM: 256-row tiles need M to be a multiple of 256");
  const uint32_t tiles = big ? (M / 256) * (N / 256) : wide ? (M / 256) * (N / 128) : (M / 128) * (N / 128);
  uint32_t splits = 1;
  if (epi == EPI_F32)
    while (tiles * splits < 512 && K / (splits * 2) >= 4 * BK && (K / (splits * 2)) % BK == 0) splits *= 2;
  return GemmPlan{wide, big, tiles, splits};
}

// Launches the product on stream s.  With plan.splits > 1 every y-slice writes
// its whole M x N partial product into `partials` (splits x M x N floats,
// dense) and a reduction sums them into g.out_f, which must then be dense.
void launch_gemm(int epi, const GemmPlan& plan, GemmArgs g, float* partials, hipStream_t s) {
  float* const final_out = g.out_f;
  if (plan.splits > 1) {
    if (g.ldf != g.N) throw Error("E_MLP_INTERNAL: split-K output must be dense");
    g.out_f = partials; g.ldf = g.N; g.split_stride = size_t(g.M) * g.N;
  }
  const dim3 grid(plan.tiles, plan.splits);
  if (plan.big) {
    // dynamic LDS beyond 64 KiB is opted into once per kernel
    static std::once_flag once[3];
    std::call_once(once[epi], [&] {
      const void* fn = epi == EPI_RELU    ? reinterpret_cast<const void*>(&gemm_nt_big_kernel<EPI_RELU>)
                       : epi == EPI_DRELU ? reinterpret_cast<const void*>(&gemm_nt_big_kernel<EPI_DRELU>)
                                          : reinterpret_cast<const void*>(&gemm_nt_big_kernel<EPI_F32>);
      RRTL_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(BIG_LDS)));
    });
    if (epi == EPI_RELU) hipLaunchKernelGGL(gemm_nt_big_kernel<EPI_RELU>, grid, dim3(GT), BIG_LDS, s, g);
    else if (epi == EPI_DRELU) hipLaunchKernelGGL(gemm_nt_big_kernel<EPI_DRELU>, grid, dim3(GT), BIG_LDS, s, g);
    else hipLaunchKernelGGL(gemm_nt_big_kernel<EPI_F32>, grid, dim3(GT), BIG_LDS, s, g);
  } else if (plan.wide) {
    if (epi == EPI_RELU) hipLaunchKernelGGL((gemm_nt_kernel<EPI_RELU, 4, 2>), grid, dim3(GT), 0, s, g);
    else if (epi == EPI_DRELU) hipLaunchKernelGGL((gemm_nt_kernel<EPI_DRELU, 4, 2>), grid, dim3(GT), 0, s, g);
    else hipLaunchKernelGGL((gemm_nt_kernel<EPI_F32, 4, 2>), grid, dim3(GT), 0, s, g);
  } else {
    if (epi == EPI_RELU) hipLaunchKernelGGL((gemm_nt_kernel<EPI_RELU, 2, 2>), grid, dim3(GT), 0, s, g);
    else if (epi == EPI_DRELU) hipLaunchKernelGGL((gemm_nt_kernel<EPI_DRELU, 2, 2>), grid, dim3(GT), 0, s, g);
    else hipLaunchKernelGGL((gemm_nt_kernel<EPI_F32, 2, 2>), grid, dim3(GT), 0, s, g);
  }
  RRTL_HIP_CHECK(hipGetLastError());
  if (plan.splits > 1) {
    const size_t n = size_t(g.M) * g.N;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(uint32_t(std::min<size_t>((n / 4 + 255) / 256, 2048))), dim3(256), 0,
                       s, partials, n, plan.splits, final_out, n);
    RRTL_HIP_CHECK(hipGetLastError());
  }
}

// ------------------------------------------------------------------ no-allocation GEMM
// Split-K partial products of rrtl_mlp_gemm_nt_noalloc, for callers that
// cannot allocate device memory (code under graph capture, serving loops).
// Every GPU the library runs on holds its own copy until the process ends.
// Each call overwrites the whole region it uses before reading it, so nothing
// needs clearing; calls that overlap on one GPU race on it.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) float scratch[kNoAllocScratchBytes / sizeof(float)];

constexpr int kMaxDevices = 64;

// The device address of `scratch` on the current device, looked up once per
// device with that device current (host code must never use the array's name
// as an address).
float* noalloc_scratch() {
  static std::mutex mu;
  static float* ptr[kMaxDevices] = {};
  int dev = 0;
  RRTL_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error("E_MLP_NOALLOC: device ordinal beyond the supported 64 devices");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(std::string("E_HIP_RUNTIME: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    ptr[dev] = static_cast<float*>(p);
  }
  return ptr[dev];
}

void gemm_nt_noalloc(const u16* a, const u16* b, uint32_t M, uint32_t N, uint32_t K, uint32_t tile, float* c,
                     hipStream_t s) {
  if (!a || !b || !c) throw Error("E_HIP_NULL: null device buffer");
  if (tile > 3) throw Error("E_MLP_GEMM: tile must be 0 (automatic), 1 (128 x 128), 2 (256 x 128) or 3 (256 x 256)");
  if (!M || !N || !K || M % 128 || N % 128 || K % BK) throw Error("E_MLP_GEMM: M and N must be multiples of 128, K of 64");
  const GemmPlan plan = plan_gemm(EPI_F32, M, N, K, tile);
  const size_t need = plan.partial_floats(M, N) * sizeof(float);
  if (need > kNoAllocScratchBytes)
    throw Error("E_MLP_NOALLOC: this product splits K " + std::to_string(plan.splits) + " ways and needs " +
                std::to_string(need) + " bytes of partial products; the preallocated scratch holds " +
                std::to_string(kNoAllocScratchBytes) + " (use rrtl_mlp_gemm_nt)");
  float* partials = plan.splits > 1 ? noalloc_scratch() : nullptr;
  GemmArgs g{a, b, M, N, K, K, K, nullptr, nullptr, nullptr, 0, nullptr, 0, c, N, K / plan.splits, 0};
  launch_gemm(EPI_F32, plan, g, partials, s);
}

// ------------------------------------------------------------------ host
template <class T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { if (p) (void)hipFree(p); }
  void alloc(size_t count) {
    if (count == 0) count = 1;
    if (p && count <= n) return;
    if (p) { (void)hipFree(p); p = nullptr; }
    RRTL_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
};

uint64_t splitmix(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  return mix64(s);
}

class Mlp {
 public:
  explicit Mlp(const rrtl_mlp_config& c) : cfg_(c) {
    L_ = c.n_layers;
    if (L_ < 2 || L_ > RRTL_MLP_MAX_LAYERS) throw Error("E_MLP_CONFIG: n_layers must be 2..8");
    if (c.dims[L_] != 1) throw Error("E_MLP_CONFIG: the last width must be 1 (one logit)");
    for (uint32_t l = 0; l < L_; l++)
      if (c.dims[l] == 0 || c.dims[l] % 128) throw Error("E_MLP_CONFIG: input and hidden widths must be multiples of 128");
    if (c.max_batch == 0 || c.max_batch % 128) throw Error("E_MLP_CONFIG: max_batch must be a positive multiple of 128");
    RRTL_HIP_CHECK(hipSetDevice(c.device));
    RRTL_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    RRTL_HIP_CHECK(hipEventCreate(&e0_));
    RRTL_HIP_CHECK(hipEventCreate(&e1_));
    // flat parameter layout: W[l] then b[l]
    size_t off = 0;
    for (uint32_t l = 0; l < L_; l++) {
      woff_[l] = off; off += size_t(c.dims[l + 1]) * c.dims[l];
      boff_[l] = off; off += c.dims[l + 1];
    }
    np_ = off;
    w_.alloc(np_); g_.alloc(np_); m_.alloc(np_); v_.alloc(np_); w16_.alloc(np_);
    size_t wt = 0;
    for (uint32_t l = 1; l + 1 < L_; l++) { wtoff_[l] = wt; wt += size_t(c.dims[l + 1]) * c.dims[l]; }
    wt16_.alloc(std::max<size_t>(wt, 1));
    const uint32_t B = c.max_batch;
    for (uint32_t l = 0; l < L_; l++) {
      H_[l].alloc(size_t(B) * c.dims[l]);
      HT_[l].alloc(size_t(B) * c.dims[l]);
      if (l > 0) { dZ_[l].alloc(size_t(B) * c.dims[l]); dZT_[l].alloc(size_t(B) * c.dims[l]); }
    }
    logit_.alloc(B); dlogit_.alloc(B); yb_.alloc(B); scal_.alloc(4);
    // U(-1/sqrt(in), 1/sqrt(in)) for weights and biases, from SplitMix64(seed)
    std::vector<float> init(np_);
    uint64_t st = c.seed;
    for (uint32_t l = 0; l < L_; l++) {
      const float bound = 1.f / std::sqrt(float(c.dims[l]));
      const size_t n = size_t(c.dims[l + 1]) * c.dims[l] + c.dims[l + 1];
      for (size_t i = 0; i < n; i++) {
        const double u = double(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
        init[woff_[l] + i] = float((2.0 * u - 1.0) * bound);
      }
    }
    set_params(init.data());
  }
  ~Mlp() {
    if (s_) (void)hipStreamSynchronize(s_);
    for (auto& e : gemm_ev_) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    if (e0_) (void)hipEventDestroy(e0_);
    if (e1_) (void)hipEventDestroy(e1_);
    if (s_) (void)hipStreamDestroy(s_);
  }

  size_t param_count() const { return np_; }
  void get_params(float* out) { copy_out(out, w_.p, np_); }
  void get_grads(float* out) { copy_out(out, g_.p, np_); }
  void set_params(const float* p) {
    RRTL_HIP_CHECK(hipMemcpyAsync(w_.p, p, np_ * 4, hipMemcpyHostToDevice, s_));
    RRTL_HIP_CHECK(hipMemsetAsync(m_.p, 0, np_ * 4, s_));
    RRTL_HIP_CHECK(hipMemsetAsync(v_.p, 0, np_ * 4, s_));
    RRTL_HIP_CHECK(hipMemsetAsync(g_.p, 0, np_ * 4, s_));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(1024), dim3(256), 0, s_, w_.p, w16_.p, uint64_t(np_));
    RRTL_HIP_CHECK(hipGetLastError());
    transpose_weights();
    t_ = 0;
    sync();
  }

  void forward_backward(const u16* x, const float* y, uint32_t n, float* loss) {
    if (n == 0 || n > cfg_.max_batch) throw Error("E_MLP_BATCH: rows must be 1..max_batch");
    begin();
    const uint32_t rows = round_rows(n), d0 = cfg_.dims[0];
    RRTL_HIP_CHECK(hipMemsetAsync(H_[0].p, 0, size_t(rows) * d0 * 2, s_));
    RRTL_HIP_CHECK(hipMemcpyAsync(H_[0].p, x, size_t(n) * d0 * 2, hipMemcpyHostToDevice, s_));
    RRTL_HIP_CHECK(hipMemsetAsync(yb_.p, 0, size_t(rows) * 4, s_));
    RRTL_HIP_CHECK(hipMemcpyAsync(yb_.p, y, size_t(n) * 4, hipMemcpyHostToDevice, s_));
    RRTL_HIP_CHECK(hipMemsetAsync(scal_.p, 0, 16, s_));
    launch_transpose(H_[0].p, d0, HT_[0].p, rows, rows, d0);
    step_fwd_bwd(rows, n, true);
    float ls = 0;
    RRTL_HIP_CHECK(hipMemcpyAsync(&ls, scal_.p, 4, hipMemcpyDeviceToHost, s_));
    end(1);
    if (loss) *loss = ls / float(n);
  }

  void forward(const u16* x, uint64_t n, float* logits) {
    begin();
    const uint32_t d0 = cfg_.dims[0];
    uint64_t steps = 0;
    for (uint64_t r0 = 0; r0 < n; r0 += cfg_.max_batch) {
      const uint32_t nb = uint32_t(std::min<uint64_t>(cfg_.max_batch, n - r0)), rows = round_rows(nb);
      RRTL_HIP_CHECK(hipMemsetAsync(H_[0].p, 0, size_t(rows) * d0 * 2, s_));
      RRTL_HIP_CHECK(hipMemcpyAsync(H_[0].p, x + r0 * d0, size_t(nb) * d0 * 2, hipMemcpyHostToDevice, s_));
      forward_rows(rows, nb, false, false);
      RRTL_HIP_CHECK(hipMemcpyAsync(logits + r0, logit_.p, size_t(nb) * 4, hipMemcpyDeviceToHost, s_));
      steps++;
    }
    end(steps);
  }

  void set_tables(const u16* ta, uint32_t na, uint32_t da, const u16* tb, uint32_t nb, uint32_t db, uint32_t de) {
    if (da + db + de != cfg_.dims[0] || da % 8 || db % 8 || de % 8)
      throw Error("E_MLP_TABLES: widths must be multiples of 8 summing to the input width");
    if (!na || !nb) throw Error("E_MLP_TABLES: empty table");
    ta_.alloc(size_t(na) * da); tb_.alloc(size_t(nb) * db);
    RRTL_HIP_CHECK(hipMemcpyAsync(ta_.p, ta, size_t(na) * da * 2, hipMemcpyHostToDevice, s_));
    RRTL_HIP_CHECK(hipMemcpyAsync(tb_.p, tb, size_t(nb) * db * 2, hipMemcpyHostToDevice, s_));
    na_ = na; nb_ = nb; da_ = da; db_ = db; de_ = de;
    sync();
  }

  void train_pairs(const uint32_t* ia, const uint32_t* ib, const u16* ex, const float* y, uint64_t n, uint32_t epochs,
                   uint32_t batch, uint64_t seed, float* loss_per_epoch) {
    if (!na_) throw Error("E_MLP_TABLES: no tables set");
    if (batch == 0 || batch % 128 || batch > cfg_.max_batch) throw Error("E_MLP_BATCH: batch must be a multiple of 128 up to max_batch");
    if (n == 0) throw Error("E_MLP_BATCH: no rows");
    upload_pairs(ia, ib, ex, y, n);
    std::vector<uint32_t> perm(n);
    perm_.alloc(n);
    begin();
    uint64_t steps = 0;
    for (uint32_t ep = 0; ep < epochs; ep++) {
      for (uint64_t i = 0; i < n; i++) perm[i] = uint32_t(i);
      uint64_t st = seed ^ (0xA5A5A5A5ull + ep);
      for (uint64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[splitmix(st) % (i + 1)]);
      RRTL_HIP_CHECK(hipMemcpyAsync(perm_.p, perm.data(), n * 4, hipMemcpyHostToDevice, s_));
      RRTL_HIP_CHECK(hipMemsetAsync(scal_.p, 0, 16, s_));
      for (uint64_t b0 = 0; b0 < n; b0 += batch) {
        const uint32_t nb = uint32_t(std::min<uint64_t>(batch, n - b0)), rows = round_rows(nb);
        gather(perm_.p, b0, nb, rows, true);
        step_fwd_bwd(rows, nb, true);
        adam();
        steps++;
      }
      float ls = 0;
      RRTL_HIP_CHECK(hipMemcpyAsync(&ls, scal_.p, 4, hipMemcpyDeviceToHost, s_));
      RRTL_HIP_CHECK(hipStreamSynchronize(s_));  // the pinned-free copy above and the permutation buffer
      if (loss_per_epoch) loss_per_epoch[ep] = ls / float(n);
    }
    end(steps);
  }

  void predict_pairs(const uint32_t* ia, const uint32_t* ib, const u16* ex, uint64_t n, float* prob) {
    if (!na_) throw Error("E_MLP_TABLES: no tables set");
    upload_pairs(ia, ib, ex, nullptr, n);
    out_.alloc(n);
    begin();
    uint64_t steps = 0;
    for (uint64_t b0 = 0; b0 < n; b0 += cfg_.max_batch) {
      const uint32_t nb = uint32_t(std::min<uint64_t>(cfg_.max_batch, n - b0)), rows = round_rows(nb);
      gather(nullptr, b0, nb, rows, false);
      forward_rows(rows, nb, false, false);
      hipLaunchKernelGGL(sigmoid_kernel, dim3((nb + 255) / 256), dim3(256), 0, s_, logit_.p, out_.p + b0, nb);
      steps++;
    }
    RRTL_HIP_CHECK(hipMemcpyAsync(prob, out_.p, n * 4, hipMemcpyDeviceToHost, s_));
    end(steps);
  }

  // C[M][N] = A[M][K] . B[N][K]^T in fp32 through the matrix-core kernel,
  // launched iters times on device copies; returns the mean launch time.
  double gemm_nt(const u16* a, const u16* b, uint32_t M, uint32_t N, uint32_t K, float* c, uint32_t iters,
                 uint32_t tile) {
    if (tile > 3) throw Error("E_MLP_GEMM: tile must be 0 (automatic), 1 (128 x 128), 2 (256 x 128) or 3 (256 x 256)");
    if (!M || !N || !K || M % 128 || N % 128 || K % BK) throw Error("E_MLP_GEMM: M and N must be multiples of 128, K of 64");
    if (iters == 0) iters = 1;
    ga_.alloc(size_t(M) * K); gb_.alloc(size_t(N) * K); gc_.alloc(size_t(M) * N);
    RRTL_HIP_CHECK(hipMemcpyAsync(ga_.p, a, size_t(M) * K * 2, hipMemcpyHostToDevice, s_));
    RRTL_HIP_CHECK(hipMemcpyAsync(gb_.p, b, size_t(N) * K * 2, hipMemcpyHostToDevice, s_));
    begin();
    for (uint32_t i = 0; i < iters; i++)
      gemm(EPI_F32, ga_.p, K, gb_.p, K, M, N, K, nullptr, nullptr, nullptr, 0, nullptr, 0, gc_.p, N, tile);
    end(iters);
    if (c) copy_out(c, gc_.p, size_t(M) * N);
    return stats_.gemm_ms / iters;
  }

  rrtl_mlp_stats stats() const { return stats_; }

 private:
  static uint32_t round_rows(uint32_t n) { return (n + 127u) / 128u * 128u; }
  void sync() { RRTL_HIP_CHECK(hipStreamSynchronize(s_)); }
  void copy_out(float* out, const float* p, size_t n) {
    RRTL_HIP_CHECK(hipMemcpyAsync(out, p, n * 4, hipMemcpyDeviceToHost, s_));
    sync();
  }

  void begin() {
    sync();
    stats_ = rrtl_mlp_stats{};
    n_gemm_ = 0;
    RRTL_HIP_CHECK(hipEventRecord(e0_, s_));
  }
  void end(uint64_t steps) {
    RRTL_HIP_CHECK(hipEventRecord(e1_, s_));
    RRTL_HIP_CHECK(hipEventSynchronize(e1_));
    float ms = 0;
    RRTL_HIP_CHECK(hipEventElapsedTime(&ms, e0_, e1_));
    stats_.kernel_ms = ms;
    double g = 0;
    for (size_t i = 0; i < n_gemm_; i++) {
      float x = 0;
      RRTL_HIP_CHECK(hipEventElapsedTime(&x, gemm_ev_[i].first, gemm_ev_[i].second));
      g += x;
    }
    stats_.gemm_ms = g;
    stats_.steps = steps;
  }

  void upload_pairs(const uint32_t* ia, const uint32_t* ib, const u16* ex, const float* y, uint64_t n) {
    for (uint64_t i = 0; i < n; i++)
      if (ia[i] >= na_ || ib[i] >= nb_) throw Error("E_MLP_TABLES: row index out of range");
    if (n > 0xFFFFFFFFull) throw Error("E_MLP_BATCH: too many rows");
    if (de_ && !ex) throw Error("E_HIP_NULL: the tables declare per-row features but none were given");
    ia_.alloc(n); ib_.alloc(n);
    RRTL_HIP_CHECK(hipMemcpyAsync(ia_.p, ia, n * 4, hipMemcpyHostToDevice, s_));
    RRTL_HIP_CHECK(hipMemcpyAsync(ib_.p, ib, n * 4, hipMemcpyHostToDevice, s_));
    if (de_) {
      ex_.alloc(n * de_);
      RRTL_HIP_CHECK(hipMemcpyAsync(ex_.p, ex, n * de_ * 2, hipMemcpyHostToDevice, s_));
    }
    if (y) {
      y_.alloc(n);
      RRTL_HIP_CHECK(hipMemcpyAsync(y_.p, y, n * 4, hipMemcpyHostToDevice, s_));
    }
    sync();
  }

  void gather(const uint32_t* perm, uint64_t base, uint32_t nb, uint32_t rows, bool labels) {
    GatherArgs a{};
    a.ta = ta_.p; a.tb = tb_.p; a.ex = de_ ? ex_.p : nullptr; a.da = da_; a.db = db_; a.de = de_;
    a.ia = ia_.p; a.ib = ib_.p; a.y = labels ? y_.p : nullptr; a.perm = perm;
    a.base = base; a.n_valid = nb; a.rows = rows;
    a.X = H_[0].p; a.yb = labels ? yb_.p : nullptr;
    const uint64_t work = uint64_t(rows) * (da_ + db_ + de_) / 8;
    hipLaunchKernelGGL(gather_kernel, dim3(uint32_t((work + 255) / 256)), dim3(256), 0, s_, a);
    RRTL_HIP_CHECK(hipGetLastError());
    if (labels) launch_transpose(H_[0].p, cfg_.dims[0], HT_[0].p, rows, rows, cfg_.dims[0]);
  }

  void gemm(int epi, const u16* A, uint32_t lda, const u16* B, uint32_t ldb, uint32_t M, uint32_t N, uint32_t K,
            const float* bias, const u16* act, u16* out, uint32_t ldo, u16* out_t, uint32_t ldt, float* out_f,
            uint32_t ldf, uint32_t tile = 0) {
    if (M % 128 || N % 128 || K % BK || lda % 8 || ldb % 8) throw Error("E_MLP_INTERNAL: GEMM shape");
    const GemmPlan plan = plan_gemm(epi, M, N, K, tile);
    GemmArgs g{A, B, M, N, K, lda, ldb, bias, act, out, ldo, out_t, ldt, out_f, ldf, K / plan.splits, 0};
    if (plan.splits > 1) gsplit_.alloc(plan.partial_floats(M, N));
    if (n_gemm_ == gemm_ev_.size()) {
      hipEvent_t a, b;
      RRTL_HIP_CHECK(hipEventCreate(&a));
      RRTL_HIP_CHECK(hipEventCreate(&b));
      gemm_ev_.push_back({a, b});
    }
    auto& ev = gemm_ev_[n_gemm_++];
    RRTL_HIP_CHECK(hipEventRecord(ev.first, s_));
    launch_gemm(epi, plan, g, plan.splits > 1 ? gsplit_.p : nullptr, s_);
    RRTL_HIP_CHECK(hipEventRecord(ev.second, s_));
    stats_.gemm_flop += 2.0 * M * N * K;
  }

  void launch_transpose(const u16* in, uint32_t ld_in, u16* out, uint32_t ld_out, uint32_t R, uint32_t C) {
    if (R % 64 || C % 64) throw Error("E_MLP_INTERNAL: transpose shape");
    hipLaunchKernelGGL(transpose_kernel, dim3((R / 64) * (C / 64)), dim3(256), 0, s_, in, ld_in, out, ld_out, R, C);
    RRTL_HIP_CHECK(hipGetLastError());
  }
  void transpose_weights() {
    for (uint32_t l = 1; l + 1 < L_; l++)
      launch_transpose(w16_.p + woff_[l], cfg_.dims[l], wt16_.p + wtoff_[l], cfg_.dims[l + 1], cfg_.dims[l + 1], cfg_.dims[l]);
  }

  // Hidden layers and the output logit over `rows` rows of H[0] (n_valid real);
  // keep_t: also write the transposed activations the weight gradients need.
  void forward_rows(uint32_t rows, uint32_t n_valid, bool keep_t, bool labels) {
    const uint32_t* d = cfg_.dims;
    for (uint32_t l = 0; l + 1 < L_; l++)
      gemm(EPI_RELU, H_[l].p, d[l], w16_.p + woff_[l], d[l], rows, d[l + 1], d[l], w_.p + boff_[l], nullptr,
           H_[l + 1].p, d[l + 1], keep_t ? HT_[l + 1].p : nullptr, rows, nullptr, 0);
    HeadArgs h{};
    h.H = H_[L_ - 1].p; h.ldh = d[L_ - 1]; h.K = d[L_ - 1]; h.rows = rows; h.n_valid = n_valid;
    h.w = w16_.p + woff_[L_ - 1]; h.b = w_.p + boff_[L_ - 1];
    h.y = labels ? yb_.p : nullptr; h.logit = logit_.p; h.dlogit = dlogit_.p; h.loss_sum = scal_.p;
    hipLaunchKernelGGL(head_fwd_kernel, dim3((rows + 63) / 64), dim3(256), 0, s_, h);
    RRTL_HIP_CHECK(hipGetLastError());
  }

  void step_fwd_bwd(uint32_t rows, uint32_t n_valid, bool labels) {
    const uint32_t* d = cfg_.dims;
    forward_rows(rows, n_valid, true, labels);
    // output layer: dZ of the last hidden layer, its weight and bias gradients
    const uint32_t H = L_ - 1, K = d[H];
    const uint64_t work = uint64_t(rows) * K / 8;
    hipLaunchKernelGGL(head_bwd_kernel, dim3(uint32_t((work + 255) / 256)), dim3(256), 0, s_, H_[H].p, K, K, rows,
                       dlogit_.p, w16_.p + woff_[H], dZ_[H].p);
    launch_transpose(dZ_[H].p, K, dZT_[H].p, rows, rows, K);
    hipLaunchKernelGGL(rowdot_kernel, dim3(K / 4), dim3(256), 0, s_, HT_[H].p, rows, rows, K, dlogit_.p, g_.p + woff_[H]);
    hipLaunchKernelGGL(vsum_kernel, dim3(1), dim3(256), 0, s_, dlogit_.p, rows, g_.p + boff_[H]);
    RRTL_HIP_CHECK(hipGetLastError());
    for (int l = int(L_) - 2; l >= 0; l--) {
      // dW[l] = dZ[l+1]^T H[l]; db[l] = row sums of dZ[l+1]^T
      gemm(EPI_F32, dZT_[l + 1].p, rows, HT_[l].p, rows, d[l + 1], d[l], rows, nullptr, nullptr, nullptr, 0, nullptr,
           0, g_.p + woff_[l], d[l]);
      hipLaunchKernelGGL(rowdot_kernel, dim3(d[l + 1] / 4), dim3(256), 0, s_, dZT_[l + 1].p, rows, rows, d[l + 1],
                         nullptr, g_.p + boff_[l]);
      RRTL_HIP_CHECK(hipGetLastError());
      if (l > 0)  // dZ[l] = (dZ[l+1] W[l]) * [H[l] > 0]
        gemm(EPI_DRELU, dZ_[l + 1].p, d[l + 1], wt16_.p + wtoff_[l], d[l + 1], rows, d[l], d[l + 1], nullptr, H_[l].p,
             dZ_[l].p, d[l], dZT_[l].p, rows, nullptr, 0);
    }
  }

  void adam() {
    t_++;
    const float c1 = float(1.0 / (1.0 - std::pow(double(cfg_.beta1), double(t_))));
    const float c2 = float(1.0 / (1.0 - std::pow(double(cfg_.beta2), double(t_))));
    hipLaunchKernelGGL(adam_kernel, dim3(1024), dim3(256), 0, s_, w_.p, g_.p, m_.p, v_.p, w16_.p, uint64_t(np_), cfg_.lr,
                       cfg_.beta1, cfg_.beta2, cfg_.eps, cfg_.weight_decay, c1, c2);
    RRTL_HIP_CHECK(hipGetLastError());
    transpose_weights();
  }

  rrtl_mlp_config cfg_;
  uint32_t L_ = 0;
  size_t np_ = 0, woff_[RRTL_MLP_MAX_LAYERS] = {}, boff_[RRTL_MLP_MAX_LAYERS] = {}, wtoff_[RRTL_MLP_MAX_LAYERS] = {};
  hipStream_t s_ = nullptr;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> gemm_ev_;
  size_t n_gemm_ = 0;
  Buf<float> w_, g_, m_, v_;
  Buf<u16> w16_, wt16_;
  Buf<u16> H_[RRTL_MLP_MAX_LAYERS], HT_[RRTL_MLP_MAX_LAYERS], dZ_[RRTL_MLP_MAX_LAYERS], dZT_[RRTL_MLP_MAX_LAYERS];
  Buf<float> logit_, dlogit_, yb_, scal_, y_, out_;
  Buf<u16> ta_, tb_, ex_, ga_, gb_;
  Buf<float> gc_, gsplit_;
  Buf<uint32_t> ia_, ib_, perm_;
  uint32_t na_ = 0, nb_ = 0, da_ = 0, db_ = 0, de_ = 0;
  uint64_t t_ = 0;
  rrtl_mlp_stats stats_{};
};

}  // namespace mlp
}  // namespace rrtl

struct rrtl_mlp {
  std::unique_ptr<rrtl::mlp::Mlp> impl;
  std::string err;
};

namespace {
thread_local std::string g_mlp_create_err;

rrtl_status mlp_classify(const std::string& m) {
  if (m.rfind("E_HIP_RUNTIME", 0) == 0) return RRTL_E_HIP;
  return RRTL_E_INVALID;
}

template <class F>
rrtl_status mlp_guard(rrtl_mlp* m, F&& f) {
  if (!m || !m->impl) return RRTL_E_INVALID;
  try {
    f(*m->impl);
    m->err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    m->err = ex.what();
    return mlp_classify(m->err);
  }
}
}  // namespace

extern "C" {

rrtl_status rrtl_mlp_create(const rrtl_mlp_config* c, rrtl_mlp** out) {
  if (!out) return RRTL_E_INVALID;
  *out = nullptr;
  try {
    if (!c || c->struct_size < sizeof(rrtl_mlp_config)) throw rrtl::Error("E_MLP_CONFIG: config is null or too small");
    auto* h = new rrtl_mlp;
    try { h->impl = std::make_unique<rrtl::mlp::Mlp>(*c); } catch (...) { delete h; throw; }
    *out = h;
    g_mlp_create_err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    g_mlp_create_err = ex.what();
    return mlp_classify(g_mlp_create_err);
  }
}
void rrtl_mlp_destroy(rrtl_mlp* m) { delete m; }
const char* rrtl_mlp_last_error(const rrtl_mlp* m) { return m ? m->err.c_str() : "E_HIP_NULL: null handle"; }
const char* rrtl_mlp_create_error(void) { return g_mlp_create_err.c_str(); }
size_t rrtl_mlp_param_count(const rrtl_mlp* m) { return m && m->impl ? m->impl->param_count() : 0; }
rrtl_status rrtl_mlp_get_params(rrtl_mlp* m, float* out) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_params(out); });
}
rrtl_status rrtl_mlp_set_params(rrtl_mlp* m, const float* p) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) { if (!p) throw rrtl::Error("E_HIP_NULL: null buffer"); x.set_params(p); });
}
rrtl_status rrtl_mlp_forward_backward(rrtl_mlp* m, const uint16_t* xin, const float* y, uint32_t n, float* loss) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) {
    if (!xin || !y) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.forward_backward(xin, y, n, loss);
  });
}
rrtl_status rrtl_mlp_get_grads(rrtl_mlp* m, float* out) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_grads(out); });
}
rrtl_status rrtl_mlp_forward(rrtl_mlp* m, const uint16_t* xin, uint64_t n, float* logits) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) {
    if (n && (!xin || !logits)) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.forward(xin, n, logits);
  });
}
rrtl_status rrtl_mlp_set_tables(rrtl_mlp* m, const uint16_t* ta, uint32_t na, uint32_t da, const uint16_t* tb,
                                uint32_t nb, uint32_t db, uint32_t de) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) {
    if (!ta || !tb) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.set_tables(ta, na, da, tb, nb, db, de);
  });
}
rrtl_status rrtl_mlp_train_pairs(rrtl_mlp* m, const uint32_t* ia, const uint32_t* ib, const uint16_t* extra,
                                 const float* y, uint64_t n, uint32_t epochs, uint32_t batch, uint64_t seed,
                                 float* loss_per_epoch) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) {
    if (!ia || !ib || !y) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.train_pairs(ia, ib, extra, y, n, epochs, batch, seed, loss_per_epoch);
  });
}
rrtl_status rrtl_mlp_predict_pairs(rrtl_mlp* m, const uint32_t* ia, const uint32_t* ib, const uint16_t* extra,
                                   uint64_t n, float* prob) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) {
    if (n && (!ia || !ib || !prob)) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.predict_pairs(ia, ib, extra, n, prob);
  });
}
rrtl_status rrtl_mlp_gemm_nt(rrtl_mlp* m, const uint16_t* a, const uint16_t* b, uint32_t M, uint32_t N, uint32_t K,
                             uint32_t tile, float* c, uint32_t iters, double* ms_per_launch) {
  return mlp_guard(m, [&](rrtl::mlp::Mlp& x) {
    if (!a || !b) throw rrtl::Error("E_HIP_NULL: null buffer");
    const double ms = x.gemm_nt(a, b, M, N, K, c, iters, tile);
    if (ms_per_launch) *ms_per_launch = ms;
  });
}
rrtl_status rrtl_mlp_gemm_nt_noalloc(const uint16_t* a_dev, const uint16_t* b_dev, uint32_t M, uint32_t N, uint32_t K,
                                     uint32_t tile, float* c_dev, void* stream) {
  try {
    rrtl::mlp::gemm_nt_noalloc(a_dev, b_dev, M, N, K, tile, c_dev, static_cast<hipStream_t>(stream));
    return RRTL_OK;
  } catch (const std::exception& ex) {
    g_mlp_create_err = ex.what();
    return mlp_classify(g_mlp_create_err);
  }
}
rrtl_status rrtl_mlp_last_stats(const rrtl_mlp* m, rrtl_mlp_stats* out) {
  if (!m || !m->impl || !out) return RRTL_E_INVALID;
  *out = m->impl->stats();
  return RRTL_OK;
}

}  // extern "C"
