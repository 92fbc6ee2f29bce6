/*
 * rrtl_mlp.h -- C ABI of the learned triage surrogate: a fully connected
 * ReLU network with one logit output, trained and evaluated on the GPU by the
 * library's own BF16 matrix-core kernels (bf16 operands, fp32 accumulation,
 * fp32 master weights, Adam).
 *
 * No counterpart in the reference: its surrogate crate
 * (crates/rrtl-surrogate) is an ONNX-runtime event predictor and does not
 * rank fault-simulation candidates.  Callers: fault-simulation drivers that
 * rank (fault, stimulus) candidates before exact simulation
 * (rrtl_amd/faultsim.py, rrtl_amd/triage.py).
 *
 * Network: dims[0] inputs -> dims[1] -> ... -> dims[n_layers] = 1, ReLU after
 * every hidden layer; loss = mean binary cross-entropy on the logit.
 * Rows go through in steps of at most max_batch (a multiple of 128).
 *
 * Parameters (get/set/grads, fp32, host): for each layer l in order, the
 * weight [dims[l+1]][dims[l]] (row-major, out x in) then the bias [dims[l+1]].
 * bf16 values are passed as their 16-bit patterns (the high half of the fp32
 * pattern, rounded to nearest even).
 *
 * Conventions as in rrtl_hip.h: every call returns an rrtl_status; the
 * message of the last failure is rrtl_mlp_last_error().
 */
#ifndef RRTL_MLP_H
#define RRTL_MLP_H

#include <stddef.h>
#include <stdint.h>

#include "rrtl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rrtl_mlp rrtl_mlp;

#define RRTL_MLP_MAX_LAYERS 8

typedef struct rrtl_mlp_config {
  uint32_t struct_size;                     /* sizeof(rrtl_mlp_config) */
  uint32_t n_layers;                        /* weight layers; dims has n_layers + 1 entries, the last is 1 */
  uint32_t dims[RRTL_MLP_MAX_LAYERS + 1];   /* hidden and input widths: multiples of 128 */
  uint32_t max_batch;                       /* rows per step, multiple of 128 */
  int32_t device;
  uint64_t seed;                            /* weight initialisation: U(-1/sqrt(in), 1/sqrt(in)) */
  float lr, beta1, beta2, eps, weight_decay; /* Adam (decoupled weight decay) */
} rrtl_mlp_config;

typedef struct rrtl_mlp_stats {
  double kernel_ms;       /* device time of the last call (HIP events on the library stream) */
  double gemm_ms;         /* of which matrix-core GEMM kernels (measured per launch) */
  double gemm_flop;       /* algorithmic GEMM flops of the last call */
  uint64_t steps;         /* optimizer steps (train) or row blocks (predict) */
} rrtl_mlp_stats;

rrtl_status rrtl_mlp_create(const rrtl_mlp_config* config, rrtl_mlp** out);
void rrtl_mlp_destroy(rrtl_mlp* m);
const char* rrtl_mlp_last_error(const rrtl_mlp* m);
const char* rrtl_mlp_create_error(void);

size_t rrtl_mlp_param_count(const rrtl_mlp* m);
rrtl_status rrtl_mlp_get_params(rrtl_mlp* m, float* out);
/* Sets the fp32 master weights (and their bf16 copies); resets Adam state. */
rrtl_status rrtl_mlp_set_params(rrtl_mlp* m, const float* params);

/* One forward and backward pass over n <= max_batch host rows x [n][dims[0]]
 * (bf16 patterns) with labels y [n]; writes the mean loss; gradients stay in
 * the library (rrtl_mlp_get_grads).  No optimizer step. */
rrtl_status rrtl_mlp_forward_backward(rrtl_mlp* m, const uint16_t* x, const float* y, uint32_t n, float* loss);
rrtl_status rrtl_mlp_get_grads(rrtl_mlp* m, float* out);
/* Logits of n host rows (any n). */
rrtl_status rrtl_mlp_forward(rrtl_mlp* m, const uint16_t* x, uint64_t n, float* logits);

/* Table rows: row r = concat(table_a[ia[r]], table_b[ib[r]], extra[r]) with
 * da + db + de == dims[0] (each a multiple of 8; de may be 0, then extra is
 * not read).  The tables live on the device and rows are assembled on the
 * fly, so a dataset of N rows costs 8 + 2 de bytes per row plus the tables. */
rrtl_status rrtl_mlp_set_tables(rrtl_mlp* m, const uint16_t* table_a, uint32_t na, uint32_t da,
                                const uint16_t* table_b, uint32_t nb, uint32_t db, uint32_t de);
/* Trains `epochs` passes over n (ia, ib, extra, y) rows in a seeded shuffled
 * order, batch rows per Adam step (multiple of 128, <= max_batch; the last
 * partial batch is used with its true size).  extra: [n][de] bf16 patterns.
 * loss_per_epoch: epochs floats or null. */
rrtl_status rrtl_mlp_train_pairs(rrtl_mlp* m, const uint32_t* ia, const uint32_t* ib, const uint16_t* extra,
                                 const float* y, uint64_t n, uint32_t epochs, uint32_t batch, uint64_t shuffle_seed,
                                 float* loss_per_epoch);
/* Sigmoid outputs for n (ia, ib, extra) rows. */
rrtl_status rrtl_mlp_predict_pairs(rrtl_mlp* m, const uint32_t* ia, const uint32_t* ib, const uint16_t* extra,
                                   uint64_t n, float* prob);
rrtl_status rrtl_mlp_last_stats(const rrtl_mlp* m, rrtl_mlp_stats* out);
/* The library's matrix-core product on its own: C[M][N] = A[M][K] . B[N][K]^T
 * (bf16 patterns in, fp32 out; M, N multiples of 128, K of 64), launched
 * `iters` times on device copies of A and B; c (null: skip the copy back)
 * receives C; ms_per_launch the mean device time of one launch (split-K
 * reduction included).  tile: 0 automatic, 1 = 128 x 128 workgroup tiles,
 * 2 = 256 x 128 (M a multiple of 256), 3 = 256 x 256 (M and N multiples of
 * 256). */
rrtl_status rrtl_mlp_gemm_nt(rrtl_mlp* m, const uint16_t* a, const uint16_t* b, uint32_t M, uint32_t N, uint32_t K,
                             uint32_t tile, float* c, uint32_t iters, double* ms_per_launch);

/* The same product on caller-owned DEVICE buffers, allocating nothing: no
 * handle, no hipMalloc/hipFree, no allocating library call.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  a_dev [M][K] and b_dev
 * [N][K] bf16 patterns, c_dev [M][N] fp32, shapes and `tile` as in
 * rrtl_mlp_gemm_nt, whose kernels and summation order it shares (results are
 * identical).  Enqueued asynchronously on `stream` (a hipStream_t; null = the
 * default stream).  Split-K partial products go to a preallocated 32 MiB
 * region per GPU (held until the process ends); a product needing more is
 * rejected with RRTL_E_INVALID ("E_MLP_NOALLOC").  Calls that overlap on one
 * GPU share that region and race.  On failure the message is in
 * rrtl_mlp_create_error() (thread-local). */
rrtl_status rrtl_mlp_gemm_nt_noalloc(const uint16_t* a_dev, const uint16_t* b_dev, uint32_t M, uint32_t N, uint32_t K,
                                     uint32_t tile, float* c_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RRTL_MLP_H */
