/*
 * rrtl_hip.h -- C ABI of the MI355X batch-simulation backend for RRTL.
 *
 * A drop-in for the reference's batch engines behind `trait SimBackend`
 * (crates/rrtl-sim-ir/src/lib.rs:5258-5270) and the runtime dispatch surface
 * `RuntimeEngine` (crates/rrtl-runtime/src/lib.rs:7785-8100).  A caller hands
 * over a packed machine program (the reference's PackedMachineProgram
 * serialised in the blob format documented in rrtl_amd/program.py) and a lane
 * count; the library owns all device memory.  Plain pointers and sizes only.
 *
 * Conventions
 *   - Every call returns RRTL_OK or an error status; the message, prefixed with
 *     an "E_..." diagnostic code like the reference's ErrorReport
 *     (crates/rrtl-ir/src/lib.rs:632-691), is available from
 *     rrtl_hip_last_error() (or rrtl_hip_create_error() when create failed).
 *   - Lane values are u32 limbs laid out [limbs][lanes] (limb-major), the
 *     per-signal slice of PackedSimulatorStorage.values (lib.rs:10045).
 *   - Memory values are [lanes][depth][limbs], the lane-major order of
 *     PackedSimulatorStorage.memories (lib.rs:10049-10054).
 *   - A handle is single-owner and not thread-safe; separate handles may be
 *     used from separate threads.  Calls are synchronous unless noted.
 */
#ifndef RRTL_HIP_H
#define RRTL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rrtl_status {
  RRTL_OK = 0,
  RRTL_E_INVALID = 1,     /* bad argument (null pointer, index or lane out of range) */
  RRTL_E_PROGRAM = 2,     /* malformed program blob */
  RRTL_E_UNSUPPORTED = 3, /* design outside this engine's scope (caller falls back) */
  RRTL_E_HIP = 4,         /* HIP runtime failure */
  RRTL_E_STATE = 5        /* call not valid in the current state (e.g. no run yet) */
} rrtl_status;

typedef struct rrtl_hip_engine rrtl_hip_engine;

/* Engine kinds (the new SimBackendKind arm; reference kinds lib.rs:3407-3425). */
enum { RRTL_HIP_AUTO = 0, RRTL_HIP_WORD = 1, RRTL_HIP_GATE = 2 };

/* Options (the HipOptions arm of RuntimeBackend, crates/rrtl-runtime/src/lib.rs:29-46). */
typedef struct rrtl_hip_options {
  uint32_t struct_size;      /* sizeof(rrtl_hip_options) */
  uint32_t engine;           /* RRTL_HIP_AUTO picks GATE for all-1-bit bitwise designs */
  int32_t device;            /* HIP device ordinal */
  uint32_t block_threads;    /* word engine workgroup size (64, 128, 192 or 256), 0 = automatic */
  uint32_t tile_words;       /* gate engine u32 lane-words per workgroup (1,2,4,8), 0 = auto */
  uint32_t ticks_per_launch; /* cycles per kernel launch in runs, 0 = whole run in one launch */
  /* Gate engine runs with more tiles than fit on the device at once: each
   * tile's cycles are cut into this many chunks handed out by a work queue, so
   * the last wave of tiles does not run alone.  0 = automatic, 1 = off.  Read
   * only when struct_size covers it. */
  uint32_t tile_chunks;
  /* Gate engine: gate positions per thread between two barriers (2, 3 or 4); a
   * position may take the previous position's result through registers.
   * 0 = automatic.  Read only when struct_size covers it. */
  uint32_t stage_gates;
  /* Word engine: how a design's records are executed.  RRTL_CODEGEN_ON
   * generates straight-line device code for the design and compiles it for the
   * GPU when the engine is created (the counterpart of the reference's
   * per-design code generator, crates/rrtl-sim-ir/src/gpu_codegen.rs:503);
   * RRTL_CODEGEN_OFF runs the record interpreter; RRTL_CODEGEN_AUTO generates
   * code for designs of at most 4000 records (and, when some records stay
   * interpreted in the generated code, at most 128 live scratch words) and
   * interprets larger ones.  Results are identical.  Read only when
   * struct_size covers it. */
  uint32_t codegen;
  /* Word engine: active lanes per 64-thread wave (4, 8, 16, 32 or 64).  Fewer
   * lanes per wave put more waves on each SIMD, which hides memory latency for
   * a batch too small to do so at 64 lanes per wave, at the price of issuing
   * each instruction for fewer lanes.  0 = automatic: 64.  With lane_waves
   * above 1 it is the lanes of one workgroup (the columns its waves share;
   * automatic: 64).  Read only when struct_size covers it. */
  uint32_t wave_lanes;
  /* Word engine runs with a halt signal: 1 = compact the lanes between
   * launches.  The run is cut into launches of ticks_per_launch cycles (when
   * that is 0: 1024, then each launch twice the last); after each launch the
   * lanes still running are packed into
   * the lowest threads, so lanes that halted free their waves and the next
   * launch runs full waves of live lanes only (the GPU counterpart of the
   * reference's work-stealing schedule over halting instances,
   * crates/rrtl-sim-ir/examples/divergent_batch.rs).  Results are identical
   * to an uncompacted run.  0 = off.  Read only when struct_size covers it. */
  uint32_t compact_lanes;
  /* Word engine with generated code: waves per 64 lanes (1, 2 or 4).  Above
   * 1, a workgroup of that many waves runs one set of 64 lanes and the
   * generated code of each evaluation is split between them: every wave runs
   * its part, values cross between waves through LDS, barriers order the
   * parts.  This fills the SIMDs of a batch too small to give each SIMD a
   * 64-lane wave of its own (several threads per lane; lanes per thread below
   * one).  0 = automatic: 4 (else 2) when the batch gives fewer than one
   * 64-lane wave per SIMD and the split is estimated to pay; 1 = off.
   * Results are identical.  Read only when struct_size covers it. */
  uint32_t lane_waves;
  /* Word engine running the record interpreter (no generated code): macro
   * records, the counterpart of the reference's fuse_macros
   * (crates/rrtl-gpu-sim/src/interp.rs:696-826).  Cones of one-limb
   * operations whose intermediate values have one reader each run as one
   * record with the intermediates in registers (at most 16 operations per
   * record); they replace the fusion menu.  0 = automatic (on), 1 = on,
   * 2 = off.  Results are identical.  Read only when struct_size covers it. */
  uint32_t interp_macros;
} rrtl_hip_options;

enum { RRTL_CODEGEN_AUTO = 0, RRTL_CODEGEN_ON = 1, RRTL_CODEGEN_OFF = 2 };

/* SimdCpuSimulator::new (lib.rs:5529): lanes == 0 is E_SIM_IR_LANES; the
 * constructor settles combinational logic (lib.rs:9058). */
rrtl_status rrtl_hip_create(const void* program_blob, size_t blob_bytes, uint32_t lanes,
                            const rrtl_hip_options* options, rrtl_hip_engine** out);
void rrtl_hip_destroy(rrtl_hip_engine* engine);
const char* rrtl_hip_last_error(const rrtl_hip_engine* engine);
const char* rrtl_hip_create_error(void); /* thread-local message of the last failed create */

uint32_t rrtl_hip_kind(const rrtl_hip_engine* engine);   /* SimBackend::kind */
uint32_t rrtl_hip_lanes(const rrtl_hip_engine* engine);  /* SimBackend::lanes */

/* set_signal_limbs / set_packed_signal_limbs (lib.rs:9909-9943): writes every
 * lane and re-settles combinational logic. */
rrtl_status rrtl_hip_set_signal_limbs(rrtl_hip_engine* e, uint32_t signal, const uint32_t* lane_limbs);
/* set_signals_replicated_raw (lib.rs:9118): writes without settling. */
rrtl_status rrtl_hip_set_signal_limbs_raw(rrtl_hip_engine* e, uint32_t signal, const uint32_t* lane_limbs);
/* get_signal_limbs (lib.rs:9177). */
rrtl_status rrtl_hip_get_signal_limbs(rrtl_hip_engine* e, uint32_t signal, uint32_t* out_lane_limbs);
/* Gate engine only: a 1-bit signal's lanes packed 32 per u32 word (ceil(lanes/32) words). */
rrtl_status rrtl_hip_get_signal_bits(rrtl_hip_engine* e, uint32_t signal, uint32_t* out_words);
/* set_memory_limbs / get_memory_limbs (lib.rs:10136-10180, RuntimeEngine :7960). */
rrtl_status rrtl_hip_set_memory_limbs(rrtl_hip_engine* e, uint32_t memory, const uint32_t* values);
rrtl_status rrtl_hip_get_memory_limbs(rrtl_hip_engine* e, uint32_t memory, uint32_t* out_values);

rrtl_status rrtl_hip_eval_combinational(rrtl_hip_engine* e);   /* lib.rs:5735 */
rrtl_status rrtl_hip_tick(rrtl_hip_engine* e);                 /* lib.rs:5741 */
rrtl_status rrtl_hip_tick_from_evaluated(rrtl_hip_engine* e);  /* tick_from_evaluated_no_post_eval, lib.rs:5748 */
rrtl_status rrtl_hip_tick_many(rrtl_hip_engine* e, uint64_t steps); /* comb-fused tick_many, lib.rs:5784-5803 */

/* snapshot_storage / restore_storage (lib.rs:9070-9103): values are
 * total_signal_words*lanes u32, memories total_memory_words*lanes u32. */
rrtl_status rrtl_hip_snapshot(rrtl_hip_engine* e, uint32_t* values, uint32_t* memories);
rrtl_status rrtl_hip_restore(rrtl_hip_engine* e, const uint32_t* values, const uint32_t* memories);

/* Stuck-at fault injection, BitParallelSimulator::set_force / clear_forces
 * (crates/rrtl-sim-ir/src/bitparallel.rs:96-101); gate engine only. */
rrtl_status rrtl_hip_set_force(rrtl_hip_engine* e, uint32_t signal, uint32_t lane, int value);
/* n set_force calls at once (the per-fault loop of fault_sim.rs:56-58); value[i] != 0 is stuck-at-1. */
rrtl_status rrtl_hip_set_forces(rrtl_hip_engine* e, const uint32_t* signal, const uint32_t* lane, const uint8_t* value,
                                size_t n);
rrtl_status rrtl_hip_clear_forces(rrtl_hip_engine* e);

rrtl_status rrtl_hip_synchronize(rrtl_hip_engine* e);         /* InterpGpuSimulator::synchronize, interp.rs:2044 */

/* ---- device-resident runs: the per-cycle replay loop of the reference's
 * harnesses (PackedSimulator replay, lib.rs:9500-9570; fault_sim.rs:66-97)
 * moved on chip.  Per cycle: drive inputs, settle, observe (PRE), edge
 * (capture, memory writes, commit, forces), then either a settle + observe
 * (POST) or nothing; the state after the run equals the reference's after its
 * trailing settle. */
enum { RRTL_IN_HOLD = 0, RRTL_IN_RANDOM = 1, RRTL_IN_CONST = 2, RRTL_IN_PULSE = 3, RRTL_IN_TRACE = 4,
       RRTL_IN_GROUP_RANDOM = 5,
       /* word engine: values shared by every lane, [cycles][uniform-trace limbs] (rrtl_hip_replay) */
       RRTL_IN_TRACE_UNIFORM = 6 };
enum { RRTL_OBSERVE_PRE = 0, RRTL_OBSERVE_POST = 1 };
typedef struct rrtl_hip_run_spec {
  uint32_t struct_size;
  uint64_t seed;
  uint64_t cycles;
  uint32_t n_inputs;
  const uint32_t* input_signal;
  const uint32_t* input_mode;   /* RRTL_IN_*: RANDOM = counter hash of (seed, signal<<8|bit, cycle, lane/64) */
  const uint64_t* input_arg;    /* CONST: value; PULSE: low 32 bits n, bit 32 value while cycle < n, bit 33 after */
  const uint32_t* trace_inputs; /* RRTL_IN_TRACE values: [cycles][trace limbs][lanes] (word engine) */
  uint32_t n_observe;
  const uint32_t* observe_signal;
  uint32_t observe_point;       /* RRTL_OBSERVE_PRE / RRTL_OBSERVE_POST */
  uint32_t record_trace;        /* also keep every observed value: [cycles][...][...] */
  uint32_t halt_signal;         /* word engine: a lane stops after the cycle this bit is 1; UINT32_MAX = none */
  /* Fields below are read only when struct_size covers them. */
  /* Gate engine fault simulation (fault_sim.rs:66-97 moved on chip): lane 0 of
   * every 64-lane group is the fault-free reference; a lane is flagged once any
   * observed value differs from its group's lane 0 (rrtl_hip_get_detect).
   * Replaces the digest of the run. */
  uint32_t detect;
  /* RRTL_IN_GROUP_RANDOM: one bit per 64-lane group g, shared by its lanes:
   * bit 0 of stim_word(seed, signal << 8, cycle, group_key[g]) (group_key null:
   * key g).  ceil(lanes/64) keys. */
  const uint64_t* group_key;
} rrtl_hip_run_spec;

typedef struct rrtl_hip_run_stats {
  double kernel_ms;       /* device time of the run's launches (HIP events on the engine stream) */
  uint64_t launches;
  uint64_t lane_cycles;   /* cycles actually simulated summed over lanes (halted lanes stop counting) */
} rrtl_hip_run_stats;

rrtl_status rrtl_hip_run(rrtl_hip_engine* e, const rrtl_hip_run_spec* spec);
rrtl_status rrtl_hip_last_run_stats(const rrtl_hip_engine* e, rrtl_hip_run_stats* out);
/* Observation digest of the last run.  Gate engine: 64 x ceil(lanes/32) u32
 * bit-sliced signature words; word engine: one u64 chain per lane. */
size_t rrtl_hip_digest_bytes(const rrtl_hip_engine* e);
rrtl_status rrtl_hip_get_digest(rrtl_hip_engine* e, void* out);
size_t rrtl_hip_trace_bytes(const rrtl_hip_engine* e);
rrtl_status rrtl_hip_get_trace(rrtl_hip_engine* e, void* out);
/* Word engine runs with a halt signal: cycles simulated per lane. */
rrtl_status rrtl_hip_get_halt_cycles(rrtl_hip_engine* e, uint64_t* out);
/* Gate engine runs with detect set: lane flags packed 32 per u32 (ceil(lanes/32) words). */
rrtl_status rrtl_hip_get_detect(rrtl_hip_engine* e, uint32_t* out_words);
/* Gate engine: [tile_words, live_slots, levels, gates] of the last compiled run. */
rrtl_status rrtl_hip_gate_tile_info(rrtl_hip_engine* e, uint32_t* out4);
/* Word engine: [generated code (1) or interpreter (0), waves per 64 lanes,
 * active lanes per wave, threads per workgroup].  No counterpart in the
 * reference: the launch shape the engine chose, for tuning and tests. */
rrtl_status rrtl_hip_word_launch_info(rrtl_hip_engine* e, uint32_t* out4);

/* ---- trace replay (SimBackend::replay_trace, crates/rrtl-sim-ir/src/lib.rs:5270-5345,
 * with ReplayOptions / ReplayReport :3463-3491, :4328-4407) moved on chip: per
 * step the inputs are set (raw), comb is evaluated, every active check compares
 * its signal with the expected value, and the edge runs with no settle after it
 * (tick_from_evaluated_no_post_eval) -- one launch for the whole plan.  Word
 * engine; checked signals up to 128 bits.  Inputs and checks are fixed lists;
 * an input that a plan step leaves unset repeats its previous value, a check
 * absent from a step is cleared in check_active. */
enum { RRTL_REPLAY_REPLICATED = 0, RRTL_REPLAY_INDEPENDENT = 1 };  /* ReplayLaneMode */
enum { RRTL_REPLAY_LANE0 = 0, RRTL_REPLAY_ALL_LANES = 1 };         /* ReplayCheckMode (Lane0Fast, AllLanes) */
typedef struct rrtl_hip_replay_spec {
  uint32_t struct_size;
  uint64_t steps;
  uint32_t lane_mode;            /* REPLICATED: one value per step for every lane; INDEPENDENT: one per lane */
  uint32_t check_mode;           /* REPLICATED + LANE0 checks lane 0 only; otherwise every lane */
  uint32_t n_inputs;
  const uint32_t* input_signal;
  const uint32_t* input_values;  /* REPLICATED [steps][input limbs]; INDEPENDENT [steps][input limbs][lanes] */
  uint32_t n_checks;
  const uint32_t* check_signal;
  const uint32_t* check_index;   /* reported check_index of each check, or NULL: its position */
  const uint32_t* expected;      /* REPLICATED [steps][check limbs]; INDEPENDENT [steps][check limbs][lanes] */
  const uint32_t* check_active;  /* [steps][ceil(n_checks / 32)] bit j = check j applies, or NULL: all */
  uint32_t max_mismatches;       /* ReplayOptions::max_mismatches */
} rrtl_hip_replay_spec;
/* ReplayMismatch: step, check_index, lane, expected and actual (u128 as 4 limbs) */
typedef struct rrtl_hip_replay_mismatch {
  uint64_t step;
  uint32_t check_index;
  uint32_t lane;
  uint32_t expected[4];
  uint32_t actual[4];
} rrtl_hip_replay_mismatch;
/* *mismatch_count = every mismatch (ReplayReport::mismatch_count); the first
 * max_mismatches in the reference's order (step, check, lane) go to
 * `mismatches` (capacity max_mismatches), their number to *n_mismatches. */
rrtl_status rrtl_hip_replay(rrtl_hip_engine* e, const rrtl_hip_replay_spec* spec, uint64_t* mismatch_count,
                            rrtl_hip_replay_mismatch* mismatches, uint32_t* n_mismatches);

/* ---- per-design code generation (gpu_codegen::emit_kernel,
 * crates/rrtl-sim-ir/src/gpu_codegen.rs:503, which returns a design's CUDA /
 * OpenCL tick kernel as source text).  The word engine's generated HIP kernel
 * "rrtl_word_jit" for `storage` (0: scratch and signals in HBM, 1: scratch in
 * LDS, 2: scratch and signals in LDS).  With storage 2 and a nonzero
 * `block_threads`, small memories are staged in LDS as the engine does for
 * workgroups of that size.  Writes at most `cap` bytes including the
 * terminating NUL and sets *len to the full length (without NUL); call with cap
 * 0 to size the buffer.  Needs no device. */
rrtl_status rrtl_hip_emit_kernel(const void* program_blob, size_t blob_bytes, uint32_t storage, uint32_t block_threads,
                                 char* out, size_t cap, size_t* len);
/* The generated kernel split between `lane_waves` (2 or 4) waves per 64
 * lanes, as the engine builds it for rrtl_hip_options.lane_waves; the first
 * line is a comment with the split's estimated speed-up over one wave, its
 * barrier phases and exchange rows.  E_HIP_OPTIONS when the design's code
 * cannot be split.  No counterpart in the reference.  Needs no device. */
rrtl_status rrtl_hip_emit_split_kernel(const void* program_blob, size_t blob_bytes, uint32_t lane_waves, char* out,
                                       size_t cap, size_t* len);
/* The record encoding the word engine's interpreter runs for a program, with
 * macro records (interp_macros 1) or with the fusion menu (2), as counts:
 * out4[0] records, out4[1] macro records, out4[2] operations inside macro
 * records, out4[3] scratch slots per lane.  Needs no device.  No counterpart
 * in the reference (its encoder reports nothing). */
rrtl_status rrtl_hip_word_program_info(const void* program_blob, size_t blob_bytes, uint32_t interp_macros,
                                       uint32_t* out4);
/* Compiles HIP source for gfx950 at run time (hiprtc) against the library's
 * device headers, as the word engine does for its generated kernels; sets
 * *code_bytes to the code object's size.  Needs no device.  On failure the
 * compiler's log is in rrtl_hip_create_error(). */
rrtl_status rrtl_hip_compile_kernel(const char* source, size_t* code_bytes);

/* ---- multi-design batches: several engines (different designs, lane counts,
 * cycle counts; all on one device) run their rrtl_hip_run specs in ONE
 * persistent, HIP-graph-captured launch whose workgroups pull (design, tile)
 * items longest first from an atomic queue.  No counterpart in the reference,
 * which runs each design on its own engine (crates/rrtl-runtime/src/lib.rs:8063).
 * The engines must outlive the batch; trace_inputs buffers of the specs must
 * stay valid while the batch exists.  Results land in each engine (digest,
 * trace, halt cycles) exactly as after rrtl_hip_run. */
typedef struct rrtl_hip_batch rrtl_hip_batch;
enum { RRTL_BATCH_PERSISTENT = 0, RRTL_BATCH_PER_DESIGN = 1, RRTL_BATCH_CONCURRENT = 2, RRTL_BATCH_STREAMS = 3 };
rrtl_status rrtl_hip_batch_create(rrtl_hip_engine* const* engines, const rrtl_hip_run_spec* specs, uint32_t n,
                                  rrtl_hip_batch** out);
/* mode RRTL_BATCH_PERSISTENT: one graph launch of a work-queue kernel generated for the batch and compiled at
 * run time, with every design's item body compiled in (gate tiles at their variant, word designs' own
 * generated code, or the record interpreter for engines without it); RRTL_BATCH_PER_DESIGN: one launch per
 * design, back to back (the comparison); RRTL_BATCH_CONCURRENT: every design on its own kernels and stream, the
 * fork/join captured into one graph and launched once; RRTL_BATCH_STREAMS: the same launches issued directly
 * on the streams. */
rrtl_status rrtl_hip_batch_run(rrtl_hip_batch* b, uint32_t mode);
rrtl_status rrtl_hip_batch_stats(const rrtl_hip_batch* b, rrtl_hip_run_stats* out);
/* The last persistent launch's shape as one line of JSON text:
 * {"grid": G, "per_cu": P, "forms": ["<how design 0's items run>", ...]}.  Copies at most cap bytes
 * (NUL-terminated) and sets *len to the text's length; a NULL out with cap 0 asks for the length only.
 * No counterpart in the reference. */
rrtl_status rrtl_hip_batch_describe(const rrtl_hip_batch* b, char* out, size_t cap, size_t* len);
const char* rrtl_hip_batch_last_error(const rrtl_hip_batch* b);
void rrtl_hip_batch_destroy(rrtl_hip_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* RRTL_HIP_H */
