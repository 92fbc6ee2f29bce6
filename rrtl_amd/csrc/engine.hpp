// Engine interface: the C++ mirror of the reference's backend trait
// `trait SimBackend` (crates/rrtl-sim-ir/src/lib.rs:5258-5270) plus the runtime
// dispatch surface of `RuntimeEngine` (crates/rrtl-runtime/src/lib.rs:7785-8100):
// signal/memory access in limbs, eval_combinational, tick,
// tick_from_evaluated_no_post_eval, tick_many, snapshot/restore in the
// PackedSimulatorStorage layout, stuck-at forces, and the device-resident run
// (stimulus in, observations out) that replaces per-cycle host round trips.
#pragma once
#include "common.hpp"

namespace rrtl {

enum EngineKind : uint32_t { ENGINE_AUTO = 0, ENGINE_WORD = 1, ENGINE_GATE = 2 };

struct EngineOptions {
  int device = 0;
  uint32_t block_threads = 0;      // 0: engine default
  uint32_t tile_words = 0;         // gate engine: u32 lane-words per workgroup tile (0: autotune)
  uint32_t ticks_per_launch = 0;   // 0: whole run in one launch
  uint32_t tile_chunks = 0;        // gate engine: work-queue chunks per tile (0: automatic, 1: off)
  uint32_t stage_gates = 0;        // gate engine: gate positions per thread per stage (0: automatic, 2, 3 or 4)
  uint32_t codegen = 0;            // word engine: CODEGEN_AUTO / CODEGEN_ON (generated kernel) or CODEGEN_OFF (interpreter)
  uint32_t wave_lanes = 0;         // word engine: active lanes per 64-thread wave (16, 32, 64; 0: automatic)
  uint32_t compact_lanes = 0;      // word engine: pack the live lanes of halting runs between launches
  uint32_t lane_waves = 0;         // word engine, generated code: waves per 64 lanes (0: automatic, 1, 2, 4)
  uint32_t interp_macros = 0;      // word engine, record interpreter: macro records (0: automatic = on, 1: on, 2: off)
};
enum : uint32_t { CODEGEN_AUTO = 0, CODEGEN_ON = 1, CODEGEN_OFF = 2 };

// One design's share of the persistent batch kernel (batch.hip), which is
// generated per batch: the design's kernel arguments, its device code and the
// statement that runs one of its work items (gate-engine tiles or word-engine
// lane blocks), how many items it has, their LDS need and an estimate of each
// item's cost for ordering.  In `call`, `U` names the job's argument union
// (U.g: gate::TileArgs, U.w: word::WordArgs), `item` the item index and `lds`
// the workgroup's dynamic LDS.
struct BatchJob {
  uint32_t kind = 0;          // ENGINE_GATE or ENGINE_WORD
  uint32_t n_items = 0;
  size_t lds_bytes = 0;
  double item_cost = 0;
  std::string code;           // definitions at namespace scope (the design's generated code), may be empty
  std::string call;           // one statement
  std::string form;           // how the items run, for launch reports
  std::vector<uint8_t> args;  // gate::TileArgs or word::WordArgs
};
// What the batch asks of a design's items: they must fit in `lds_budget`
// bytes of LDS per 256-thread workgroup, and their definitions go in
// namespace `ns` (unique within the batch).
struct BatchRequest {
  size_t lds_budget = 65536;
  std::string ns;
};

class Engine {
 public:
  Engine(Program prog, uint32_t lanes, const EngineOptions& opt) : P(std::move(prog)), lanes_(lanes), opt_(opt) {}
  virtual ~Engine() = default;
  virtual uint32_t kind() const = 0;
  uint32_t lanes() const { return lanes_; }
  const Program& program() const { return P; }

  // Writes lane values ([limbs][lanes] u32); `settle` re-evaluates comb like
  // set_packed_signal_limbs (lib.rs:9909-9943).
  virtual void set_signal_limbs(uint32_t sig, const uint32_t* lane_limbs, bool settle) = 0;
  virtual void get_signal_limbs(uint32_t sig, uint32_t* out) = 0;
  // Memory values are [lanes][depth][limbs] (set_memory_limbs, lib.rs:10136).
  virtual void set_memory_limbs(uint32_t mem, const uint32_t* vals, bool settle) = 0;
  virtual void get_memory_limbs(uint32_t mem, uint32_t* out) = 0;
  virtual void eval_combinational() = 0;
  virtual void tick() = 0;
  virtual void tick_from_evaluated() = 0;
  virtual void tick_many(uint64_t steps) = 0;
  // values: word-major (offset+limb)*lanes+lane; memories: lane-major (lib.rs:10045-10054)
  virtual void snapshot(uint32_t* values, uint32_t* memories) = 0;
  virtual void restore(const uint32_t* values, const uint32_t* memories) = 0;
  virtual void set_force(uint32_t sig, uint32_t lane, int value) = 0;
  // n forces at once, in order, as n set_force calls
  virtual void set_forces(const uint32_t* sig, const uint32_t* lane, const uint8_t* value, size_t n) {
    for (size_t i = 0; i < n; i++) set_force(sig[i], lane[i], value[i]);
  }
  virtual void clear_forces() = 0;
  virtual void run(const RunSpec& spec) = 0;
  // Results of the last run: digest (gate engine [64][words32] u32 bit-sliced,
  // word engine [lanes] u64), trace, per-lane halt cycles.
  virtual size_t digest_bytes() const = 0;
  virtual void get_digest(void* out) = 0;
  virtual size_t trace_bytes() const = 0;
  virtual void get_trace(void* out) = 0;
  virtual void get_halt_cycles(uint64_t* out) { (void)out; throw Error("E_HIP_UNSUPPORTED: engine has no halting"); }
  // Runs with detect set: lane flags packed 32 per u32 word (ceil(lanes/32) words).
  virtual void get_detect(uint32_t* out) { (void)out; throw Error("E_HIP_UNSUPPORTED: engine has no fault detection"); }
  // Runs with expected values: the total mismatch count and the first `max`
  // mismatches in (step, check, lane) order.
  virtual uint64_t get_mismatches(uint32_t max, std::vector<Mismatch>& out) {
    (void)max; (void)out;
    throw Error("E_HIP_UNSUPPORTED: trace replay with checks runs on the word engine");
  }
  virtual void synchronize() { RRTL_HIP_CHECK(hipStreamSynchronize(stream_)); }
  // Multi-design batches: upload a run and return its share of the generated
  // batch kernel; reset its outputs before every launch; book-keep after it.
  virtual BatchJob batch_prepare(const RunSpec& spec, const BatchRequest& req) = 0;
  virtual void batch_reset(hipStream_t s) = 0;
  virtual void batch_finish(const RunSpec& spec) = 0;
  // Concurrent batches: prepare a run of the engine's own kernels (uploads;
  // returns when they are done), enqueue those kernels on `s` without copies,
  // fills or host waits (so the launches can be captured into a graph;
  // returns the launch count), book-keep after they ran.  batch_reset clears
  // the run's outputs before each launch.
  virtual void async_prepare(const RunSpec& spec) = 0;
  virtual uint64_t async_enqueue(hipStream_t s) = 0;
  virtual void async_finish(const RunSpec& spec) = 0;
  const RunStats& last_stats() const { return stats_; }
  hipStream_t stream() const { return stream_; }

 protected:
  Program P;
  uint32_t lanes_;
  EngineOptions opt_;
  hipStream_t stream_ = nullptr;
  RunStats stats_;
};

Engine* make_gate_engine(Program prog, uint32_t lanes, const EngineOptions& opt);
Engine* make_word_engine(Program prog, uint32_t lanes, const EngineOptions& opt);

}  // namespace rrtl
