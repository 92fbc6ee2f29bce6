// Multi-design batches: many designs with uneven lane counts and cycle counts
// in one persistent, HIP-graph-captured launch.
//
// Each design contributes work items -- gate-engine tiles (W lane-words of one
// bit-sliced netlist, the whole run on chip) or word-engine lane blocks (64 or
// 256 lanes of one word-level design).  Items are ordered most expensive
// first by an estimate of their cost and handed out by an atomic counter, so
// every workgroup keeps pulling work until the queue drains: short designs no
// longer leave the GPU idle while a long one runs, and a design too small to
// fill 256 CUs shares the chip with the others.
//
// The persistent kernel is generated per batch and compiled at run time: each
// design's item body is compiled in at its own shape -- the gate tile body at
// the variant (tile words, positions per stage, addressing) its tile program
// needs, a word design's own generated straight-line code (or the record
// interpreter when the engine runs without it) -- behind one switch on the
// job index.  The workgroup's LDS is the largest item's, and word designs take
// the form that fits within the gate tiles' LDS (batch_prepare), so the tiles
// keep the residency of their own kernel; the register budget is the tile
// kernel's (four waves per SIMD) when the batch has gate items.  The counter
// reset and the batch kernel are captured once into a graph and replayed.
// The alternative the batch is measured against is one launch per design on
// one stream.
//
// The concurrent form keeps every design on its own kernels (the gate tile
// kernel, the word engine's generated or interpreted kernel) and overlaps them
// instead: each design's launches go on a stream of its own, forked from and
// joined to the batch stream, and the whole fork/join is captured into one
// graph that a single graph launch replays.
//
// Reference context: the reference runs each design on its own engine and the
// runtime loops over shards (crates/rrtl-runtime/src/lib.rs:7446, :8063); it has
// no cross-design launch.
#include "../../include/rrtl_hip.h"
#include "engine.hpp"
#include "batch_device.hpp"
#include "rtc.hpp"
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>

namespace rrtl {
namespace {

using batch::BatchArgs;
using batch::JobDev;

// Clears the work counter; a kernel (not a memset node) so the graph replays it.
__global__ void batch_reset_kernel(uint32_t* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

// Source of the persistent kernel "rrtl_batch_jit" for the batch's jobs.
std::string batch_kernel_source(const std::vector<BatchJob>& jobs, bool gate_budget) {
  std::string src = "#include \"batch_device.hpp\"\n\n";
  for (const BatchJob& J : jobs) src += J.code;
  src += std::string("\nextern \"C\" __global__ void __launch_bounds__(256") +
         (gate_budget ? ", " + std::to_string(gate::TILE_WAVES_PER_SIMD) : std::string()) +
         ") rrtl_batch_jit(rrtl::batch::BatchArgs B) {\n"
         "  extern __shared__ __align__(16) uint32_t lds[];\n"
         "  rrtl::batch::batch_loop(B, [&](uint32_t job, uint32_t item, const rrtl::batch::JobDev::U& U) {\n"
         "    switch (job) {\n";
  for (size_t j = 0; j < jobs.size(); j++)
    src += "      case " + std::to_string(j) + ": { " + jobs[j].call + " break; }\n";
  src += "      default: break;\n    }\n  });\n}\n";
  return src;
}

}  // namespace

class Batch {
 public:
  Batch(std::vector<Engine*> engines, std::vector<RunSpec> specs) : E(std::move(engines)), S(std::move(specs)) {
    if (E.empty()) throw Error("E_HIP_BATCH: empty batch");
    RRTL_HIP_CHECK(hipGetDevice(&device_));
    RRTL_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    RRTL_HIP_CHECK(hipEventCreate(&ev0_));
    RRTL_HIP_CHECK(hipEventCreate(&ev1_));
    RRTL_HIP_CHECK(hipMalloc(&counter_, sizeof(uint32_t)));
    hipDeviceProp_t prop;
    RRTL_HIP_CHECK(hipGetDeviceProperties(&prop, device_));
    cus_ = prop.multiProcessorCount;
    lds_max_ = prop.sharedMemPerBlock;
  }
  ~Batch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (cexec_) (void)hipGraphExecDestroy(cexec_);
    if (cgraph_) (void)hipGraphDestroy(cgraph_);
    for (auto e : join_) (void)hipEventDestroy(e);
    if (fork_) (void)hipEventDestroy(fork_);
    for (auto st : side_) (void)hipStreamDestroy(st);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (jobs_) (void)hipFree(jobs_);
    if (items_) (void)hipFree(items_);
    if (counter_) (void)hipFree(counter_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  // One persistent launch for every design.
  void run_persistent() {
    // gate tiles first: their LDS is the budget the word designs' items must fit
    std::vector<BatchJob> jobs(E.size());
    size_t budget = 0;
    bool any_gate = false;
    for (size_t j = 0; j < E.size(); j++)
      if (E[j]->kind() == ENGINE_GATE) {
        jobs[j] = E[j]->batch_prepare(S[j], BatchRequest{});
        budget = std::max(budget, jobs[j].lds_bytes);
        any_gate = true;
      }
    if (!any_gate) budget = 64 * 1024;
    for (size_t j = 0; j < E.size(); j++)
      if (E[j]->kind() != ENGINE_GATE) {
        BatchRequest req;
        req.lds_budget = budget;
        req.ns = "d" + std::to_string(j);
        jobs[j] = E[j]->batch_prepare(S[j], req);
      }
    std::vector<JobDev> dev(E.size());
    std::vector<std::pair<double, uint2>> order;
    size_t lds = 0;
    forms_.clear();
    for (size_t j = 0; j < E.size(); j++) {
      const BatchJob& J = jobs[j];
      std::memset(&dev[j], 0, sizeof(JobDev));
      std::memcpy(&dev[j].u, J.args.data(), std::min(J.args.size(), sizeof(JobDev::U)));
      lds = std::max(lds, J.lds_bytes);
      forms_.push_back(J.form);
      for (uint32_t i = 0; i < J.n_items; i++) order.push_back({J.item_cost, make_uint2(uint32_t(j), i)});
    }
    if (lds > std::min<size_t>(lds_max_, 64 * 1024))
      throw Error("E_HIP_TOO_LARGE: a batch item needs more than 64 KiB of LDS");
    std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<uint2> items(order.size());
    for (size_t k = 0; k < order.size(); k++) items[k] = order[k].second;
    const std::string src = batch_kernel_source(jobs, any_gate);
    const bool same = src == last_src_ && items == last_items_ && dev.size() == last_jobs_.size() &&
                      std::memcmp(dev.data(), last_jobs_.data(), dev.size() * sizeof(JobDev)) == 0 && lds == last_lds_;
    if (!same) rebuild(src, dev, items, lds);
    for (Engine* e : E) e->batch_reset(stream_);
    RRTL_HIP_CHECK(hipEventRecord(ev0_, stream_));
    RRTL_HIP_CHECK(hipGraphLaunch(exec_, stream_));
    RRTL_HIP_CHECK(hipEventRecord(ev1_, stream_));
    RRTL_HIP_CHECK(hipEventSynchronize(ev1_));
    float ms = 0;
    RRTL_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
    finish(ms, 1);
  }

  // The comparison: one launch per design, back to back.
  void run_per_design() {
    double ms = 0;
    uint64_t launches = 0;
    for (size_t j = 0; j < E.size(); j++) {
      E[j]->run(S[j]);
      ms += E[j]->last_stats().kernel_ms;
      launches += E[j]->last_stats().launches;
    }
    stats_ = RunStats{};
    stats_.kernel_ms = ms;
    stats_.launches = launches;
    for (size_t j = 0; j < E.size(); j++) stats_.lane_cycles += E[j]->last_stats().lane_cycles;
  }

  // Every design on its own kernels and stream, overlapped: one graph launch
  // of the captured fork/join (graph), or the same launches issued directly.
  void run_concurrent(bool graph) {
    for (size_t j = 0; j < E.size(); j++) E[j]->async_prepare(S[j]);
    if (side_.empty()) {
      side_.resize(E.size());
      join_.resize(E.size());
      for (size_t j = 0; j < E.size(); j++) {
        RRTL_HIP_CHECK(hipStreamCreateWithFlags(&side_[j], hipStreamNonBlocking));
        RRTL_HIP_CHECK(hipEventCreateWithFlags(&join_[j], hipEventDisableTiming));
      }
      RRTL_HIP_CHECK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    }
    uint64_t launches = 0;
    auto issue = [&]() {
      RRTL_HIP_CHECK(hipEventRecord(fork_, stream_));
      for (size_t j = 0; j < E.size(); j++) {
        RRTL_HIP_CHECK(hipStreamWaitEvent(side_[j], fork_, 0));
        launches += E[j]->async_enqueue(side_[j]);
        RRTL_HIP_CHECK(hipEventRecord(join_[j], side_[j]));
        RRTL_HIP_CHECK(hipStreamWaitEvent(stream_, join_[j], 0));
      }
    };
    if (graph) {
      // the launches depend on the prepared runs; capture afresh each time
      if (cexec_) { RRTL_HIP_CHECK(hipGraphExecDestroy(cexec_)); cexec_ = nullptr; }
      if (cgraph_) { RRTL_HIP_CHECK(hipGraphDestroy(cgraph_)); cgraph_ = nullptr; }
      RRTL_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      issue();
      RRTL_HIP_CHECK(hipStreamEndCapture(stream_, &cgraph_));
      RRTL_HIP_CHECK(hipGraphInstantiate(&cexec_, cgraph_, nullptr, nullptr, 0));
    }
    for (Engine* e : E) e->batch_reset(stream_);
    RRTL_HIP_CHECK(hipEventRecord(ev0_, stream_));
    if (graph) {
      RRTL_HIP_CHECK(hipGraphLaunch(cexec_, stream_));
    } else {
      issue();
    }
    RRTL_HIP_CHECK(hipEventRecord(ev1_, stream_));
    RRTL_HIP_CHECK(hipEventSynchronize(ev1_));
    float ms = 0;
    RRTL_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
    stats_ = RunStats{};
    stats_.kernel_ms = ms;
    stats_.launches = graph ? 1 : launches;
    for (size_t j = 0; j < E.size(); j++) {
      E[j]->async_finish(S[j]);
      stats_.lane_cycles += E[j]->last_stats().lane_cycles;
    }
  }

  const RunStats& stats() const { return stats_; }
  uint32_t grid() const { return grid_; }
  uint32_t per_cu() const { return per_cu_; }
  const std::vector<std::string>& forms() const { return forms_; }

 private:
  void rebuild(const std::string& src, const std::vector<JobDev>& jobs, const std::vector<uint2>& items, size_t lds) {
    if (exec_) { RRTL_HIP_CHECK(hipGraphExecDestroy(exec_)); exec_ = nullptr; }
    if (graph_) { RRTL_HIP_CHECK(hipGraphDestroy(graph_)); graph_ = nullptr; }
    if (jobs_) { RRTL_HIP_CHECK(hipFree(jobs_)); jobs_ = nullptr; }
    if (items_) { RRTL_HIP_CHECK(hipFree(items_)); items_ = nullptr; }
    fn_ = rtc_kernel(src, "rrtl_batch_jit", device_);
    RRTL_HIP_CHECK(hipMalloc(&jobs_, jobs.size() * sizeof(JobDev)));
    RRTL_HIP_CHECK(hipMalloc(&items_, std::max<size_t>(1, items.size()) * sizeof(uint2)));
    RRTL_HIP_CHECK(hipMemcpy(jobs_, jobs.data(), jobs.size() * sizeof(JobDev), hipMemcpyHostToDevice));
    if (!items.empty()) RRTL_HIP_CHECK(hipMemcpy(items_, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
    int per_cu = 0;
    RRTL_HIP_CHECK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn_, 256, lds));
    per_cu_ = uint32_t(std::max(1, per_cu));
    grid_ = per_cu_ * uint32_t(cus_);
    grid_ = std::min<uint32_t>(grid_, std::max<uint32_t>(1, uint32_t(items.size())));
    BatchArgs B{jobs_, items_, uint32_t(items.size()), counter_};
    void* params[] = {&B};
    RRTL_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream_, counter_);
    RRTL_HIP_CHECK(hipModuleLaunchKernel(fn_, grid_, 1, 1, 256, 1, 1, uint32_t(lds), stream_, params, nullptr));
    RRTL_HIP_CHECK(hipStreamEndCapture(stream_, &graph_));
    RRTL_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    last_src_ = src; last_jobs_ = jobs; last_items_ = items; last_lds_ = lds;
  }

  void finish(double ms, uint64_t launches) {
    stats_ = RunStats{};
    stats_.kernel_ms = ms;
    stats_.launches = launches;
    for (size_t j = 0; j < E.size(); j++) {
      E[j]->batch_finish(S[j]);
      stats_.lane_cycles += E[j]->last_stats().lane_cycles;
    }
  }

  std::vector<Engine*> E;
  std::vector<RunSpec> S;
  int device_ = 0, cus_ = 256;
  size_t lds_max_ = 65536;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  std::vector<hipStream_t> side_;       // concurrent form: one stream per design
  std::vector<hipEvent_t> join_;
  hipEvent_t fork_ = nullptr;
  hipGraph_t cgraph_ = nullptr;
  hipGraphExec_t cexec_ = nullptr;
  JobDev* jobs_ = nullptr;
  uint2* items_ = nullptr;
  uint32_t* counter_ = nullptr;
  uint32_t grid_ = 0, per_cu_ = 0;
  hipFunction_t fn_ = nullptr;          // the generated persistent kernel (module cached by source)
  std::vector<std::string> forms_;      // per design: how its items run in the persistent kernel
  std::string last_src_;
  std::vector<JobDev> last_jobs_;
  std::vector<uint2> last_items_;
  size_t last_lds_ = 0;
  RunStats stats_;
};

Batch* make_batch(std::vector<Engine*> engines, std::vector<RunSpec> specs) {
  return new Batch(std::move(engines), std::move(specs));
}
void batch_run(Batch* b, uint32_t mode) {
  if (mode == RRTL_BATCH_PERSISTENT) b->run_persistent();
  else if (mode == RRTL_BATCH_PER_DESIGN) b->run_per_design();
  else if (mode == RRTL_BATCH_CONCURRENT) b->run_concurrent(true);
  else if (mode == RRTL_BATCH_STREAMS) b->run_concurrent(false);
  else throw Error("E_HIP_BATCH: unknown batch mode");
}
const RunStats& batch_stats(const Batch* b) { return b->stats(); }
std::string batch_describe(const Batch* b) {
  std::string t = "{\"grid\": " + std::to_string(b->grid()) + ", \"per_cu\": " + std::to_string(b->per_cu()) +
                  ", \"forms\": [";
  for (size_t j = 0; j < b->forms().size(); j++) t += (j ? ", \"" : "\"") + b->forms()[j] + "\"";
  return t + "]}";
}
void batch_destroy(Batch* b) { delete b; }

}  // namespace rrtl
