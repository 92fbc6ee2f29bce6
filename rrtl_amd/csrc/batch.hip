// Multi-design batches: many designs with uneven lane counts and cycle counts
// in one persistent, HIP-graph-captured launch.
//
// Each design contributes work items -- gate-engine tiles (W lane-words of one
// bit-sliced netlist, the whole run on chip) or word-engine lane blocks (256
// lanes of one word-level design).  Items are ordered longest first by an
// estimate of their cost and handed out by an atomic counter, so every
// workgroup keeps pulling work until the queue drains: short designs no longer
// leave the GPU idle while a long one runs, and a design too small to fill 256
// CUs shares the chip with the others.  The counter reset and the batch kernel
// are captured once into a graph and replayed.  The alternative the batch is
// measured against is one launch per design on one stream.
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
#include "gate_device.hpp"
#include "word_device.hpp"
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>

namespace rrtl {
namespace {

struct JobDev {
  uint32_t kind, variant, pad0, pad1;
  union U {
    gate::TileArgs g;
    word::WordArgs w;
  } u;
};

struct BatchArgs {
  const JobDev* jobs;
  const uint2* items;   // (job, item) longest first
  uint32_t n_items;
  uint32_t* counter;
};

// Clears the work counter; a kernel (not a memset node) so the graph replays it.
__global__ void batch_reset_kernel(uint32_t* counter) {
  if (threadIdx.x == 0) *counter = 0;
}

__global__ void __launch_bounds__(256) batch_kernel(BatchArgs B) {
  extern __shared__ __align__(16) uint32_t lds[];
  __shared__ uint32_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(B.counter, 1u);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);  // uniform exit branch
    __syncthreads();
    if (it >= B.n_items) break;  // uniform: every wave of the workgroup leaves together
    const uint2 wi = B.items[it];
    const JobDev& J = B.jobs[wi.x];
    if (J.kind == ENGINE_GATE) {
      const uint32_t v = J.variant;
      gate::with_tile_variant(v & 0xFFu, (v & gate::GATE_VARIANT_BYTE_ADDR) != 0, v >> gate::GATE_VARIANT_M_SHIFT,
                              [&](auto w, auto ba, auto m) {
                                gate::tile_body<decltype(w)::value, decltype(ba)::value, decltype(m)::value>(J.u.g, wi.y, lds);
                              });
    } else if (J.variant == word::WM_ALL_LDS) {
      word::word_body<word::WM_ALL_LDS>(J.u.w, wi.y);
    } else if (J.variant == word::WM_SCRATCH_LDS) {
      word::word_body<word::WM_SCRATCH_LDS>(J.u.w, wi.y);
    } else {
      word::word_body<word::WM_HBM>(J.u.w, wi.y);
    }
    __syncthreads();  // LDS is reused by the next item
  }
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
    std::vector<JobDev> jobs(E.size());
    std::vector<std::pair<double, uint2>> order;
    size_t lds = 0;
    for (size_t j = 0; j < E.size(); j++) {
      BatchJob J = E[j]->batch_prepare(S[j]);
      std::memset(&jobs[j], 0, sizeof(JobDev));
      jobs[j].kind = J.kind; jobs[j].variant = J.variant;
      std::memcpy(&jobs[j].u, J.args.data(), std::min(J.args.size(), sizeof(JobDev::U)));
      lds = std::max(lds, J.lds_bytes);
      for (uint32_t i = 0; i < J.n_items; i++) order.push_back({J.item_cost, make_uint2(uint32_t(j), i)});
    }
    if (lds > lds_max_) throw Error("E_HIP_TOO_LARGE: a batch item needs more LDS than a workgroup has");
    std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<uint2> items(order.size());
    for (size_t k = 0; k < order.size(); k++) items[k] = order[k].second;
    const bool same = items == last_items_ && jobs.size() == last_jobs_.size() &&
                      std::memcmp(jobs.data(), last_jobs_.data(), jobs.size() * sizeof(JobDev)) == 0 && lds == last_lds_;
    if (!same) rebuild(jobs, items, lds);
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

 private:
  void rebuild(const std::vector<JobDev>& jobs, const std::vector<uint2>& items, size_t lds) {
    if (exec_) { RRTL_HIP_CHECK(hipGraphExecDestroy(exec_)); exec_ = nullptr; }
    if (graph_) { RRTL_HIP_CHECK(hipGraphDestroy(graph_)); graph_ = nullptr; }
    if (jobs_) { RRTL_HIP_CHECK(hipFree(jobs_)); jobs_ = nullptr; }
    if (items_) { RRTL_HIP_CHECK(hipFree(items_)); items_ = nullptr; }
    RRTL_HIP_CHECK(hipMalloc(&jobs_, jobs.size() * sizeof(JobDev)));
    RRTL_HIP_CHECK(hipMalloc(&items_, std::max<size_t>(1, items.size()) * sizeof(uint2)));
    RRTL_HIP_CHECK(hipMemcpy(jobs_, jobs.data(), jobs.size() * sizeof(JobDev), hipMemcpyHostToDevice));
    if (!items.empty()) RRTL_HIP_CHECK(hipMemcpy(items_, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
    if (lds > 65536)
      RRTL_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&batch_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    int per_cu = 0;
    RRTL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel, 256, lds));
    grid_ = uint32_t(std::max(1, per_cu) * cus_);
    grid_ = std::min<uint32_t>(grid_, std::max<uint32_t>(1, uint32_t(items.size())));
    BatchArgs B{jobs_, items_, uint32_t(items.size()), counter_};
    RRTL_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, stream_, counter_);
    hipLaunchKernelGGL(batch_kernel, dim3(grid_), dim3(256), lds, stream_, B);
    RRTL_HIP_CHECK(hipStreamEndCapture(stream_, &graph_));
    RRTL_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
    last_jobs_ = jobs; last_items_ = items; last_lds_ = lds;
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
  uint32_t grid_ = 0;
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
uint32_t batch_grid(const Batch* b) { return b->grid(); }
void batch_destroy(Batch* b) { delete b; }

}  // namespace rrtl
