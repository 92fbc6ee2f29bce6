// Shared host/device helpers for the HIP engines.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "program.hpp"

namespace rrtl {

#define RRTL_HIP_CHECK(expr)                                                                 \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      throw ::rrtl::Error(std::string("E_HIP_RUNTIME: ") + #expr + ": " + hipGetErrorString(_e)); \
  } while (0)

// mix64, stim_key, stim_from_key, lane_digest and the stimulus modes: device_defs.hpp

struct RunSpec {
  uint64_t seed = 0, cycles = 0;
  std::vector<uint32_t> in_sig, in_mode;
  std::vector<uint64_t> in_arg;
  std::vector<uint32_t> obs_sig;
  bool observe_post = false;
  bool trace = false;
  uint32_t halt_sig = 0xFFFFFFFFu;
  const uint32_t* trace_inputs = nullptr;  // IN_TRACE values, host [cycles][n_trace_limbs][lanes]
  bool detect = false;                     // flag lanes whose observed values leave lane 0 of their 64-lane group
  std::vector<uint64_t> group_key;         // IN_GROUP_RANDOM key per 64-lane group (empty: the group index)
  // Trace replay (word engine): values shared by all lanes, checks of the
  // observed signals, and the mismatch records kept per lane.
  const uint32_t* trace_uniform = nullptr; // IN_TRACE_UNIFORM values, host [cycles][n_uniform_limbs]
  const uint32_t* expect = nullptr;        // host [cycles][observed limbs] (or [...][lanes] when expect_per_lane)
  bool expect_per_lane = false;
  const uint32_t* check_active = nullptr;  // host [cycles][ceil(observed / 32)] bits, or null: every check
  bool check_lane0 = false;                // only lane 0 compares
  uint32_t max_mismatches = 0;             // records kept per lane
  bool final_eval = true;                  // false: no settle after the last cycle (replay ends on the edge)
};

// One mismatch of a replay: the cycle (step), the check's position among the
// observed signals, the lane and the actual value's first four limbs.
struct Mismatch {
  uint64_t step = 0;
  uint32_t check = 0, lane = 0;
  uint32_t actual[4] = {0, 0, 0, 0};
};

struct RunStats {
  double kernel_ms = 0;        // device time of the timed launches (HIP events)
  uint64_t launches = 0;
  uint64_t lane_cycles = 0;    // cycles actually simulated, summed over lanes
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
  // Keeps the allocation when it is already large enough, so device pointers
  // handed to earlier launches (or captured graphs) stay valid across runs.
  void alloc(size_t count) {
    if (count == 0) count = 1;
    if (p && count <= n) return;
    release();
    RRTL_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) RRTL_HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void zero(hipStream_t s) { RRTL_HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
};

}  // namespace rrtl
