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

// SplitMix64 finaliser (Steele, Lea, Flood, OOPSLA 2014) used as a counter-based
// hash for on-device stimulus: bit b of input s at cycle c for lane l is bit
// (l mod 64) of  stim_word(seed, (s << 8) | b, c, l / 64).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t stim_key(uint64_t seed, uint64_t stream, uint64_t cycle) {
  return mix64(seed ^ ((stream + 1) * 0x9E3779B97F4A7C15ull) ^ ((cycle + 1) * 0xD1B54A32D192ED03ull));
}
__host__ __device__ inline uint64_t stim_from_key(uint64_t key, uint64_t word64) {
  return mix64(key + (word64 + 1) * 0x9E3779B97F4A7C15ull);
}
// per-lane observation chain of the word engine
__host__ __device__ inline uint64_t lane_digest(uint64_t d, uint32_t limb, uint32_t j, uint32_t k) {
  return mix64(d ^ uint64_t(limb) ^ (uint64_t(j * 64u + k + 1u) << 32));
}

// Stimulus modes (rrtl_hip_input_mode in the C header).
enum : uint32_t { IN_HOLD = 0, IN_RANDOM = 1, IN_CONST = 2, IN_PULSE = 3, IN_TRACE = 4, IN_GROUP_RANDOM = 5 };

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
