// Definitions shared by the host code, the HIP kernels and the kernels the
// word engine generates and compiles at run time: machine-op and block
// numbering, stimulus modes and the counter hash.  No standard-library
// dependency, so the run-time compiler can include it as is.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#else
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
typedef __hip_internal::size_t size_t;
#endif

namespace rrtl {

// PackedInstrKind (crates/rrtl-sim-ir/src/lib.rs:266-303), numbered in declaration order
enum Op : uint32_t {
  OP_LIT, OP_SIGNAL, OP_NOT, OP_AND, OP_OR, OP_XOR, OP_ADD, OP_SUB, OP_MUL, OP_EQ, OP_NE,
  OP_LT, OP_MUX, OP_SLICE, OP_ZEXT, OP_SEXT, OP_TRUNC, OP_CAST, OP_CONCAT, OP_MEMREAD, OP_COUNT
};
enum EffectKind : uint32_t { EF_STORE = 0, EF_CAPTURE = 1, EF_MEMWRITE = 2 };
enum SignalKind : uint32_t { K_INPUT = 0, K_OUTPUT = 1, K_WIRE = 2, K_REG = 3 };
enum BlockId { B_ASYNC = 0, B_COMB = 1, B_NEXT = 2, B_COMMIT = 3 };

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
enum : uint32_t { IN_HOLD = 0, IN_RANDOM = 1, IN_CONST = 2, IN_PULSE = 3, IN_TRACE = 4, IN_GROUP_RANDOM = 5,
                  IN_TRACE_UNIFORM = 6 };

}  // namespace rrtl
