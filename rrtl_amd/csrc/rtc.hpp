// Run-time compilation of generated device code for gfx950 (hiprtc), with the
// library's device headers supplied from text embedded at build time.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

namespace rrtl {

// The kernel `name` of `source` compiled for `device`.  Compiled modules are
// cached per (device, source) for the life of the process; a compile error
// throws E_HIP_RUNTIME with the compiler's log.  Thread-safe.
hipFunction_t rtc_kernel(const std::string& source, const char* name, int device);

// Code object of `source` compiled for gfx950 (no device needed); a compile
// error throws E_HIP_RUNTIME with the compiler's log.
std::vector<char> rtc_compile(const std::string& source);

// Seconds spent compiling (cache misses only) since the process started.
double rtc_compile_seconds();

}  // namespace rrtl
