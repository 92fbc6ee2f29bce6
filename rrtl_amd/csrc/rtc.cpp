// hiprtc front end for generated kernels.  The device headers the generated
// sources include (device_defs.hpp, word_device.hpp) are compiled into the
// library as text (rtc_headers.inc, written by the build), so a generated
// kernel is built against exactly the definitions the library itself uses.
#include "rtc.hpp"
#include "common.hpp"
#include <hip/hiprtc.h>
#include <chrono>
#include <map>
#include <mutex>
#include <utility>
#include <vector>
#include "rtc_headers.inc"  // kRtcHeaderNames, kRtcHeaderTexts, kRtcHeaderCount

namespace rrtl {
namespace {

std::mutex g_mu;
std::map<std::pair<int, std::string>, hipModule_t> g_modules;  // never unloaded
double g_compile_s = 0;

}  // namespace

std::vector<char> rtc_compile(const std::string& source) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source.c_str(), "rrtl_generated.hip", kRtcHeaderCount, kRtcHeaderTexts,
                          kRtcHeaderNames) != HIPRTC_SUCCESS)
    throw Error("E_HIP_RUNTIME: hiprtcCreateProgram failed");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(std::string("E_HIP_RUNTIME: run-time compile of a generated kernel failed: ") +
                hiprtcGetErrorString(r) + "\n" + log.substr(0, 4000));
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  std::vector<char> code(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return code;
}

hipFunction_t rtc_kernel(const std::string& source, const char* name, int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(device, source);
  auto it = g_modules.find(key);
  if (it == g_modules.end()) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<char> code = rtc_compile(source);
    g_compile_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // load into `device`'s context, leaving the caller's current device as it was
    int prev = 0;
    RRTL_HIP_CHECK(hipGetDevice(&prev));
    RRTL_HIP_CHECK(hipSetDevice(device));
    hipModule_t mod;
    const hipError_t st = hipModuleLoadData(&mod, code.data());
    RRTL_HIP_CHECK(hipSetDevice(prev));
    RRTL_HIP_CHECK(st);
    it = g_modules.emplace(std::move(key), mod).first;
  }
  hipFunction_t fn;
  RRTL_HIP_CHECK(hipModuleGetFunction(&fn, it->second, name));
  return fn;
}

double rtc_compile_seconds() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_compile_s;
}

}  // namespace rrtl
