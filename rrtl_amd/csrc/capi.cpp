// C ABI over the engines (include/rrtl_hip.h).  Every entry point catches C++
// exceptions and turns them into a status plus an "E_..." message.
#include "../../include/rrtl_hip.h"
#include <cstddef>
#include "engine.hpp"
#include "rtc.hpp"
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>

namespace rrtl {
void gate_engine_get_words(Engine* e, uint32_t sig, uint32_t* out);
void gate_engine_tile_info(Engine* e, uint32_t* info);
void word_engine_launch_info(Engine* e, uint32_t* info);
class Batch;
Batch* make_batch(std::vector<Engine*> engines, std::vector<RunSpec> specs);
void batch_run(Batch* b, uint32_t mode);
const RunStats& batch_stats(const Batch* b);
std::string batch_describe(const Batch* b);
void batch_destroy(Batch* b);
std::string word_kernel_source(const Program& prog, int mode, uint32_t block);
std::string word_split_kernel_source(const Program& prog, uint32_t lane_waves);
void word_program_info(const Program& prog, bool macros, uint32_t* out4);
}

struct rrtl_hip_batch {
  rrtl::Batch* impl = nullptr;
  std::string err;
  ~rrtl_hip_batch() { if (impl) rrtl::batch_destroy(impl); }
};

struct rrtl_hip_engine {
  std::unique_ptr<rrtl::Engine> impl;
  std::string err;
};

namespace {
thread_local std::string g_create_err;

rrtl_status classify(const std::string& m) {
  if (m.rfind("E_HIP_RUNTIME", 0) == 0) return RRTL_E_HIP;
  if (m.rfind("E_PROGRAM", 0) == 0 || m.rfind("E_SIM_IR_MACHINE", 0) == 0) return RRTL_E_PROGRAM;
  if (m.rfind("E_BITPARALLEL", 0) == 0 || m.rfind("E_HIP_UNSUPPORTED", 0) == 0 || m.rfind("E_HIP_WIDE", 0) == 0 ||
      m.rfind("E_HIP_TOO_LARGE", 0) == 0 || m.rfind("E_HIP_COMB_ORDER", 0) == 0)
    return RRTL_E_UNSUPPORTED;
  if (m.rfind("E_HIP_NO_RUN", 0) == 0) return RRTL_E_STATE;
  return RRTL_E_INVALID;
}

template <class F>
rrtl_status guard(rrtl_hip_engine* e, F&& f) {
  if (!e || !e->impl) return RRTL_E_INVALID;
  try {
    f(*e->impl);
    e->err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    e->err = ex.what();
    return classify(e->err);
  }
}

rrtl::RunSpec to_spec(const rrtl_hip_run_spec* s, uint32_t lanes) {
  if (!s || s->struct_size < offsetof(rrtl_hip_run_spec, detect))
    throw rrtl::Error("E_HIP_RUN_SPEC: run spec is null or too small");
  rrtl::RunSpec r;
  r.seed = s->seed; r.cycles = s->cycles;
  if (s->n_inputs && (!s->input_signal || !s->input_mode)) throw rrtl::Error("E_HIP_RUN_SPEC: input arrays are null");
  if (s->n_observe && !s->observe_signal) throw rrtl::Error("E_HIP_RUN_SPEC: observe array is null");
  r.in_sig.assign(s->input_signal, s->input_signal + s->n_inputs);
  r.in_mode.assign(s->input_mode, s->input_mode + s->n_inputs);
  if (s->input_arg) r.in_arg.assign(s->input_arg, s->input_arg + s->n_inputs);
  else r.in_arg.assign(s->n_inputs, 0);
  for (uint32_t m : r.in_mode) if (m > RRTL_IN_GROUP_RANDOM) throw rrtl::Error("E_HIP_RUN_SPEC: unknown input mode");
  r.obs_sig.assign(s->observe_signal, s->observe_signal + s->n_observe);
  r.observe_post = s->observe_point == RRTL_OBSERVE_POST;
  r.trace = s->record_trace != 0;
  r.halt_sig = s->halt_signal;
  r.trace_inputs = s->trace_inputs;
  if (s->struct_size >= offsetof(rrtl_hip_run_spec, detect) + sizeof(uint32_t)) r.detect = s->detect != 0;
  if (s->struct_size >= offsetof(rrtl_hip_run_spec, group_key) + sizeof(void*) && s->group_key) {
    // one key per 64-lane group of the engine
    r.group_key.assign(s->group_key, s->group_key + (uint64_t(lanes) + 63) / 64);
  }
  return r;
}
}  // namespace

extern "C" {

rrtl_status rrtl_hip_create(const void* blob, size_t bytes, uint32_t lanes, const rrtl_hip_options* options,
                            rrtl_hip_engine** out) {
  if (!out) return RRTL_E_INVALID;
  *out = nullptr;
  try {
    if (lanes == 0) throw rrtl::Error("E_SIM_IR_LANES: packed simulator requires at least one lane");
    rrtl::EngineOptions opt;
    uint32_t kind = RRTL_HIP_AUTO;
    if (options) {
      if (options->struct_size < offsetof(rrtl_hip_options, tile_chunks))
        throw rrtl::Error("E_HIP_OPTIONS: options struct is too small");
      kind = options->engine; opt.device = options->device; opt.block_threads = options->block_threads;
      opt.tile_words = options->tile_words; opt.ticks_per_launch = options->ticks_per_launch;
      if (options->struct_size >= offsetof(rrtl_hip_options, tile_chunks) + sizeof(uint32_t))
        opt.tile_chunks = options->tile_chunks;
      if (options->struct_size >= offsetof(rrtl_hip_options, stage_gates) + sizeof(uint32_t))
        opt.stage_gates = options->stage_gates;
      if (options->struct_size >= offsetof(rrtl_hip_options, codegen) + sizeof(uint32_t))
        opt.codegen = options->codegen;
      if (options->struct_size >= offsetof(rrtl_hip_options, wave_lanes) + sizeof(uint32_t))
        opt.wave_lanes = options->wave_lanes;
      if (options->struct_size >= offsetof(rrtl_hip_options, compact_lanes) + sizeof(uint32_t))
        opt.compact_lanes = options->compact_lanes;
      if (options->struct_size >= offsetof(rrtl_hip_options, lane_waves) + sizeof(uint32_t))
        opt.lane_waves = options->lane_waves;
      if (options->struct_size >= offsetof(rrtl_hip_options, interp_macros) + sizeof(uint32_t))
        opt.interp_macros = options->interp_macros;
    }
    if (opt.interp_macros > 2) throw rrtl::Error("E_HIP_OPTIONS: interp_macros must be 0, 1 or 2");
    if (opt.lane_waves > 4 || opt.lane_waves == 3)
      throw rrtl::Error("E_HIP_OPTIONS: lane_waves must be 0, 1, 2 or 4");
    if (opt.block_threads && (opt.block_threads % 64 || opt.block_threads > 256))
      throw rrtl::Error("E_HIP_OPTIONS: block_threads must be 64, 128, 192 or 256");
    if (opt.tile_words && opt.tile_words != 1 && opt.tile_words != 2 && opt.tile_words != 4 && opt.tile_words != 8)
      throw rrtl::Error("E_HIP_OPTIONS: tile_words must be 1, 2, 4 or 8");
    if (opt.codegen > rrtl::CODEGEN_OFF) throw rrtl::Error("E_HIP_OPTIONS: codegen must be 0, 1 or 2");
    if (opt.wave_lanes && (opt.wave_lanes < 4 || opt.wave_lanes > 64 || (opt.wave_lanes & (opt.wave_lanes - 1))))
      throw rrtl::Error("E_HIP_OPTIONS: wave_lanes must be 4, 8, 16, 32 or 64");
    if (opt.stage_gates && (opt.stage_gates < 2 || opt.stage_gates > 4))
      throw rrtl::Error("E_HIP_OPTIONS: stage_gates must be 2, 3 or 4");
    rrtl::Program prog = rrtl::Program::parse(blob, bytes);
    if (kind == RRTL_HIP_AUTO) kind = prog.is_gate_level() ? RRTL_HIP_GATE : RRTL_HIP_WORD;
    auto* h = new rrtl_hip_engine;
    try {
      if (kind == RRTL_HIP_GATE) h->impl.reset(rrtl::make_gate_engine(std::move(prog), lanes, opt));
      else if (kind == RRTL_HIP_WORD) h->impl.reset(rrtl::make_word_engine(std::move(prog), lanes, opt));
      else throw rrtl::Error("E_HIP_OPTIONS: unknown engine kind");
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    g_create_err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    g_create_err = ex.what();
    return classify(g_create_err);
  }
}

static rrtl_status emit_source(const std::string& src, char* out, size_t cap, size_t* len) {
  *len = src.size();
  if (cap) {
    const size_t n = std::min(cap - 1, src.size());
    std::memcpy(out, src.data(), n);
    out[n] = '\0';
  }
  g_create_err.clear();
  return RRTL_OK;
}

rrtl_status rrtl_hip_emit_split_kernel(const void* blob, size_t bytes, uint32_t lane_waves, char* out, size_t cap,
                                       size_t* len) {
  try {
    if (!blob || !len || (cap && !out)) throw rrtl::Error("E_HIP_NULL: null argument");
    return emit_source(rrtl::word_split_kernel_source(rrtl::Program::parse(blob, bytes), lane_waves), out, cap, len);
  } catch (const std::exception& ex) {
    g_create_err = ex.what();
    return classify(g_create_err);
  }
}

rrtl_status rrtl_hip_emit_kernel(const void* blob, size_t bytes, uint32_t storage, uint32_t block_threads, char* out,
                                 size_t cap, size_t* len) {
  try {
    if (!blob || !len || (cap && !out)) throw rrtl::Error("E_HIP_NULL: null argument");
    const std::string src =
        rrtl::word_kernel_source(rrtl::Program::parse(blob, bytes), int(storage), block_threads);
    *len = src.size();
    if (cap) {
      const size_t n = std::min(cap - 1, src.size());
      std::memcpy(out, src.data(), n);
      out[n] = '\0';
    }
    g_create_err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    g_create_err = ex.what();
    return classify(g_create_err);
  }
}

rrtl_status rrtl_hip_word_program_info(const void* blob, size_t bytes, uint32_t interp_macros, uint32_t* out4) {
  try {
    if (!blob || !out4) throw rrtl::Error("E_HIP_NULL: null argument");
    if (interp_macros < 1 || interp_macros > 2) throw rrtl::Error("E_HIP_OPTIONS: interp_macros must be 1 or 2");
    rrtl::word_program_info(rrtl::Program::parse(blob, bytes), interp_macros == 1, out4);
    g_create_err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    g_create_err = ex.what();
    return classify(g_create_err);
  }
}

rrtl_status rrtl_hip_compile_kernel(const char* source, size_t* code_bytes) {
  try {
    if (!source || !code_bytes) throw rrtl::Error("E_HIP_NULL: null argument");
    *code_bytes = rrtl::rtc_compile(source).size();
    g_create_err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    g_create_err = ex.what();
    return classify(g_create_err);
  }
}

void rrtl_hip_destroy(rrtl_hip_engine* e) { delete e; }
const char* rrtl_hip_last_error(const rrtl_hip_engine* e) { return e ? e->err.c_str() : "E_HIP_NULL: null engine"; }
const char* rrtl_hip_create_error(void) { return g_create_err.c_str(); }
uint32_t rrtl_hip_kind(const rrtl_hip_engine* e) { return e && e->impl ? e->impl->kind() : 0; }
uint32_t rrtl_hip_lanes(const rrtl_hip_engine* e) { return e && e->impl ? e->impl->lanes() : 0; }

rrtl_status rrtl_hip_set_signal_limbs(rrtl_hip_engine* e, uint32_t s, const uint32_t* v) {
  return guard(e, [&](rrtl::Engine& x) { if (!v) throw rrtl::Error("E_HIP_NULL: null buffer"); x.set_signal_limbs(s, v, true); });
}
rrtl_status rrtl_hip_set_signal_limbs_raw(rrtl_hip_engine* e, uint32_t s, const uint32_t* v) {
  return guard(e, [&](rrtl::Engine& x) { if (!v) throw rrtl::Error("E_HIP_NULL: null buffer"); x.set_signal_limbs(s, v, false); });
}
rrtl_status rrtl_hip_get_signal_limbs(rrtl_hip_engine* e, uint32_t s, uint32_t* out) {
  return guard(e, [&](rrtl::Engine& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_signal_limbs(s, out); });
}
rrtl_status rrtl_hip_get_signal_bits(rrtl_hip_engine* e, uint32_t s, uint32_t* out) {
  return guard(e, [&](rrtl::Engine& x) {
    if (!out) throw rrtl::Error("E_HIP_NULL: null buffer");
    if (x.kind() != RRTL_HIP_GATE) throw rrtl::Error("E_HIP_UNSUPPORTED: packed bits need the gate engine");
    rrtl::gate_engine_get_words(&x, s, out);
  });
}
rrtl_status rrtl_hip_set_memory_limbs(rrtl_hip_engine* e, uint32_t m, const uint32_t* v) {
  return guard(e, [&](rrtl::Engine& x) { if (!v) throw rrtl::Error("E_HIP_NULL: null buffer"); x.set_memory_limbs(m, v, true); });
}
rrtl_status rrtl_hip_get_memory_limbs(rrtl_hip_engine* e, uint32_t m, uint32_t* out) {
  return guard(e, [&](rrtl::Engine& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_memory_limbs(m, out); });
}
rrtl_status rrtl_hip_eval_combinational(rrtl_hip_engine* e) { return guard(e, [](rrtl::Engine& x) { x.eval_combinational(); }); }
rrtl_status rrtl_hip_tick(rrtl_hip_engine* e) { return guard(e, [](rrtl::Engine& x) { x.tick(); }); }
rrtl_status rrtl_hip_tick_from_evaluated(rrtl_hip_engine* e) { return guard(e, [](rrtl::Engine& x) { x.tick_from_evaluated(); }); }
rrtl_status rrtl_hip_tick_many(rrtl_hip_engine* e, uint64_t n) { return guard(e, [&](rrtl::Engine& x) { x.tick_many(n); }); }
rrtl_status rrtl_hip_snapshot(rrtl_hip_engine* e, uint32_t* v, uint32_t* m) {
  return guard(e, [&](rrtl::Engine& x) {
    if (!v || (!m && x.program().total_memory_words)) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.snapshot(v, m);
  });
}
rrtl_status rrtl_hip_restore(rrtl_hip_engine* e, const uint32_t* v, const uint32_t* m) {
  return guard(e, [&](rrtl::Engine& x) {
    if (!v || (!m && x.program().total_memory_words)) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.restore(v, m);
  });
}
rrtl_status rrtl_hip_set_force(rrtl_hip_engine* e, uint32_t s, uint32_t lane, int value) {
  return guard(e, [&](rrtl::Engine& x) { x.set_force(s, lane, value); });
}
rrtl_status rrtl_hip_set_forces(rrtl_hip_engine* e, const uint32_t* s, const uint32_t* lane, const uint8_t* value, size_t n) {
  return guard(e, [&](rrtl::Engine& x) {
    if (n && (!s || !lane || !value)) throw rrtl::Error("E_HIP_NULL: null buffer");
    x.set_forces(s, lane, value, n);
  });
}
rrtl_status rrtl_hip_clear_forces(rrtl_hip_engine* e) { return guard(e, [](rrtl::Engine& x) { x.clear_forces(); }); }
rrtl_status rrtl_hip_synchronize(rrtl_hip_engine* e) { return guard(e, [](rrtl::Engine& x) { x.synchronize(); }); }
rrtl_status rrtl_hip_run(rrtl_hip_engine* e, const rrtl_hip_run_spec* spec) {
  return guard(e, [&](rrtl::Engine& x) { x.run(to_spec(spec, x.lanes())); });
}
rrtl_status rrtl_hip_last_run_stats(const rrtl_hip_engine* e, rrtl_hip_run_stats* out) {
  if (!e || !e->impl || !out) return RRTL_E_INVALID;
  const rrtl::RunStats& s = e->impl->last_stats();
  out->kernel_ms = s.kernel_ms; out->launches = s.launches; out->lane_cycles = s.lane_cycles;
  return RRTL_OK;
}
size_t rrtl_hip_digest_bytes(const rrtl_hip_engine* e) { return e && e->impl ? e->impl->digest_bytes() : 0; }
rrtl_status rrtl_hip_get_digest(rrtl_hip_engine* e, void* out) {
  return guard(e, [&](rrtl::Engine& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_digest(out); });
}
size_t rrtl_hip_trace_bytes(const rrtl_hip_engine* e) { return e && e->impl ? e->impl->trace_bytes() : 0; }
rrtl_status rrtl_hip_get_trace(rrtl_hip_engine* e, void* out) {
  return guard(e, [&](rrtl::Engine& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_trace(out); });
}
rrtl_status rrtl_hip_get_halt_cycles(rrtl_hip_engine* e, uint64_t* out) {
  return guard(e, [&](rrtl::Engine& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_halt_cycles(out); });
}
rrtl_status rrtl_hip_get_detect(rrtl_hip_engine* e, uint32_t* out) {
  return guard(e, [&](rrtl::Engine& x) { if (!out) throw rrtl::Error("E_HIP_NULL: null buffer"); x.get_detect(out); });
}
namespace {
// Trace replay as the reference's default SimBackend::replay_trace
// (crates/rrtl-sim-ir/src/lib.rs:5270-5345): per step the inputs are set raw,
// comb is evaluated, every active check is read back and compared per lane,
// and tick_from_evaluated_no_post_eval advances; mismatches are counted and
// the first max_mismatches kept in (step, check, lane) order.  The gate engine
// runs it through its API calls (each step a few launches and copies); unlike
// the reference's generic loop it also takes per-lane (independent) inputs on
// any number of lanes, as the word engine's on-chip replay does.
void replay_host_loop(rrtl::Engine& x, const rrtl_hip_replay_spec* sp, uint64_t* mismatch_count,
                      rrtl_hip_replay_mismatch* mismatches, uint32_t* n_mismatches) {
  const rrtl::Program& P = x.program();
  const bool indep = sp->lane_mode == RRTL_REPLAY_INDEPENDENT;
  const uint64_t lanes = x.lanes();
  const uint64_t check_lanes = (!indep && sp->check_mode == RRTL_REPLAY_LANE0) ? 1 : lanes;
  uint32_t in_limbs = 0, check_limbs = 0;
  for (uint32_t i = 0; i < sp->n_inputs; i++) {
    if (sp->input_signal[i] >= P.signals.size()) throw rrtl::Error("E_SIM_IR_REPLAY_SIGNAL: input signal out of range");
    in_limbs += P.signals[sp->input_signal[i]].limbs;
  }
  for (uint32_t j = 0; j < sp->n_checks; j++) {
    if (sp->check_signal[j] >= P.signals.size()) throw rrtl::Error("E_SIM_IR_REPLAY_SIGNAL: check signal out of range");
    if (P.signals[sp->check_signal[j]].limbs > 4) throw rrtl::Error("E_HIP_WIDE: checked signals are limited to 128 bits");
    check_limbs += P.signals[sp->check_signal[j]].limbs;
  }
  const uint32_t check_words = (sp->n_checks + 31) / 32;
  std::vector<uint32_t> buf, got;
  uint64_t count = 0;
  uint32_t kept = 0;
  for (uint64_t step = 0; step < sp->steps; step++) {
    uint32_t off = 0;
    for (uint32_t i = 0; i < sp->n_inputs; i++) {
      const uint32_t lb = P.signals[sp->input_signal[i]].limbs;
      buf.assign(size_t(lb) * lanes, 0u);
      for (uint32_t l = 0; l < lb; l++)
        for (uint64_t ln = 0; ln < lanes; ln++)
          buf[l * lanes + ln] = indep ? sp->input_values[((step * in_limbs) + off + l) * lanes + ln]
                                      : sp->input_values[step * in_limbs + off + l];
      x.set_signal_limbs(sp->input_signal[i], buf.data(), false);
      off += lb;
    }
    x.eval_combinational();
    off = 0;
    for (uint32_t j = 0; j < sp->n_checks; j++) {
      const uint32_t sig = sp->check_signal[j], lb = P.signals[sig].limbs;
      const bool active = !sp->check_active || ((sp->check_active[step * check_words + j / 32] >> (j % 32)) & 1u);
      if (active) {
        got.assign(size_t(lb) * lanes, 0u);
        x.get_signal_limbs(sig, got.data());
        for (uint64_t ln = 0; ln < check_lanes; ln++) {
          bool same = true;
          uint32_t exp[4] = {0, 0, 0, 0}, act[4] = {0, 0, 0, 0};
          for (uint32_t l = 0; l < lb; l++) {
            exp[l] = indep ? sp->expected[((step * check_limbs) + off + l) * lanes + ln]
                           : sp->expected[step * check_limbs + off + l];
            act[l] = got[l * lanes + ln];
            same = same && exp[l] == act[l];
          }
          if (same) continue;
          count++;
          if (kept < sp->max_mismatches) {
            rrtl_hip_replay_mismatch& o = mismatches[kept++];
            o.step = step;
            o.check_index = sp->check_index ? sp->check_index[j] : j;
            o.lane = uint32_t(ln);
            for (uint32_t l = 0; l < 4; l++) { o.expected[l] = exp[l]; o.actual[l] = act[l]; }
          }
        }
      }
      off += lb;
    }
    x.tick_from_evaluated();
  }
  *mismatch_count = count;
  *n_mismatches = kept;
}
}  // namespace

rrtl_status rrtl_hip_replay(rrtl_hip_engine* e, const rrtl_hip_replay_spec* sp, uint64_t* mismatch_count,
                            rrtl_hip_replay_mismatch* mismatches, uint32_t* n_mismatches) {
  return guard(e, [&](rrtl::Engine& x) {
    if (!sp || sp->struct_size < sizeof(rrtl_hip_replay_spec)) throw rrtl::Error("E_HIP_RUN_SPEC: replay spec is null or too small");
    if (!mismatch_count || !n_mismatches || (sp->max_mismatches && !mismatches))
      throw rrtl::Error("E_HIP_NULL: null report buffer");
    if (sp->lane_mode > RRTL_REPLAY_INDEPENDENT || sp->check_mode > RRTL_REPLAY_ALL_LANES)
      throw rrtl::Error("E_HIP_RUN_SPEC: unknown replay lane or check mode");
    if ((sp->n_inputs && (!sp->input_signal || !sp->input_values)) || (sp->n_checks && (!sp->check_signal || !sp->expected)))
      throw rrtl::Error("E_HIP_NULL: null input or check array");
    if (x.kind() == RRTL_HIP_GATE) {
      replay_host_loop(x, sp, mismatch_count, mismatches, n_mismatches);
      return;
    }
    const rrtl::Program& P = x.program();
    const bool indep = sp->lane_mode == RRTL_REPLAY_INDEPENDENT;
    rrtl::RunSpec r;
    r.cycles = sp->steps;
    for (uint32_t i = 0; i < sp->n_inputs; i++) {
      if (sp->input_signal[i] >= P.signals.size()) throw rrtl::Error("E_SIM_IR_REPLAY_SIGNAL: input signal out of range");
      r.in_sig.push_back(sp->input_signal[i]);
      r.in_mode.push_back(indep ? rrtl::IN_TRACE : rrtl::IN_TRACE_UNIFORM);
      r.in_arg.push_back(0);
    }
    if (indep) r.trace_inputs = sp->input_values; else r.trace_uniform = sp->input_values;
    std::vector<uint32_t> limb_at;  // first limb of each check in a step's expected row
    uint32_t check_limbs = 0;
    for (uint32_t j = 0; j < sp->n_checks; j++) {
      if (sp->check_signal[j] >= P.signals.size()) throw rrtl::Error("E_SIM_IR_REPLAY_SIGNAL: check signal out of range");
      r.obs_sig.push_back(sp->check_signal[j]);
      limb_at.push_back(check_limbs);
      check_limbs += P.signals[sp->check_signal[j]].limbs;
    }
    static const uint32_t kNoChecks = 0;
    r.expect = sp->n_checks ? sp->expected : &kNoChecks;
    r.expect_per_lane = indep;
    r.check_active = sp->check_active;
    r.check_lane0 = !indep && sp->check_mode == RRTL_REPLAY_LANE0;
    r.max_mismatches = sp->max_mismatches;
    r.final_eval = false;
    x.run(r);
    std::vector<rrtl::Mismatch> mm;
    *mismatch_count = x.get_mismatches(sp->max_mismatches, mm);
    const uint64_t lanes = x.lanes();
    for (size_t k = 0; k < mm.size(); k++) {
      rrtl_hip_replay_mismatch& o = mismatches[k];
      const rrtl::Mismatch& m = mm[k];
      o.step = m.step;
      o.check_index = sp->check_index ? sp->check_index[m.check] : m.check;
      o.lane = m.lane;
      const uint32_t limbs = P.signals[sp->check_signal[m.check]].limbs;
      for (uint32_t l = 0; l < 4; l++) {
        const uint64_t at = m.step * check_limbs + limb_at[m.check] + l;
        o.expected[l] = l >= limbs ? 0u : indep ? sp->expected[at * lanes + m.lane] : sp->expected[at];
        o.actual[l] = m.actual[l];
      }
    }
    *n_mismatches = uint32_t(mm.size());
  });
}
rrtl_status rrtl_hip_gate_tile_info(rrtl_hip_engine* e, uint32_t* out4) {
  return guard(e, [&](rrtl::Engine& x) {
    if (!out4) throw rrtl::Error("E_HIP_NULL: null buffer");
    if (x.kind() != RRTL_HIP_GATE) throw rrtl::Error("E_HIP_UNSUPPORTED: not a gate engine");
    rrtl::gate_engine_tile_info(&x, out4);
  });
}

rrtl_status rrtl_hip_word_launch_info(rrtl_hip_engine* e, uint32_t* out4) {
  return guard(e, [&](rrtl::Engine& x) {
    if (!out4) throw rrtl::Error("E_HIP_NULL: null buffer");
    if (x.kind() != RRTL_HIP_WORD) throw rrtl::Error("E_HIP_UNSUPPORTED: not a word engine");
    rrtl::word_engine_launch_info(&x, out4);
  });
}

rrtl_status rrtl_hip_batch_create(rrtl_hip_engine* const* engines, const rrtl_hip_run_spec* specs, uint32_t n,
                                  rrtl_hip_batch** out) {
  if (!out) return RRTL_E_INVALID;
  *out = nullptr;
  try {
    if (!engines || !specs || n == 0) throw rrtl::Error("E_HIP_BATCH: empty batch");
    std::vector<rrtl::Engine*> es;
    std::vector<rrtl::RunSpec> ss;
    for (uint32_t i = 0; i < n; i++) {
      if (!engines[i] || !engines[i]->impl) throw rrtl::Error("E_HIP_NULL: null engine in batch");
      es.push_back(engines[i]->impl.get());
      ss.push_back(to_spec(&specs[i], engines[i]->impl->lanes()));
    }
    auto* b = new rrtl_hip_batch;
    try { b->impl = rrtl::make_batch(std::move(es), std::move(ss)); } catch (...) { delete b; throw; }
    *out = b;
    g_create_err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    g_create_err = ex.what();
    return classify(g_create_err);
  }
}
rrtl_status rrtl_hip_batch_run(rrtl_hip_batch* b, uint32_t mode) {
  if (!b || !b->impl) return RRTL_E_INVALID;
  try {
    if (mode > RRTL_BATCH_STREAMS) throw rrtl::Error("E_HIP_BATCH: unknown batch mode");
    rrtl::batch_run(b->impl, mode);
    b->err.clear();
    return RRTL_OK;
  } catch (const std::exception& ex) {
    b->err = ex.what();
    return classify(b->err);
  }
}
rrtl_status rrtl_hip_batch_stats(const rrtl_hip_batch* b, rrtl_hip_run_stats* out) {
  if (!b || !b->impl || !out) return RRTL_E_INVALID;
  const rrtl::RunStats& s = rrtl::batch_stats(b->impl);
  out->kernel_ms = s.kernel_ms; out->launches = s.launches; out->lane_cycles = s.lane_cycles;
  return RRTL_OK;
}
rrtl_status rrtl_hip_batch_describe(const rrtl_hip_batch* b, char* out, size_t cap, size_t* len) {
  if (!b || !b->impl || !len || (!out && cap)) return RRTL_E_INVALID;
  const std::string t = rrtl::batch_describe(b->impl);
  *len = t.size();
  if (out && cap) {
    const size_t n = std::min(cap - 1, t.size());
    std::memcpy(out, t.data(), n);
    out[n] = 0;
  }
  return RRTL_OK;
}
const char* rrtl_hip_batch_last_error(const rrtl_hip_batch* b) { return b ? b->err.c_str() : "E_HIP_NULL: null batch"; }
void rrtl_hip_batch_destroy(rrtl_hip_batch* b) { delete b; }

}  // extern "C"
