// Packed machine program: parser for the versioned blob handed across the C ABI.
//
// The in-memory form mirrors the reference's PackedMachineProgram
// (crates/rrtl-sim-ir/src/lib.rs:67-323): signals with word-major u32-limb
// layouts, memories, and four blocks (async_reset_comb, comb, tick_next,
// tick_commit) of packets, each an SSA instruction list followed by effects.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <stdexcept>
#include "device_defs.hpp"

namespace rrtl {

// Op, EffectKind, SignalKind, BlockId: device_defs.hpp

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

inline uint32_t limbs_of(uint32_t w) { uint32_t l = (w + 31) / 32; return l ? l : 1; }
inline uint32_t final_mask(uint32_t w) { uint32_t r = w % 32; return r ? ((1u << r) - 1) : 0xFFFFFFFFu; }

struct Signal { uint32_t width, sgn, offset, limbs, kind; std::string name; };
struct Memory { uint32_t width, sgn, depth, addr_width, offset, limbs; std::string name; };
struct Instr { uint32_t op, dst, width, sgn; std::vector<uint32_t> args; };
struct Reset { uint32_t signal, kind, polarity; std::vector<uint32_t> value; };
struct Effect {
  uint32_t kind, dst, value;          // STORE/CAPTURE: dst signal, value id; MEMWRITE: dst = memory
  bool has_reset = false; Reset reset;
  uint32_t enable = 0, addr = 0, data = 0;
};
struct Packet { std::vector<Instr> instrs; std::vector<Effect> effects; };
struct Block { uint32_t n_values = 0; std::vector<Packet> packets; std::vector<uint32_t> vwidth; };

struct Program {
  std::vector<Signal> signals;
  std::vector<Memory> memories;
  Block blocks[4];
  uint32_t total_signal_words = 0, total_memory_words = 0;

  static Program parse(const void* data, size_t nbytes) {
    if (!data || nbytes < 32) throw Error("E_PROGRAM_TRUNCATED: program blob is too short");
    size_t n = nbytes / 4, p = 0;
    std::vector<uint32_t> w(n);
    std::memcpy(w.data(), data, n * 4);
    auto need = [&](size_t k) { if (p + k > n) throw Error("E_PROGRAM_TRUNCATED: program blob ends early"); };
    need(8);
    if (w[0] != 0x4C545252u || w[1] != 0x31475250u) throw Error("E_PROGRAM_MAGIC: not an RRTL packed program blob");
    if (w[2] != 1) throw Error("E_PROGRAM_VERSION: unsupported program blob version");
    Program P;
    uint32_t nsig = w[3], nmem = w[4];
    P.total_signal_words = w[5]; P.total_memory_words = w[6];
    uint32_t name_bytes = w[7];
    p = 8;
    need(size_t(5) * nsig);
    P.signals.resize(nsig);
    for (auto& s : P.signals) {
      s.width = w[p]; s.sgn = w[p + 1]; s.offset = w[p + 2]; s.limbs = w[p + 3]; s.kind = w[p + 4]; p += 5;
      if (s.width == 0 || s.limbs != limbs_of(s.width) || uint64_t(s.offset) + s.limbs > P.total_signal_words)
        throw Error("E_PROGRAM_SIGNAL: inconsistent signal layout");
    }
    need(size_t(6) * nmem);
    P.memories.resize(nmem);
    for (auto& m : P.memories) {
      m.width = w[p]; m.sgn = w[p + 1]; m.depth = w[p + 2]; m.addr_width = w[p + 3]; m.offset = w[p + 4]; m.limbs = w[p + 5]; p += 6;
      if (m.width == 0 || m.limbs != limbs_of(m.width) || uint64_t(m.offset) + uint64_t(m.depth) * m.limbs > P.total_memory_words)
        throw Error("E_PROGRAM_MEMORY: inconsistent memory layout");
    }
    for (int b = 0; b < 4; b++) {
      Block& bk = P.blocks[b];
      need(2);
      bk.n_values = w[p]; uint32_t npk = w[p + 1]; p += 2;
      bk.vwidth.assign(bk.n_values, 0);
      bk.packets.resize(npk);
      for (auto& pk : bk.packets) {
        need(2);
        uint32_t ni = w[p], ne = w[p + 1]; p += 2;
        pk.instrs.resize(ni);
        for (auto& in : pk.instrs) {
          need(5);
          in.op = w[p]; in.dst = w[p + 1]; in.width = w[p + 2]; in.sgn = w[p + 3]; uint32_t na = w[p + 4]; p += 5;
          need(na);
          in.args.assign(w.begin() + p, w.begin() + p + na); p += na;
          if (in.op >= OP_COUNT || in.dst >= bk.n_values || in.width == 0) throw Error("E_PROGRAM_INSTR: malformed instruction");
          bk.vwidth[in.dst] = in.width;
        }
        pk.effects.resize(ne);
        for (auto& ef : pk.effects) {
          need(1);
          ef.kind = w[p++];
          if (ef.kind == EF_STORE) { need(2); ef.dst = w[p]; ef.value = w[p + 1]; p += 2; }
          else if (ef.kind == EF_CAPTURE) {
            need(7);
            ef.dst = w[p]; ef.value = w[p + 1]; ef.has_reset = w[p + 2] != 0;
            ef.reset.signal = w[p + 3]; ef.reset.kind = w[p + 4]; ef.reset.polarity = w[p + 5];
            uint32_t nl = w[p + 6]; p += 7; need(nl);
            ef.reset.value.assign(w.begin() + p, w.begin() + p + nl); p += nl;
          } else if (ef.kind == EF_MEMWRITE) {
            need(4); ef.dst = w[p]; ef.enable = w[p + 1]; ef.addr = w[p + 2]; ef.data = w[p + 3]; p += 4;
          } else throw Error("E_PROGRAM_EFFECT: unknown effect kind");
        }
      }
    }
    P.validate_refs();
    // names (optional)
    size_t nb = std::min<size_t>(name_bytes, nbytes - p * 4);
    const char* names = static_cast<const char*>(data) + p * 4;
    size_t at = 0;
    for (size_t i = 0; i < P.signals.size() + P.memories.size() && at < nb; i++) {
      size_t len = strnlen(names + at, nb - at);
      std::string s(names + at, len);
      if (i < P.signals.size()) P.signals[i].name = s; else P.memories[i - P.signals.size()].name = s;
      at += len + 1;
    }
    return P;
  }

  // Every value is defined before use; every signal/memory index is in range.
  void validate_refs() const {
    for (const Block& bk : blocks) {
      std::vector<uint8_t> defined(bk.n_values, 0);
      auto use = [&](uint32_t v) {
        if (v >= bk.n_values || !defined[v]) throw Error("E_SIM_IR_MACHINE_USE_BEFORE_DEF: machine value used before it is defined");
      };
      for (const Packet& pk : bk.packets) {
        for (const Instr& in : pk.instrs) {
          switch (in.op) {
            case OP_LIT: break;
            case OP_SIGNAL: if (in.args.size() != 1 || in.args[0] >= signals.size()) throw Error("E_PROGRAM_SIGNAL_REF: bad signal index"); break;
            case OP_NOT: case OP_ZEXT: case OP_SEXT: case OP_TRUNC: case OP_CAST:
              if (in.args.size() != 1) throw Error("E_PROGRAM_INSTR: arity"); use(in.args[0]); break;
            case OP_SLICE: if (in.args.size() != 2) throw Error("E_PROGRAM_INSTR: arity"); use(in.args[0]); break;
            case OP_LT: if (in.args.size() != 3) throw Error("E_PROGRAM_INSTR: arity"); use(in.args[0]); use(in.args[1]); break;
            case OP_MUX: if (in.args.size() != 3) throw Error("E_PROGRAM_INSTR: arity"); for (auto a : in.args) use(a); break;
            case OP_CONCAT: if (in.args.empty()) throw Error("E_PROGRAM_INSTR: arity"); for (auto a : in.args) use(a); break;
            case OP_MEMREAD:
              if (in.args.size() != 2 || in.args[0] >= memories.size()) throw Error("E_PROGRAM_MEMORY_REF: bad memory index");
              use(in.args[1]); break;
            default: if (in.args.size() != 2) throw Error("E_PROGRAM_INSTR: arity"); use(in.args[0]); use(in.args[1]); break;
          }
        }
        for (const Instr& in : pk.instrs) defined[in.dst] = 1;
        for (const Effect& ef : pk.effects) {
          if (ef.kind == EF_MEMWRITE) {
            if (ef.dst >= memories.size()) throw Error("E_PROGRAM_MEMORY_REF: bad memory index");
            use(ef.enable); use(ef.addr); use(ef.data);
          } else {
            if (ef.dst >= signals.size()) throw Error("E_PROGRAM_SIGNAL_REF: bad signal index");
            use(ef.value);
            if (ef.has_reset && ef.reset.signal >= signals.size()) throw Error("E_PROGRAM_SIGNAL_REF: bad reset signal");
          }
        }
      }
    }
  }

  bool is_gate_level(std::string* why = nullptr) const {
    for (const auto& s : signals) if (s.width != 1) { if (why) *why = "signal `" + s.name + "` is wider than 1 bit"; return false; }
    if (!memories.empty()) { if (why) *why = "memories are not supported by the bit-sliced gate engine"; return false; }
    for (const Block& bk : blocks)
      for (const Packet& pk : bk.packets) {
        for (const Instr& in : pk.instrs)
          if (!(in.op == OP_AND || in.op == OP_OR || in.op == OP_XOR || in.op == OP_NOT || in.op == OP_MUX ||
                in.op == OP_SIGNAL || in.op == OP_LIT)) { if (why) *why = "op is not a gate-level bitwise op"; return false; }
        for (const Effect& ef : pk.effects) if (ef.kind == EF_MEMWRITE) { if (why) *why = "memory writes are not supported"; return false; }
      }
    return true;
  }
};

}  // namespace rrtl
