// Word-level engine: one lane per thread over u32 limbs, the MI355X counterpart
// of the reference's per-lane machine semantics (PackedSimulator,
// crates/rrtl-sim-ir/src/lib.rs:9033-10255; the host batch engine
// SimdCpuSimulator :5375-7093 computes the same thing instruction-outer).
//
// Semantics: eval_combinational = Comb then AsyncResetComb (:9278-9281);
// tick_from_evaluated_no_post_eval = capture tick_next (reset value when the
// reset is asserted), run tick_commit (memory writes from pre-commit state),
// commit (:9613-9703); tick_many(n) = eval, then n x {edge, eval} (:9295-9307).
// Instruction results follow eval_machine_instr_lane (:9796-9907) and the limb
// helpers (:12942-13196): masked to width, ripple add/sub, truncated schoolbook
// multiply, Eq/Ne at the wider operand, Lt at the lhs width (signed by sign bit),
// Mux on bit 0, MSB-first Concat, out-of-range memory reads give 0.
//
// Design as data: the host compiles each block into fixed 8-word records read
// with uniform (scalar) loads.  Operands name one of three spaces: per-lane
// scratch slots (SSA values, reused by liveness), per-lane signal storage in
// HBM (word-major (offset+limb)*lanes + lane, the reference's storage layout),
// or a uniform constant pool.  Scratch lives in LDS when a workgroup's share
// fits, else in HBM laid out slot-major so every access is coalesced.  One
// launch runs a whole multi-cycle run: stimulus from a counter hash or a
// caller trace, observations into a per-lane digest or a trace, per-lane halt.
#include "engine.hpp"
#include "word_device.hpp"
#include "rtc.hpp"
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <numeric>
#include <tuple>
#include <cstring>
#include <functional>
#include <map>
#include <unordered_set>
#include <unordered_map>

namespace rrtl {
namespace {

using namespace word;

struct WordProgram {
  std::vector<Rec> blk[4];         // async, comb, next (captures), commit (mem writes + reg commits)
  std::vector<uint32_t> cpool;
  uint32_t scratch_slots = 0;
  std::vector<uint32_t> mem_table; // per memory: offset, depth, limbs, width
};

struct Operand { uint32_t code; uint32_t limbs; uint32_t width; };

class WordCompiler {
 public:
  // macros: cones of single-reader one-limb operations become macro records
  // (interpreter only; replaces the fusion menu, as the reference runs
  // fuse_macros on a menu-off encoding, interp.rs:691-695)
  explicit WordCompiler(const Program& P, bool macros = false) : P(P), macros_(macros) {
    fuse_ = !macros;
    for (auto& m : P.memories) out.mem_table.insert(out.mem_table.end(), {m.offset, m.depth, m.limbs, m.width});
  }
  WordProgram compile() {
    // staging for captures lives after every block's value slots
    lower(B_ASYNC); lower(B_COMB); lower(B_NEXT); lower(B_COMMIT);
    uint32_t stage_base = max_slot_;
    uint32_t stage = stage_base;
    for (Rec& r : out.blk[B_NEXT])
      if ((r.w[0] & 0xFF) == R_CAPTURE) { r.w[1] = stage; stage += (r.w[0] >> 8) & 0xFF; }
    for (Rec& r : out.blk[B_COMMIT])
      if ((r.w[0] & 0xFF) == R_COMMIT) r.w[2] = SP_SCRATCH | (stage_base + r.w[2]);
    out.scratch_slots = std::max<uint32_t>(1, stage);
    return std::move(out);
  }

 private:
  uint32_t const_of(const std::vector<uint32_t>& limbs, uint32_t width) {
    uint32_t n = limbs_of(width);
    std::vector<uint32_t> v(n, 0);
    for (uint32_t l = 0; l < n && l < limbs.size(); l++) v[l] = limbs[l];
    v[n - 1] &= final_mask(width);
    auto key = std::make_pair(width, v);
    auto it = cmemo_.find(key);
    if (it != cmemo_.end()) return it->second;
    uint32_t at = uint32_t(out.cpool.size());
    out.cpool.insert(out.cpool.end(), v.begin(), v.end());
    cmemo_[key] = at;
    return at;
  }

  // Fusion plan of a block (the reference interpreter's menu, interp.rs:393-495,
  // Not-rooted cells first): a one-limb AND/OR/XOR/ADD/MUL (or an already fused
  // ANDOR/ORAND under a NOT) whose value has exactly one reader, an operation
  // of the same width, is folded into that reader's record and gets no slot.
  // Equal widths keep the result exact: the dropped inner mask only clears
  // bits the reader's own mask clears, and carries and products only move
  // bits upward.
  struct Fused { uint32_t fop, a, b, c, d = NONE; };
  void plan_fusion(const Block& bk, std::vector<uint8_t>& absorbed, std::unordered_map<uint32_t, Fused>& fused) {
    std::vector<const Instr*> def(bk.n_values, nullptr);
    std::vector<uint32_t> uses(bk.n_values, 0);
    for (const Packet& pk : bk.packets) {
      for (const Instr& in : pk.instrs) {
        def[in.dst] = &in;
        if (in.op == OP_LIT || in.op == OP_SIGNAL) continue;
        if (in.op == OP_MEMREAD) { uses[in.args[1]]++; continue; }
        const size_t na = in.op == OP_SLICE ? 1 : in.op == OP_LT ? 2 : in.args.size();
        for (size_t k = 0; k < na; k++) uses[in.args[k]]++;
      }
      for (const Effect& ef : pk.effects) {
        if (ef.kind == EF_MEMWRITE) { uses[ef.enable]++; uses[ef.addr]++; uses[ef.data]++; }
        else uses[ef.value]++;
      }
    }
    absorbed.assign(bk.n_values, 0);
    auto one_limb = [&](const Instr& in) { return limbs_of(in.width) == 1; };
    // a plain inner value of operation `op` that `u` may absorb
    auto plain = [&](uint32_t v, uint32_t op, const Instr& u) {
      const Instr* d = def[v];
      return d && d->op == op && uses[v] == 1 && !absorbed[v] && !fused.count(v) && d->width == u.width && one_limb(*d);
    };
    for (const Packet& pk : bk.packets)
      for (const Instr& u : pk.instrs) {
        if (!one_limb(u)) continue;
        const uint32_t d = u.dst;
        if (u.op == OP_NOT) {
          const uint32_t x = u.args[0];
          auto it = fused.find(x);
          if (it != fused.end() && uses[x] == 1 && def[x]->width == u.width &&
              (it->second.fop == FOP_ANDOR || it->second.fop == FOP_ORAND)) {
            // OR(AND, c) / AND(OR, c) under a NOT: AOI22 / OAI22 when c is itself a
            // single-reader AND / OR, else AOI21 / OAI21
            const bool aoi = it->second.fop == FOP_ANDOR;
            const uint32_t c = it->second.c;
            if (plain(c, aoi ? uint32_t(OP_AND) : uint32_t(OP_OR), u)) {
              fused[d] = {aoi ? uint32_t(FOP_AOI22) : uint32_t(FOP_OAI22), it->second.a, it->second.b,
                          def[c]->args[0], def[c]->args[1]};
              absorbed[c] = 1;
            } else {
              fused[d] = {aoi ? uint32_t(FOP_AOI21) : uint32_t(FOP_OAI21), it->second.a, it->second.b, c};
            }
            fused.erase(it);
            absorbed[x] = 1;
            continue;
          }
          for (uint32_t op : {uint32_t(OP_AND), uint32_t(OP_OR), uint32_t(OP_XOR)})
            if (plain(x, op, u)) {
              fused[d] = {op == OP_AND ? uint32_t(FOP_NAND) : op == OP_OR ? uint32_t(FOP_NOR) : uint32_t(FOP_XNOR),
                          def[x]->args[0], def[x]->args[1], NONE};
              absorbed[x] = 1;
              break;
            }
          continue;
        }
        // (reader, inner) pairs: OR(AND), AND(OR), XOR(XOR), ADD(ADD), ADD(MUL),
        // SUB(MUL) with the product as the minuend (SUB does not commute)
        static const uint32_t pairs[6][3] = {{OP_OR, OP_AND, FOP_ANDOR}, {OP_AND, OP_OR, FOP_ORAND},
                                             {OP_XOR, OP_XOR, FOP_XOR3}, {OP_ADD, OP_ADD, FOP_ADD3},
                                             {OP_ADD, OP_MUL, FOP_MULADD}, {OP_SUB, OP_MUL, FOP_MULSUB}};
        bool done = false;
        for (const auto& pr : pairs) {
          if (done || u.op != pr[0]) continue;
          for (int k = 0; k < (u.op == OP_SUB ? 1 : 2) && !done; k++) {
            const uint32_t x = u.args[k], other = u.args[1 - k];
            if (x == other || !plain(x, pr[1], u)) continue;
            fused[d] = {pr[2], def[x]->args[0], def[x]->args[1], other};
            absorbed[x] = 1;
            done = true;
          }
        }
      }
  }

  // Operand fields of the record of `in` (every operation but CONCAT and
  // MEMREAD, which keep lists and memory indices): codes in w[2..4], limb
  // counts in w[7], the width an operation reads at in w[6].
  template <class F>
  static void fill_operands(const Instr& in, Rec& r, F&& opnd) {
    switch (in.op) {
      case OP_NOT: case OP_ZEXT: case OP_TRUNC: case OP_CAST: case OP_SEXT:
        r.w[2] = opnd(in.args[0]).code; r.w[7] = opnd(in.args[0]).limbs; r.w[6] = opnd(in.args[0]).width; break;
      case OP_SLICE:
        r.w[2] = opnd(in.args[0]).code; r.w[7] = opnd(in.args[0]).limbs; r.w[6] = in.args[1]; break;
      case OP_LT:
        r.w[2] = opnd(in.args[0]).code; r.w[3] = opnd(in.args[1]).code;
        r.w[7] = opnd(in.args[0]).limbs | (opnd(in.args[1]).limbs << 8) | (in.args[2] ? 1u << 24 : 0);
        r.w[6] = opnd(in.args[0]).width; break;
      case OP_MUX:
        r.w[2] = opnd(in.args[1]).code; r.w[3] = opnd(in.args[2]).code; r.w[4] = opnd(in.args[0]).code;
        r.w[7] = opnd(in.args[1]).limbs | (opnd(in.args[2]).limbs << 8); break;
      default:  // binary
        r.w[2] = opnd(in.args[0]).code; r.w[3] = opnd(in.args[1]).code;
        r.w[7] = opnd(in.args[0]).limbs | (opnd(in.args[1]).limbs << 8);
        r.w[6] = std::max(opnd(in.args[0]).width, opnd(in.args[1]).width);
        break;
    }
  }

  // Operands a one-limb record of `in` reads, in record order.
  static std::vector<uint32_t> rec_args(const Instr& in) {
    switch (in.op) {
      case OP_NOT: case OP_ZEXT: case OP_TRUNC: case OP_CAST: case OP_SEXT: case OP_SLICE: return {in.args[0]};
      case OP_MUX: return {in.args[1], in.args[2], in.args[0]};
      default: return {in.args[0], in.args[1]};
    }
  }
  static bool macro_op(uint32_t op) {
    switch (op) {
      case OP_NOT: case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: case OP_EQ:
      case OP_NE: case OP_LT: case OP_MUX: case OP_SLICE: case OP_ZEXT: case OP_SEXT: case OP_TRUNC: case OP_CAST:
        return true;
      default: return false;
    }
  }

  // Macro plan of a block: walking back from the last operation, every
  // one-limb operation (of one-limb operands) not yet absorbed roots a cone
  // that takes in each operand produced by such an operation with exactly one
  // reader, depth first, up to MACRO_REGS operations.  The cone runs at the
  // root's position, so an operand is taken in only if no store between its
  // producer and the root writes a signal the producer reads; operands from
  // outside the cone stay live until the root (ext_last).
  void plan_macros(const Block& bk, int b, std::vector<uint8_t>& absorbed,
                   std::unordered_map<uint32_t, std::vector<uint32_t>>& macros,
                   std::unordered_map<uint32_t, int64_t>& ext_last) {
    std::vector<const Instr*> def(bk.n_values, nullptr);
    std::vector<uint32_t> uses(bk.n_values, 0);
    std::vector<const Instr*> reader(bk.n_values, nullptr);
    std::vector<int64_t> at(bk.n_values, -1);
    std::unordered_map<uint32_t, std::vector<int64_t>> stores;  // signal -> positions of stores to it
    std::vector<const Instr*> order;
    int64_t pos = 0;
    for (const Packet& pk : bk.packets) {
      for (const Instr& in : pk.instrs) {
        def[in.dst] = &in;
        at[in.dst] = pos++;
        order.push_back(&in);
        if (in.op == OP_LIT || in.op == OP_SIGNAL) continue;
        if (in.op == OP_MEMREAD) { uses[in.args[1]]++; continue; }
        const size_t na = in.op == OP_SLICE ? 1 : in.op == OP_LT ? 2 : in.args.size();
        for (size_t k = 0; k < na; k++) { uses[in.args[k]]++; reader[in.args[k]] = &in; }
      }
      for (const Effect& ef : pk.effects) {
        if (ef.kind == EF_MEMWRITE) { uses[ef.enable] += 2; uses[ef.addr] += 2; uses[ef.data] += 2; }
        else uses[ef.value] += 2;  // a value an effect reads is never taken into a cone
        const bool writes = (ef.kind == EF_STORE && b != B_NEXT) ||
                            (ef.kind == EF_CAPTURE && b == B_ASYNC && ef.has_reset && ef.reset.kind == 1);
        if (writes) stores[ef.dst].push_back(pos);
        pos++;
      }
    }
    auto one_limb = [&](uint32_t v) { return def[v] && limbs_of(def[v]->width) == 1; };
    auto simple = [&](const Instr& in) {
      if (!macro_op(in.op) || limbs_of(in.width) != 1) return false;
      for (uint32_t a : rec_args(in)) if (!one_limb(a)) return false;
      return true;
    };
    auto hazard_free = [&](uint32_t v, int64_t root_at) {
      for (uint32_t a : rec_args(*def[v])) {
        if (def[a]->op != OP_SIGNAL) continue;
        auto it = stores.find(def[a]->args[0]);
        if (it == stores.end()) continue;
        for (int64_t q : it->second) if (q > at[v] && q < root_at) return false;
      }
      return true;
    };
    for (auto it = order.rbegin(); it != order.rend(); ++it) {
      const Instr& root = **it;
      if (absorbed[root.dst] || !simple(root)) continue;
      const int64_t root_at = at[root.dst];
      std::vector<uint32_t> cone;
      // depth: operations of the current path still waiting to be appended
      std::function<void(uint32_t, uint32_t)> take = [&](uint32_t v, uint32_t depth) {
        for (uint32_t a : rec_args(*def[v])) {
          const bool inl = uses[a] == 1 && reader[a] == def[v] && simple(*def[a]) && !absorbed[a];
          if (inl && cone.size() + depth + 1 <= MACRO_REGS && hazard_free(a, root_at)) take(a, depth + 1);
        }
        cone.push_back(v);
      };
      take(root.dst, 1);
      if (cone.size() < 2) continue;
      for (size_t k = 0; k + 1 < cone.size(); k++) {
        absorbed[cone[k]] = 1;
        for (uint32_t a : rec_args(*def[cone[k]]))
          if (std::find(cone.begin(), cone.end(), a) == cone.end()) {
            auto& e = ext_last[a];
            e = std::max(e, root_at);
          }
      }
      macros[root.dst] = std::move(cone);
    }
  }

  void lower(int b) {
    const Block& bk = P.blocks[b];
    std::vector<Operand> val(bk.n_values, {NONE, 0, 0});
    std::vector<uint8_t> absorbed;
    std::unordered_map<uint32_t, Fused> fused;
    std::unordered_map<uint32_t, std::vector<uint32_t>> macros;
    std::unordered_map<uint32_t, int64_t> ext_last;
    if (fuse_) plan_fusion(bk, absorbed, fused);
    else absorbed.assign(bk.n_values, 0);
    if (macros_) plan_macros(bk, b, absorbed, macros, ext_last);
    std::vector<const Instr*> defs(bk.n_values, nullptr);
    for (const Packet& pk : bk.packets)
      for (const Instr& in : pk.instrs) defs[in.dst] = &in;
    // liveness: last record index that reads each value
    std::vector<int64_t> last_use(bk.n_values, -1);
    int64_t pos = 0;
    for (const Packet& pk : bk.packets) {
      for (const Instr& in : pk.instrs) {
        if (in.op == OP_LIT || in.op == OP_SIGNAL) { pos++; continue; }
        size_t na = in.args.size();
        if (in.op == OP_SLICE || in.op == OP_LT) na = in.op == OP_SLICE ? 1 : 2;
        for (size_t k = 0; k < na; k++) {
          uint32_t a = (in.op == OP_MEMREAD) ? in.args[1] : in.args[k];
          last_use[a] = pos;
          if (in.op == OP_MEMREAD) break;
        }
        // a fused record reads its inner operation's operands
        auto fz = fused.find(in.dst);
        if (fz != fused.end())
          for (uint32_t a : {fz->second.a, fz->second.b, fz->second.c, fz->second.d}) if (a != NONE) last_use[a] = pos;
        pos++;
      }
      for (const Effect& ef : pk.effects) {
        if (ef.kind == EF_MEMWRITE) { last_use[ef.enable] = pos; last_use[ef.addr] = pos; last_use[ef.data] = pos; }
        else last_use[ef.value] = pos;
        pos++;
      }
    }
    for (const auto& e : ext_last) last_use[e.first] = std::max(last_use[e.first], e.second);
    // linear-scan slot allocation over limb ranges
    std::multimap<int64_t, std::pair<uint32_t, uint32_t>> live;  // last use -> (slot, limbs)
    std::vector<std::vector<uint32_t>> free_by_limbs(MAXL + 1);
    uint32_t next_slot = 0;
    auto alloc = [&](uint32_t limbs, int64_t at, int64_t last) -> uint32_t {
      while (!live.empty() && live.begin()->first < at) {
        free_by_limbs[live.begin()->second.second].push_back(live.begin()->second.first);
        live.erase(live.begin());
      }
      uint32_t s;
      if (!free_by_limbs[limbs].empty()) { s = free_by_limbs[limbs].back(); free_by_limbs[limbs].pop_back(); }
      else { s = next_slot; next_slot += limbs; }
      live.emplace(std::max(last, at), std::make_pair(s, limbs));
      return s;
    };
    pos = 0;
    auto& recs = out.blk[b];
    for (const Packet& pk : bk.packets) {
      for (const Instr& in : pk.instrs) {
        const uint32_t L = limbs_of(in.width);
        if (L > MAXL) throw Error("E_HIP_WIDE: values wider than 256 bits are not supported on the GPU");
        if (in.op == OP_LIT) {
          std::vector<uint32_t> lv(in.args.begin(), in.args.end());
          val[in.dst] = {SP_CONST | const_of(lv, in.width), L, in.width};
          pos++; continue;
        }
        if (in.op == OP_SIGNAL) {
          const Signal& s = P.signals[in.args[0]];
          if (s.width == in.width) { val[in.dst] = {SP_SIGNAL | s.offset, s.limbs, s.width}; pos++; continue; }
          uint32_t slot = alloc(L, pos, last_use[in.dst]);
          Rec r{};
          r.w[0] = R_SIGLOAD | (L << 8); r.w[1] = slot; r.w[2] = SP_SIGNAL | s.offset; r.w[5] = in.width;
          r.w[7] = s.limbs;
          recs.push_back(r);
          val[in.dst] = {SP_SCRATCH | slot, L, in.width};
          pos++; continue;
        }
        if (absorbed[in.dst]) { pos++; continue; }  // folded into its one reader's record
        Rec r{};
        r.w[0] = in.op | (L << 8);
        r.w[5] = in.width;
        auto opnd = [&](uint32_t v) -> const Operand& { return val[v]; };
        auto fz = fused.find(in.dst);
        if (fz != fused.end()) {
          const Fused& f = fz->second;
          r.w[0] = f.fop | (1u << 8);
          r.w[2] = opnd(f.a).code; r.w[3] = opnd(f.b).code; r.w[4] = f.c == NONE ? 0u : opnd(f.c).code;
          r.w[7] = opnd(f.a).limbs | (opnd(f.b).limbs << 8);
          r.w[6] = f.d != NONE ? opnd(f.d).code : std::max(opnd(f.a).width, opnd(f.b).width);
          uint32_t slot = alloc(L, pos, last_use[in.dst]);
          r.w[1] = slot;
          recs.push_back(r);
          val[in.dst] = {SP_SCRATCH | slot, L, in.width};
          pos++;
          continue;
        }
        auto mc = macros.find(in.dst);
        if (mc != macros.end()) {
          // the cone's operations as 8-word records in the constant pool,
          // each writing its own register; operands inside the cone read it
          const std::vector<uint32_t>& cone = mc->second;
          std::unordered_map<uint32_t, Operand> local;
          const uint32_t at = uint32_t(out.cpool.size());
          for (uint32_t k = 0; k < cone.size(); k++) {
            const Instr& si = *defs[cone[k]];
            auto sub_opnd = [&](uint32_t v) -> const Operand& {
              auto lt = local.find(v);
              return lt != local.end() ? lt->second : val[v];
            };
            Rec sr{};
            sr.w[0] = si.op | (1u << 8);
            sr.w[5] = si.width;
            fill_operands(si, sr, sub_opnd);
            sr.w[1] = k;
            out.cpool.insert(out.cpool.end(), sr.w, sr.w + 8);
            local[cone[k]] = {SP_LOCAL | k, 1, si.width};
          }
          r.w[0] = R_MACRO | (1u << 8);
          r.w[2] = at; r.w[3] = uint32_t(cone.size());
          uint32_t slot = alloc(L, pos, last_use[in.dst]);
          r.w[1] = slot;
          recs.push_back(r);
          val[in.dst] = {SP_SCRATCH | slot, L, in.width};
          pos++;
          continue;
        }
        switch (in.op) {
          case OP_CONCAT: {
            uint32_t at = uint32_t(out.cpool.size());
            for (uint32_t a : in.args) { out.cpool.push_back(opnd(a).code); out.cpool.push_back(opnd(a).width | (opnd(a).limbs << 16)); }
            r.w[2] = at; r.w[3] = uint32_t(in.args.size()); break;
          }
          case OP_MEMREAD:
            r.w[6] = in.args[0]; r.w[2] = opnd(in.args[1]).code; r.w[7] = opnd(in.args[1]).limbs; break;
          default:
            fill_operands(in, r, opnd);
            break;
        }
        uint32_t slot = alloc(L, pos, last_use[in.dst]);
        r.w[1] = slot;
        recs.push_back(r);
        val[in.dst] = {SP_SCRATCH | slot, L, in.width};
        pos++;
      }
      for (const Effect& ef : pk.effects) {
        Rec r{};
        if (ef.kind == EF_STORE && b != B_NEXT) {
          const Signal& s = P.signals[ef.dst];
          r.w[0] = R_STORE | (s.limbs << 8); r.w[1] = s.offset; r.w[2] = val[ef.value].code;
          r.w[5] = s.width; r.w[7] = val[ef.value].limbs;
          recs.push_back(r);
        } else if (ef.kind == EF_CAPTURE && b == B_ASYNC) {
          if (!ef.has_reset || ef.reset.kind != 1) { pos++; continue; }
          const Signal& s = P.signals[ef.dst];
          r.w[0] = R_ASYNC_CAPTURE | (s.limbs << 8); r.w[1] = s.offset; r.w[2] = val[ef.value].code;
          r.w[3] = SP_SIGNAL | P.signals[ef.reset.signal].offset; r.w[5] = s.width;
          r.w[6] = ef.reset.polarity; r.w[7] = val[ef.value].limbs;
          recs.push_back(r);
        } else if (ef.kind == EF_CAPTURE && b == B_NEXT) {
          const Signal& s = P.signals[ef.dst];
          r.w[0] = R_CAPTURE | (s.limbs << 8); r.w[2] = val[ef.value].code; r.w[5] = s.width;
          r.w[7] = val[ef.value].limbs;
          if (ef.has_reset) {
            r.w[3] = SP_SIGNAL | P.signals[ef.reset.signal].offset;
            r.w[4] = SP_CONST | const_of(ef.reset.value, s.width);
            r.w[6] = ef.reset.polarity | 2u;
          }
          recs.push_back(r);
          Rec c{};
          c.w[0] = R_COMMIT | (s.limbs << 8); c.w[1] = s.offset; c.w[2] = capture_limbs_; c.w[5] = s.width;
          capture_limbs_ += s.limbs;
          commits_.push_back(c);
        } else if (ef.kind == EF_MEMWRITE && b == B_COMMIT) {
          const Memory& m = P.memories[ef.dst];
          r.w[0] = R_MEMWRITE | (m.limbs << 8); r.w[1] = ef.dst; r.w[2] = val[ef.enable].code;
          r.w[3] = val[ef.addr].code; r.w[4] = val[ef.data].code; r.w[5] = m.width;
          r.w[7] = val[ef.addr].limbs | (val[ef.data].limbs << 8);
          recs.push_back(r);
        }
        pos++;
      }
    }
    max_slot_ = std::max(max_slot_, next_slot);
    if (b == B_COMMIT) out.blk[B_COMMIT].insert(out.blk[B_COMMIT].end(), commits_.begin(), commits_.end());
  }

  const Program& P;
  WordProgram out;
  bool macros_ = false;  // macro records (plan_macros)
  bool fuse_ = true;     // fused records (plan_fusion)
  std::map<std::pair<uint32_t, std::vector<uint32_t>>, uint32_t> cmemo_;
  std::vector<Rec> commits_;
  uint32_t capture_limbs_ = 0, max_slot_ = 0;
};


// ------------------------------------------------------------------ generated code
// A design's records rewritten as straight-line device code and compiled for
// gfx950 at run time (the MI355X counterpart of the reference's per-design GPU
// code generator, crates/rrtl-gpu-sim/src/gpu_codegen.rs).  One-limb records
// become one C++ statement each with their value in a register and literals
// inlined; a signal word is loaded once per evaluation and forwarded after a
// store; captures stay in registers until the commit.  Records with results
// wider than one limb keep the interpreter's code: they are emitted as
// exec_rec calls on a literal record, which the compiler specialises, with
// their values in scratch slots.  Memories given an LDS row (mem_row) are
// staged in LDS for the launch -- copied in by enter(), back by leave() -- and
// accessed there.  The executor relies on word_body running
// B_COMB always followed by B_ASYNC and B_NEXT always followed by B_COMMIT: it
// runs each pair as one function (jit_eval, jit_edge).
// Where a generated kernel finds its state in LDS.  With a nonzero stride (the
// launch's workgroup size, WM_ALL_LDS) signal and staged-memory words are
// addressed as constant offsets from the thread's LDS column: no per-access
// address arithmetic.  mem_row: per memory, its first row in the staged-memory
// area, or -1 (HBM).
struct JitLayout {
  std::vector<int32_t> mem_row;
  uint32_t stride = 0, slots = 0, sig_words = 0;
};

class WordJit {
 public:
  explicit WordJit(const WordProgram& WP, JitLayout lay = {}) : WP(WP), lay_(std::move(lay)) {
    mem_row_ = lay_.mem_row;
    mem_row_.resize(WP.mem_table.size() / 4, -1);
    begin(fn_[0]); emit_block(B_COMB); scr_.clear(); emit_block(B_ASYNC);
    begin(fn_[1]); emit_block(B_NEXT); scr_.clear(); emit_block(B_COMMIT);
  }
  // Scratch slots the generated code needs (only records kept in interpreted form use them).
  bool uses_scratch() const { return generic_; }
  uint32_t generic_records() const { return n_generic_; }
  // The evaluations split between `waves` waves of one lane set (WM_ALL_LDS
  // with 64-lane columns); empty plans when the code cannot be split.
  struct Split {
    bool ok = false;
    uint32_t phases = 0, xslots = 0;     // barrier phases, exchange rows
    double est = 0, serial = 0;          // estimated time of the split and of one wave (cost units)
    std::vector<int32_t> wave, phase;    // per statement (-1: a signal load, placed per wave)
    std::vector<std::vector<int32_t>> load_phase;  // [wave][statement]: phase of a signal load in that wave, or -1
    std::vector<int32_t> xslot;          // per statement: exchange row of an exported value, or -1
  };
  // Plans for both evaluations; estimated speed-up over one wave.
  double plan_waves(uint32_t waves) {
    double t1 = 0, tg = 0;
    xslots_ = 0;
    for (int f = 0; f < 2; f++) {
      split_[f] = split(fn_[f], waves);
      if (!split_[f].ok) return 0;
      t1 += split_[f].serial; tg += split_[f].est;
      xslots_ = std::max(xslots_, split_[f].xslots);
    }
    waves_ = waves;
    return tg > 0 ? t1 / tg : 0;
  }
  uint32_t exchange_rows() const { return xslots_; }
  uint32_t phases(int f) const { return split_[f].phases; }
  // The generated evaluations as device code in namespace rrtl::word::<ns>:
  // jit_eval, jit_edge and the executor JitExec that word_body runs them
  // through; after plan_waves(G) they run G waves per 64 lanes, exchanging
  // values through LDS rows from xrow0 on.
  std::string definitions(uint32_t xrow0, const std::string& ns) const {
    const uint32_t G = waves_;
    std::string fn[2];
    for (int f = 0; f < 2; f++) fn[f] = G > 1 ? emit_split(fn_[f], split_[f], G, xrow0) : emit_serial(fn_[f]);
    return "namespace rrtl {\nnamespace word {\nnamespace " + ns + " {\n\n"
           "template <class Cx>\n__device__ __forceinline__ void jit_eval(const Cx& C) {\n" + fn[0] + "}\n\n"
           "template <class Cx>\n__device__ __forceinline__ void jit_edge(const Cx& C) {\n" + fn[1] + "}\n\n"
           "struct JitExec {\n"
           "  template <class Cx>\n"
           "  __device__ __forceinline__ static void run(const WordArgs&, uint32_t id, const Cx& C) {\n"
           "    if (id == B_COMB) jit_eval(C);\n"
           "    else if (id == B_COMMIT) jit_edge(C);\n"
           "  }\n" + staging() +
           "};\n\n}  // namespace " + ns + "\n}  // namespace word\n}  // namespace rrtl\n";
  }
  uint32_t waves() const { return waves_; }
  // Translation unit with the kernel "rrtl_word_jit" for one storage mode.
  std::string source(int mode, uint32_t xrow0 = 0) const {
    return "#include \"word_device.hpp\"\n\n" + definitions(xrow0, "jit") +
           "\nextern \"C\" __global__ void __launch_bounds__(256) rrtl_word_jit(rrtl::word::WordArgs A) {\n"
           "  rrtl::word::word_body<" + std::to_string(mode) + ", rrtl::word::jit::JitExec, " + std::to_string(waves_) +
           ">(A, blockIdx.x);\n}\n";
  }

 private:
  static std::string hex(uint32_t v) {
    char b[16];
    snprintf(b, sizeof b, "0x%08xu", v);
    return b;
  }
  static std::string num(uint32_t v) { return std::to_string(v) + "u"; }
  // enter()/leave(): copy the LDS-staged memories in from HBM and back out.
  std::string staging() const {
    std::string in, out;
    for (size_t m = 0; m < mem_row_.size(); m++) {
      if (mem_row_[m] < 0) continue;
      const uint32_t* mt = &WP.mem_table[4 * m];
      const std::string n = num(mt[1] * mt[2]), row = num(uint32_t(mem_row_[m])), base = num(mt[0]);
      in += "    for (uint32_t w = 0; w < " + n + "; w++) C.lmem(" + row + ", w) = C.mem(" + base + ", w);\n";
      out += "    for (uint32_t w = 0; w < " + n + "; w++) C.mem(" + base + ", w) = C.lmem(" + row + ", w);\n";
    }
    return "  template <class Cx>\n  __device__ __forceinline__ static void enter(const Cx& C) {\n" + in + "  }\n"
           "  template <class Cx>\n  __device__ __forceinline__ static void leave(const Cx& C) {\n" + out + "  }\n";
  }
  // Word `word` (an expression) of memory m, in LDS or HBM.
  std::string mem_ref(uint32_t m, const std::string& word) const {
    const uint32_t* mt = &WP.mem_table[4 * size_t(m)];
    if (mem_row_[m] >= 0 && lay_.stride)
      return "L[" + num((lay_.slots + lay_.sig_words + uint32_t(mem_row_[m])) * lay_.stride) + " + (" + word + ") * " +
             num(lay_.stride) + "]";
    if (mem_row_[m] >= 0) return "C.lmem(" + num(uint32_t(mem_row_[m])) + ", " + word + ")";
    return "C.mem(" + num(mt[0]) + ", " + word + ")";
  }
  static std::string masked(const std::string& e, uint32_t width) {
    return width % 32 ? "(" + e + " & " + hex(final_mask(width)) + ")" : e;
  }
  // One generated statement (or a guarded group of them) with what the wave
  // split needs to know: the value it defines, the signal words it loads or
  // stores, the memory it accesses.
  enum : int { S_PURE = 0, S_LOAD = 1, S_STORE = 2, S_MEM = 3, S_GENERIC = 4 };
  struct Stmt {
    std::string text, def;
    bool is_bool = false;
    int kind = S_PURE;
    std::vector<uint32_t> words;
    int32_t mem = -1;
  };
  void begin(std::vector<Stmt>& out) {
    st_ = &out; sig_.clear(); scr_.clear(); stage_.clear();
  }
  std::string header() const { return lay_.stride ? "  uint32_t* const L = lds_words() + C.col;\n" : ""; }
  std::string emit_serial(const std::vector<Stmt>& S) const {
    std::string out = header();
    for (auto& x : S) out += x.text;
    return out;
  }
  // Signal word w as an lvalue expression.
  std::string sig_ref(uint32_t w) const {
    if (lay_.stride) return "L[" + num((lay_.slots + w) * lay_.stride) + "]";
    return "";
  }
  void store_sig(uint32_t off, uint32_t l, const std::string& v) {
    if (lay_.stride) line(sig_ref(off + l) + " = " + v + ";");
    else line("C.stsig(" + num(off) + ", " + num(l) + ", " + v + ");");
    if (!grouping_) st_->back().kind = S_STORE;
    st_->back().words.push_back(off + l);
  }
  void line(const std::string& s) {
    if (grouping_) { st_->back().text += "  " + s + "\n"; return; }
    Stmt x;
    x.text = "  " + s + "\n";
    st_->push_back(std::move(x));
  }
  // A guarded group: the lines up to end_group() form one statement of `kind`.
  void begin_group(const std::string& s, int kind, int32_t mem = -1) {
    line(s);
    st_->back().kind = kind; st_->back().mem = mem;
    grouping_ = true;
  }
  void end_group() { st_->back().text += "  }\n"; grouping_ = false; }
  std::string fresh(char p) { return p + std::to_string(tmp_++); }
  std::string def(char p, const std::string& e) {
    std::string n = fresh(p);
    line("const uint32_t " + n + " = " + e + ";");
    st_->back().def = n;
    return n;
  }
  std::string def_bool(char p, const std::string& e) {
    std::string n = fresh(p);
    line("const bool " + n + " = " + e + ";");
    st_->back().def = n; st_->back().is_bool = true;
    return n;
  }
  uint32_t cp(uint32_t i) const {
    if (i >= WP.cpool.size()) throw Error("E_PROGRAM: constant pool index out of range");
    return WP.cpool[i];
  }
  const uint32_t* mtab(uint32_t m) const {
    if (4 * size_t(m) + 4 > WP.mem_table.size()) throw Error("E_PROGRAM: memory index out of range");
    return &WP.mem_table[4 * size_t(m)];
  }

  // Limb `limb` of operand `code` as an expression (loads are emitted first).
  std::string opnd(uint32_t code, uint32_t limb) {
    const uint32_t idx = code & ~SP_MASK;
    switch (code & SP_MASK) {
      case SP_CONST: return hex(cp(idx + limb));
      case SP_SIGNAL: {
        const uint32_t w = idx + limb;
        auto it = sig_.find(w);
        if (it != sig_.end()) return it->second;
        const std::string n = def('s', lay_.stride ? sig_ref(w) : "C.ld(" + hex(SP_SIGNAL | w) + ", 0)");
        st_->back().kind = S_LOAD; st_->back().words = {w};
        return sig_[w] = n;
      }
      default: {
        auto it = scr_.find(idx + limb);
        if (it != scr_.end()) return it->second;
        const std::string n = def('t', "C.ld(" + hex(idx + limb) + ", 0)");
        st_->back().kind = S_GENERIC;
        return n;
      }
    }
  }
  std::string reset_cond(uint32_t code, uint32_t pol) {
    const std::string r = opnd(code, 0);
    return def_bool('r', "(" + r + " & 1u) " + (pol == 0 ? "!=" : "==") + " 0u");
  }
  // In-range test of an address operand of `n` limbs against `depth`.
  std::string addr_ok(uint32_t code, uint32_t n, uint32_t depth, std::string& lo) {
    lo = opnd(code, 0);
    std::string ok = "(" + lo + " < " + num(depth) + ")";
    for (uint32_t l = 1; l < n; l++) ok += " && (" + opnd(code, l) + " == 0u)";
    return ok;
  }

  // One-limb result expression, or "" when the record keeps interpreted form.
  std::string expr1(const Rec& R) {
    const uint32_t op = R.w[0] & 0xFF;
    switch (op) {
      case OP_NOT: return "~" + opnd(R.w[2], 0);
      case FOP_NAND: case FOP_NOR: case FOP_XNOR: case FOP_ANDOR: case FOP_ORAND: case FOP_XOR3: case FOP_ADD3:
      case FOP_MULADD: case FOP_AOI21: case FOP_OAI21: case FOP_MULSUB: case FOP_AOI22: case FOP_OAI22: {
        const std::string a = opnd(R.w[2], 0), b = opnd(R.w[3], 0);
        switch (op) {
          case FOP_NAND: return "~(" + a + " & " + b + ")";
          case FOP_NOR: return "~(" + a + " | " + b + ")";
          case FOP_XNOR: return "~(" + a + " ^ " + b + ")";
          default: break;
        }
        const std::string c = opnd(R.w[4], 0);
        switch (op) {
          case FOP_ANDOR: return "((" + a + " & " + b + ") | " + c + ")";
          case FOP_ORAND: return "((" + a + " | " + b + ") & " + c + ")";
          case FOP_XOR3: return "(" + a + " ^ " + b + " ^ " + c + ")";
          case FOP_ADD3: return "(" + a + " + " + b + " + " + c + ")";
          case FOP_MULADD: return "(" + a + " * " + b + " + " + c + ")";
          case FOP_AOI21: return "~((" + a + " & " + b + ") | " + c + ")";
          case FOP_MULSUB: return "(" + a + " * " + b + " - " + c + ")";
          case FOP_AOI22: return "~((" + a + " & " + b + ") | (" + c + " & " + opnd(R.w[6], 0) + "))";
          case FOP_OAI22: return "~((" + a + " | " + b + ") & (" + c + " | " + opnd(R.w[6], 0) + "))";
          default: return "~((" + a + " | " + b + ") & " + c + ")";
        }
      }
      case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: {
        static const char* sym[] = {"&", "|", "^", "+", "-", "*"};
        const std::string a = opnd(R.w[2], 0), b = opnd(R.w[3], 0);
        return "(" + a + " " + sym[op - OP_AND] + " " + b + ")";
      }
      case OP_MUX: {
        const std::string c = opnd(R.w[4], 0), a = opnd(R.w[2], 0), b = opnd(R.w[3], 0);
        return "((" + c + " & 1u) ? " + a + " : " + b + ")";
      }
      case OP_ZEXT: case OP_TRUNC: case OP_CAST: case R_SIGLOAD: return opnd(R.w[2], 0);
      case OP_EQ: case OP_NE: {
        const uint32_t la = R.w[7] & 0xFF, lb = (R.w[7] >> 8) & 0xFF, n = std::max(la, lb);
        std::string e;
        for (uint32_t l = 0; l < n; l++) {
          const std::string a = l < la ? opnd(R.w[2], l) : "0u", b = l < lb ? opnd(R.w[3], l) : "0u";
          const std::string t = "uint32_t(" + a + " == " + b + ")";
          e = e.empty() ? t : e + " & " + t;
        }
        if (e.empty()) e = "1u";
        return op == OP_EQ ? "(" + e + ")" : "((" + e + ") ^ 1u)";
      }
      case OP_LT: {
        const uint32_t wl = R.w[6];
        if (wl == 0 || wl > 32) return "";
        const std::string a = opnd(R.w[2], 0), b = opnd(R.w[3], 0);
        if (R.w[7] >> 24) {  // signed: the wl-bit values moved to the top and compared as int
          const std::string sh = num(32 - wl);
          return "uint32_t(int(" + a + " << " + sh + ") < int(" + b + " << " + sh + "))";
        }
        return "uint32_t(" + masked(a, wl) + " < " + masked(b, wl) + ")";
      }
      case OP_SLICE: {
        const uint32_t n = R.w[7], lsb = R.w[6], li = lsb >> 5, sh = lsb & 31;
        const std::string lo = li < n ? opnd(R.w[2], li) : "0u";
        if (!sh) return lo;
        const std::string hi = li + 1 < n ? opnd(R.w[2], li + 1) : "";
        if (hi.empty()) return "(" + lo + " >> " + num(sh) + ")";
        return "((" + lo + " >> " + num(sh) + ") | (" + hi + " << " + num(32 - sh) + "))";
      }
      case OP_SEXT: {
        const uint32_t iw = R.w[6];
        if (iw == 0) return "";
        const std::string x = opnd(R.w[2], 0);
        if (iw >= 32) return x;
        const std::string sh = num(32 - iw);
        return "uint32_t(int(" + x + " << " + sh + ") >> " + sh + ")";
      }
      case OP_CONCAT: {
        std::string e;
        uint32_t off = 0;
        for (int k = int(R.w[3]) - 1; k >= 0; k--) {
          const uint32_t code = cp(R.w[2] + 2 * k), wd = cp(R.w[2] + 2 * k + 1) & 0xFFFF;
          std::string t = masked(opnd(code, 0), wd);
          if (off >= 32) return "";
          if (off) t = "(" + t + " << " + num(off) + ")";
          e = e.empty() ? t : e + " | " + t;
          off += wd;
        }
        return e.empty() ? "0u" : "(" + e + ")";
      }
      case OP_MEMREAD: {
        const uint32_t* mt = mtab(R.w[6]);
        if (mt[1] == 0) return "0u";
        std::string lo;
        const std::string ok = addr_ok(R.w[2], R.w[7], mt[1], lo);
        const std::string o = def_bool('o', ok);
        // out-of-range lanes read word 0 (in bounds) and take 0
        const std::string m = def('m', mem_ref(R.w[6], "uint32_t(" + o + " ? " + lo + " : 0u) * " + num(mt[2])));
        st_->back().kind = S_MEM; st_->back().mem = int32_t(R.w[6]);
        return "(" + o + " ? " + (mt[2] == 1 ? masked(m, mt[3]) : m) + " : 0u)";
      }
      default: return "";
    }
  }

  void spill(uint32_t code, uint32_t limbs) {
    if ((code & SP_MASK) != SP_SCRATCH) return;
    for (uint32_t l = 0; l < limbs; l++) {
      auto it = scr_.find((code & ~SP_MASK) + l);
      if (it != scr_.end()) { line("C.st(" + hex(it->first) + ", 0, " + it->second + ");"); st_->back().kind = S_GENERIC; }
    }
  }
  void emit_generic(const Rec& R) {
    const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF;
    // register-held operands go to their slots first
    switch (op) {
      case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: case OP_EQ: case OP_NE: case OP_LT:
        spill(R.w[2], R.w[7] & 0xFF); spill(R.w[3], (R.w[7] >> 8) & 0xFF); break;
      case OP_MUX:
        spill(R.w[2], R.w[7] & 0xFF); spill(R.w[3], (R.w[7] >> 8) & 0xFF); spill(R.w[4], 1); break;
      case OP_CONCAT:
        for (uint32_t k = 0; k < R.w[3]; k++) spill(cp(R.w[2] + 2 * k), cp(R.w[2] + 2 * k + 1) >> 16);
        break;
      default: spill(R.w[2], R.w[7]); break;
    }
    std::string w;
    for (int k = 0; k < 8; k++) w += (k ? ", " : "") + hex(R.w[k]);
    line("{ const Rec R = {{" + w + "}}; exec_rec(C, R); }");
    st_->back().kind = S_GENERIC;
    for (uint32_t l = 0; l < L; l++) scr_.erase(R.w[1] + l);
    generic_ = true;
    n_generic_++;
  }

  void emit_effect(const Rec& R) {
    const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF;
    switch (op) {
      case R_STORE:
        for (uint32_t l = 0; l < L; l++) {
          std::string v = l < R.w[7] ? opnd(R.w[2], l) : "0u";
          if (l + 1 == L) v = masked(v, R.w[5]);
          const std::string n = def('w', v);
          store_sig(R.w[1], l, n);
          sig_[R.w[1] + l] = n;
        }
        break;
      case R_ASYNC_CAPTURE: {
        const std::string r = reset_cond(R.w[3], R.w[6]);
        std::vector<std::string> v(L);
        for (uint32_t l = 0; l < L; l++) {
          v[l] = l < R.w[7] ? opnd(R.w[2], l) : "0u";
          if (l + 1 == L) v[l] = masked(v[l], R.w[5]);
        }
        begin_group("if (" + r + ") {", S_STORE);
        for (uint32_t l = 0; l < L; l++) store_sig(R.w[1], l, v[l]);
        end_group();
        for (uint32_t l = 0; l < L; l++) sig_.erase(R.w[1] + l);
        break;
      }
      case R_CAPTURE: {
        const std::string r = (R.w[6] & 2u) ? reset_cond(R.w[3], R.w[6] & 1u) : "";
        for (uint32_t l = 0; l < L; l++) {
          std::string v = l < R.w[7] ? opnd(R.w[2], l) : "0u";
          if (!r.empty()) v = "(" + r + " ? " + opnd(R.w[4], l) + " : " + v + ")";
          if (l + 1 == L) v = masked(v, R.w[5]);
          stage_[R.w[1] + l] = def('c', v);
        }
        break;
      }
      case R_COMMIT:
        for (uint32_t l = 0; l < L; l++) {
          auto it = stage_.find((R.w[2] & ~SP_MASK) + l);
          if (it == stage_.end()) throw Error("E_PROGRAM: register commit without a capture");
          store_sig(R.w[1], l, it->second);
          sig_[R.w[1] + l] = it->second;
        }
        break;
      case R_MEMWRITE: {
        const uint32_t* mt = mtab(R.w[1]);
        if (mt[1] == 0) break;
        const uint32_t la = R.w[7] & 0xFF, ld = (R.w[7] >> 8) & 0xFF;
        const std::string en = opnd(R.w[2], 0);
        std::string lo;
        const std::string ok = addr_ok(R.w[3], la, mt[1], lo);
        std::vector<std::string> v(L);
        for (uint32_t l = 0; l < L; l++) {
          v[l] = l < ld ? opnd(R.w[4], l) : "0u";
          if (l + 1 == L) v[l] = masked(v[l], mt[3]);
        }
        begin_group("if ((" + en + " & 1u) && " + ok + ") {", S_MEM, int32_t(R.w[1]));
        for (uint32_t l = 0; l < L; l++)
          line("  " + mem_ref(R.w[1], lo + " * " + num(mt[2]) + " + " + num(l)) + " = " + v[l] + ";");
        end_group();
        break;
      }
      default: throw Error("E_PROGRAM: unknown effect record");
    }
  }

  // Cost units of a statement (about one VALU instruction each).
  static double cost_of(const Stmt& x) {
    if (x.kind == S_LOAD) return 1;
    if (x.kind == S_STORE) return double(x.words.size());
    if (x.kind == S_MEM) return 4;
    double c = 1;
    const std::string& t = x.text;
    for (size_t i = 0; i < t.size(); i++) {
      const char ch = t[i];
      if (ch == '&' || ch == '|' || ch == '^' || ch == '+' || ch == '*' || ch == '~' || ch == '?' || ch == '-') c++;
      else if ((ch == '<' || ch == '>') && !(i + 1 < t.size() && t[i + 1] == ch)) c++;
      else if (ch == '=' && i > 0 && (t[i - 1] == '=' || t[i - 1] == '!')) c++;
    }
    return c;
  }
  // Names of generated values a statement mentions (fresh() names: a letter
  // and digits, not inside a longer identifier or a literal).
  static std::vector<std::string> names_in(const std::string& t) {
    std::vector<std::string> out;
    for (size_t i = 0; i < t.size(); i++) {
      const char ch = t[i];
      if (ch < 'a' || ch > 'z') continue;
      if (i > 0 && (std::isalnum(static_cast<unsigned char>(t[i - 1])) || t[i - 1] == '_' || t[i - 1] == '.')) continue;
      size_t j = i + 1;
      while (j < t.size() && std::isdigit(static_cast<unsigned char>(t[j]))) j++;
      if (j == i + 1 || (j < t.size() && (std::isalnum(static_cast<unsigned char>(t[j])) || t[j] == '_'))) continue;
      out.push_back(t.substr(i, j - i));
      i = j - 1;
    }
    return out;
  }

  // Splits one evaluation between G waves in barrier phases (bulk-synchronous
  // list scheduling).  Statements run in program order within a wave; a value
  // used by another wave is written to an LDS exchange row in its phase and
  // read in a later one; signal loads are repeated in every wave that needs
  // them; all accesses of one memory stay in one wave; a signal store runs in
  // a phase after every other wave's last load of that word (and after the
  // earlier stores of it), so no wave reads a word another wave has already
  // rewritten.  Budgets of work per wave and phase are tried from one phase
  // per wave down to many; the estimate (per phase the busiest wave's work
  // plus its exchange accesses, plus a barrier) picks the plan.  Code with
  // interpreted records or with a load after a store of the same word cannot
  // be split.
  static Split split(const std::vector<Stmt>& S, uint32_t G) {
    // Cost units of a barrier, of an exchange access and the largest cone
    // recomputed per wave (picorv32 at 16,384 lanes, 4 waves: 4.61e9
    // lane-cycles/s at 100 / 8 against 4.40e9 at 24 / 4 and 3.52e9 unsplit).
    constexpr double kBarrier = 100, kXchg = 1, kRecompute = 8;
    Split best;
    const size_t n = S.size();
    std::unordered_map<std::string, uint32_t> def_of;
    std::vector<std::vector<uint32_t>> uses(n);
    std::unordered_set<uint32_t> stored;
    double serial = 0;
    for (size_t i = 0; i < n; i++) {
      const Stmt& x = S[i];
      if (x.kind == S_GENERIC) return best;
      if (x.kind == S_LOAD && stored.count(x.words[0])) return best;
      if (x.kind == S_STORE) for (uint32_t w : x.words) stored.insert(w);
      for (auto& nm : names_in(x.text)) {
        if (nm == x.def) continue;
        auto it = def_of.find(nm);
        if (it != def_of.end()) uses[i].push_back(it->second);
      }
      std::sort(uses[i].begin(), uses[i].end());
      uses[i].erase(std::unique(uses[i].begin(), uses[i].end()), uses[i].end());
      if (!x.def.empty()) def_of[x.def] = uint32_t(i);
      serial += cost_of(x);
    }
    best.serial = serial;
    // Cheap values of signal loads alone are recomputed in every wave that
    // uses them (like the loads) instead of crossing between waves.
    std::vector<uint8_t> rep(n, 0);
    std::vector<double> cone(n, 0);
    for (size_t i = 0; i < n; i++) {
      if (S[i].kind != S_PURE || S[i].def.empty()) continue;
      double c = cost_of(S[i]);
      bool ok = true;
      for (uint32_t q : uses[i]) {
        if (S[q].kind == S_LOAD) continue;
        if (!rep[q]) { ok = false; break; }
        c += cone[q];
      }
      if (ok && c <= kRecompute) { rep[i] = 1; cone[i] = c; }
    }
    auto is_work = [&](size_t i) { return (S[i].kind == S_PURE && !rep[i]) || S[i].kind == S_MEM; };
    // work predecessors: values used, and the previous access of the same memory
    std::vector<std::vector<uint32_t>> pred(n), succ(n);
    std::unordered_map<int32_t, uint32_t> last_mem;
    int32_t n_mem = 0;
    for (size_t i = 0; i < n; i++) {
      if (!is_work(i)) continue;
      for (uint32_t q : uses[i]) if (is_work(q)) pred[i].push_back(q);
      if (S[i].kind == S_MEM) {
        n_mem = std::max(n_mem, S[i].mem + 1);
        auto it = last_mem.find(S[i].mem);
        if (it != last_mem.end()) pred[i].push_back(it->second);
        last_mem[S[i].mem] = uint32_t(i);
      }
      for (uint32_t q : pred[i]) succ[q].push_back(uint32_t(i));
    }
    std::vector<double> cost(n), bl(n, 0);
    double total = 0;
    for (size_t i = 0; i < n; i++) { cost[i] = cost_of(S[i]); if (is_work(i)) total += cost[i]; }
    for (size_t k = n; k-- > 0;) {
      if (!is_work(k)) continue;
      double m = 0;
      for (uint32_t u : succ[k]) m = std::max(m, bl[u]);
      bl[k] = cost[k] + m;
    }
    std::vector<double> budgets;
    for (uint32_t k = 1; k <= 48; k++) budgets.push_back(std::max(4.0, total / G / k));
    for (double B : budgets) {
      Split W;
      W.ok = true; W.serial = serial;
      W.wave.assign(n, -1); W.phase.assign(n, -1);
      std::vector<uint32_t> left(n, 0), ready;
      size_t nwork = 0;
      for (size_t i = 0; i < n; i++) {
        if (!is_work(i)) continue;
        nwork++;
        left[i] = uint32_t(pred[i].size());
        if (!left[i]) ready.push_back(uint32_t(i));
      }
      std::vector<int32_t> pin(size_t(n_mem), -1);
      std::vector<double> load(G, 0);
      int32_t p = 0;
      size_t placed = 0;
      while (placed < nwork) {
        std::sort(ready.begin(), ready.end(), [&](uint32_t a, uint32_t b) { return bl[a] != bl[b] ? bl[a] > bl[b] : a < b; });
        int32_t bi = -1, bg = -1;
        for (size_t r = 0; r < ready.size() && bi < 0; r++) {
          const uint32_t i = ready[r];
          double bkey = 0;
          for (uint32_t g = 0; g < G; g++) {
            if (S[i].kind == S_MEM && pin[S[i].mem] >= 0 && pin[S[i].mem] != int32_t(g)) continue;
            int32_t er = 0, away = 0;
            for (uint32_t q : pred[i]) {
              er = std::max(er, W.wave[q] == int32_t(g) ? W.phase[q] : W.phase[q] + 1);
              away += W.wave[q] != int32_t(g);
            }
            if (er > p) continue;
            if (load[g] > 0 && load[g] + cost[i] > B) continue;
            const double key = load[g] + 2 * kXchg * away;
            if (bg < 0 || key < bkey) { bg = int32_t(g); bkey = key; }
          }
          if (bg >= 0) bi = int32_t(r);
        }
        if (bi < 0) { p++; std::fill(load.begin(), load.end(), 0.0); continue; }
        const uint32_t i = ready[size_t(bi)];
        ready.erase(ready.begin() + bi);
        W.wave[i] = bg; W.phase[i] = p; load[size_t(bg)] += cost[i];
        if (S[i].kind == S_MEM) pin[S[i].mem] = bg;
        placed++;
        for (uint32_t u : succ[i]) if (--left[u] == 0) ready.push_back(u);
      }
      // signal loads and recomputed values: in each wave at the phase of their first user there
      W.load_phase.assign(G, std::vector<int32_t>(n, -1));
      std::function<void(uint32_t, int32_t, int32_t)> need_load = [&](uint32_t l, int32_t g, int32_t ph) {
        int32_t& lp = W.load_phase[size_t(g)][l];
        if (lp >= 0 && lp <= ph) return;
        lp = ph;
        if (rep[l]) for (uint32_t q : uses[l]) need_load(q, g, ph);
      };
      for (size_t i = 0; i < n; i++)
        if (is_work(i))
          for (uint32_t q : uses[i]) if (S[q].kind == S_LOAD || rep[q]) need_load(q, W.wave[i], W.phase[i]);
      // stores: in the wave of the latest value they store; the loads and
      // recomputed values a store uses itself (register copies) are taken in
      // phase 0, before any store, so every load is known before the stores
      // are ordered after them
      std::vector<int32_t> store_wave(n, 0);
      for (size_t i = 0; i < n; i++) {
        if (S[i].kind != S_STORE) continue;
        int32_t g = 0, gp = -1;
        for (uint32_t q : uses[i]) if (is_work(q) && W.phase[q] > gp) { gp = W.phase[q]; g = W.wave[q]; }
        store_wave[i] = g;
        for (uint32_t q : uses[i]) if (S[q].kind == S_LOAD || rep[q]) need_load(q, g, 0);
      }
      std::unordered_map<uint32_t, std::vector<int32_t>> last_load;  // word -> per wave: last load phase
      auto note_loads = [&]() {
        last_load.clear();
        for (uint32_t g = 0; g < G; g++)
          for (size_t l = 0; l < n; l++)
            if (W.load_phase[g][l] >= 0 && S[l].kind == S_LOAD) {
              auto& v = last_load[S[l].words[0]];
              if (v.empty()) v.assign(G, -1);
              v[g] = std::max(v[g], W.load_phase[g][l]);
            }
      };
      note_loads();
      std::unordered_map<uint32_t, std::pair<int32_t, int32_t>> last_store;  // word -> (wave, phase)
      for (size_t i = 0; i < n; i++) {
        if (S[i].kind != S_STORE) continue;
        const int32_t g = store_wave[i];
        int32_t ph = 0;
        for (uint32_t q : uses[i]) if (is_work(q)) ph = std::max(ph, W.wave[q] == g ? W.phase[q] : W.phase[q] + 1);
        for (uint32_t w : S[i].words) {
          auto it = last_load.find(w);
          if (it != last_load.end())
            for (uint32_t h = 0; h < G; h++)
              if (it->second[h] >= 0) ph = std::max(ph, int32_t(h) == g ? it->second[h] : it->second[h] + 1);
          auto js = last_store.find(w);
          if (js != last_store.end()) ph = std::max(ph, js->second.first == g ? js->second.second : js->second.second + 1);
        }
        W.wave[i] = g; W.phase[i] = ph;
        for (uint32_t w : S[i].words) last_store[w] = {g, ph};
      }
      int32_t P = 0;
      for (size_t i = 0; i < n; i++) P = std::max(P, W.phase[i] + 1);
      for (uint32_t g = 0; g < G; g++) for (size_t l = 0; l < n; l++) P = std::max(P, W.load_phase[g][l] + 1);
      W.phases = uint32_t(std::max(P, 1));
      // exchange rows: a value used in another wave, live from its phase to its last reader's
      std::vector<int32_t> last_read(n, -1);
      std::vector<std::vector<double>> busy(G, std::vector<double>(W.phases, 0));
      for (size_t i = 0; i < n; i++) {
        if (W.wave[i] >= 0) busy[size_t(W.wave[i])][size_t(W.phase[i])] += cost[i];
        if (W.wave[i] < 0) continue;
        std::vector<uint8_t> imported(G, 0);
        for (uint32_t q : uses[i]) {
          if (!is_work(q) || W.wave[q] == W.wave[i]) continue;
          last_read[q] = std::max(last_read[q], W.phase[i]);
        }
      }
      for (uint32_t g = 0; g < G; g++)
        for (size_t l = 0; l < n; l++) if (W.load_phase[g][l] >= 0) busy[g][size_t(W.load_phase[g][l])] += cost[l];
      // imports: one read per (value, reading wave)
      for (uint32_t g = 0; g < G; g++) {
        std::vector<int32_t> first(n, -1);
        for (size_t i = 0; i < n; i++) {
          if (W.wave[i] != int32_t(g)) continue;
          for (uint32_t q : uses[i])
            if (is_work(q) && W.wave[q] != int32_t(g) && (first[q] < 0 || W.phase[i] < first[q])) first[q] = W.phase[i];
        }
        for (size_t q = 0; q < n; q++) if (first[q] >= 0) busy[g][size_t(first[q])] += kXchg;
      }
      W.xslot.assign(n, -1);
      std::vector<int32_t> row_free;  // per row: first phase it may be written again
      for (int32_t ph = 0; ph < int32_t(W.phases); ph++)
        for (size_t q = 0; q < n; q++) {
          if (last_read[q] < 0 || W.phase[q] != ph) continue;
          busy[size_t(W.wave[q])][size_t(ph)] += kXchg;
          int32_t r = -1;
          for (size_t k = 0; k < row_free.size() && r < 0; k++) if (row_free[k] <= ph) r = int32_t(k);
          if (r < 0) { r = int32_t(row_free.size()); row_free.push_back(0); }
          row_free[size_t(r)] = last_read[q] + 1;
          W.xslot[q] = r;
        }
      W.xslots = uint32_t(row_free.size());
      double est = 0;
      for (uint32_t ph = 0; ph < W.phases; ph++) {
        double m = 0;
        for (uint32_t g = 0; g < G; g++) m = std::max(m, busy[g][ph]);
        est += m + kBarrier;
      }
      W.est = est;
      if (!best.ok || W.est < best.est) best = std::move(W);
    }
    verify_split(S, best, G, uses, rep);
    return best;
  }

  // Checks a split against the sequential program: every value is computed
  // before its use (in the same wave, or in an earlier phase of another wave
  // through an exchange row); any two accesses of one signal word of which
  // one is a store, and any two accesses of one memory, keep their program
  // order (within a wave by phase and position, across waves by phase);
  // exchange rows are not rewritten before their last read.  A violation is
  // an internal error.
  static void verify_split(const std::vector<Stmt>& S, const Split& W, uint32_t G,
                           const std::vector<std::vector<uint32_t>>& uses, const std::vector<uint8_t>& rep) {
    const size_t n = S.size();
    auto fail = [&](const std::string& what, size_t i) {
      throw Error("E_HIP_INTERNAL: wave split: " + what + " at statement " + std::to_string(i) + ": " + S[i].text);
    };
    struct Occ { uint32_t g; int32_t ph; uint32_t i; };
    std::vector<std::vector<Occ>> occ(n);
    for (size_t i = 0; i < n; i++) {
      if (W.wave[i] >= 0) occ[i].push_back({uint32_t(W.wave[i]), W.phase[i], uint32_t(i)});
      else
        for (uint32_t g = 0; g < G; g++)
          if (W.load_phase[g][i] >= 0) occ[i].push_back({g, W.load_phase[g][i], uint32_t(i)});
    }
    auto before = [](const Occ& a, const Occ& b) { return a.ph < b.ph || (a.ph == b.ph && a.i < b.i); };
    std::vector<int32_t> last_read(n, -1);
    for (size_t i = 0; i < n; i++)
      for (const Occ& o : occ[i])
        for (uint32_t q : uses[i]) {
          bool ok = false;
          if (W.wave[q] < 0) {
            for (const Occ& d : occ[q]) ok |= d.g == o.g && before(d, o);
          } else if (uint32_t(W.wave[q]) == o.g) {
            ok = before(occ[q][0], o);
          } else {
            ok = W.phase[q] < o.ph && W.xslot[q] >= 0;
            last_read[q] = std::max(last_read[q], o.ph);
          }
          if (!ok) fail("value " + S[q].def + " not available", i);
        }
    std::unordered_map<uint64_t, std::vector<std::pair<Occ, bool>>> acc;  // (kind, word or memory) -> accesses
    for (size_t i = 0; i < n; i++) {
      if (S[i].kind == S_LOAD || S[i].kind == S_STORE)
        for (uint32_t w : S[i].words)
          for (const Occ& o : occ[i]) acc[uint64_t(w)].push_back({o, S[i].kind == S_STORE});
      if (S[i].kind == S_MEM)
        for (const Occ& o : occ[i]) acc[(1ull << 40) | uint32_t(S[i].mem)].push_back({o, true});
    }
    for (auto& kv : acc) {
      const auto& v = kv.second;
      for (size_t a = 0; a < v.size(); a++)
        for (size_t b = 0; b < v.size(); b++) {
          const Occ &x = v[a].first, &y = v[b].first;
          if (x.i >= y.i || !(v[a].second || v[b].second)) continue;
          const bool ok = x.g == y.g ? before(x, y) : x.ph < y.ph;
          if (!ok) fail(kv.first >> 40 ? "memory access order" : "signal access order", y.i);
        }
    }
    std::unordered_map<int32_t, std::vector<uint32_t>> rows;
    for (size_t q = 0; q < n; q++) if (W.xslot[q] >= 0) rows[W.xslot[q]].push_back(uint32_t(q));
    for (auto& kv : rows)
      for (uint32_t a : kv.second)
        for (uint32_t b : kv.second)
          if (a != b && W.phase[a] <= W.phase[b] && last_read[a] >= W.phase[b]) fail("exchange row reused while live", b);
    (void)rep;
  }

  // jit_eval / jit_edge of the G-wave kernel: per wave its statements phase by
  // phase in program order (exchange reads before first use, exchange writes
  // after the definition), a barrier between phases, and one after the last
  // phase so the next step sees every wave's stores.
  std::string emit_split(const std::vector<Stmt>& S, const Split& W, uint32_t G, uint32_t xrow0) const {
    const size_t n = S.size();
    std::vector<std::vector<uint32_t>> uses(n);
    std::unordered_map<std::string, uint32_t> def_of;
    for (size_t i = 0; i < n; i++) {
      for (auto& nm : names_in(S[i].text)) {
        if (nm == S[i].def) continue;
        auto it = def_of.find(nm);
        if (it != def_of.end()) uses[i].push_back(it->second);
      }
      if (!S[i].def.empty()) def_of[S[i].def] = uint32_t(i);
    }
    auto xref = [&](uint32_t q) { return "L[" + num((xrow0 + uint32_t(W.xslot[q])) * lay_.stride) + "]"; };
    std::string out = header();
    for (uint32_t g = 0; g < G; g++) {
      out += "  if (C.wave == " + num(g) + ") {\n";
      std::vector<uint8_t> have(n, 0);
      for (uint32_t ph = 0; ph < W.phases; ph++) {
        for (size_t i = 0; i < n; i++) {
          const bool mine = W.wave[i] == int32_t(g) && W.phase[i] == int32_t(ph);
          const bool load = W.wave[i] < 0 && W.load_phase[g][i] == int32_t(ph);
          if (!mine && !load) continue;
          for (uint32_t q : uses[i]) {
            if (have[q] || W.wave[q] < 0 || W.wave[q] == int32_t(g) || W.xslot[q] < 0) continue;
            out += S[q].is_bool ? "    const bool " + S[q].def + " = " + xref(q) + " != 0u;\n"
                                : "    const uint32_t " + S[q].def + " = " + xref(q) + ";\n";
            have[q] = 1;
          }
          for (size_t a = 0, b; a < S[i].text.size(); a = b + 1) {
            b = S[i].text.find('\n', a);
            out += "  " + S[i].text.substr(a, b - a + 1);
          }
          if (mine && W.xslot[i] >= 0)
            out += "    " + xref(uint32_t(i)) + " = " + (S[i].is_bool ? "uint32_t(" + S[i].def + ")" : S[i].def) + ";\n";
        }
        if (ph + 1 < W.phases) out += "    __syncthreads();\n";
      }
      out += "  }\n";
    }
    out += "  __syncthreads();\n";
    return out;
  }

  void emit_block(int b) {
    for (const Rec& R : WP.blk[b]) {
      const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF;
      if (op == R_MACRO) throw Error("E_HIP_INTERNAL: macro records are for the interpreter only");
      if (is_effect_op(op)) { emit_effect(R); continue; }
      std::string e = L == 1 ? expr1(R) : "";
      if (e.empty()) { emit_generic(R); continue; }
      scr_[R.w[1]] = def('v', masked(e, R.w[5]));
    }
  }

  const WordProgram& WP;
  JitLayout lay_;
  std::vector<int32_t> mem_row_;  // per memory: first LDS row, or -1 (HBM)
  std::vector<Stmt> fn_[2];       // jit_eval, jit_edge
  Split split_[2];
  uint32_t waves_ = 1, xslots_ = 0;
  std::vector<Stmt>* st_ = nullptr;
  bool grouping_ = false;
  std::unordered_map<uint32_t, std::string> sig_, scr_, stage_;  // word -> register holding it
  uint32_t tmp_ = 0, n_generic_ = 0;
  bool generic_ = false;
};

// Memories a generated kernel stages in LDS (WM_ALL_LDS): small ones (a
// register file, a FIFO's storage), smallest first, while a block of `block`
// lanes keeps its scratch, signals and staged memories within `limit` bytes
// (64 KiB by default).  Returns
// each memory's first LDS row (-1: stays in HBM); *used = staged words per lane.
std::vector<int32_t> lds_memory_rows(const Program& P, uint32_t slots, uint32_t block, uint32_t* used,
                                     size_t limit = 64 * 1024) {
  std::vector<uint32_t> order(P.memories.size());
  std::iota(order.begin(), order.end(), 0u);
  auto words = [&](uint32_t m) { return uint64_t(P.memories[m].depth) * P.memories[m].limbs; };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return words(a) < words(b); });
  std::vector<int32_t> row(P.memories.size(), -1);
  const uint64_t fixed = uint64_t(slots) + P.total_signal_words;
  *used = 0;
  for (uint32_t m : order) {
    const uint64_t w = words(m);
    if (w == 0 || w > 1024 || (fixed + *used + w) * block * 4 > limit) continue;
    row[m] = int32_t(*used);
    *used += uint32_t(w);
  }
  return row;
}

// ------------------------------------------------------------------ lane compaction
// Between the launches of a halting run (rrtl_hip_options.compact_lanes).
// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2), so a
// lane stays in the partition (workgroup id mod 8) it started in: partition x
// owns workgroups x, x + 8, ... and its slot j runs lane map[x * cap + j] for
// j < live[x].  After a launch each partition keeps its lanes whose halt flag
// is still clear, in their order (a stable scan), so the live lanes fill the
// partition's first workgroups, the lanes sharing a cache line of word-major
// state stay together on one XCD, and the other workgroups exit at once.
constexpr uint32_t kParts = 8;

__global__ void lane_map_init_kernel(uint32_t* map, uint32_t* live, uint32_t lanes, uint32_t per_block,
                                     uint32_t blocks, uint32_t cap) {
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  if (b < blocks && t < per_block) {
    const uint32_t part = b % kParts, j = (b / kParts) * per_block + t, lane = b * per_block + t;
    map[size_t(part) * cap + j] = lane < lanes ? lane : 0xFFFFFFFFu;
  }
  if (b == 0 && t < kParts) live[t] = ((blocks + kParts - 1 - t) / kParts) * per_block;
}

// The kept lanes of a partition go to its slots in their order, but never onto
// fewer waves than the partition's XCD runs at once (one per SIMD, `floor_slots`
// slots): while the live lanes fit, they are spread over those waves (the first
// q slots of each), so a memory-bound design keeps every CU's memory path busy;
// only beyond that are the waves packed full.  Two passes: count, then place.
__global__ void __launch_bounds__(1024) lane_compact_kernel(const uint32_t* map_in, const uint32_t* live_in,
                                                            const uint64_t* halt, uint32_t* map_out,
                                                            uint32_t* live_out, uint32_t cap, uint32_t lanes,
                                                            uint32_t wave_slots, uint32_t floor_slots) {
  __shared__ uint32_t wave_keep[16];
  const uint32_t part = blockIdx.x, n = live_in[part];
  const uint32_t* in = map_in + size_t(part) * cap;
  uint32_t* out = map_out + size_t(part) * cap;
  const uint32_t w = threadIdx.x >> 6, lid = threadIdx.x & 63u;
  auto keep_of = [&](uint32_t j) {
    const uint32_t lane = j < n ? in[j] : 0xFFFFFFFFu;
    return lane < lanes && !(halt[lane] & 1u);
  };
  // pass 1: how many lanes stay
  uint32_t kept = 0;
  for (uint32_t c0 = 0; c0 < n; c0 += 1024u) {
    const uint64_t m = __builtin_amdgcn_ballot_w64(keep_of(c0 + threadIdx.x));
    if (lid == 0) wave_keep[w] = uint32_t(__popcll(m));
    __syncthreads();
    for (uint32_t k = 0; k < 16u; k++) kept += wave_keep[k];
    __syncthreads();
  }
  // slots of the next launch: whole waves, at least floor_slots, at most the current count
  uint32_t slots = (kept + wave_slots - 1) / wave_slots * wave_slots;
  slots = std::min(std::max(slots, floor_slots), (n + wave_slots - 1) / wave_slots * wave_slots);
  slots = std::min(slots, cap);
  const uint32_t waves = std::max(1u, slots / wave_slots);
  const uint32_t q = std::max(1u, (kept + waves - 1) / waves);  // live lanes per wave
  for (uint32_t j = threadIdx.x; j < slots; j += 1024u) out[j] = 0xFFFFFFFFu;
  __syncthreads();
  // pass 2: place kept lane k at the k % q-th slot of wave k / q
  uint32_t base = 0;
  for (uint32_t c0 = 0; c0 < n; c0 += 1024u) {
    const uint32_t j = c0 + threadIdx.x;
    const bool keep = keep_of(j);
    const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
    if (lid == 0) wave_keep[w] = uint32_t(__popcll(m));
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t k = 0; k < 16u; k++) {
      const uint32_t c = wave_keep[k];
      before += k < w ? c : 0u;
      total += c;
    }
    if (keep) {
      const uint32_t k = base + before + uint32_t(__popcll(m & ((uint64_t(1) << lid) - 1u)));
      out[(k / q) * wave_slots + k % q] = in[j];
    }
    __syncthreads();
    base += total;
  }
  if (threadIdx.x == 0) live_out[part] = slots;
}

// ------------------------------------------------------------------ engine
class WordEngine final : public Engine {
 public:
  WordEngine(Program prog, uint32_t lanes, const EngineOptions& opt) : Engine(std::move(prog), lanes, opt) {
    if (lanes == 0) throw Error("E_SIM_IR_LANES: packed simulator requires at least one lane");
    WP = WordCompiler(P).compile();
    // Automatic code generation is for designs whose straight-line code stays
    // small (compile time grows faster than linearly with its length); larger
    // designs run the record interpreter.  Generated code keeps its values in
    // registers and needs the scratch slots only for records it leaves in
    // interpreted form, so the scratch limit applies to those designs alone.
    size_t n_rec = 0;
    for (auto& b : WP.blk) n_rec += b.size();
    std::unique_ptr<WordJit> jit;
    if (opt.codegen == CODEGEN_ON || (opt.codegen == CODEGEN_AUTO && n_rec <= 4000)) jit.reset(new WordJit(WP));
    if (jit && opt.codegen == CODEGEN_AUTO && jit->uses_scratch() && WP.scratch_slots > 128) jit.reset();
    // the interpreter runs macro records in place of the fusion menu
    if (!jit && opt.interp_macros != 2) WP = WordCompiler(P, true).compile();
    run_slots_ = jit && !jit->uses_scratch() ? 0 : WP.scratch_slots;
    RRTL_HIP_CHECK(hipSetDevice(opt.device));
    hipDeviceProp_t prop;
    RRTL_HIP_CHECK(hipGetDeviceProperties(&prop, opt.device));
    lds_max_ = prop.sharedMemPerBlock;
    RRTL_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    values_.alloc(size_t(P.total_signal_words) * lanes); values_.zero(stream_);
    mems_.alloc(size_t(P.total_memory_words) * lanes); mems_.zero(stream_);
    // One lane per thread: blocks of 256 lanes, smaller (down to one wave)
    // when that is what it takes to give every CU a block.  Waves stay full (64
    // lanes) unless the options ask for fewer: thinner waves for batches that
    // leave SIMDs idle lost on every design measured with the current kernels
    // (the README counter at 4,096 lanes 8 %, the RV32I core at 16,384 lanes
    // 3.5 %, the mixed batch's counter and FIFO 20-34 %, picorv32 at 16 lanes
    // per wave 2.4e9 against 3.5e9 lane-cycles/s).
    const uint32_t cus = uint32_t(std::max(1, prop.multiProcessorCount));
    cus_ = cus;
    // Several waves per 64 lanes when the batch gives fewer than one 64-lane
    // wave per SIMD (four per CU): the generated code of each evaluation is
    // split between the waves of a workgroup, which share the lanes' LDS
    // columns (WordJit::split).  Automatic when the split is estimated to pay
    // a fifth or more.
    std::unique_ptr<WordJit> split_jit;
    if (jit && !jit->uses_scratch() && opt.lane_waves != 1) {
      const uint64_t blocks64 = (uint64_t(lanes) + 63) / 64, simds = 4ull * cus;
      uint32_t G = opt.lane_waves;
      if (!G) G = blocks64 * 4 <= simds ? 4 : (blocks64 * 2 <= simds ? 2 : 1);
      if (G > 1) {
        // lanes per workgroup (the LDS columns the waves share): 64, or fewer
        // when asked for (wave_lanes), which puts more workgroups on a CU
        const uint32_t S = opt.wave_lanes ? opt.wave_lanes : 64u;
        JitLayout lay;
        uint32_t mem_words = 0;
        lay.mem_row = lds_memory_rows(P, 0, S, &mem_words);
        lay.stride = S; lay.slots = 0; lay.sig_words = P.total_signal_words;
        std::unique_ptr<WordJit> sj(new WordJit(WP, std::move(lay)));
        const double gain = sj->plan_waves(G);
        const size_t rows = size_t(P.total_signal_words) + mem_words + sj->exchange_rows();
        const bool fits = rows * S * 4 <= std::min<size_t>(lds_max_, 64 * 1024);
        if (gain > 0 && fits && (opt.lane_waves || gain >= 1.2)) {
          lane_waves_ = G; lds_mem_words_ = mem_words; split_jit = std::move(sj);
          split_lds_ = rows * S * 4;
          split_lanes_ = S;
        } else if (opt.lane_waves) {
          throw Error(gain > 0 ? "E_HIP_OPTIONS: lane_waves: the split design does not fit in LDS"
                               : "E_HIP_OPTIONS: lane_waves: the generated code cannot be split between waves");
        }
      }
    } else if (opt.lane_waves > 1) {
      throw Error("E_HIP_OPTIONS: lane_waves needs generated code without interpreted records");
    }
    wave_lanes_ = opt.wave_lanes ? opt.wave_lanes : 64u;  // 4, 8, 16, 32 or 64
    if (lane_waves_ > 1) {
      wave_lanes_ = split_lanes_;
      block_ = 64 * lane_waves_;
    } else if (wave_lanes_ < 64) {
      block_ = 64;
    } else if (opt.block_threads) {
      block_ = opt.block_threads;
    } else {
      const uint32_t per_cu = (lanes + cus - 1) / cus;
      block_ = std::min<uint32_t>(256, std::max<uint32_t>(64, (per_cu + 63) / 64 * 64));
    }
    mode_ = lane_waves_ > 1 ? WM_ALL_LDS : lds_mode(block_, run_slots_);
    use_lds_ = mode_ != WM_HBM;
    if (lane_waves_ > 1) {
      run_slots_ = 0;
      jit_fn_ = rtc_kernel(split_jit->source(WM_ALL_LDS, P.total_signal_words + lds_mem_words_), "rrtl_word_jit",
                           opt.device);
      jit.reset();
    }
    if (jit) {
      // the final layout: LDS-staged memories and constant LDS offsets
      JitLayout lay;
      if (mode_ == WM_ALL_LDS) {
        lay.mem_row = lds_memory_rows(P, run_slots_, block_, &lds_mem_words_);
        lay.stride = block_; lay.slots = run_slots_; lay.sig_words = P.total_signal_words;
      }
      jit.reset(new WordJit(WP, std::move(lay)));
    }
    if (jit) jit_fn_ = rtc_kernel(jit->source(mode_), "rrtl_word_jit", opt.device);
    if (lane_waves_ > 1 && split_lds_ > 65536)
      throw Error("E_HIP_INTERNAL: split kernel LDS above 64 KiB");
    if (!use_lds_) { scratch_.alloc(size_t(WP.scratch_slots) * lanes); scratch_.zero(stream_); }
    for (int b = 0; b < 4; b++) {
      std::vector<uint32_t> flat;
      for (auto& r : WP.blk[b]) flat.insert(flat.end(), r.w, r.w + 8);
      blk_[b].upload(flat, stream_);
    }
    cpool_.upload(WP.cpool.empty() ? std::vector<uint32_t>{0} : WP.cpool, stream_);
    mtab_.upload(WP.mem_table.empty() ? std::vector<uint32_t>{0} : WP.mem_table, stream_);
    RRTL_HIP_CHECK(hipEventCreate(&ev0_)); RRTL_HIP_CHECK(hipEventCreate(&ev1_));
    launch(F_EVAL_FIRST, 0, nullptr);
    synchronize();
  }
  ~WordEngine() override {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    values_.release(); mems_.release(); scratch_.release(); cpool_.release(); mtab_.release();
    for (auto& b : blk_) b.release();
    for (auto& b : run_bufs_) b.release();
    digest_.release(); trace_.release(); halt_.release(); mm_count_.release(); mm_rec_.release();
    lane_map_.release(); live_.release();
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  uint32_t kind() const override { return ENGINE_WORD; }

  void set_signal_limbs(uint32_t sig, const uint32_t* lane_limbs, bool settle) override {
    const Signal& s = sig_at(sig);
    std::vector<uint32_t> rows(size_t(s.limbs) * lanes_);
    for (uint32_t l = 0; l < s.limbs; l++)
      for (uint32_t lane = 0; lane < lanes_; lane++) {
        uint32_t v = lane_limbs[size_t(l) * lanes_ + lane];
        rows[size_t(l) * lanes_ + lane] = l + 1 == s.limbs ? (v & final_mask(s.width)) : v;
      }
    RRTL_HIP_CHECK(hipMemcpyAsync(values_.p + size_t(s.offset) * lanes_, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, stream_));
    if (settle) eval_combinational(); else synchronize();
  }
  void get_signal_limbs(uint32_t sig, uint32_t* out) override {
    const Signal& s = sig_at(sig);
    RRTL_HIP_CHECK(hipMemcpyAsync(out, values_.p + size_t(s.offset) * lanes_, size_t(s.limbs) * lanes_ * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t lane = 0; lane < lanes_; lane++) out[size_t(s.limbs - 1) * lanes_ + lane] &= final_mask(s.width);
  }
  void set_memory_limbs(uint32_t mem, const uint32_t* vals, bool settle) override {
    const Memory& m = mem_at(mem);
    std::vector<uint32_t> rows(size_t(m.depth) * m.limbs * lanes_);
    for (uint32_t lane = 0; lane < lanes_; lane++)
      for (uint32_t d = 0; d < m.depth; d++)
        for (uint32_t l = 0; l < m.limbs; l++) {
          uint32_t v = vals[(size_t(lane) * m.depth + d) * m.limbs + l];
          if (l + 1 == m.limbs) v &= final_mask(m.width);
          rows[(size_t(d) * m.limbs + l) * lanes_ + lane] = v;
        }
    RRTL_HIP_CHECK(hipMemcpyAsync(mems_.p + size_t(m.offset) * lanes_, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, stream_));
    if (settle) eval_combinational(); else synchronize();
  }
  void get_memory_limbs(uint32_t mem, uint32_t* out) override {
    const Memory& m = mem_at(mem);
    std::vector<uint32_t> rows(size_t(m.depth) * m.limbs * lanes_);
    RRTL_HIP_CHECK(hipMemcpyAsync(rows.data(), mems_.p + size_t(m.offset) * lanes_, rows.size() * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t lane = 0; lane < lanes_; lane++)
      for (uint32_t d = 0; d < m.depth; d++)
        for (uint32_t l = 0; l < m.limbs; l++) out[(size_t(lane) * m.depth + d) * m.limbs + l] = rows[(size_t(d) * m.limbs + l) * lanes_ + lane];
  }
  void eval_combinational() override { launch(F_EVAL_FIRST, 0, nullptr); synchronize(); }
  void tick() override { tick_many(1); }
  void tick_from_evaluated() override { launch(F_EDGE, 1, nullptr); synchronize(); }
  void tick_many(uint64_t steps) override {
    if (!steps) return;
    RunSpec s; s.cycles = steps;
    launch(F_EVAL_FIRST | F_EDGE | F_EVAL_POST, steps, &s, false);
  }
  void run(const RunSpec& spec) override { launch(run_flags(spec), spec.cycles, &spec, true); }

  // The design's items in the generated batch kernel, in the first form that
  // fits the batch's LDS budget per 256-thread workgroup:
  //   generated code, 256 lanes per item, signals (and small memories) in LDS;
  //   generated code split between four waves per 64 lanes (values crossing
  //     waves through LDS rows), small memories staged in LDS or not;
  //   generated code with the state in HBM (scratch in LDS if it fits);
  // and the record interpreter when the engine runs without generated code.
  BatchJob batch_prepare(const RunSpec& spec, const BatchRequest& req) override {
    WordArgs A = make_args(run_flags(spec), spec.cycles, &spec, true);
    const size_t budget = std::min(req.lds_budget, lds_max_);
    const uint32_t sig = P.total_signal_words;
    size_t recs = 0;
    for (auto& b : WP.blk) recs += b.size();
    BatchJob J;
    J.kind = ENGINE_WORD;
    int mode = WM_HBM;
    uint32_t G = 1, per_item = 256, slots = 0;
    std::string exec = "rrtl::word::InterpExec";
    bool chosen = false;
    if (jit_fn_) {
      const bool scr = WordJit(WP).uses_scratch();
      slots = scr ? WP.scratch_slots : 0;
      if (size_t(slots + sig) * 256 * 4 <= budget) {
        JitLayout lay;
        uint32_t mw = 0;
        lay.mem_row = lds_memory_rows(P, slots, 256, &mw, budget);
        lay.stride = 256; lay.slots = slots; lay.sig_words = sig;
        J.code = WordJit(WP, std::move(lay)).definitions(0, req.ns);
        mode = WM_ALL_LDS;
        J.lds_bytes = size_t(slots + sig + mw) * 256 * 4;
        // a design whose lanes fit one 64-lane wave per CU runs one wave per
        // item instead: four times the items, each on a CU of its own, and a
        // long run no longer shares a CU's issue slots with three more of its
        // waves (with more lanes, one-wave items would leave workgroup slots
        // holding a single wave: the mixed batch's 63,654-lane counter ran
        // 1.7x slower that way)
        if ((uint64_t(lanes_) + 63) / 64 <= cus_) {
          per_item = 64;
          J.form = "word generated, 64 lanes per item, state in LDS";
        } else {
          J.form = "word generated, 256 lanes per item, state in LDS";
        }
        chosen = true;
      }
      for (int stage = 1; !chosen && !scr && stage >= 0; stage--) {
        JitLayout lay;
        uint32_t mw = 0;
        if (stage) lay.mem_row = lds_memory_rows(P, 0, 64, &mw, budget);
        lay.stride = 64; lay.slots = 0; lay.sig_words = sig;
        WordJit j(WP, std::move(lay));
        if (j.plan_waves(4) <= 0) break;
        const size_t rows = size_t(sig) + mw + j.exchange_rows();
        if (rows * 64 * 4 > budget) continue;
        J.code = j.definitions(sig + mw, req.ns);
        mode = WM_ALL_LDS; G = 4; per_item = 64;
        J.lds_bytes = rows * 64 * 4;
        J.form = std::string("word generated, four waves per 64 lanes, ") +
                 (mw ? "small memories staged in LDS" : "memories in HBM");
        chosen = true;
      }
      if (!chosen) {
        mode = size_t(slots) * 256 * 4 <= budget ? (slots ? WM_SCRATCH_LDS : WM_HBM) : WM_HBM;
        J.code = WordJit(WP).definitions(0, req.ns);
        J.lds_bytes = mode == WM_SCRATCH_LDS ? size_t(slots) * 256 * 4 : 0;
        J.form = "word generated, 256 lanes per item, state in HBM";
        chosen = true;
      }
      exec = "rrtl::word::" + req.ns + "::JitExec";
    } else {
      slots = WP.scratch_slots;
      mode = size_t(slots + sig) * 256 * 4 <= budget ? WM_ALL_LDS
             : size_t(slots) * 256 * 4 <= budget       ? WM_SCRATCH_LDS
                                                        : WM_HBM;
      J.lds_bytes = lds_bytes(mode, 256, slots);
      J.form = "word interpreter, 256 lanes per item";
    }
    if (mode == WM_HBM && slots && !scratch_.p) { scratch_.alloc(size_t(WP.scratch_slots) * lanes_); scratch_.zero(stream_); }
    A.n_scratch = slots;
    // every thread of an item runs a lane (G = 1, 256 lanes), one wave of it
    // (G = 1, 64 lanes), or 64 lanes per item (G = 4)
    A.block_lanes = (G == 1 && per_item == 64) ? 64u : 0u;
    A.scratch = mode == WM_HBM ? scratch_.p : nullptr;
    J.call = "rrtl::word::word_body<" + std::to_string(mode) + ", " + exec + ", " + std::to_string(G) + ">(U.w, item);";
    J.n_items = (lanes_ + per_item - 1) / per_item;
    // cost of one item: its waves' record walks per cycle (a split item runs
    // about twice as fast as one wave, state in HBM about half as fast)
    const double waves = double(per_item) / 64.0 / (G > 1 ? 2.0 : 1.0) * (mode == WM_HBM ? 2.0 : 1.0);
    J.item_cost = double(spec.cycles) * double(recs) * waves * (jit_fn_ ? 1.0 : 20.0);
    J.args.resize(sizeof(WordArgs));
    std::memcpy(J.args.data(), &A, sizeof(WordArgs));
    batch_spec_halt_ = has_halt_;
    batch_digest_ = has_digest_;
    RRTL_HIP_CHECK(hipStreamSynchronize(stream_));
    return J;
  }
  void batch_reset(hipStream_t s) override {
    if (batch_digest_) RRTL_HIP_CHECK(hipMemsetAsync(digest_.p, 0, size_t(lanes_) * 8, s));
    if (batch_spec_halt_) RRTL_HIP_CHECK(hipMemsetAsync(halt_.p, 0, size_t(lanes_) * 8, s));
  }
  void batch_finish(const RunSpec& spec) override {
    stats_ = RunStats{};
    stats_.lane_cycles = batch_spec_halt_ ? halted_lane_cycles() : uint64_t(lanes_) * spec.cycles;
  }
  void snapshot(uint32_t* values, uint32_t* memories) override {
    RRTL_HIP_CHECK(hipMemcpyAsync(values, values_.p, size_t(P.total_signal_words) * lanes_ * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t m = 0; m < P.memories.size(); m++) {
      const Memory& mm = P.memories[m];
      std::vector<uint32_t> lm(size_t(lanes_) * mm.depth * mm.limbs);
      get_memory_limbs(m, lm.data());
      for (uint32_t lane = 0; lane < lanes_; lane++)
        for (size_t k = 0; k < size_t(mm.depth) * mm.limbs; k++)
          memories[size_t(lane) * P.total_memory_words + mm.offset + k] = lm[size_t(lane) * mm.depth * mm.limbs + k];
    }
  }
  void restore(const uint32_t* values, const uint32_t* memories) override {
    RRTL_HIP_CHECK(hipMemcpyAsync(values_.p, values, size_t(P.total_signal_words) * lanes_ * 4, hipMemcpyHostToDevice, stream_));
    synchronize();
    for (uint32_t m = 0; m < P.memories.size(); m++) {
      const Memory& mm = P.memories[m];
      std::vector<uint32_t> lm(size_t(lanes_) * mm.depth * mm.limbs);
      for (uint32_t lane = 0; lane < lanes_; lane++)
        for (size_t k = 0; k < size_t(mm.depth) * mm.limbs; k++)
          lm[size_t(lane) * mm.depth * mm.limbs + k] = memories[size_t(lane) * P.total_memory_words + mm.offset + k];
      set_memory_limbs(m, lm.data(), false);
    }
  }
  void set_force(uint32_t, uint32_t, int) override { throw Error("E_HIP_UNSUPPORTED: stuck-at forces need the bit-sliced gate engine"); }
  void clear_forces() override {}
  size_t digest_bytes() const override { return has_digest_ ? size_t(lanes_) * 8 : 0; }
  void get_digest(void* out) override {
    if (!has_digest_) throw Error("E_HIP_NO_RUN: no run has produced a digest");
    RRTL_HIP_CHECK(hipMemcpyAsync(out, digest_.p, size_t(lanes_) * 8, hipMemcpyDeviceToHost, stream_));
    synchronize();
  }
  size_t trace_bytes() const override { return trace_bytes_; }
  void get_trace(void* out) override {
    if (!trace_bytes_) throw Error("E_HIP_NO_RUN: no run has produced a trace");
    RRTL_HIP_CHECK(hipMemcpyAsync(out, trace_.p, trace_bytes_, hipMemcpyDeviceToHost, stream_));
    synchronize();
  }
  void get_halt_cycles(uint64_t* out) override {
    if (!has_halt_) throw Error("E_HIP_NO_RUN: no run with a halt signal");
    RRTL_HIP_CHECK(hipMemcpyAsync(out, halt_.p, size_t(lanes_) * 8, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t l = 0; l < lanes_; l++) out[l] >>= 1;
  }
  uint64_t get_mismatches(uint32_t max, std::vector<Mismatch>& out) override {
    if (!mm_count_.p || !mm_lanes_) throw Error("E_HIP_NO_RUN: no run has checked expected values");
    std::vector<uint32_t> cnt(lanes_);
    RRTL_HIP_CHECK(hipMemcpy(cnt.data(), mm_count_.p, size_t(lanes_) * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> rec(size_t(mm_lanes_) * mm_max_ * 6);
    if (!rec.empty()) RRTL_HIP_CHECK(hipMemcpy(rec.data(), mm_rec_.p, rec.size() * 4, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    out.clear();
    for (uint32_t lane = 0; lane < mm_lanes_; lane++) {
      total += cnt[lane];
      for (uint32_t r = 0; r < std::min(cnt[lane], mm_max_); r++) {
        const uint32_t* x = &rec[(size_t(lane) * mm_max_ + r) * 6];
        Mismatch m;
        m.step = x[0]; m.check = x[1]; m.lane = lane;
        std::copy(x + 2, x + 6, m.actual);
        out.push_back(m);
      }
    }
    // the reference's report order: step, then check, then lane; each lane's
    // first records contain the overall first ones
    std::sort(out.begin(), out.end(), [](const Mismatch& a, const Mismatch& b) {
      return std::make_tuple(a.step, a.check, a.lane) < std::make_tuple(b.step, b.check, b.lane);
    });
    if (out.size() > max) out.resize(max);
    return total;
  }
  uint32_t scratch_slots() const { return run_slots_; }
  bool compiled() const { return jit_fn_ != nullptr; }
  void launch_info(uint32_t* out) const {
    out[0] = jit_fn_ ? 1u : 0u; out[1] = lane_waves_; out[2] = wave_lanes_; out[3] = block_;
  }
  bool scratch_in_lds() const { return use_lds_; }

 private:
  const Signal& sig_at(uint32_t s) const { if (s >= P.signals.size()) throw Error("E_SIM_IR_SIGNAL: signal index out of range"); return P.signals[s]; }
  const Memory& mem_at(uint32_t m) const { if (m >= P.memories.size()) throw Error("E_SIM_IR_MEMORY: memory index out of range"); return P.memories[m]; }

  template <class T>
  T* upload_run(size_t idx, const std::vector<T>& v) {
    auto& b = run_bufs_[idx];
    size_t bytes = std::max<size_t>(8, v.size() * sizeof(T));
    b.alloc((bytes + 3) / 4);
    if (!v.empty()) RRTL_HIP_CHECK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
    return reinterpret_cast<T*>(b.p);
  }

  // LDS bytes a block of `block` lanes with `slots` scratch slots (and, in
  // WM_ALL_LDS, `mem_words` words of LDS-staged memories) needs in `mode`.
  // Lanes one workgroup runs.
  uint32_t lanes_per_block() const { return lane_waves_ > 1 ? split_lanes_ : (wave_lanes_ < 64 ? wave_lanes_ : block_); }
  size_t lds_bytes(int mode, uint32_t block, uint32_t slots, uint32_t mem_words = 0) const {
    if (mode == WM_HBM) return 0;
    const size_t words = slots + (mode == WM_ALL_LDS ? size_t(P.total_signal_words) + mem_words : 0);
    return words * block * 4;
  }
  // Signals join scratch in LDS when both fit in 64 KiB per block.
  int lds_mode(uint32_t block, uint32_t slots) const {
    if (lds_bytes(WM_ALL_LDS, block, slots) <= 64 * 1024) return WM_ALL_LDS;
    if (lds_bytes(WM_SCRATCH_LDS, block, slots) <= 64 * 1024) return WM_SCRATCH_LDS;
    return WM_HBM;
  }

  WordArgs make_args(uint32_t flags, uint64_t cycles, const RunSpec* spec, bool outputs) {
    WordArgs A{};
    A.n_scratch = run_slots_; A.n_sig_words = P.total_signal_words;
    A.block_lanes = wave_lanes_ < 64 ? wave_lanes_ : 0;
    A.values = values_.p; A.mems = mems_.p; A.scratch = use_lds_ ? nullptr : scratch_.p;
    for (int b = 0; b < 4; b++) { A.blk[b] = reinterpret_cast<const Rec*>(blk_[b].p); A.nrec[b] = uint32_t(WP.blk[b].size()); }
    A.cpool = cpool_.p; A.mem_table = mtab_.p; A.lanes = lanes_; A.flags = flags; A.cycles = cycles;
    A.halt_offset = NONE;
    if (spec && outputs) {
      if (spec->detect) throw Error("E_HIP_UNSUPPORTED: fault detection runs on the bit-sliced gate engine");
      for (uint32_t m : spec->in_mode)
        if (m == IN_GROUP_RANDOM) throw Error("E_HIP_UNSUPPORTED: group-shared random inputs run on the bit-sliced gate engine");
      A.seed = spec->seed;
      // Input descriptors: (offset, limbs, width, mode | trace limb base << 8).
      // Constant inputs driven by one entry go first (written once per launch),
      // the others keep their order; held inputs write nothing.
      std::vector<uint32_t> in_desc, var_desc; std::vector<uint64_t> in_arg, var_arg;
      uint32_t trace_limbs = 0, uni_limbs = 0;
      std::unordered_map<uint32_t, uint32_t> entries;
      for (uint32_t sg : spec->in_sig) entries[sg]++;
      for (size_t i = 0; i < spec->in_sig.size(); i++) {
        const Signal& s = sig_at(spec->in_sig[i]);
        uint32_t mode = spec->in_mode[i];
        uint64_t arg = spec->in_arg.empty() ? 0 : spec->in_arg[i];
        if (mode > IN_TRACE_UNIFORM) throw Error("E_HIP_RUN_SPEC: unknown input mode");
        uint32_t base = 0;
        if (mode == IN_RANDOM) {
          if (s.width > 256) throw Error("E_HIP_WIDE: random stimulus is limited to 256-bit inputs");
          arg = uint64_t(spec->in_sig[i]) << 40;
        }
        if (mode == IN_TRACE) { base = trace_limbs; trace_limbs += s.limbs; }
        if (mode == IN_TRACE_UNIFORM) { base = uni_limbs; uni_limbs += s.limbs; }
        if (mode == IN_HOLD) continue;
        if (base >= (1u << 24)) throw Error("E_HIP_RUN_SPEC: too many trace input limbs");
        const bool once = mode == IN_CONST && entries[spec->in_sig[i]] == 1;
        auto& d = once ? in_desc : var_desc;
        d.insert(d.end(), {s.offset, s.limbs, s.width, mode | (base << 8)});
        (once ? in_arg : var_arg).push_back(arg);
      }
      A.n_in_once = uint32_t(in_desc.size() / 4);
      in_desc.insert(in_desc.end(), var_desc.begin(), var_desc.end());
      in_arg.insert(in_arg.end(), var_arg.begin(), var_arg.end());
      A.in_desc = upload_run(0, in_desc); A.in_arg = upload_run(1, in_arg); A.n_in = uint32_t(in_desc.size() / 4);
      if (trace_limbs) {
        if (!spec->trace_inputs) throw Error("E_HIP_RUN_SPEC: trace inputs requested without a trace buffer");
        std::vector<uint32_t> tin(spec->trace_inputs, spec->trace_inputs + size_t(cycles) * trace_limbs * lanes_);
        A.trace_in = upload_run(2, tin); A.trace_in_limbs = trace_limbs;
      }
      if (uni_limbs) {
        if (!spec->trace_uniform) throw Error("E_HIP_RUN_SPEC: uniform trace inputs requested without their values");
        std::vector<uint32_t> tu(spec->trace_uniform, spec->trace_uniform + size_t(cycles) * uni_limbs);
        A.trace_uni = upload_run(4, tu); A.trace_uni_limbs = uni_limbs;
      }
      std::vector<uint32_t> obs;
      uint32_t obs_limbs = 0;
      for (uint32_t s : spec->obs_sig) { const Signal& g = sig_at(s); obs.insert(obs.end(), {g.offset, g.limbs}); obs_limbs += g.limbs; }
      A.obs_desc = upload_run(3, obs); A.n_obs = uint32_t(spec->obs_sig.size()); A.n_obs_limbs = obs_limbs;
      has_digest_ = A.n_obs > 0;
      if (has_digest_) { digest_.alloc(lanes_); digest_.zero(stream_); A.digest = digest_.p; }
      trace_bytes_ = 0;
      if (spec->trace && A.n_obs) { trace_bytes_ = size_t(cycles) * obs_limbs * lanes_ * 4; trace_.alloc(trace_bytes_ / 4); A.trace = trace_.p; }
      mm_lanes_ = 0;
      if (spec->expect) {
        for (uint32_t s : spec->obs_sig)
          if (sig_at(s).limbs > 4) throw Error("E_HIP_WIDE: checked signals are limited to 128 bits");
        const size_t per = size_t(obs_limbs) * (spec->expect_per_lane ? lanes_ : 1u);
        std::vector<uint32_t> ex(spec->expect, spec->expect + size_t(cycles) * per);
        A.expect = upload_run(5, ex); A.expect_per_lane = spec->expect_per_lane ? 1u : 0u;
        if (spec->check_active) {
          A.check_words = (A.n_obs + 31) / 32;
          std::vector<uint32_t> ca(spec->check_active, spec->check_active + size_t(cycles) * A.check_words);
          A.check_active = upload_run(6, ca);
        }
        A.check_lanes = spec->check_lane0 ? 1u : 0xFFFFFFFFu;
        mm_lanes_ = spec->check_lane0 ? 1u : lanes_;
        if (cycles >= (uint64_t(1) << 32)) throw Error("E_HIP_RUN_SPEC: a checked run is limited to 2^32 - 1 steps");
        // a lane cannot mismatch more often than steps x checks
        mm_max_ = uint32_t(std::min<uint64_t>(spec->max_mismatches, cycles * A.n_obs));
        if (uint64_t(mm_lanes_) * mm_max_ * 24 > (uint64_t(1) << 31))
          throw Error("E_HIP_TOO_LARGE: mismatch records for every checked lane exceed 2 GiB; lower max_mismatches");
        mm_count_.alloc(lanes_); mm_count_.zero(stream_);
        mm_rec_.alloc(std::max<size_t>(1, size_t(mm_lanes_) * mm_max_ * 6));
        A.mm_count = mm_count_.p; A.mm_rec = mm_rec_.p; A.mm_max = mm_max_;
      }
      has_halt_ = spec->halt_sig != NONE;
      if (has_halt_) {
        const Signal& h = sig_at(spec->halt_sig);
        A.halt_offset = h.offset;
        halt_.alloc(lanes_); halt_.zero(stream_);
        A.halt_cycles = halt_.p;
      }
    }
    return A;
  }

  uint32_t run_flags(const RunSpec& spec) const {
    uint32_t flags = F_EVAL_PRE | F_EDGE;
    bool need_post = spec.observe_post || (spec.halt_sig != NONE && P.signals.at(spec.halt_sig).kind != K_REG);
    if (need_post) flags |= F_EVAL_POST; else if (spec.final_eval) flags |= F_FINAL_EVAL;
    if (spec.observe_post) flags |= F_OBS_POST;
    return flags;
  }

  // Slots per XCD partition in a lane map of a compacting run.
  uint32_t compact_cap() const {
    const uint32_t per_block = lanes_per_block();
    const uint32_t grid = (lanes_ + per_block - 1) / per_block;
    return (grid + kParts - 1) / kParts * per_block;
  }
  // The two lane maps and live counts of a compacting run, allocated ahead of
  // enqueue() (whose launches a batch captures into a graph).
  void reserve_compaction() {
    if (!opt_.compact_lanes) return;
    lane_map_.alloc(size_t(compact_cap()) * kParts * 2);
    live_.alloc(kParts * 2);
  }

  // The run's kernels on stream s (kernels only: no copies, fills, allocations
  // or host waits, so the sequence can be captured into a graph).  Returns the launches.
  uint64_t enqueue(const WordArgs& A, uint64_t cycles, bool outputs, hipStream_t s) {
    // lane compaction: halting runs cut into launches, live lanes packed between them
    const bool compact = opt_.compact_lanes && outputs && A.halt_offset != NONE && A.halt_cycles;
    uint64_t chunk = (opt_.ticks_per_launch && outputs) ? opt_.ticks_per_launch : std::max<uint64_t>(cycles, 1);
    // automatic launches of a compacting run: 1,024 cycles, then each twice the
    // last, so a long tail of a few live lanes costs a handful of launches
    const bool doubling = compact && !opt_.ticks_per_launch;
    if (doubling) chunk = 1024;
    const uint32_t per_block = lanes_per_block();
    const uint32_t grid = (lanes_ + per_block - 1) / per_block;
    // the engine's own launches (batch items size theirs in batch_prepare)
    const size_t lds = lane_waves_ > 1 ? split_lds_ : lds_bytes(mode_, block_, run_slots_, lds_mem_words_);
    const uint32_t cap = compact_cap();  // slots per partition
    if (compact) {
      // allocated by reserve_compaction() before any launch is issued: these launches may be captured
      if (lane_map_.n < size_t(cap) * kParts * 2 || live_.n < kParts * 2)
        throw Error("E_HIP_INTERNAL: lane maps not reserved");
      hipLaunchKernelGGL(lane_map_init_kernel, dim3(grid), dim3(std::max<uint32_t>(per_block, kParts)), 0, s,
                         lane_map_.p, live_.p, lanes_, per_block, grid, cap);
      RRTL_HIP_CHECK(hipGetLastError());
    }
    uint64_t done = 0, launches = 0;
    do {
      WordArgs B = A;
      B.cycle_base = done;
      B.cycles = std::min<uint64_t>(chunk, cycles - done);
      if (done) B.flags &= ~uint32_t(F_EVAL_FIRST);
      // with compaction a lane may leave after any launch: every launch ends with the final settle
      if (done + B.cycles < cycles && !compact) B.flags &= ~uint32_t(F_FINAL_EVAL);
      const uint32_t cur = uint32_t(launches & 1u);
      if (compact) {
        B.lane_map = lane_map_.p + size_t(cur) * cap * kParts; B.live = live_.p + cur * kParts; B.map_cap = cap;
      }
      if (A.trace) B.trace = A.trace + done * A.n_obs_limbs * lanes_;
      if (A.trace_in) B.trace_in = A.trace_in + done * A.trace_in_limbs * lanes_;
      if (A.trace_uni) B.trace_uni = A.trace_uni + done * A.trace_uni_limbs;
      if (A.expect) B.expect = A.expect + done * A.n_obs_limbs * (A.expect_per_lane ? lanes_ : 1u);
      if (A.check_active) B.check_active = A.check_active + done * A.check_words;
      if (jit_fn_) {
        void* params[] = {&B};
        RRTL_HIP_CHECK(hipModuleLaunchKernel(jit_fn_, grid, 1, 1, block_, 1, 1, uint32_t(lds), s, params, nullptr));
      } else if (mode_ == WM_ALL_LDS) hipLaunchKernelGGL(word_kernel<WM_ALL_LDS>, dim3(grid), dim3(block_), lds, s, B);
      else if (mode_ == WM_SCRATCH_LDS) hipLaunchKernelGGL(word_kernel<WM_SCRATCH_LDS>, dim3(grid), dim3(block_), lds, s, B);
      else hipLaunchKernelGGL(word_kernel<WM_HBM>, dim3(grid), dim3(block_), 0, s, B);
      RRTL_HIP_CHECK(hipGetLastError());
      if (compact && done + B.cycles < cycles) {
        const uint32_t nxt = cur ^ 1u;
        const uint32_t wave_slots = std::min<uint32_t>(per_block, 64u);
        // a wave per SIMD (a workgroup per CU when its waves share the lanes)
        const uint32_t floor_slots = std::max<uint32_t>(1u, cus_ / kParts) * (lane_waves_ > 1 ? 1u : 4u) * wave_slots;
        hipLaunchKernelGGL(lane_compact_kernel, dim3(kParts), dim3(1024), 0, s, lane_map_.p + size_t(cur) * cap * kParts,
                           live_.p + cur * kParts, A.halt_cycles, lane_map_.p + size_t(nxt) * cap * kParts,
                           live_.p + nxt * kParts, cap, lanes_, wave_slots, floor_slots);
        RRTL_HIP_CHECK(hipGetLastError());
      }
      launches++;
      done += B.cycles;
      if (doubling) chunk *= 2;
    } while (done < cycles);
    return launches;
  }
  // Simulated lane-cycles of the last run: per-lane halt counts when halting.
  uint64_t halted_lane_cycles() {
    std::vector<uint64_t> h(lanes_);
    RRTL_HIP_CHECK(hipMemcpy(h.data(), halt_.p, size_t(lanes_) * 8, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (auto x : h) total += x >> 1;
    return total;
  }

  void launch(uint32_t flags, uint64_t cycles, const RunSpec* spec, bool outputs = false) {
    const WordArgs A = make_args(flags, cycles, spec, outputs);
    reserve_compaction();
    stats_ = RunStats{};
    RRTL_HIP_CHECK(hipEventRecord(ev0_, stream_));
    stats_.launches = enqueue(A, cycles, outputs, stream_);
    RRTL_HIP_CHECK(hipEventRecord(ev1_, stream_));
    RRTL_HIP_CHECK(hipEventSynchronize(ev1_));
    float ms = 0;
    RRTL_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
    stats_.kernel_ms = ms;
    stats_.lane_cycles = uint64_t(lanes_) * cycles;
    if (has_halt_ && spec && outputs) stats_.lane_cycles = halted_lane_cycles();
  }

  void async_prepare(const RunSpec& spec) override {
    async_args_ = make_args(run_flags(spec), spec.cycles, &spec, true);
    async_cycles_ = spec.cycles;
    reserve_compaction();
    batch_spec_halt_ = has_halt_;
    batch_digest_ = has_digest_;
    RRTL_HIP_CHECK(hipStreamSynchronize(stream_));
  }
  uint64_t async_enqueue(hipStream_t s) override { return enqueue(async_args_, async_cycles_, true, s); }
  void async_finish(const RunSpec& spec) override {
    stats_ = RunStats{};
    stats_.lane_cycles = batch_spec_halt_ ? halted_lane_cycles() : uint64_t(lanes_) * spec.cycles;
  }

  WordProgram WP;
  WordArgs async_args_{};               // the run of a concurrent batch
  uint64_t async_cycles_ = 0;
  uint32_t run_slots_ = 0;              // scratch slots of the engine's own launches
  uint32_t lds_mem_words_ = 0;          // words per lane of memories staged in LDS by the generated kernel
  hipFunction_t jit_fn_ = nullptr;      // the design's generated kernel (null: record interpreter)
  size_t lds_max_ = 65536;
  uint32_t block_ = 256;
  uint32_t wave_lanes_ = 64;            // active lanes per 64-thread wave (16, 32 or 64)
  uint32_t lane_waves_ = 1;             // waves per 64 lanes (generated code split between them)
  uint32_t split_lanes_ = 64;           // split kernel: lanes per workgroup (its LDS columns)
  size_t split_lds_ = 0;                // LDS per workgroup of the split kernel
  int mode_ = WM_SCRATCH_LDS;
  bool use_lds_ = true, has_digest_ = false, has_halt_ = false, batch_spec_halt_ = false, batch_digest_ = false;
  size_t trace_bytes_ = 0;
  DevBuf<uint32_t> values_, mems_, scratch_, cpool_, mtab_, blk_[4], run_bufs_[8], trace_, mm_count_, mm_rec_;
  uint32_t mm_lanes_ = 0, mm_max_ = 0;  // lanes that checked in the last replay, records kept per lane
  DevBuf<uint64_t> digest_, halt_;
  DevBuf<uint32_t> lane_map_, live_;    // lane compaction: two lane maps and their live counts
  uint32_t cus_ = 256;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace

std::string word_kernel_source(const Program& prog, int mode, uint32_t block) {
  if (mode < WM_HBM || mode > WM_ALL_LDS) throw Error("E_HIP_OPTIONS: storage mode must be 0, 1 or 2");
  if (block % 64 || block > 256) throw Error("E_HIP_OPTIONS: block_threads must be 0, 64, 128, 192 or 256");
  const WordProgram wp = WordCompiler(prog).compile();
  if (mode == WM_ALL_LDS && block) {
    JitLayout lay;
    lay.slots = WordJit(wp).uses_scratch() ? wp.scratch_slots : 0;
    uint32_t used = 0;
    lay.mem_row = lds_memory_rows(prog, lay.slots, block, &used);
    lay.stride = block; lay.sig_words = prog.total_signal_words;
    return WordJit(wp, std::move(lay)).source(mode);
  }
  return WordJit(wp).source(mode);
}

// The kernel the engine generates for `lane_waves` waves per 64 lanes, with a
// first comment line giving the split's estimated speed-up over one wave, its
// barrier phases and exchange rows.
std::string word_split_kernel_source(const Program& prog, uint32_t lane_waves) {
  if (lane_waves != 2 && lane_waves != 4) throw Error("E_HIP_OPTIONS: lane_waves must be 2 or 4");
  const WordProgram wp = WordCompiler(prog).compile();
  if (WordJit(wp).uses_scratch()) throw Error("E_HIP_OPTIONS: lane_waves needs generated code without interpreted records");
  JitLayout lay;
  uint32_t mem_words = 0;
  lay.mem_row = lds_memory_rows(prog, 0, 64, &mem_words);
  lay.stride = 64; lay.sig_words = prog.total_signal_words;
  WordJit j(wp, std::move(lay));
  const double gain = j.plan_waves(lane_waves);
  if (gain <= 0) throw Error("E_HIP_OPTIONS: lane_waves: the generated code cannot be split between waves");
  char head[160];
  snprintf(head, sizeof head, "// split: %u waves, estimated speed-up %.3f, exchange rows %u, phases %u/%u\n",
           lane_waves, gain, j.exchange_rows(), j.phases(0), j.phases(1));
  return head + j.source(WM_ALL_LDS, prog.total_signal_words + mem_words);
}

void word_program_info(const Program& prog, bool macros, uint32_t* out4) {
  const WordProgram wp = WordCompiler(prog, macros).compile();
  uint32_t recs = 0, mrec = 0, mops = 0;
  for (auto& b : wp.blk)
    for (const Rec& r : b) {
      recs++;
      if ((r.w[0] & 0xFF) == R_MACRO) { mrec++; mops += r.w[3]; }
    }
  out4[0] = recs; out4[1] = mrec; out4[2] = mops; out4[3] = wp.scratch_slots;
}

Engine* make_word_engine(Program prog, uint32_t lanes, const EngineOptions& opt) {
  return new WordEngine(std::move(prog), lanes, opt);
}

void word_engine_launch_info(Engine* e, uint32_t* info) { static_cast<WordEngine*>(e)->launch_info(info); }

}  // namespace rrtl
