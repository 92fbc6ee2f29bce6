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
#include <algorithm>
#include <map>
#include <unordered_map>

namespace rrtl {
namespace {

constexpr uint32_t MAXL = 8;           // widest value: 256 bits
constexpr uint32_t SP_SCRATCH = 0u << 30, SP_SIGNAL = 1u << 30, SP_CONST = 2u << 30, SP_MASK = 3u << 30;
constexpr uint32_t NONE = 0xFFFFFFFFu;

// record ops: machine ops keep their numbers, effects follow
enum : uint32_t { R_STORE = 32, R_ASYNC_CAPTURE, R_CAPTURE, R_COMMIT, R_MEMWRITE, R_SIGLOAD };

struct Rec { uint32_t w[8]; };  // op|limbs<<8, dst, a, b, c, width, aux0, aux1

struct WordProgram {
  std::vector<Rec> blk[4];         // async, comb, next (captures), commit (mem writes + reg commits)
  std::vector<uint32_t> cpool;
  uint32_t scratch_slots = 0;
  std::vector<uint32_t> mem_table; // per memory: offset, depth, limbs, width
};

struct Operand { uint32_t code; uint32_t limbs; uint32_t width; };

class WordCompiler {
 public:
  explicit WordCompiler(const Program& P) : P(P) {
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

  void lower(int b) {
    const Block& bk = P.blocks[b];
    std::vector<Operand> val(bk.n_values, {NONE, 0, 0});
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
        pos++;
      }
      for (const Effect& ef : pk.effects) {
        if (ef.kind == EF_MEMWRITE) { last_use[ef.enable] = pos; last_use[ef.addr] = pos; last_use[ef.data] = pos; }
        else last_use[ef.value] = pos;
        pos++;
      }
    }
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
        Rec r{};
        r.w[0] = in.op | (L << 8);
        r.w[5] = in.width;
        auto opnd = [&](uint32_t v) -> const Operand& { return val[v]; };
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
          case OP_CONCAT: {
            uint32_t at = uint32_t(out.cpool.size());
            for (uint32_t a : in.args) { out.cpool.push_back(opnd(a).code); out.cpool.push_back(opnd(a).width | (opnd(a).limbs << 16)); }
            r.w[2] = at; r.w[3] = uint32_t(in.args.size()); break;
          }
          case OP_MEMREAD:
            r.w[6] = in.args[0]; r.w[2] = opnd(in.args[1]).code; r.w[7] = opnd(in.args[1]).limbs; break;
          default:  // binary
            r.w[2] = opnd(in.args[0]).code; r.w[3] = opnd(in.args[1]).code;
            r.w[7] = opnd(in.args[0]).limbs | (opnd(in.args[1]).limbs << 8);
            r.w[6] = std::max(opnd(in.args[0]).width, opnd(in.args[1]).width);
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
  std::map<std::pair<uint32_t, std::vector<uint32_t>>, uint32_t> cmemo_;
  std::vector<Rec> commits_;
  uint32_t capture_limbs_ = 0, max_slot_ = 0;
};

// ------------------------------------------------------------------ kernel
enum : uint32_t { F_EVAL_FIRST = 1, F_EVAL_PRE = 2, F_EVAL_POST = 4, F_FINAL_EVAL = 8, F_OBS_POST = 16, F_EDGE = 32 };

struct WordArgs {
  uint32_t* values; uint32_t* mems; uint32_t* scratch;
  const Rec* blk[4]; uint32_t nrec[4];
  const uint32_t* cpool; const uint32_t* mem_table;
  uint32_t lanes, flags;
  uint64_t cycles, cycle_base, seed;
  // stimulus: per input (offset, limbs, width, mode), arg, stream base
  const uint32_t* in_desc; const uint64_t* in_arg; uint32_t n_in;
  const uint32_t* trace_in; uint32_t trace_in_limbs;      // [cycles][trace_in_limbs][lanes]
  const uint32_t* obs_desc; uint32_t n_obs, n_obs_limbs;  // per observed: (offset, limbs)
  uint64_t* digest; uint32_t* trace;                      // [lanes] / [cycles][n_obs_limbs][lanes]
  uint32_t halt_offset; uint64_t* halt_cycles;            // NONE = no halting
};

struct Ctx {
  const WordArgs* A;
  uint32_t lane;
  uint32_t* scr; uint32_t sstride;
  __device__ __forceinline__ uint32_t ld(uint32_t code, uint32_t limb) const {
    const uint32_t idx = code & ~SP_MASK;
    switch (code & SP_MASK) {
      case SP_SCRATCH: return scr[size_t(idx + limb) * sstride];
      case SP_SIGNAL: return A->values[size_t(idx + limb) * A->lanes + lane];
      default: return A->cpool[idx + limb];
    }
  }
  // operand of `have` limbs zero-extended to `n`
  __device__ __forceinline__ void ldv(uint32_t code, uint32_t have, uint32_t n, uint32_t* v) const {
    for (uint32_t l = 0; l < n; l++) v[l] = l < have ? ld(code, l) : 0u;
  }
  __device__ __forceinline__ void st(uint32_t slot, uint32_t limb, uint32_t v) const { scr[size_t(slot + limb) * sstride] = v; }
  __device__ __forceinline__ void stsig(uint32_t off, uint32_t limb, uint32_t v) const { A->values[size_t(off + limb) * A->lanes + lane] = v; }
};

__device__ __forceinline__ uint32_t fmask(uint32_t w) { uint32_t r = w & 31; return r ? ((1u << r) - 1) : 0xFFFFFFFFu; }
__device__ __forceinline__ uint32_t get_bit(const uint32_t* v, uint32_t n, uint32_t bit) { uint32_t l = bit >> 5; return l < n ? (v[l] >> (bit & 31)) & 1u : 0u; }
__device__ __forceinline__ uint64_t decode_usize(const uint32_t* v, uint32_t n) {
  uint64_t out = v[0] | (n > 1 ? uint64_t(v[1]) << 32 : 0);
  for (uint32_t l = 2; l < n; l++) if (v[l]) return ~0ull;
  return out;
}

__device__ void exec_rec(const Ctx& C, const Rec& R) {
  const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF, width = R.w[5];
  uint32_t a[MAXL], b[MAXL], r[MAXL];
  if (L == 1 && op <= OP_MEMREAD && op != OP_CONCAT && op != OP_MEMREAD && op != OP_SLICE && op != OP_EQ && op != OP_NE &&
      op != OP_LT && op != OP_SEXT) {
    // one-limb fast path
    const uint32_t x = C.ld(R.w[2], 0);
    uint32_t v;
    switch (op) {
      case OP_NOT: v = ~x; break;
      case OP_AND: v = x & C.ld(R.w[3], 0); break;
      case OP_OR: v = x | C.ld(R.w[3], 0); break;
      case OP_XOR: v = x ^ C.ld(R.w[3], 0); break;
      case OP_ADD: v = x + C.ld(R.w[3], 0); break;
      case OP_SUB: v = x - C.ld(R.w[3], 0); break;
      case OP_MUL: v = x * C.ld(R.w[3], 0); break;
      case OP_MUX: v = (C.ld(R.w[4], 0) & 1u) ? x : C.ld(R.w[3], 0); break;
      default: v = x; break;  // ZEXT / TRUNC / CAST
    }
    C.st(R.w[1], 0, v & fmask(width));
    return;
  }
  switch (op) {
    case R_SIGLOAD: {
      C.ldv(R.w[2], R.w[7], L, r);
      break;
    }
    case OP_NOT: C.ldv(R.w[2], R.w[7], L, a); for (uint32_t l = 0; l < L; l++) r[l] = ~a[l]; break;
    case OP_AND: case OP_OR: case OP_XOR: case OP_ADD: case OP_SUB: case OP_MUL: {
      C.ldv(R.w[2], R.w[7] & 0xFF, L, a);
      C.ldv(R.w[3], (R.w[7] >> 8) & 0xFF, L, b);
      if (op == OP_AND) for (uint32_t l = 0; l < L; l++) r[l] = a[l] & b[l];
      else if (op == OP_OR) for (uint32_t l = 0; l < L; l++) r[l] = a[l] | b[l];
      else if (op == OP_XOR) for (uint32_t l = 0; l < L; l++) r[l] = a[l] ^ b[l];
      else if (op == OP_ADD) { uint64_t c = 0; for (uint32_t l = 0; l < L; l++) { uint64_t t = uint64_t(a[l]) + b[l] + c; r[l] = uint32_t(t); c = t >> 32; } }
      else if (op == OP_SUB) { uint64_t br = 0; for (uint32_t l = 0; l < L; l++) { uint64_t s = uint64_t(b[l]) + br, x = a[l]; r[l] = uint32_t(x - s); br = x < s; } }
      else {
        for (uint32_t l = 0; l < L; l++) r[l] = 0;
        for (uint32_t j = 0; j < L; j++) {
          uint64_t c = 0;
          for (uint32_t k = 0; j + k < L; k++) { uint64_t t = uint64_t(a[j]) * b[k] + r[j + k] + c; r[j + k] = uint32_t(t); c = t >> 32; }
        }
      }
      break;
    }
    case OP_EQ: case OP_NE: {
      uint32_t la = R.w[7] & 0xFF, lb = (R.w[7] >> 8) & 0xFF, n = la > lb ? la : lb;
      uint32_t eq = 1;
      for (uint32_t l = 0; l < n; l++) eq &= (l < la ? C.ld(R.w[2], l) : 0u) == (l < lb ? C.ld(R.w[3], l) : 0u);
      r[0] = op == OP_EQ ? eq : (eq ^ 1u);
      for (uint32_t l = 1; l < L; l++) r[l] = 0;
      break;
    }
    case OP_LT: {
      const uint32_t wl = R.w[6], n = (wl + 31) / 32;
      C.ldv(R.w[2], R.w[7] & 0xFF, n, a);
      C.ldv(R.w[3], (R.w[7] >> 8) & 0xFF, n, b);
      a[n - 1] &= fmask(wl); b[n - 1] &= fmask(wl);
      uint32_t res = 0, decided = 0;
      if (R.w[7] >> 24) {
        uint32_t sa = get_bit(a, n, wl - 1), sb = get_bit(b, n, wl - 1);
        if (sa != sb) { res = sa; decided = 1; }
      }
      for (int l = int(n) - 1; l >= 0 && !decided; l--) if (a[l] != b[l]) { res = a[l] < b[l]; decided = 1; }
      r[0] = res;
      for (uint32_t l = 1; l < L; l++) r[l] = 0;
      break;
    }
    case OP_MUX: {
      const bool c = C.ld(R.w[4], 0) & 1u;
      C.ldv(c ? R.w[2] : R.w[3], c ? (R.w[7] & 0xFF) : ((R.w[7] >> 8) & 0xFF), L, r);
      break;
    }
    case OP_SLICE: {
      const uint32_t n = R.w[7], lsb = R.w[6];
      C.ldv(R.w[2], n, n, a);
      for (uint32_t l = 0; l < L; l++) {
        const uint32_t bit = lsb + 32 * l, li = bit >> 5, sh = bit & 31;
        uint32_t lo = li < n ? a[li] : 0u, hi = li + 1 < n ? a[li + 1] : 0u;
        r[l] = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
      }
      break;
    }
    case OP_ZEXT: case OP_TRUNC: case OP_CAST: C.ldv(R.w[2], R.w[7], L, r); break;
    case OP_SEXT: {
      const uint32_t iw = R.w[6], n = R.w[7];
      C.ldv(R.w[2], n, L, r);
      const uint32_t sign = get_bit(r, L, iw - 1);
      // clear above iw, then fill with the sign
      for (uint32_t l = 0; l < L; l++) {
        uint32_t lo_bit = 32 * l;
        if (lo_bit >= iw) r[l] = sign ? ~0u : 0u;
        else if (lo_bit + 32 > iw) {
          uint32_t keep = (1u << (iw - lo_bit)) - 1;
          r[l] = (r[l] & keep) | (sign ? ~keep : 0u);
        }
      }
      break;
    }
    case OP_CONCAT: {
      for (uint32_t l = 0; l < L; l++) r[l] = 0;
      uint32_t off = 0;
      for (int k = int(R.w[3]) - 1; k >= 0; k--) {
        const uint32_t code = C.A->cpool[R.w[2] + 2 * k], wd = C.A->cpool[R.w[2] + 2 * k + 1] & 0xFFFF,
                       n = C.A->cpool[R.w[2] + 2 * k + 1] >> 16;
        for (uint32_t l = 0; l < n; l++) {
          uint32_t v = C.ld(code, l);
          if (l == n - 1) v &= fmask(wd);
          const uint32_t bit = off + 32 * l, li = bit >> 5, sh = bit & 31;
          if (li < L) r[li] |= v << sh;
          if (sh && li + 1 < L) r[li + 1] |= v >> (32 - sh);
        }
        off += wd;
      }
      break;
    }
    case OP_MEMREAD: {
      const uint32_t* mt = &C.A->mem_table[4 * R.w[6]];
      C.ldv(R.w[2], R.w[7], R.w[7], a);
      const uint64_t addr = decode_usize(a, R.w[7]);
      for (uint32_t l = 0; l < L; l++) r[l] = 0;
      if (addr < mt[1]) {
        for (uint32_t l = 0; l < mt[2] && l < L; l++)
          r[l] = C.A->mems[(size_t(mt[0]) + addr * mt[2] + l) * C.A->lanes + C.lane];
        if (mt[2] <= L) r[mt[2] - 1] &= fmask(mt[3]);
      }
      break;
    }
  }
  r[L - 1] &= fmask(width);
  for (uint32_t l = 0; l < L; l++) C.st(R.w[1], l, r[l]);
}

__device__ __forceinline__ bool reset_asserted(const Ctx& C, uint32_t rsig_code, uint32_t pol) {
  const uint32_t v = C.ld(rsig_code, 0) & 1u;
  return pol == 0 ? v != 0 : v == 0;
}

__device__ void exec_effect(const Ctx& C, const Rec& R) {
  const uint32_t op = R.w[0] & 0xFF, L = (R.w[0] >> 8) & 0xFF;
  uint32_t v[MAXL];
  switch (op) {
    case R_STORE:
      C.ldv(R.w[2], R.w[7], L, v);
      v[L - 1] &= fmask(R.w[5]);
      for (uint32_t l = 0; l < L; l++) C.stsig(R.w[1], l, v[l]);
      break;
    case R_ASYNC_CAPTURE:
      if (reset_asserted(C, R.w[3], R.w[6])) {
        C.ldv(R.w[2], R.w[7], L, v);
        v[L - 1] &= fmask(R.w[5]);
        for (uint32_t l = 0; l < L; l++) C.stsig(R.w[1], l, v[l]);
      }
      break;
    case R_CAPTURE:
      if ((R.w[6] & 2u) && reset_asserted(C, R.w[3], R.w[6] & 1u)) for (uint32_t l = 0; l < L; l++) v[l] = C.ld(R.w[4], l);
      else C.ldv(R.w[2], R.w[7], L, v);
      v[L - 1] &= fmask(R.w[5]);
      for (uint32_t l = 0; l < L; l++) C.st(R.w[1], l, v[l]);
      break;
    case R_COMMIT:
      for (uint32_t l = 0; l < L; l++) C.stsig(R.w[1], l, C.ld(R.w[2], l));
      break;
    case R_MEMWRITE: {
      if (!(C.ld(R.w[2], 0) & 1u)) break;
      const uint32_t* mt = &C.A->mem_table[4 * R.w[1]];
      uint32_t la = R.w[7] & 0xFF, ld = (R.w[7] >> 8) & 0xFF;
      C.ldv(R.w[3], la, la, v);
      const uint64_t addr = decode_usize(v, la);
      if (addr >= mt[1]) break;
      C.ldv(R.w[4], ld, L, v);
      v[L - 1] &= fmask(mt[3]);
      for (uint32_t l = 0; l < L; l++) C.A->mems[(size_t(mt[0]) + addr * mt[2] + l) * C.A->lanes + C.lane] = v[l];
      break;
    }
  }
}

__device__ void exec_block(const Ctx& C, int b) {
  const Rec* recs = C.A->blk[b];
  const uint32_t n = C.A->nrec[b];
  for (uint32_t i = 0; i < n; i++) {
    const Rec R = recs[i];  // uniform: every lane runs the same record
    if ((R.w[0] & 0xFF) >= R_STORE && (R.w[0] & 0xFF) != R_SIGLOAD) exec_effect(C, R);
    else exec_rec(C, R);
  }
}

__device__ __forceinline__ void eval_comb(const Ctx& C) { exec_block(C, B_COMB); exec_block(C, B_ASYNC); }
__device__ __forceinline__ void edge(const Ctx& C) { exec_block(C, B_NEXT); exec_block(C, B_COMMIT); }

__device__ void apply_stimulus(const Ctx& C, uint64_t c, uint64_t cyc) {
  const WordArgs& A = *C.A;
  uint32_t trace_limb = 0;
  for (uint32_t i = 0; i < A.n_in; i++) {
    const uint32_t off = A.in_desc[4 * i], limbs = A.in_desc[4 * i + 1], width = A.in_desc[4 * i + 2], mode = A.in_desc[4 * i + 3];
    const uint64_t arg = A.in_arg[i];
    const uint32_t sig_index = uint32_t(arg >> 40);  // stream id for IN_RANDOM
    for (uint32_t l = 0; l < limbs; l++) {
      uint32_t v = 0;
      if (mode == IN_RANDOM) {
        const uint32_t nb = min(32u, width - 32 * l);
        for (uint32_t bit = 0; bit < nb; bit++) {
          const uint64_t stream = (uint64_t(sig_index) << 8) | (32 * l + bit);
          const uint64_t w = stim_from_key(stim_key(A.seed, stream, cyc), C.lane >> 6);
          v |= uint32_t((w >> (C.lane & 63)) & 1u) << bit;
        }
      } else if (mode == IN_CONST) {
        v = l == 0 ? uint32_t(arg) : (l == 1 ? uint32_t(arg >> 32) : 0u);
      } else if (mode == IN_PULSE) {
        v = l == 0 ? uint32_t(cyc < (arg & 0xFFFFFFFFull) ? (arg >> 32) & 1 : (arg >> 33) & 1) : 0u;
      } else if (mode == IN_TRACE) {
        v = A.trace_in[(size_t(c) * A.trace_in_limbs + trace_limb + l) * A.lanes + C.lane];
      } else {
        continue;
      }
      if (l == limbs - 1) v &= fmask(width);
      C.stsig(off, l, v);
    }
    if (mode == IN_TRACE) trace_limb += limbs;
  }
}

__device__ void observe(const Ctx& C, uint64_t c, uint64_t& d) {
  const WordArgs& A = *C.A;
  uint32_t k = 0;
  for (uint32_t j = 0; j < A.n_obs; j++) {
    const uint32_t off = A.obs_desc[2 * j], limbs = A.obs_desc[2 * j + 1];
    for (uint32_t l = 0; l < limbs; l++, k++) {
      const uint32_t v = A.values[size_t(off + l) * A.lanes + C.lane];
      if (A.trace) A.trace[(size_t(c) * A.n_obs_limbs + k) * A.lanes + C.lane] = v;
      if (A.digest) d = lane_digest(d, v, j, l);
    }
  }
}

template <bool kLds>
__global__ void __launch_bounds__(256) word_kernel(WordArgs A) {
  extern __shared__ uint32_t lds[];
  const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= A.lanes) return;  // no barriers below: lanes are independent
  Ctx C;
  C.A = &A; C.lane = lane;
  if (kLds) { C.scr = lds + threadIdx.x; C.sstride = blockDim.x; }
  else { C.scr = A.scratch + lane; C.sstride = A.lanes; }
  uint64_t d = A.digest ? A.digest[lane] : 0;
  uint64_t done = 0;
  bool halted = false;
  if (A.halt_cycles) { done = A.halt_cycles[lane] >> 1; halted = A.halt_cycles[lane] & 1; }
  if (A.flags & F_EVAL_FIRST) eval_comb(C);
  for (uint64_t c = 0; c < A.cycles && !halted; c++) {
    const uint64_t cyc = A.cycle_base + c;
    if (A.n_in) apply_stimulus(C, c, cyc);
    if (A.flags & F_EVAL_PRE) eval_comb(C);
    if (A.n_obs && !(A.flags & F_OBS_POST)) observe(C, c, d);
    if (A.flags & F_EDGE) edge(C);
    if (A.flags & F_EVAL_POST) eval_comb(C);
    if (A.n_obs && (A.flags & F_OBS_POST)) observe(C, c, d);
    done++;
    if (A.halt_offset != NONE && (A.values[size_t(A.halt_offset) * A.lanes + lane] & 1u)) halted = true;
  }
  if (A.flags & F_FINAL_EVAL) eval_comb(C);
  if (A.digest) A.digest[lane] = d;
  if (A.halt_cycles) A.halt_cycles[lane] = (done << 1) | (halted ? 1u : 0u);
}

// ------------------------------------------------------------------ engine
class WordEngine final : public Engine {
 public:
  WordEngine(Program prog, uint32_t lanes, const EngineOptions& opt) : Engine(std::move(prog), lanes, opt) {
    if (lanes == 0) throw Error("E_SIM_IR_LANES: packed simulator requires at least one lane");
    WP = WordCompiler(P).compile();
    RRTL_HIP_CHECK(hipSetDevice(opt.device));
    hipDeviceProp_t prop;
    RRTL_HIP_CHECK(hipGetDeviceProperties(&prop, opt.device));
    lds_max_ = prop.sharedMemPerBlock;
    RRTL_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    values_.alloc(size_t(P.total_signal_words) * lanes); values_.zero(stream_);
    mems_.alloc(size_t(P.total_memory_words) * lanes); mems_.zero(stream_);
    block_ = opt.block_threads ? opt.block_threads : 256;
    use_lds_ = size_t(WP.scratch_slots) * block_ * 4 <= 64 * 1024;
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
    digest_.release(); trace_.release(); halt_.release();
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
  void run(const RunSpec& spec) override {
    uint32_t flags = F_EVAL_PRE | F_EDGE;
    bool need_post = spec.observe_post || (spec.halt_sig != NONE && P.signals.at(spec.halt_sig).kind != K_REG);
    if (need_post) flags |= F_EVAL_POST; else flags |= F_FINAL_EVAL;
    if (spec.observe_post) flags |= F_OBS_POST;
    launch(flags, spec.cycles, &spec, true);
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
  uint32_t scratch_slots() const { return WP.scratch_slots; }
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

  void launch(uint32_t flags, uint64_t cycles, const RunSpec* spec, bool outputs = false) {
    WordArgs A{};
    A.values = values_.p; A.mems = mems_.p; A.scratch = use_lds_ ? nullptr : scratch_.p;
    for (int b = 0; b < 4; b++) { A.blk[b] = reinterpret_cast<const Rec*>(blk_[b].p); A.nrec[b] = uint32_t(WP.blk[b].size()); }
    A.cpool = cpool_.p; A.mem_table = mtab_.p; A.lanes = lanes_; A.flags = flags; A.cycles = cycles;
    A.halt_offset = NONE;
    if (spec && outputs) {
      A.seed = spec->seed;
      std::vector<uint32_t> in_desc; std::vector<uint64_t> in_arg;
      uint32_t trace_limbs = 0;
      for (size_t i = 0; i < spec->in_sig.size(); i++) {
        const Signal& s = sig_at(spec->in_sig[i]);
        uint32_t mode = spec->in_mode[i];
        uint64_t arg = spec->in_arg.empty() ? 0 : spec->in_arg[i];
        if (mode == IN_RANDOM) {
          if (s.width > 256) throw Error("E_HIP_WIDE: random stimulus is limited to 256-bit inputs");
          arg = uint64_t(spec->in_sig[i]) << 40;
        }
        if (mode == IN_TRACE) trace_limbs += s.limbs;
        in_desc.insert(in_desc.end(), {s.offset, s.limbs, s.width, mode});
        in_arg.push_back(arg);
      }
      A.in_desc = upload_run(0, in_desc); A.in_arg = upload_run(1, in_arg); A.n_in = uint32_t(in_desc.size() / 4);
      if (trace_limbs) {
        if (!spec->trace_inputs) throw Error("E_HIP_RUN_SPEC: trace inputs requested without a trace buffer");
        std::vector<uint32_t> tin(spec->trace_inputs, spec->trace_inputs + size_t(cycles) * trace_limbs * lanes_);
        A.trace_in = upload_run(2, tin); A.trace_in_limbs = trace_limbs;
      }
      std::vector<uint32_t> obs;
      uint32_t obs_limbs = 0;
      for (uint32_t s : spec->obs_sig) { const Signal& g = sig_at(s); obs.insert(obs.end(), {g.offset, g.limbs}); obs_limbs += g.limbs; }
      A.obs_desc = upload_run(3, obs); A.n_obs = uint32_t(spec->obs_sig.size()); A.n_obs_limbs = obs_limbs;
      has_digest_ = A.n_obs > 0;
      if (has_digest_) { digest_.alloc(lanes_); digest_.zero(stream_); A.digest = digest_.p; }
      trace_bytes_ = 0;
      if (spec->trace && A.n_obs) { trace_bytes_ = size_t(cycles) * obs_limbs * lanes_ * 4; trace_.alloc(trace_bytes_ / 4); A.trace = trace_.p; }
      has_halt_ = spec->halt_sig != NONE;
      if (has_halt_) {
        const Signal& h = sig_at(spec->halt_sig);
        A.halt_offset = h.offset;
        halt_.alloc(lanes_); halt_.zero(stream_);
        A.halt_cycles = halt_.p;
      }
    }
    const uint64_t chunk = (opt_.ticks_per_launch && outputs) ? opt_.ticks_per_launch : std::max<uint64_t>(cycles, 1);
    const uint32_t grid = (lanes_ + block_ - 1) / block_;
    const size_t lds = use_lds_ ? size_t(WP.scratch_slots) * block_ * 4 : 0;
    stats_ = RunStats{};
    RRTL_HIP_CHECK(hipEventRecord(ev0_, stream_));
    uint64_t done = 0;
    do {
      WordArgs B = A;
      B.cycle_base = done;
      B.cycles = std::min<uint64_t>(chunk, cycles - done);
      if (done) B.flags &= ~uint32_t(F_EVAL_FIRST);
      if (done + B.cycles < cycles) B.flags &= ~uint32_t(F_FINAL_EVAL);
      if (A.trace) B.trace = A.trace + done * A.n_obs_limbs * lanes_;
      if (A.trace_in) B.trace_in = A.trace_in + done * A.trace_in_limbs * lanes_;
      if (use_lds_) hipLaunchKernelGGL(word_kernel<true>, dim3(grid), dim3(block_), lds, stream_, B);
      else hipLaunchKernelGGL(word_kernel<false>, dim3(grid), dim3(block_), 0, stream_, B);
      RRTL_HIP_CHECK(hipGetLastError());
      stats_.launches++;
      done += B.cycles;
    } while (done < cycles);
    RRTL_HIP_CHECK(hipEventRecord(ev1_, stream_));
    RRTL_HIP_CHECK(hipEventSynchronize(ev1_));
    float ms = 0;
    RRTL_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
    stats_.kernel_ms = ms;
    stats_.lane_cycles = uint64_t(lanes_) * cycles;
    if (has_halt_ && spec && outputs) {
      std::vector<uint64_t> h(lanes_);
      RRTL_HIP_CHECK(hipMemcpy(h.data(), halt_.p, size_t(lanes_) * 8, hipMemcpyDeviceToHost));
      uint64_t total = 0;
      for (auto x : h) total += x >> 1;
      stats_.lane_cycles = total;
    }
  }

  WordProgram WP;
  size_t lds_max_ = 65536;
  uint32_t block_ = 256;
  bool use_lds_ = true, has_digest_ = false, has_halt_ = false;
  size_t trace_bytes_ = 0;
  DevBuf<uint32_t> values_, mems_, scratch_, cpool_, mtab_, blk_[4], run_bufs_[4], trace_;
  DevBuf<uint64_t> digest_, halt_;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace

Engine* make_word_engine(Program prog, uint32_t lanes, const EngineOptions& opt) {
  return new WordEngine(std::move(prog), lanes, opt);
}

}  // namespace rrtl
