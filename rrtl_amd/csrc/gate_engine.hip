// Bit-sliced gate engine: 32 lanes per u32 word (two words per reference u64),
// the MI355X counterpart of rrtl_sim_ir::bitparallel::BitParallelSimulator
// (crates/rrtl-sim-ir/src/bitparallel.rs:31-355).
//
// Semantics (BitParallelSimulator): settle = async-reset-comb then comb, each
// packet's stores visible to later packets; tick = settle, capture tick_next with
// the reset mux (a & rv) | (~a & v), commit, re-apply stuck-at forces, settle;
// a forced (signal, lane) is re-clamped after every store of that signal and
// after the commit (set_force/clamp_signal, :96-131).  On async-reset stores the
// register keeps its value on lanes whose reset is not asserted (the reference's
// :281-287 writes the literal there; PackedSimulator and the scalar Simulator
// keep the register, and so does this engine).
//
// Device layout.  HBM holds every node row-major: node n's lanes are
// state[n * nwords + w], one u32 per 32 lanes (node = signal index, then the two
// constants, then compiler temporaries).  The timed path is `tile_kernel`: one
// workgroup owns a tile of W u32 lane-words for the whole run, keeps only the
// live values of the levelized gate list in LDS (slots reused by liveness, dead
// logic dropped), generates its primary-input bits on chip from a counter hash,
// folds observed outputs into a per-lane signature, and touches HBM only at the
// start and end of the run.  Gates of one level run in parallel over
// (gate, word) pairs; a workgroup barrier separates levels.
#include "engine.hpp"
#include <cstdlib>
#include "gate_device.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <queue>
#include <unordered_set>
#include <unordered_map>

namespace rrtl {
namespace {  // anon

using namespace gate;

// Clears the work-queue control words before a queue launch (a kernel, so a
// captured graph could replay it): the item counter and the per-tile chunk
// counts always, the error flag ctl[1] according to err_mode (0: keep, so an
// error raised by an earlier launch of the run survives to the end of the run;
// 1: clear, the run's first launch; 2: set, the host's error injection).
__global__ void queue_reset_kernel(uint32_t* ctl, uint32_t n, uint32_t err_mode) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (i != 1) ctl[i] = 0u;
    else if (err_mode) ctl[1] = err_mode == 2 ? 1u : 0u;
  }
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) p[i] = v;
}
void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  const uint32_t grid = uint32_t(std::min<uint64_t>((n + 255) / 256, 1024));
  hipLaunchKernelGGL(fill_u32_kernel, dim3(std::max(grid, 1u)), dim3(256), 0, s, p, n, v);
  RRTL_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ compile
struct GateNet {
  uint32_t nsig = 0, n_nodes = 0, c0 = 0, c1 = 0;
  std::vector<CGate> settle;     // async-reset gates then comb gates, reference order
  std::vector<CGate> next;       // tick_next value gates
  std::vector<CCommit> commits;
  std::vector<uint8_t> written;  // signal written by a settle gate
};

GateNet build_net(const Program& P) {
  GateNet N;
  N.nsig = uint32_t(P.signals.size());
  N.c0 = N.nsig; N.c1 = N.nsig + 1;
  uint32_t next_temp = N.nsig + 2;
  auto lower_block = [&](const Block& bk, int kind, std::vector<CGate>& out) {
    std::vector<uint32_t> node(bk.n_values, NONE), uses(bk.n_values, 0);
    std::vector<int64_t> gate_of(bk.n_values, -1);
    size_t k0 = 0;
    for (size_t k = 0; k < bk.packets.size(); k++) {
      if (bk.packets[k].effects.empty() && k + 1 < bk.packets.size()) continue;
      std::unordered_set<uint32_t> read_sig;
      for (size_t q = k0; q <= k; q++) {
        for (const Instr& in : bk.packets[q].instrs) {
          if (in.op == OP_SIGNAL) read_sig.insert(in.args[0]);
          else if (in.op != OP_LIT) for (uint32_t a : in.args) uses[a]++;
        }
        for (const Effect& ef : bk.packets[q].effects) uses[ef.value]++;
      }
      const size_t gstart = out.size();
      auto temp = [&]() { return next_temp++; };
      for (size_t q = k0; q <= k; q++) {
        for (const Instr& in : bk.packets[q].instrs) {
          switch (in.op) {
            case OP_SIGNAL: node[in.dst] = in.args[0]; break;
            case OP_LIT: node[in.dst] = (!in.args.empty() && (in.args[0] & 1)) ? N.c1 : N.c0; break;
            case OP_NOT: {
              int64_t g = gate_of[in.args[0]];
              if (g >= int64_t(gstart) && uses[in.args[0]] == 1 && out[g].b != N.c1) {
                out[g].op ^= GI_O;
                node[in.dst] = out[g].dst; gate_of[in.dst] = g;
              } else {
                uint32_t t = temp();
                out.push_back({G_NOT, node[in.args[0]], N.c1, t});
                node[in.dst] = t; gate_of[in.dst] = int64_t(out.size()) - 1;
              }
              break;
            }
            case OP_AND: case OP_OR: case OP_XOR: {
              uint32_t t = temp();
              out.push_back({in.op == OP_AND ? G_AND : in.op == OP_OR ? G_OR : G_XOR, node[in.args[0]], node[in.args[1]], t});
              node[in.dst] = t; gate_of[in.dst] = int64_t(out.size()) - 1;
              break;
            }
            case OP_MUX: {  // (c & t) | (e & ~c)
              uint32_t c = node[in.args[0]], tv = node[in.args[1]], ev = node[in.args[2]];
              uint32_t t1 = temp(), t2 = temp(), t3 = temp();
              out.push_back({G_AND, c, tv, t1});
              out.push_back({G_ANDN, ev, c, t2});
              out.push_back({G_OR, t1, t2, t3});
              node[in.dst] = t3; gate_of[in.dst] = int64_t(out.size()) - 1;
              break;
            }
            default: throw Error("E_BITPARALLEL: op is not a gate-level bitwise op");
          }
        }
      }
      for (const Effect& ef : bk.packets[k].effects) {
        if (kind != 2 && ef.kind == EF_STORE) {
          int64_t g = gate_of[ef.value];
          bool is_temp = node[ef.value] >= N.nsig + 2;
          if (g >= int64_t(gstart) && is_temp && uses[ef.value] == 1 && !read_sig.count(ef.dst)) out[g].dst = ef.dst;
          else out.push_back({G_COPY, node[ef.value], N.c1, ef.dst});
        } else if (kind == 0 && ef.kind == EF_CAPTURE) {
          if (!ef.has_reset || ef.reset.kind != 1) continue;
          // reg = asserted ? value : reg, asserted = rsig (active high) or ~rsig
          uint32_t r = ef.dst, v = node[ef.value], rs = ef.reset.signal;
          uint32_t t1 = temp(), t2 = temp();
          if (ef.reset.polarity == 0) { out.push_back({G_AND, rs, v, t1}); out.push_back({G_ANDN, r, rs, t2}); }
          else { out.push_back({G_ANDN, v, rs, t1}); out.push_back({G_AND, r, rs, t2}); }
          out.push_back({G_OR, t1, t2, r});
        } else if (kind == 2 && ef.kind == EF_CAPTURE) {
          uint32_t rbit = (!ef.reset.value.empty() && (ef.reset.value[0] & 1)) ? 1u : 0u;
          N.commits.push_back({ef.dst, node[ef.value], ef.has_reset ? ef.reset.signal : NONE, ef.reset.polarity, rbit,
                               ef.has_reset ? 1u : 0u});
        }
      }
      for (size_t q = k0; q <= k; q++) for (const Instr& in : bk.packets[q].instrs) uses[in.dst] = 0;
      k0 = k + 1;
    }
  };
  lower_block(P.blocks[B_ASYNC], 0, N.settle);
  lower_block(P.blocks[B_COMB], 1, N.settle);
  lower_block(P.blocks[B_NEXT], 2, N.next);
  N.n_nodes = next_temp;
  N.written.assign(N.nsig, 0);
  // Read-before-write inside one settle would make the result depend on the
  // packet order (a combinational cycle): the levelized engines reject it.
  // Registers are exempt: their only writers in a settle are async-reset
  // rewrites, each of which reads the register's committed value before
  // writing it, as the reference's packet order does (bitparallel.rs:256-300).
  std::vector<int64_t> first_w(N.n_nodes, -1);
  for (size_t i = 0; i < N.settle.size(); i++) {
    uint32_t d = N.settle[i].dst;
    if (d < N.nsig) {
      if (first_w[d] >= 0 && P.signals[d].kind != K_REG) throw Error("E_HIP_COMB_ORDER: signal `" + P.signals[d].name + "` is driven twice");
      if (first_w[d] < 0) first_w[d] = int64_t(i);
      N.written[d] = 1;
    }
  }
  for (size_t i = 0; i < N.settle.size(); i++)
    for (uint32_t s : {N.settle[i].a, N.settle[i].b})
      if (s < N.nsig && first_w[s] > int64_t(i) && P.signals[s].kind != K_REG)
        throw Error("E_HIP_COMB_ORDER: signal `" + P.signals[s].name + "` is read before it is driven (combinational cycle)");
  return N;
}

// One compiled tile program: stages of up to M gates per thread over LDS
// slots, one barrier per stage.
struct TileProgram {
  uint32_t M = 1;                        // gate positions per thread per stage
  uint32_t n_cycle_stages = 0, n_post_stages = 0;
  // [stage][position][thread] x 2 words: a | b << 16, dst | op << 16 (slot
  // indices; sink slots are sink + k; a is unused at positions > 0, which
  // take operand a from the previous position); op = ia | ib << 1 | fix << 2 | x << 3
  std::vector<uint32_t> desc;
  std::vector<uint32_t> stage_fix;       // per stage: bit k = some gate at position k applies its inversion
  uint32_t n_gates = 0;                  // gates evaluated per cycle (cycle and post programs)
  uint32_t direct_commits = 0;           // registers whose next value is computed into their own slot
  std::vector<uint32_t> load_slot, load_node;     // persistent slots loaded at start, stored at end
  std::vector<uint32_t> stim_slot, stim_mode, stim_stream;
  std::vector<uint64_t> stim_arg;
  std::vector<uint32_t> commit;          // 4 words: reg_slot, val_slot, rsig_slot|NONE, flags(pol | rbit<<1 | has<<2)
  std::vector<uint32_t> obs_slot;        // slot | stored value inverted << 31
  uint32_t n_slots = 0, sink = 0, c0 = 0, c1 = 0;  // live slots [0, n_slots), sink slots [sink, sink + 64)
  // per node: persistent slot, or NONE
  std::unordered_map<uint32_t, uint32_t> persist;
  // forced-node clamp positions: node -> (stage position, slot) for each instance
  std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>> clamp_sites;
};


// Stage schedule of a gate subset.  A stage gives every thread up to M gate
// positions, evaluated in order between two barriers.  Position 0 takes a
// ready gate (every input from an earlier stage), most urgent first (smallest
// as-late-as-possible level); position k > 0 takes a chained gate: one whose
// only same-stage input is the thread's gate at position k - 1, read through
// exactly one operand and passed in registers (or stays empty).  deps[i] are ordering edges (read-after-write through
// prod_a / prod_b, plus write-after-read/write on rewritten nodes).
struct StageSchedule {
  uint32_t stages = 0;
  std::vector<uint32_t> stage, depth, thread;  // per gate of the subset
  std::vector<uint8_t> chained, chain_b;       // chained through operand a (or b)
};

StageSchedule stage_schedule(size_t n, const std::vector<std::vector<uint32_t>>& deps, const std::vector<uint32_t>& prod_a,
                             const std::vector<uint32_t>& prod_b, const std::vector<uint8_t>& in_set,
                             const std::vector<uint32_t>& asap, const std::vector<uint8_t>& no_chain_from, uint32_t M) {
  StageSchedule S;
  S.stage.assign(n, 0); S.depth.assign(n, 0); S.thread.assign(n, 0); S.chained.assign(n, 0); S.chain_b.assign(n, 0);
  std::vector<uint32_t> alap(n, 0), pending(n, 0);
  std::vector<std::vector<uint32_t>> users(n);
  uint32_t span = 0;
  for (size_t i = 0; i < n; i++) if (in_set[i]) span = std::max(span, asap[i]);
  for (size_t i = 0; i < n; i++) {
    if (!in_set[i]) continue;
    std::vector<uint32_t> ds(deps[i]);
    std::sort(ds.begin(), ds.end());
    ds.erase(std::unique(ds.begin(), ds.end()), ds.end());
    for (uint32_t j : ds) if (in_set[j]) { users[j].push_back(uint32_t(i)); pending[i]++; }
  }
  for (size_t k = n; k-- > 0;) {
    if (!in_set[k]) continue;
    uint32_t a = span;
    for (uint32_t u : users[k]) a = std::min(a, alap[u] - 1);
    alap[k] = a;
  }
  using Key = std::pair<uint32_t, uint32_t>;  // (alap, index), smallest first
  std::priority_queue<Key, std::vector<Key>, std::greater<Key>> ready;
  for (size_t i = 0; i < n; i++) if (in_set[i] && pending[i] == 0) ready.push({alap[i], uint32_t(i)});
  std::vector<uint8_t> done(n, 0);
  std::vector<uint32_t> grid(size_t(TILE_THREADS) * M), placed;
  for (uint32_t s = 1; !ready.empty(); s++) {
    std::fill(grid.begin(), grid.end(), NONE);
    placed.clear();
    auto put = [&](uint32_t g, uint32_t t, uint32_t k) {
      grid[size_t(t) * M + k] = g; done[g] = 1; placed.push_back(g);
      S.stage[g] = s; S.depth[g] = k; S.thread[g] = t;
    };
    for (uint32_t t = 0; t < TILE_THREADS && !ready.empty(); t++) { put(ready.top().second, t, 0); ready.pop(); }
    for (uint32_t k = 1; k < M; k++) {
      for (uint32_t t = 0; t < TILE_THREADS; t++) {
        const uint32_t p = grid[size_t(t) * M + k - 1];
        if (p != NONE && !no_chain_from[p]) {
          uint32_t best = NONE;
          for (uint32_t u : users[p]) {
            if (done[u]) continue;
            const bool ra = prod_a[u] == p, rb = prod_b[u] == p;
            if (ra == rb) continue;  // p reaches u through both operands or through neither
            bool ok = true;
            for (uint32_t j : deps[u])
              if (j != p && in_set[j] && (!done[j] || S.stage[j] >= s)) { ok = false; break; }
            if (ok && (best == NONE || alap[u] < alap[best])) best = u;
          }
          if (best != NONE) {
            put(best, t, k);
            S.chained[best] = 1; S.chain_b[best] = prod_b[best] == p;
            continue;
          }
        }
      }
    }
    for (uint32_t g : placed)
      for (uint32_t u : users[g])
        if (--pending[u] == 0 && !done[u]) ready.push({alap[u], u});
    S.stages = s;
  }
  return S;
}

// Order of per-thread table entries (entry i runs on thread i % 256 in round
// i / 256) for the per-cycle tail: each half-wave's 64-bit reads land on
// distinct slot residues mod 32 and each quarter-wave's 64-bit writes on
// distinct residues mod 16 where the entries allow it (the bank groups of
// gfx950's LDS for those accesses).  Greedy, position by position: the entry
// whose residues are least used so far in its half-wave and quarter-wave,
// writes weighted as their higher cost.  rslot empty: the entries only write.
std::vector<uint32_t> bank_order(const std::vector<uint32_t>& rslot, const std::vector<uint32_t>& wslot) {
  const size_t n = wslot.size();
  const bool reads = !rslot.empty();
  {  // entries that write one slot twice keep the table order
    std::vector<uint32_t> w(wslot);
    std::sort(w.begin(), w.end());
    if (std::adjacent_find(w.begin(), w.end()) != w.end()) {
      std::vector<uint32_t> id(n);
      for (size_t i = 0; i < n; i++) id[i] = uint32_t(i);
      return id;
    }
  }
  std::vector<std::vector<uint32_t>> bucket(32 * 16);  // [read residue][write residue]
  for (size_t i = n; i-- > 0;) bucket[(reads ? rslot[i] % 32 : 0) * 16 + wslot[i] % 16].push_back(uint32_t(i));
  std::vector<uint32_t> order;
  order.reserve(n);
  uint32_t rc[32] = {}, wc[16] = {};
  for (size_t p = 0; p < n; p++) {
    if (p % 32 == 0) std::fill(rc, rc + 32, 0u);
    if (p % 16 == 0) std::fill(wc, wc + 16, 0u);
    size_t best = 0;
    uint32_t best_cost = ~0u;
    for (size_t b = 0; b < bucket.size(); b++) {
      if (bucket[b].empty()) continue;
      const uint32_t cost = 3 * wc[b % 16] + rc[b / 16];
      if (cost < best_cost) { best_cost = cost; best = b; }
    }
    order.push_back(bucket[best].back());
    bucket[best].pop_back();
    rc[best / 16]++; wc[best % 16]++;
  }
  return order;
}

TileProgram compile_tile(const GateNet& N, const Program& P, const RunSpec& spec, const std::vector<uint8_t>& forced,
                         uint32_t M) {
  TileProgram T;
  T.M = M;
  const uint32_t nsig = N.nsig;
  std::vector<const CGate*> gl;
  for (auto& g : N.settle) gl.push_back(&g);
  for (auto& g : N.next) gl.push_back(&g);
  const size_t ng = gl.size();
  // Ordering edges in list order: a read depends on the node's writer at that
  // point (prod_a / prod_b); a write (an async-reset rewrite of a register, or
  // any node written twice) also orders after the earlier readers and the
  // earlier writer of that node.
  std::vector<int64_t> writer(N.n_nodes, -1);
  std::vector<uint32_t> lvl(ng, 0), prod_a(ng, NONE), prod_b(ng, NONE);
  std::vector<std::vector<uint32_t>> deps(ng);
  std::vector<std::vector<uint32_t>> readers(N.n_nodes);
  for (size_t i = 0; i < ng; i++) {
    if (writer[gl[i]->a] >= 0) prod_a[i] = uint32_t(writer[gl[i]->a]);
    if (writer[gl[i]->b] >= 0) prod_b[i] = uint32_t(writer[gl[i]->b]);
    for (uint32_t s : {gl[i]->a, gl[i]->b}) {
      if (writer[s] >= 0) deps[i].push_back(uint32_t(writer[s]));
      readers[s].push_back(uint32_t(i));
    }
    const uint32_t d = gl[i]->dst;
    for (uint32_t r : readers[d]) if (r != i) deps[i].push_back(r);
    readers[d].clear();
    if (writer[d] >= 0) deps[i].push_back(uint32_t(writer[d]));
    uint32_t l = 0;
    for (uint32_t j : deps[i]) l = std::max(l, lvl[j]);
    lvl[i] = l + 1;  // as soon as possible
    writer[d] = int64_t(i);
  }
  // cycle program: gates reaching a commit, a PRE observation or a forced register
  std::vector<uint8_t> need(ng, 0);
  std::vector<uint32_t> stack;
  auto mark_gate = [&](std::vector<uint8_t>& set, uint32_t g) { if (g != NONE && !set[g]) { set[g] = 1; stack.push_back(g); } };
  auto mark_node = [&](std::vector<uint8_t>& set, uint32_t n) { if (writer[n] >= 0) mark_gate(set, uint32_t(writer[n])); };
  for (auto& c : N.commits) { mark_node(need, c.val); if (c.rsig != NONE) mark_node(need, c.rsig); }
  if (!spec.observe_post) for (uint32_t s : spec.obs_sig) mark_node(need, s);
  for (uint32_t s = 0; s < nsig; s++) if (forced[s] && P.signals[s].kind == K_REG) mark_node(need, s);
  while (!stack.empty()) { uint32_t i = stack.back(); stack.pop_back(); mark_gate(need, prod_a[i]); mark_gate(need, prod_b[i]); }
  // post program: cone of the observed signals within the settle, re-evaluated after the commit
  std::vector<uint8_t> need2(ng, 0);
  if (spec.observe_post) {
    for (uint32_t s : spec.obs_sig) if (writer[s] >= 0 && size_t(writer[s]) < N.settle.size()) mark_gate(need2, uint32_t(writer[s]));
    while (!stack.empty()) {
      uint32_t i = stack.back(); stack.pop_back();
      for (uint32_t p : {prod_a[i], prod_b[i]}) if (p != NONE && p < N.settle.size()) mark_gate(need2, p);
    }
  }
  // persistent nodes: signals not written by the settle (inputs, registers,
  // host-driven wires) that are read, committed, observed or stimulated
  std::vector<uint8_t> pers(N.n_nodes, 0);
  auto want_persistent = [&](uint32_t n) { if (n < nsig && (!N.written[n] || P.signals[n].kind == K_REG)) pers[n] = 1; };
  for (size_t i = 0; i < ng; i++) if (need[i] || need2[i]) { want_persistent(gl[i]->a); want_persistent(gl[i]->b); want_persistent(gl[i]->dst); }
  for (auto& c : N.commits) { pers[c.reg] = 1; want_persistent(c.val); if (c.rsig != NONE) want_persistent(c.rsig); }
  for (uint32_t s : spec.obs_sig) want_persistent(s);
  for (uint32_t s : spec.in_sig) pers[s] = 1;
  // no chaining out of a gate whose value must sit in LDS before its readers
  // run (a clamped force) or that writes a persistent node
  std::vector<uint8_t> no_chain(ng, 0);
  for (size_t i = 0; i < ng; i++) {
    const uint32_t d = gl[i]->dst;
    if (pers[d] || (d < nsig && forced[d])) no_chain[i] = 1;
  }
  const StageSchedule CS = stage_schedule(ng, deps, prod_a, prod_b, need, lvl, no_chain, M);
  const StageSchedule PS = stage_schedule(ng, deps, prod_a, prod_b, need2, lvl, no_chain, M);
  const uint32_t L = CS.stages, L2 = spec.observe_post ? PS.stages : 0;
  // slot 0 / 1: constants; then persistent nodes
  T.c0 = 0; T.c1 = 1;
  uint32_t next_slot = 2;
  for (uint32_t n = 0; n < nsig; n++) if (pers[n]) T.persist[n] = next_slot++;
  const uint32_t first_free = next_slot;
  // timeline positions: cycle stages 1..L, PRE observe L+1, commit L+2, post stages L+3..
  const uint32_t POS_OBS_PRE = L + 1, POS_COMMIT = L + 2, POS_POST0 = L + 2;
  struct Inst { uint32_t gate, pos, slot, depth, thread; bool post, chained, chain_b; };
  std::vector<Inst> insts;
  std::vector<uint32_t> inst_of_cycle(ng, NONE), inst_of_post(ng, NONE);
  for (size_t i = 0; i < ng; i++)
    if (need[i]) {
      inst_of_cycle[i] = uint32_t(insts.size());
      insts.push_back({uint32_t(i), CS.stage[i], NONE, CS.depth[i], CS.thread[i], false, CS.chained[i] != 0, CS.chain_b[i] != 0});
    }
  for (size_t i = 0; i < ng; i++)
    if (need2[i]) {
      inst_of_post[i] = uint32_t(insts.size());
      insts.push_back({uint32_t(i), POS_POST0 + PS.stage[i], NONE, PS.depth[i], PS.thread[i], true, PS.chained[i] != 0,
                       PS.chain_b[i] != 0});
    }
  T.n_gates = uint32_t(insts.size());
  auto inst_of = [&](uint32_t gate, bool post) { return post ? inst_of_post[gate] : inst_of_cycle[gate]; };
  // last LDS read of each instance (chained reads come from registers)
  std::vector<uint32_t> last(insts.size(), 0);
  for (auto& in : insts) {
    for (int op = 0; op < 2; op++) {
      const uint32_t p = op == 0 ? prod_a[in.gate] : prod_b[in.gate];
      if (p == NONE) continue;
      const uint32_t node = op == 0 ? gl[in.gate]->a : gl[in.gate]->b;
      if (pers[node]) continue;
      if (in.chained && (op == 1) == in.chain_b) continue;
      const uint32_t si = inst_of(p, in.post);
      if (si != NONE) last[si] = std::max(last[si], in.pos);
    }
  }
  auto src_inst = [&](uint32_t node, bool post) -> uint32_t {
    if (writer[node] < 0) return NONE;
    return inst_of(uint32_t(writer[node]), post);
  };
  for (auto& c : N.commits) {
    for (uint32_t s : {c.val, c.rsig}) {
      if (s == NONE || pers[s]) continue;
      uint32_t si = src_inst(s, false);
      if (si != NONE) last[si] = std::max(last[si], POS_COMMIT);
    }
  }
  for (uint32_t s : spec.obs_sig) {
    if (pers[s]) continue;
    uint32_t si = src_inst(s, spec.observe_post);
    if (si != NONE) last[si] = std::max(last[si], spec.observe_post ? POS_POST0 + L2 + 1 : POS_OBS_PRE);
  }
  // operand slots of an instance (producer instance at that point of the list)
  auto operand_slot = [&](const Inst& in, int op) -> uint32_t {
    const uint32_t node = op == 0 ? gl[in.gate]->a : gl[in.gate]->b;
    if (node == N.c0) return T.c0;
    if (node == N.c1) return T.c1;
    if (pers[node]) return T.persist.at(node);
    const uint32_t p = op == 0 ? prod_a[in.gate] : prod_b[in.gate];
    const uint32_t si = p == NONE ? NONE : inst_of(p, in.post);
    if (si == NONE) throw Error("E_HIP_INTERNAL: node has no producer in the tile program");
    if (insts[si].slot == NONE) throw Error("E_HIP_INTERNAL: operand read before its producer's stage");
    return insts[si].slot;
  };
  auto slot_of = [&](uint32_t node, bool post) -> uint32_t {
    if (node == N.c0) return T.c0;
    if (node == N.c1) return T.c1;
    if (node < nsig && pers[node]) return T.persist.at(node);
    const uint32_t si = src_inst(node, post);
    if (si == NONE || insts[si].slot == NONE) throw Error("E_HIP_INTERNAL: node has no producer in the tile program");
    return insts[si].slot;
  };
  // Allocation, stage by stage, with LDS banks in mind.  A wavefront's LDS
  // access is served per half-wave (32 lanes); the lanes of a half conflict
  // when distinct slots share (slot mod 32) (at W = 2 a slot is an 8-byte pair
  // of the 64 banks; every W keeps the residue -> bank-group map), and the
  // access takes as many passes as the most shared residue.  So per stage:
  //  * each thread's gates move together to one of the 8 half-waves, chosen
  //    where their a and b residues (per position) are least used so far
  //    (distinct slots only: equal slots are a broadcast);
  //  * per position, the outputs of a half take distinct residues -- residues
  //    with free slots first (fullest first), else a fresh slot of the residue
  //    whose fresh run is shortest; the lanes without an output write sink
  //    slots of the residues left over, so every write of a half hits 32 bank
  //    groups.
  // A value read only through the register chain gets no slot (it is written
  // to a sink slot).  Slots are released after their last read (linear scan);
  // a fresh slot is taken only when no slot of an unused residue is free.
  constexpr uint32_t RES = 32, HALVES = TILE_THREADS / RES;
  uint32_t max_pos = 0;
  for (auto& in : insts) max_pos = std::max(max_pos, in.pos);
  std::vector<std::vector<uint32_t>> at_pos(size_t(max_pos) + 1);
  for (uint32_t ii = 0; ii < insts.size(); ii++) at_pos[insts[ii].pos].push_back(ii);
  std::vector<std::vector<uint32_t>> free_res(RES);
  std::vector<uint32_t> fresh(RES);
  for (uint32_t r = 0; r < RES; r++) fresh[r] = first_free + ((r + RES - first_free % RES) % RES);
  uint32_t slot_end = first_free;
  std::multimap<uint32_t, uint32_t> releases;  // last-read pos -> slot
  // Borrowed register slots.  A register's slot is rewritten by the commit at
  // the end of every cycle, so between the last read of its value in the
  // cycle program and the commit it can hold a cycle-program value that dies
  // before the commit (last read at a cycle stage, not at the observation or
  // the commit).  Not for registers read by the commit itself (staged), by the
  // observation, forced, or written by a gate.
  std::vector<uint8_t> reg_borrow(N.n_nodes, 0);
  for (auto& c : N.commits) if (P.signals[c.reg].kind == K_REG && !forced[c.reg]) reg_borrow[c.reg] = 1;
  for (auto& c : N.commits) { reg_borrow[c.val] = 0; if (c.rsig != NONE) reg_borrow[c.rsig] = 0; }
  for (uint32_t s : spec.obs_sig) reg_borrow[s] = 0;
  for (size_t i = 0; i < ng; i++) reg_borrow[gl[i]->dst] = 0;
  std::vector<uint32_t> reg_last(N.n_nodes, 0);
  for (auto& in : insts) {
    if (in.post) continue;
    for (int op = 0; op < 2; op++) {
      const uint32_t node = op == 0 ? gl[in.gate]->a : gl[in.gate]->b;
      if (node < nsig && pers[node]) reg_last[node] = std::max(reg_last[node], in.pos);
    }
  }
  // A register's next value may be written straight into the register's own
  // slot when the register's last read comes before it and nothing else
  // reads that slot at the commit (the value is the D of this one commit, no
  // reset signal, not observed): the commit then has nothing to copy.
  std::vector<uint32_t> direct_reg(N.n_nodes, NONE), val_uses(N.n_nodes, 0);
  std::vector<uint8_t> observed_node(N.n_nodes, 0);
  for (uint32_t s : spec.obs_sig) observed_node[s] = 1;
  for (auto& c : N.commits) { val_uses[c.val]++; if (c.rsig != NONE) val_uses[c.rsig] += 2; }
  for (auto& c : N.commits)
    if (val_uses[c.val] == 1 && reg_borrow[c.reg] && !observed_node[c.val] && c.val != c.reg) direct_reg[c.val] = c.reg;
  std::vector<std::vector<uint32_t>> borrow_res(RES);
  std::vector<uint8_t> slot_borrowed(first_free, 0);
  std::multimap<uint32_t, uint32_t> borrow_releases;  // free after this pos -> slot
  for (auto& kv : T.persist)
    if (reg_borrow[kv.first]) { slot_borrowed[kv.second] = 1; borrow_releases.emplace(reg_last[kv.first], kv.second); }
  std::vector<uint32_t> lane_of(insts.size(), NONE);
  std::vector<uint32_t> write_mask;            // per (position, depth, half): residues written
  std::vector<uint32_t> mask_base(size_t(max_pos) + 2, 0);
  for (uint32_t pos = 1; pos <= max_pos + 1; pos++) mask_base[pos] = mask_base[pos - 1] + M * HALVES;
  write_mask.assign(mask_base[max_pos + 1] + M * HALVES, 0);
  // LDS writes are served per quarter-wave over 32 banks: at W = 2 the 16
  // lanes of a quarter write without conflict when their slots differ mod 16.
  // Per (position, depth, half, quarter): residues mod 16 written.
  constexpr uint32_t QRES = 16;
  std::vector<uint32_t> write_mask_q(write_mask.size() * 2, 0);
  std::vector<uint32_t> cnt_a(size_t(M) * HALVES * RES), cnt_b(size_t(M) * HALVES * RES);
  std::vector<std::vector<uint32_t>> half_slots(size_t(M) * HALVES);
  std::vector<std::vector<uint32_t>> by_thread(TILE_THREADS);
  std::vector<uint32_t> half_fill(HALVES), half_of(TILE_THREADS, 0);
  for (uint32_t pos = 1; pos <= max_pos; pos++) {
    while (!releases.empty() && releases.begin()->first < pos) {
      const uint32_t sl = releases.begin()->second;
      (sl < first_free && slot_borrowed[sl] ? borrow_res : free_res)[sl % RES].push_back(sl);
      releases.erase(releases.begin());
    }
    while (!borrow_releases.empty() && borrow_releases.begin()->first < pos) {
      borrow_res[borrow_releases.begin()->second % RES].push_back(borrow_releases.begin()->second);
      borrow_releases.erase(borrow_releases.begin());
    }
    const std::vector<uint32_t>& v = at_pos[pos];
    if (v.empty()) continue;
    for (auto& b : by_thread) b.clear();
    for (uint32_t ii : v) by_thread[insts[ii].thread].push_back(ii);
    uint32_t nthreads = 0;
    for (auto& b : by_thread) if (!b.empty()) nthreads++;
    const uint32_t cap = (nthreads + HALVES - 1) / HALVES;
    std::fill(cnt_a.begin(), cnt_a.end(), 0u); std::fill(cnt_b.begin(), cnt_b.end(), 0u);
    for (auto& hs : half_slots) hs.clear();
    std::fill(half_fill.begin(), half_fill.end(), 0u);
    for (uint32_t t = 0; t < TILE_THREADS; t++) {
      if (by_thread[t].empty()) continue;
      // reads of this thread per position
      uint32_t best_h = 0, best_key = ~0u;
      for (uint32_t h = 0; h < HALVES; h++) {
        if (half_fill[h] >= cap) continue;
        uint32_t key = half_fill[h];
        for (uint32_t ii : by_thread[t]) {
          const Inst& in = insts[ii];
          const auto& hs = half_slots[in.depth * HALVES + h];
          for (int op = 0; op < 2; op++) {
            if (in.chained && (op == 1) == in.chain_b) continue;
            const uint32_t s = operand_slot(in, op);
            if (std::find(hs.begin(), hs.end(), s) != hs.end()) continue;
            // a later position's one read is issued as its b read
            const uint32_t c = (op == 0 && in.depth == 0 ? cnt_a : cnt_b)[(in.depth * HALVES + h) * RES + s % RES];
            key += c * 1024;
          }
        }
        if (key < best_key) { best_key = key; best_h = h; }
      }
      for (uint32_t ii : by_thread[t]) {
        const Inst& in = insts[ii];
        auto& hs = half_slots[in.depth * HALVES + best_h];
        for (int op = 0; op < 2; op++) {
          if (in.chained && (op == 1) == in.chain_b) continue;
          const uint32_t s = operand_slot(in, op);
          if (std::find(hs.begin(), hs.end(), s) != hs.end()) continue;
          (op == 0 && in.depth == 0 ? cnt_a : cnt_b)[(in.depth * HALVES + best_h) * RES + s % RES]++;
          hs.push_back(s);
        }
      }
      half_of[t] = best_h;
      half_fill[best_h]++;
    }
    // Local search over the greedy placement: swap two threads of different
    // halves when that lowers sum over (half, read instruction, residue) of
    // (distinct slots)^2, a smooth stand-in for the passes each read takes.
    // Read instructions: 0 = operand a at position 0, 1 + k = operand b at
    // position k.  Deterministic (seeded by the stage).
    {
      std::vector<std::vector<std::pair<uint32_t, uint32_t>>> reads(TILE_THREADS);  // (instruction, slot)
      std::vector<uint32_t> ths;
      for (uint32_t t = 0; t < TILE_THREADS; t++) {
        if (by_thread[t].empty()) continue;
        ths.push_back(t);
        for (uint32_t ii : by_thread[t]) {
          const Inst& in = insts[ii];
          for (int op = 0; op < 2; op++) {
            if (in.chained && (op == 1) == in.chain_b) continue;
            reads[t].push_back({op == 0 && in.depth == 0 ? 0u : 1u + in.depth, operand_slot(in, op)});
          }
        }
      }
      const uint32_t NI = M + 1;
      std::vector<std::unordered_map<uint32_t, uint32_t>> refs(size_t(HALVES) * NI);
      std::vector<int32_t> dcnt(size_t(HALVES) * NI * RES, 0);
      auto move = [&](uint32_t t, uint32_t h, int sign) -> int64_t {
        int64_t d = 0;
        for (auto& r : reads[t]) {
          const size_t hi = size_t(h) * NI + r.first;
          uint32_t& rc = refs[hi][r.second];
          const bool change = sign > 0 ? rc++ == 0 : --rc == 0;
          if (!change) continue;
          int32_t& c = dcnt[hi * RES + r.second % RES];
          d += sign > 0 ? 2 * int64_t(c) + 1 : 1 - 2 * int64_t(c);
          c += sign;
        }
        return d;
      };
      for (uint32_t t : ths) move(t, half_of[t], +1);
      if (ths.size() > 1) {
        uint64_t rs = 0x9E3779B97F4A7C15ull * (pos + 1);
        auto rng = [&]() {  // xorshift64*
          rs ^= rs >> 12; rs ^= rs << 25; rs ^= rs >> 27;
          return uint32_t((rs * 0x2545F4914F6CDD1Dull) >> 32);
        };
        const uint32_t tries = uint32_t(ths.size()) * 40;
        for (uint32_t it = 0; it < tries; it++) {
          const uint32_t t1 = ths[rng() % ths.size()], t2 = ths[rng() % ths.size()];
          const uint32_t h1 = half_of[t1], h2 = half_of[t2];
          if (h1 == h2) continue;
          int64_t d = move(t1, h1, -1) + move(t2, h2, -1);
          d += move(t1, h2, +1) + move(t2, h1, +1);
          if (d < 0) { half_of[t1] = h2; half_of[t2] = h1; continue; }
          move(t1, h2, -1); move(t2, h1, -1); move(t1, h1, +1); move(t2, h2, +1);
        }
      }
      std::fill(half_fill.begin(), half_fill.end(), 0u);
      for (uint32_t t : ths) {
        const uint32_t h = half_of[t];
        for (uint32_t ii : by_thread[t]) lane_of[ii] = h * RES + half_fill[h];
        half_fill[h]++;
      }
    }
    // outputs, position by position within the stage
    for (uint32_t k = 0; k < M; k++)
      for (uint32_t ii : v) {
        Inst& in = insts[ii];
        if (in.depth != k) continue;
        const size_t mi = mask_base[pos] + k * HALVES + lane_of[ii] / RES;
        uint32_t& taken = write_mask[mi];
        uint32_t& taken_q = write_mask_q[mi * 2 + (lane_of[ii] % RES) / QRES];
        const uint32_t d = gl[in.gate]->dst;
        if (d < nsig && pers[d]) {
          in.slot = T.persist.at(d);
          taken |= 1u << (in.slot % RES); taken_q |= 1u << (in.slot % QRES);
          continue;
        }
        if (last[ii] == 0) { in.slot = NONE - 1; continue; }  // read only through the register chain
        if (!in.post && direct_reg[d] != NONE && src_inst(d, false) == ii) {
          const uint32_t reg = direct_reg[d], qs = T.persist.at(reg);
          auto& br = borrow_res[qs % RES];
          auto it = std::find(br.begin(), br.end(), qs);
          if (reg_last[reg] < in.pos && it != br.end() && !(taken >> (qs % RES) & 1u) && !(taken_q >> (qs % QRES) & 1u)) {
            br.erase(it);
            in.slot = qs;
            taken |= 1u << (qs % RES); taken_q |= 1u << (qs % QRES);
            releases.emplace(std::max(last[ii], in.pos), qs);
            continue;
          }
        }
        const bool may_borrow = !in.post && last[ii] <= L;
        auto avail = [&](uint32_t c) { return free_res[c].size() + (may_borrow ? borrow_res[c].size() : 0); };
        // a residue unused by the other outputs of the half (mod 32) and of the
        // quarter (mod 16); failing that, of the half only; else accept a conflict
        uint32_t r = RES;
        for (int strict = 1; strict >= 0 && r == RES; strict--)
          for (uint32_t c = 0; c < RES; c++) {
            if (taken >> c & 1u) continue;
            if (strict && (taken_q >> (c % QRES) & 1u)) continue;
            if (r == RES) { r = c; continue; }
            const size_t fc = avail(c), fr = avail(r);
            if (fc > fr || (fc == 0 && fr == 0 && fresh[c] < fresh[r])) r = c;
          }
        if (r == RES) r = 0;
        taken |= 1u << r; taken_q |= 1u << (r % QRES);
        if (may_borrow && !borrow_res[r].empty()) { in.slot = borrow_res[r].back(); borrow_res[r].pop_back(); }
        else if (!free_res[r].empty()) { in.slot = free_res[r].back(); free_res[r].pop_back(); }
        else { in.slot = fresh[r]; fresh[r] += RES; }
        slot_end = std::max(slot_end, in.slot + 1);
        releases.emplace(std::max(last[ii], in.pos), in.slot);
      }
  }
  T.n_slots = slot_end;
  // Output polarity.  A gate stores (x ? A ^ B : A & B) without its output
  // inversion; the stored value is the true value XOR pol.  Readers fold the
  // polarity of what they read into their input inversions (an AND-type reads
  // a ^ ia ^ pol_a; an XOR-type carries pol_a ^ pol_b into its own polarity),
  // the commit and the observation undo it, and only a gate whose slot must
  // hold the true value (a persistent node, a clamped force) applies its
  // inversion: `fix`.  eff_op = the op bits the kernel uses (ia, ib, fix, x).
  std::vector<uint8_t> pol(insts.size(), 0);
  std::vector<uint32_t> eff_op(insts.size(), 0);
  {
    std::vector<uint32_t> order(insts.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
      const Inst &a = insts[x], &b = insts[y];
      return std::make_tuple(a.pos, a.depth) < std::make_tuple(b.pos, b.depth);
    });
    auto operand_pol = [&](const Inst& in, int op) -> uint32_t {
      const uint32_t node = op == 0 ? gl[in.gate]->a : gl[in.gate]->b;
      if (node == N.c0 || node == N.c1 || pers[node]) return 0;
      const uint32_t p = op == 0 ? prod_a[in.gate] : prod_b[in.gate];
      const uint32_t si = p == NONE ? NONE : inst_of(p, in.post);
      if (si == NONE) throw Error("E_HIP_INTERNAL: node has no producer in the tile program");
      return pol[si];
    };
    for (uint32_t ii : order) {
      const Inst& in = insts[ii];
      const uint32_t op = gl[in.gate]->op, d = gl[in.gate]->dst;
      const uint32_t pa = operand_pol(in, 0), pb = operand_pol(in, 1), io = (op >> 2) & 1u;
      uint32_t ia = 0, ib = 0, tpol;
      if (op & GI_X) {
        tpol = io ^ pa ^ pb;
      } else {
        ia = (op & 1u) ^ pa; ib = ((op >> 1) & 1u) ^ pb;
        tpol = io;
      }
      const bool must_be_true = (d < nsig && (pers[d] || forced[d]));
      const uint32_t fix = must_be_true ? tpol : 0u;
      pol[ii] = uint8_t(tpol ^ fix);
      eff_op[ii] = ia | (ib << 1) | (fix << 2) | (op & GI_X);
    }
  }
  auto value_pol = [&](uint32_t node, bool post) -> uint32_t {
    if (node == N.c0 || node == N.c1 || (node < nsig && pers[node])) return 0;
    const uint32_t si = src_inst(node, post);
    return si == NONE ? 0u : pol[si];
  };
  // commits: a value or reset signal that is itself a committed register would
  // be overwritten mid-commit; stage it through a copy in an extra stage that
  // runs after every cycle stage.
  std::unordered_set<uint32_t> committed;
  for (auto& c : N.commits) committed.insert(c.reg);
  std::vector<uint32_t> staging;  // (src slot, temp slot) pairs
  for (auto& c : N.commits) {
    uint32_t vs = slot_of(c.val, false);
    uint32_t rs = c.rsig == NONE ? NONE : slot_of(c.rsig, false);
    auto stage = [&](uint32_t slot) {
      uint32_t t = T.n_slots++;
      staging.push_back(slot); staging.push_back(t);
      return t;
    };
    const uint32_t vpol = value_pol(c.val, false), rpol = c.rsig == NONE ? 0u : value_pol(c.rsig, false);
    // the next value already sits in the register's slot with its true polarity
    if (vs == T.persist.at(c.reg) && !c.has_reset && vpol == 0) { T.direct_commits++; continue; }
    if (committed.count(c.val)) vs = stage(vs);
    if (c.rsig != NONE && committed.count(c.rsig)) rs = stage(rs);
    // flags: reset polarity (with the stored polarity of the reset signal
    // folded in) | reset value << 1 | has reset << 2 | stored value inverted << 3
    T.commit.insert(T.commit.end(), {T.persist.at(c.reg), vs, rs == NONE ? 0u : rs,
                                     (c.pol ^ rpol) | (c.rbit << 1) | (c.has_reset << 2) | (vpol << 3)});
  }
  T.sink = (T.n_slots + RES - 1) & ~(RES - 1);
  if (T.sink + SINK_SLOTS > 65535) throw Error("E_HIP_TOO_LARGE: gate program needs more than 65535 live slots");
  // Kernel stages: cycle stages, staging stages (copies at position 0), post stages.
  const uint32_t n_staging = uint32_t((staging.size() / 2 + TILE_THREADS - 1) / TILE_THREADS);
  T.n_cycle_stages = L + n_staging;
  T.n_post_stages = L2;
  const uint32_t KS = T.n_cycle_stages + T.n_post_stages;
  const size_t row_words = size_t(M) * TILE_THREADS * 2;
  T.desc.assign(size_t(KS) * row_words, NONE);
  T.stage_fix.assign(KS, 0);
  auto entry = [&](uint32_t kst, uint32_t k, uint32_t lane) -> uint32_t* { return &T.desc[kst * row_words + (size_t(k) * TILE_THREADS + lane) * 2]; };
  std::vector<uint32_t> kstage_mask(size_t(KS) * M * HALVES, 0);
  for (uint32_t ii = 0; ii < insts.size(); ii++) {
    const Inst& in = insts[ii];
    const CGate& g = *gl[in.gate];
    const uint32_t kst = in.post ? T.n_cycle_stages + (in.pos - POS_POST0 - 1) : in.pos - 1;
    uint32_t a = in.chained && !in.chain_b ? NONE : operand_slot(in, 0);
    uint32_t b = in.chained && in.chain_b ? NONE : operand_slot(in, 1);
    uint32_t op = eff_op[ii];
    if (in.chained && in.chain_b) {  // the chained operand goes to a
      std::swap(a, b);
      op = (op & ~(GI_A | GI_B)) | ((op & GI_A) ? GI_B : 0u) | ((op & GI_B) ? GI_A : 0u);
    }
    if (in.depth > 0 && !in.chained) throw Error("E_HIP_INTERNAL: unchained gate after position 0");
    const uint32_t d = in.slot == NONE - 1 ? NONE : in.slot;
    uint32_t* e = entry(kst, in.depth, lane_of[ii]);
    e[0] = (a == NONE ? 0xFFFFu : a) | ((b == NONE ? 0xFFFFu : b) << 16);
    e[1] = (d == NONE ? 0xFFFFu : d) | (op << 16);
    if (op & GI_O) T.stage_fix[kst] |= 1u << in.depth;
    kstage_mask[(size_t(kst) * M + in.depth) * HALVES + lane_of[ii] / RES] =
        write_mask[mask_base[in.pos] + in.depth * HALVES + lane_of[ii] / RES];
    if (g.dst < nsig && forced[g.dst] && d != NONE) T.clamp_sites.push_back({g.dst, {kst + 1, d}});
  }
  for (size_t q = 0; q < staging.size() / 2; q++) {
    const uint32_t kst = L + uint32_t(q / TILE_THREADS), lane = uint32_t(q % TILE_THREADS);
    uint32_t* e = entry(kst, 0, lane);
    e[0] = staging[2 * q] | (T.c1 << 16);
    e[1] = staging[2 * q + 1] | (G_COPY << 16);
    kstage_mask[(size_t(kst) * M) * HALVES + lane / RES] |= 1u << (staging[2 * q + 1] % RES);
  }
  // Lanes without a gate at a position, and outputs read only through the
  // register chain, write the sink slot of a residue no output of their half
  // uses; operand fields not read (a chained operand, an idle lane) repeat the
  // first real read of the half (a broadcast) or constant 0.
  for (uint32_t kst = 0; kst < KS; kst++)
    for (uint32_t k = 0; k < M; k++)
      for (uint32_t h = 0; h < HALVES; h++) {
        uint32_t taken = kstage_mask[(size_t(kst) * M + k) * HALVES + h];
        uint32_t rep_a = T.c0, rep_b = T.c0;
        bool have_a = false, have_b = false;
        for (uint32_t l = h * RES; l < h * RES + RES; l++) {
          const uint32_t* e = entry(kst, k, l);
          if (e[0] == NONE) continue;
          if (!have_a && (e[0] & 0xFFFFu) != 0xFFFFu) { rep_a = e[0] & 0xFFFFu; have_a = true; }
          if (!have_b && (e[0] >> 16) != 0xFFFFu) { rep_b = e[0] >> 16; have_b = true; }
        }
        // residues mod 16 the real outputs of each quarter write
        uint32_t qtaken[2] = {0u, 0u};
        for (uint32_t l = h * RES; l < h * RES + RES; l++) {
          const uint32_t* e = entry(kst, k, l);
          if (e[0] != NONE && (e[1] & 0xFFFFu) != 0xFFFFu) qtaken[(l % RES) / QRES] |= 1u << ((e[1] & 0xFFFFu) % QRES);
        }
        // a sink residue unused in the half (mod 32) and in the lane's quarter (mod 16)
        auto sink_slot = [&](uint32_t l) {
          uint32_t& tq = qtaken[(l % RES) / QRES];
          uint32_t r = RES;
          for (uint32_t c = 0; c < RES && r == RES; c++) if (!(taken >> c & 1u) && !(tq >> (c % QRES) & 1u)) r = c;
          for (uint32_t c = 0; c < RES && r == RES; c++) if (!(taken >> c & 1u)) r = c;
          if (r == RES) r = 0;
          taken |= 1u << r; tq |= 1u << (r % QRES);
          return T.sink + r;
        };
        for (uint32_t l = h * RES; l < h * RES + RES; l++) {
          uint32_t* e = entry(kst, k, l);
          if (e[0] == NONE) { e[0] = rep_a | (rep_b << 16); e[1] = sink_slot(l) | (G_AND << 16); continue; }
          if ((e[0] & 0xFFFFu) == 0xFFFFu) e[0] = (e[0] & 0xFFFF0000u) | rep_a;
          if ((e[0] >> 16) == 0xFFFFu) e[0] = (e[0] & 0xFFFFu) | (rep_b << 16);
          if ((e[1] & 0xFFFFu) == 0xFFFFu) e[1] = (e[1] & 0xFFFF0000u) | sink_slot(l);
        }
      }
  // persistent load/store lists
  for (auto& kv : T.persist) {
    T.load_slot.push_back(kv.second); T.load_node.push_back(kv.first);
    if (P.signals[kv.first].kind == K_REG && forced[kv.first]) T.clamp_sites.push_back({kv.first, {CLAMP_COMMIT, kv.second}});
  }
  // stimulus
  for (size_t i = 0; i < spec.in_sig.size(); i++) {
    if (spec.in_mode[i] == IN_HOLD) continue;
    if (spec.in_mode[i] == IN_TRACE || spec.in_mode[i] == IN_TRACE_UNIFORM)
      throw Error("E_HIP_UNSUPPORTED: trace-driven inputs are not supported by the gate engine run");
    if (spec.in_mode[i] > IN_GROUP_RANDOM) throw Error("E_HIP_RUN_SPEC: unknown input mode");
    T.stim_slot.push_back(T.persist.at(spec.in_sig[i]));
    T.stim_mode.push_back(spec.in_mode[i]);
    T.stim_stream.push_back(spec.in_sig[i] << 8);
    T.stim_arg.push_back(spec.in_arg.empty() ? 0 : spec.in_arg[i]);
  }
  {  // which thread drives which input: conflict-free writes (entries are independent)
    const std::vector<uint32_t> ord = bank_order({}, T.stim_slot);
    auto permute = [&](auto& v) { auto c = v; for (size_t i = 0; i < ord.size(); i++) v[i] = c[ord[i]]; };
    permute(T.stim_slot); permute(T.stim_mode); permute(T.stim_stream); permute(T.stim_arg);
  }
  for (uint32_t s : spec.obs_sig) T.obs_slot.push_back(slot_of(s, spec.observe_post) | (value_pol(s, spec.observe_post) << 31));
  return T;
}


// Exact, unhurried kernels for the API-level operations (eval_combinational,
// tick, tick_from_evaluated): one thread per u32 lane-word walks the whole gate
// list in reference order over the HBM rows, clamping forced nodes as it goes.
struct HbmArgs {
  uint32_t* state; uint64_t nwords;
  const uint4* gates; uint32_t n_gates;     // (op, a, b, dst) node ids
  const uint4* commits; uint32_t n_commits; uint32_t stage_row;
  const uint32_t* fstart; const uint32_t* fent;  // per node: entries (word, and, or)
  uint32_t n_nodes;
};

__device__ __forceinline__ void hbm_clamp(const HbmArgs& A, uint32_t node, uint64_t w) {
  if (!A.fstart) return;
  for (uint32_t i = A.fstart[node]; i < A.fstart[node + 1]; i++) {
    const uint32_t* f = &A.fent[size_t(i) * 3];
    if (f[0] != w) continue;
    uint32_t* v = &A.state[size_t(node) * A.nwords + w];
    *v = (*v & f[1]) | f[2];
  }
}

__global__ void hbm_gates_kernel(HbmArgs A) {
  const uint64_t w = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= A.nwords) return;
  for (uint32_t i = 0; i < A.n_gates; i++) {
    const uint4 g = A.gates[i];
    uint32_t a = A.state[size_t(g.y) * A.nwords + w], b = A.state[size_t(g.z) * A.nwords + w];
    A.state[size_t(g.w) * A.nwords + w] = gate_eval(g.x, a, b);
    hbm_clamp(A, g.w, w);
  }
}

__global__ void hbm_commit_kernel(HbmArgs A) {
  const uint64_t w = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= A.nwords) return;
  for (uint32_t k = 0; k < A.n_commits; k++) {
    const uint4 cm = A.commits[k];  // reg, val, rsig, flags
    uint32_t v = A.state[size_t(cm.y) * A.nwords + w];
    if (cm.w & 4) {
      uint32_t a = A.state[size_t(cm.z) * A.nwords + w];
      if (cm.w & 1) a = ~a;
      uint32_t rv = (cm.w & 2) ? ~0u : 0u;
      v = (a & rv) | (~a & v);
    }
    A.state[size_t(A.stage_row + k) * A.nwords + w] = v;
  }
  for (uint32_t k = 0; k < A.n_commits; k++)
    A.state[size_t(A.commits[k].x) * A.nwords + w] = A.state[size_t(A.stage_row + k) * A.nwords + w];
  // apply_forces: every forced node is re-clamped after the commit
  if (A.fstart)
    for (uint32_t n = 0; n < A.n_nodes; n++)
      if (A.fstart[n + 1] != A.fstart[n]) hbm_clamp(A, n, w);
}

// ------------------------------------------------------------------ engine
struct Force { uint32_t sig, lane, value; };

class GateEngine final : public Engine {
 public:
  GateEngine(Program prog, uint32_t lanes, const EngineOptions& opt) : Engine(std::move(prog), lanes, opt) {
    std::string why;
    if (!P.is_gate_level(&why)) throw Error("E_BITPARALLEL: " + why);
    N = build_net(P);
    nwords_ = (uint64_t(lanes) + 31) / 32;
    RRTL_HIP_CHECK(hipSetDevice(opt.device));
    hipDeviceProp_t prop;
    RRTL_HIP_CHECK(hipGetDeviceProperties(&prop, opt.device));
    lds_max_ = prop.sharedMemPerBlock;
    cus_ = uint32_t(prop.multiProcessorCount);
    int wall_khz = 0;
    RRTL_HIP_CHECK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, opt.device));
    wall_khz_ = wall_khz > 0 ? uint64_t(wall_khz) : 100000u;
    // Test hook: RRTL_HIP_INJECT_QUEUE_ERROR=k raises the work queue's error
    // flag from the host before the run's k-th queue launch (0-based), so a
    // test can check that an error in any launch fails the whole run.
    if (const char* inj = std::getenv("RRTL_HIP_INJECT_QUEUE_ERROR")) inject_queue_error_ = uint64_t(std::strtoull(inj, nullptr, 10));
    RRTL_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    stage_row_ = N.n_nodes;
    rows_ = N.n_nodes + uint32_t(N.commits.size());
    state_.alloc(size_t(rows_) * nwords_);
    state_.zero(stream_);
    RRTL_HIP_CHECK(hipMemsetAsync(state_.p + size_t(N.c1) * nwords_, 0xFF, nwords_ * 4, stream_));
    std::vector<uint4> sg, ng, cm;
    for (auto& g : N.settle) sg.push_back(make_uint4(g.op, g.a, g.b, g.dst));
    for (auto& g : N.next) ng.push_back(make_uint4(g.op, g.a, g.b, g.dst));
    for (auto& c : N.commits) cm.push_back(make_uint4(c.reg, c.val, c.rsig == NONE ? 0 : c.rsig, c.pol | (c.rbit << 1) | (c.has_reset << 2)));
    settle_g_.upload(sg, stream_); next_g_.upload(ng, stream_); commit_g_.upload(cm, stream_);
    RRTL_HIP_CHECK(hipEventCreate(&ev0_)); RRTL_HIP_CHECK(hipEventCreate(&ev1_));
    eval_combinational();
    synchronize();
  }
  ~GateEngine() override {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    state_.release(); settle_g_.release(); next_g_.release(); commit_g_.release();
    fstart_.release(); fent_.release(); tile_force_.release(); tile_fstart_.release(); digest_.release(); trace_.release();
    detect_.release(); group_key_.release();
    for (auto& b : prog_bufs_) b.release();
    if (stream_) (void)hipStreamDestroy(stream_);
  }
  uint32_t kind() const override { return ENGINE_GATE; }

  void set_signal_limbs(uint32_t sig, const uint32_t* lane_limbs, bool settle) override {
    check_sig(sig);
    refresh_if_stale();
    std::vector<uint32_t> row(nwords_, 0);
    for (uint32_t l = 0; l < lanes_; l++) if (lane_limbs[l] & 1) row[l / 32] |= 1u << (l % 32);
    RRTL_HIP_CHECK(hipMemcpyAsync(state_.p + size_t(sig) * nwords_, row.data(), nwords_ * 4, hipMemcpyHostToDevice, stream_));
    synchronize();
    if (settle) eval_combinational();
  }
  void get_signal_limbs(uint32_t sig, uint32_t* out) override {
    check_sig(sig);
    refresh_if_stale();
    std::vector<uint32_t> row(nwords_);
    RRTL_HIP_CHECK(hipMemcpyAsync(row.data(), state_.p + size_t(sig) * nwords_, nwords_ * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t l = 0; l < lanes_; l++) out[l] = (row[l / 32] >> (l % 32)) & 1u;
  }
  void get_signal_words(uint32_t sig, uint32_t* out) {
    check_sig(sig);
    refresh_if_stale();
    RRTL_HIP_CHECK(hipMemcpyAsync(out, state_.p + size_t(sig) * nwords_, nwords_ * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
  }
  void set_memory_limbs(uint32_t, const uint32_t*, bool) override { throw Error("E_BITPARALLEL: the gate engine has no memories"); }
  void get_memory_limbs(uint32_t, uint32_t*) override { throw Error("E_BITPARALLEL: the gate engine has no memories"); }

  void eval_combinational() override {
    launch_hbm_gates(settle_g_, uint32_t(N.settle.size()));
    wires_stale_ = false;
  }
  void tick() override { eval_combinational(); edge(); eval_combinational(); }
  void tick_from_evaluated() override { refresh_if_stale(); edge(); }
  void tick_many(uint64_t steps) override {
    if (!steps) return;
    RunSpec s;
    s.cycles = steps;
    run_impl(s, false);
  }
  void run(const RunSpec& spec) override { run_impl(spec, true); }

  void snapshot(uint32_t* values, uint32_t*) override {
    refresh_if_stale();
    std::vector<uint32_t> rows(size_t(N.nsig) * nwords_);
    RRTL_HIP_CHECK(hipMemcpyAsync(rows.data(), state_.p, rows.size() * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (uint32_t s = 0; s < N.nsig; s++)
      for (uint32_t l = 0; l < lanes_; l++)
        values[size_t(P.signals[s].offset) * lanes_ + l] = (rows[size_t(s) * nwords_ + l / 32] >> (l % 32)) & 1u;
  }
  void restore(const uint32_t* values, const uint32_t*) override {
    std::vector<uint32_t> rows(size_t(N.nsig) * nwords_, 0);
    for (uint32_t s = 0; s < N.nsig; s++)
      for (uint32_t l = 0; l < lanes_; l++)
        if (values[size_t(P.signals[s].offset) * lanes_ + l] & 1) rows[size_t(s) * nwords_ + l / 32] |= 1u << (l % 32);
    RRTL_HIP_CHECK(hipMemcpyAsync(state_.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, stream_));
    synchronize();
    wires_stale_ = false;
  }
  void set_force(uint32_t sig, uint32_t lane, int value) override {
    check_sig(sig);
    if (lane >= lanes_) throw Error("E_SIM_IR_LANE: lane is out of range");
    if (P.signals[sig].kind == K_INPUT) throw Error("E_HIP_FORCE_INPUT: inputs are driven by the caller; force the net they feed instead");
    forces_.push_back({sig, lane, value ? 1u : 0u});
    forces_dirty_ = true;
    forces_version_++;
  }
  void set_forces(const uint32_t* sig, const uint32_t* lane, const uint8_t* value, size_t n) override {
    for (size_t i = 0; i < n; i++) {
      check_sig(sig[i]);
      if (lane[i] >= lanes_) throw Error("E_SIM_IR_LANE: lane is out of range");
      if (P.signals[sig[i]].kind == K_INPUT) throw Error("E_HIP_FORCE_INPUT: inputs are driven by the caller; force the net they feed instead");
    }
    forces_.reserve(forces_.size() + n);
    for (size_t i = 0; i < n; i++) forces_.push_back({sig[i], lane[i], value[i] ? 1u : 0u});
    forces_dirty_ = true;
    forces_version_++;
  }
  void clear_forces() override { forces_.clear(); forces_dirty_ = true; forces_version_++; }

  void get_detect(uint32_t* out) override {
    if (!has_detect_) throw Error("E_HIP_NO_RUN: no run has produced detection flags");
    RRTL_HIP_CHECK(hipMemcpyAsync(out, detect_.p, nwords_ * 4, hipMemcpyDeviceToHost, stream_));
    synchronize();
  }

  size_t digest_bytes() const override { return has_digest_ ? size_t(64) * nwords_ * 4 : 0; }
  void get_digest(void* out) override {
    if (!has_digest_) throw Error("E_HIP_NO_RUN: no run has produced a digest");
    RRTL_HIP_CHECK(hipMemcpyAsync(out, digest_.p, digest_bytes(), hipMemcpyDeviceToHost, stream_));
    synchronize();
  }
  size_t trace_bytes() const override { return trace_bytes_; }
  void get_trace(void* out) override {
    if (!trace_bytes_) throw Error("E_HIP_NO_RUN: no run has produced a trace");
    RRTL_HIP_CHECK(hipMemcpyAsync(out, trace_.p, trace_bytes_, hipMemcpyDeviceToHost, stream_));
    synchronize();
  }
  uint32_t tile_words() const { return W_; }
  uint32_t tile_slots() const { return last_slots_; }
  uint32_t tile_levels() const { return last_levels_; }
  uint32_t tile_gates() const { return last_gates_; }
  uint64_t nwords() const { return nwords_; }

 private:
  void check_sig(uint32_t sig) const { if (sig >= N.nsig) throw Error("E_SIM_IR_SIGNAL: signal index out of range"); }

  void refresh_if_stale() { if (wires_stale_) eval_combinational(); }

  HbmArgs hbm_args() {
    sync_forces();
    HbmArgs A{};
    A.state = state_.p; A.nwords = nwords_;
    A.commits = commit_g_.p; A.n_commits = uint32_t(N.commits.size()); A.stage_row = stage_row_;
    A.fstart = forces_.empty() ? nullptr : fstart_.p; A.fent = fent_.p; A.n_nodes = N.n_nodes;
    return A;
  }
  void launch_hbm_gates(DevBuf<uint4>& g, uint32_t n) {
    HbmArgs A = hbm_args();
    A.gates = g.p; A.n_gates = n;
    uint32_t bs = 64, grid = uint32_t((nwords_ + bs - 1) / bs);
    if (n) hipLaunchKernelGGL(hbm_gates_kernel, dim3(grid), dim3(bs), 0, stream_, A);
    RRTL_HIP_CHECK(hipGetLastError());
    synchronize();
  }
  void edge() {
    launch_hbm_gates(next_g_, uint32_t(N.next.size()));
    HbmArgs A = hbm_args();
    uint32_t bs = 64, grid = uint32_t((nwords_ + bs - 1) / bs);
    hipLaunchKernelGGL(hbm_commit_kernel, dim3(grid), dim3(bs), 0, stream_, A);
    RRTL_HIP_CHECK(hipGetLastError());
    synchronize();
  }

  // Forces merged per (key, u32 lane-word) into (and, or) masks, later forces
  // winning: entries sorted by key, word and arrival order, then folded.
  struct MergedForce { uint64_t key; uint32_t word, andm, orm; };
  template <class KeyOf>
  std::vector<MergedForce> merge_forces(KeyOf key_of) const {
    struct E { uint64_t key; uint32_t word, idx; };
    std::vector<E> e;
    e.reserve(forces_.size());
    for (uint32_t i = 0; i < forces_.size(); i++) {
      uint64_t k;
      uint32_t w;
      if (key_of(forces_[i], k, w)) e.push_back({k, w, i});
    }
    std::sort(e.begin(), e.end(), [](const E& a, const E& b) {
      return a.key != b.key ? a.key < b.key : a.word != b.word ? a.word < b.word : a.idx < b.idx;
    });
    std::vector<MergedForce> out;
    for (size_t i = 0; i < e.size(); i++) {
      if (i == 0 || e[i].key != e[i - 1].key || e[i].word != e[i - 1].word) out.push_back({e[i].key, e[i].word, ~0u, 0u});
      const Force& f = forces_[e[i].idx];
      const uint32_t bit = 1u << (f.lane % 32);
      out.back().andm &= ~bit;
      if (f.value) out.back().orm |= bit; else out.back().orm &= ~bit;
    }
    return out;
  }

  // per-node force lists for the HBM kernels
  void sync_forces() {
    if (!forces_dirty_) return;
    std::vector<uint32_t> start(N.n_nodes + 1, 0);
    auto m = merge_forces([](const Force& f, uint64_t& k, uint32_t& w) { k = f.sig; w = f.lane / 32; return true; });
    std::vector<uint32_t> ent;
    ent.reserve(m.size() * 3);
    for (auto& x : m) { start[x.key + 1]++; ent.insert(ent.end(), {x.word, x.andm, x.orm}); }
    for (uint32_t n = 0; n < N.n_nodes; n++) start[n + 1] += start[n];
    fstart_.upload(start, stream_);
    fent_.upload(ent, stream_);
    forces_dirty_ = false;
    synchronize();
  }

  template <class T>
  T* upload_prog(size_t idx, const std::vector<T>& v) {
    auto& b = prog_bufs_[idx];
    size_t bytes = std::max<size_t>(1, v.size() * sizeof(T));
    b.alloc((bytes + 3) / 4);
    if (!v.empty()) RRTL_HIP_CHECK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
    return reinterpret_cast<T*>(b.p);
  }

  // LDS of one tile: live slots, the sink slots of gate-less threads, the
  // 64-word signature.
  static uint32_t sink_base(uint32_t n_slots) { return (n_slots + 31u) & ~31u; }  // residue of sink + lane = lane
  static size_t tile_lds_bytes(uint32_t n_slots, uint32_t w) {
    return size_t(sink_base(n_slots) + SINK_SLOTS + 64) * w * 4;
  }

  uint32_t choose_tile_words(uint32_t n_slots) const {
    if (opt_.tile_words) return opt_.tile_words;
    // The W >= 2 that keeps the most workgroups resident: tiles of the run,
    // capped by the CUs times the workgroups whose LDS fits a CU (at most
    // four: the one-pass kernel's register budget); ties go to the wider tile,
    // which spends fewer descriptor fetches and address operations per lane.
    uint32_t best = 0;
    uint64_t best_res = 0;
    for (uint32_t w : {2u, 4u, 8u}) {
      const uint64_t per_cu = std::min<uint64_t>(lds_max_ / tile_lds_bytes(n_slots, w), TILE_WAVES_PER_SIMD);
      const uint64_t res = std::min<uint64_t>((nwords_ + w - 1) / w, per_cu * cus_);
      if (per_cu > 0 && res >= best_res) { best = w; best_res = res; }
    }
    return best ? best : 1;
  }

  // W lane-words per tile; byte_addr: descriptors hold LDS byte offsets.
  struct Prepared {
    TileArgs A;
    uint32_t W = 0, M = 2, tiles = 0;
    size_t lds = 0;
    bool byte_addr = false;
    std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>> clamp_sites;  // forced node -> (position, slot)
    uint64_t force_version = ~0ull;                                               // forces the entries were built from
  };

  // The compiled tile program depends on which inputs are driven, what is
  // observed and where forces sit; repeated runs with the same shape reuse it.
  std::string plan_key(const RunSpec& spec) const {
    std::string k;
    auto put = [&](uint64_t v) { k.append(reinterpret_cast<const char*>(&v), 8); };
    put(spec.in_sig.size());
    for (size_t i = 0; i < spec.in_sig.size(); i++) { put(spec.in_sig[i]); put(spec.in_mode[i]); put(spec.in_arg.empty() ? 0 : spec.in_arg[i]); }
    put(spec.obs_sig.size());
    for (uint32_t s : spec.obs_sig) put(s);
    put(spec.observe_post); put(opt_.tile_words); put(opt_.stage_gates);
    // which signals carry a force shapes the program; the lanes do not
    std::vector<uint8_t> forced(N.nsig, 0);
    for (auto& f : forces_) forced[f.sig] = 1;
    for (uint32_t s = 0; s < N.nsig; s++) if (forced[s]) put(s);
    return k;
  }

  Prepared prepare(const RunSpec& spec, bool with_outputs) {
    if (spec.expect) throw Error("E_HIP_UNSUPPORTED: trace replay with checks runs on the word engine");
    for (uint32_t s : spec.obs_sig) check_sig(s);
    for (uint32_t s : spec.in_sig) check_sig(s);
    if (spec.halt_sig != NONE) throw Error("E_HIP_UNSUPPORTED: the gate engine does not halt lanes");
    // No settle of stale wires here: a tile run loads only nodes the settle
    // never writes (inputs, registers, host-driven wires) and recomputes every
    // wire it reads, so stale wire rows stay stale until an API read.
    const std::string key = plan_key(spec);
    Prepared R;
    if (!(key == plan_key_ && have_plan_)) {
      plan_ = compile_and_upload(spec);
      plan_key_ = key; have_plan_ = true;
    }
    if (plan_.force_version != forces_version_) {
      upload_force_entries(plan_);
      plan_.force_version = forces_version_;
    }
    R = plan_;
    TileArgs& A = R.A;
    // detection: lanes compared with lane 0 of their 64-lane group, which
    // needs both u32 words of every group in one tile
    const bool detect = with_outputs && spec.detect;
    if (spec.detect && spec.obs_sig.empty()) throw Error("E_HIP_RUN_SPEC: detection needs observed signals");
    if (detect && R.W < 2) throw Error("E_HIP_UNSUPPORTED: detection needs tiles of at least two lane-words (64 lanes)");
    has_detect_ = detect;
    if (detect) {
      detect_.alloc(nwords_);
      fill_u32(detect_.p, nwords_, 0u, stream_);
    }
    A.detect = detect ? detect_.p : nullptr;
    const uint64_t groups = (nwords_ + 1) / 2;
    if (!spec.group_key.empty()) {
      if (spec.group_key.size() < groups) throw Error("E_HIP_RUN_SPEC: group_key needs one key per 64-lane group");
      group_key_.alloc(groups);
      RRTL_HIP_CHECK(hipMemcpyAsync(group_key_.p, spec.group_key.data(), groups * 8, hipMemcpyHostToDevice, stream_));
    }
    A.group_key = spec.group_key.empty() ? nullptr : group_key_.p;
    has_digest_ = with_outputs && !spec.obs_sig.empty() && !detect;
    if (has_digest_) { digest_.alloc(size_t(64) * nwords_); digest_.zero(stream_); }
    A.digest = has_digest_ ? digest_.p : nullptr;
    if (!with_outputs) A.n_obs = 0;
    trace_bytes_ = 0;
    if (with_outputs && spec.trace && !spec.obs_sig.empty()) {
      trace_bytes_ = size_t(spec.cycles) * spec.obs_sig.size() * nwords_ * 4;
      trace_.alloc(trace_bytes_ / 4);
    }
    A.trace = trace_bytes_ ? trace_.p : nullptr;
    A.seed = spec.seed; A.observe_post = spec.observe_post ? 1 : 0;
    A.cycle_base = 0; A.cycles = spec.cycles;
    return R;
  }

  Prepared compile_and_upload(const RunSpec& spec) {
    std::vector<uint8_t> forced(N.nsig, 0);
    for (auto& f : forces_) forced[f.sig] = 1;
    // Gate positions per thread per stage: the M whose stage count x (fixed
    // stage cost + per-position cost) is lowest, unless the options fix it.
    // With four workgroups per CU the barrier waits are hidden and a stage
    // costs what its positions cost: on the 20k-gate netlist 35 stages of 2
    // positions (70 per cycle) run 12 % faster than 26 of 3 (78) and 25 %
    // faster than 22 of 4 (88).
    TileProgram T;
    if (opt_.stage_gates) {
      T = compile_tile(N, P, spec, forced, opt_.stage_gates);
    } else {
      double best = 0;
      for (uint32_t m : {2u, 3u, 4u}) {
        TileProgram C = compile_tile(N, P, spec, forced, m);
        const double cost = double(C.n_cycle_stages + C.n_post_stages) * (0.25 + m);
        if (m == 2 || cost < best) { best = cost; T = std::move(C); }
      }
    }
    uint32_t W = choose_tile_words(T.n_slots);
    if (W != 1 && W != 2 && W != 4 && W != 8) throw Error("E_HIP_OPTIONS: tile_words must be 1, 2, 4 or 8");
    size_t lds = tile_lds_bytes(T.n_slots, W);
    if (lds > lds_max_) {
      W = 1; lds = tile_lds_bytes(T.n_slots, 1);
      if (lds > lds_max_) throw Error("E_HIP_TOO_LARGE: live gate state does not fit in LDS for one 32-lane tile");
    }
    const uint32_t sink = T.sink, n_slots = sink + SINK_SLOTS;
    if (sink != sink_base(T.n_slots)) throw Error("E_HIP_INTERNAL: sink slot base");
    const bool byte_addr = lds <= 65536;
    const uint32_t M = T.M;
    W_ = W; M_ = M; last_slots_ = T.n_slots; last_levels_ = T.n_cycle_stages + T.n_post_stages; last_gates_ = T.n_gates;
    const uint32_t tiles = uint32_t((nwords_ + W - 1) / W);
    TileArgs A{};
    A.state = state_.p; A.nwords = nwords_;
    // descriptors compiled as [stage][position][thread], uploaded as
    // [stage][thread][position] (a thread's positions in one fetch), with
    // STAGE_AHEAD rows past the last stage; every field is checked against
    // the tile's slots
    const uint32_t KS = T.n_cycle_stages + T.n_post_stages;
    const size_t stride = size_t(M) * TILE_THREADS;
    auto enc = [&](uint32_t slot) {
      if (slot >= n_slots) throw Error("E_HIP_INTERNAL: descriptor slot out of range");
      return byte_addr ? slot * W * 4 : slot;
    };
    std::vector<uint2> desc(size_t(KS + STAGE_AHEAD) * stride);
    auto at = [&](size_t i) -> uint2& {  // compiled index -> uploaded index
      return desc[i - i % stride + (i % TILE_THREADS) * M + (i % stride) / TILE_THREADS];
    };
    for (size_t i = 0; i < desc.size(); i++) {
      if (i < size_t(KS) * stride) {
        const uint32_t x = T.desc[2 * i], y = T.desc[2 * i + 1];
        const uint32_t op = (y >> 16) & 0xFu;
        if (byte_addr) {
          // position 0: a | b << 16; position 1: the thread's mode bits (xor,
          // fix per position) | b << 16; later positions: b << 16; the input
          // inversions as 0x00 / 0xFF bytes of y
          const size_t k = (i % stride) / TILE_THREADS, row = i - (i % stride), lane = i % TILE_THREADS;
          uint32_t low = 0;
          if (k == 0) {
            low = enc(x & 0xFFFFu);
          } else if (k == 1) {
            for (uint32_t q = 0; q < M; q++) {
              const uint32_t oq = (T.desc[2 * (row + q * TILE_THREADS + lane) + 1] >> 16) & 0xFu;
              low |= (((oq >> 3) & 1u) | (((oq >> 2) & 1u) << 1)) << (2 * q);
            }
          }
          at(i) = make_uint2(low | (enc(x >> 16) << 16),
                             enc(y & 0xFFFFu) | ((op & GI_A) ? 0xFF0000u : 0u) | ((op & GI_B) ? 0xFF000000u : 0u));
        } else {
          at(i) = make_uint2(enc(x & 0xFFFFu) | (enc(x >> 16) << 16), enc(y & 0xFFFFu) | (op << 16));
        }
      } else {
        at(i) = make_uint2(enc(T.c0) | (enc(T.c0) << 16), enc(sink + uint32_t(i % 32)));
      }
    }
    // the rows past the last stage repeat stage (row % 2): a cycle's last
    // fetches then leave the next cycle's first two descriptor sets in place
    for (uint32_t q = 0; q < STAGE_AHEAD; q++) {
      const size_t src = (KS + q) % 2;
      if (src < KS) std::copy(desc.begin() + src * stride, desc.begin() + (src + 1) * stride, desc.begin() + (KS + q) * stride);
    }
    A.desc = upload_prog(0, desc);
    A.any_fix = 0;
    for (uint32_t f : T.stage_fix) A.any_fix |= f;
    A.n_cycle_stages = T.n_cycle_stages; A.n_post_stages = T.n_post_stages;
    A.load_slot = upload_prog(2, T.load_slot); A.load_node = upload_prog(3, T.load_node); A.n_load = uint32_t(T.load_slot.size());
    A.stim_slot = upload_prog(4, T.stim_slot); A.stim_mode = upload_prog(5, T.stim_mode);
    A.stim_stream = upload_prog(6, T.stim_stream); A.stim_arg = upload_prog(7, T.stim_arg); A.n_stim = uint32_t(T.stim_slot.size());
    // commits: reset-free ones at byte offsets go to the register-held list
    // (value inversion in bit 0 of the value offset: offsets are multiples of
    // 4), the others to the general table
    std::vector<uint4> cm;
    std::vector<uint32_t> cf;
    const size_t fast_cap = size_t(FAST_COMMITS_PER_THREAD) * TILE_THREADS;
    std::vector<uint32_t> cf_reg, cf_val, cm_reg, cm_val;
    for (size_t k = 0; k < T.commit.size() / 4; k++) {
      const uint32_t reg = T.commit[4 * k], val = T.commit[4 * k + 1], rs = T.commit[4 * k + 2], fl = T.commit[4 * k + 3];
      if (byte_addr && !(fl & 4u) && cf.size() < fast_cap) {
        cf.push_back(enc(reg) | ((enc(val) | ((fl >> 3) & 1u)) << 16));
        cf_reg.push_back(reg); cf_val.push_back(val);
      } else {
        cm.push_back(make_uint4(reg, val, rs, fl));
        cm_reg.push_back(reg); cm_val.push_back(val);
      }
    }
    // commits of one list are independent (no value slot is a committed
    // register's), so their order is free: spread them over the LDS banks
    auto reorder = [](auto& v, const std::vector<uint32_t>& ord) {
      auto c = v; for (size_t i = 0; i < ord.size(); i++) v[i] = c[ord[i]];
    };
    reorder(cf, bank_order(cf_val, cf_reg));
    reorder(cm, bank_order(cm_val, cm_reg));
    A.commit = upload_prog(8, cm); A.n_commit = uint32_t(cm.size());
    A.commit_fast = upload_prog(13, cf); A.n_commit_fast = uint32_t(cf.size());
    A.obs_slot = upload_prog(9, T.obs_slot); A.n_obs = uint32_t(T.obs_slot.size());
    // One barrier phase for the PRE observation, the commit and the next
    // cycle's inputs when no observed slot is committed or driven, no
    // committed value or reset is read from a driven input, and there is no
    // post program.
    {
      std::unordered_set<uint32_t> written(T.stim_slot.begin(), T.stim_slot.end());
      bool ok = !spec.observe_post && T.n_post_stages == 0;
      for (size_t k = 0; ok && k < T.commit.size() / 4; k++) {
        const uint32_t val = T.commit[4 * k + 1], rs = T.commit[4 * k + 2], fl = T.commit[4 * k + 3];
        if (written.count(val) || ((fl & 4u) && written.count(rs))) ok = false;
      }
      for (size_t k = 0; k < T.commit.size() / 4; k++) written.insert(T.commit[4 * k]);
      for (uint32_t o : T.obs_slot) if (written.count(o & 0x7FFFFFFFu)) ok = false;
      A.merge_tail = ok ? 1u : 0u;
    }
    A.n_slots = n_slots; A.c0 = T.c0; A.c1 = T.c1;
    Prepared R;
    R.A = A; R.W = W; R.M = M; R.tiles = tiles; R.lds = lds; R.byte_addr = byte_addr;
    R.clamp_sites = std::move(T.clamp_sites);
    return R;
  }

  // Per-tile force clamp entries (pos, slot, word, and, or), merged per LDS
  // word; within a tile the stage clamps come first (ordered by position) and
  // the commit clamps (position CLAMP_COMMIT) last, from force_mid on.
  void upload_force_entries(Prepared& R) {
    TileArgs& A = R.A;
    A.force = A.force_start = A.force_mid = nullptr;
    if (forces_.empty()) return;
    const uint32_t W = R.W, tiles = R.tiles;
    // clamp sites ranked by (position, slot); each forced signal's site ranks
    std::vector<std::pair<uint32_t, uint32_t>> site;  // (position, slot) by rank
    std::unordered_map<uint32_t, std::vector<uint32_t>> ranks_of;
    {
      std::vector<uint32_t> ord(R.clamp_sites.size());
      for (uint32_t i = 0; i < ord.size(); i++) ord[i] = i;
      std::sort(ord.begin(), ord.end(), [&](uint32_t x, uint32_t y) { return R.clamp_sites[x].second < R.clamp_sites[y].second; });
      for (uint32_t r = 0; r < ord.size(); r++) {
        site.push_back(R.clamp_sites[ord[r]].second);
        ranks_of[R.clamp_sites[ord[r]].first].push_back(r);
      }
    }
    // one item per (force, site), keyed tile | rank, then word, then arrival
    struct E { uint64_t key; uint32_t word, idx; };
    std::vector<E> e;
    e.reserve(forces_.size());
    for (uint32_t i = 0; i < forces_.size(); i++) {
      auto it = ranks_of.find(forces_[i].sig);
      if (it == ranks_of.end()) continue;
      const uint32_t gw = forces_[i].lane / 32;
      for (uint32_t r : it->second) e.push_back({(uint64_t(gw / W) << 32) | r, gw % W, i});
    }
    std::sort(e.begin(), e.end(), [](const E& x, const E& y) {
      return x.key != y.key ? x.key < y.key : x.word != y.word ? x.word < y.word : x.idx < y.idx;
    });
    std::vector<uint32_t> fent, fstart(tiles + 1, 0), fmid(tiles, 0);
    for (size_t i = 0; i < e.size(); i++) {
      if (i == 0 || e[i].key != e[i - 1].key || e[i].word != e[i - 1].word) {
        const auto& ps = site[uint32_t(e[i].key)];
        fent.insert(fent.end(), {ps.first, ps.second, e[i].word, ~0u, 0u});
        fstart[uint32_t(e[i].key >> 32) + 1]++;
      }
      const Force& f = forces_[e[i].idx];
      uint32_t* x = &fent[fent.size() - 5];
      const uint32_t bit = 1u << (f.lane % 32);
      x[3] &= ~bit;
      if (f.value) x[4] |= bit; else x[4] &= ~bit;
    }
    if (fent.empty()) return;
    for (uint32_t t = 0; t < tiles; t++) fstart[t + 1] += fstart[t];
    for (uint32_t t = 0; t < tiles; t++) {
      uint32_t m = fstart[t];
      while (m < fstart[t + 1] && fent[size_t(m) * 5] != CLAMP_COMMIT) m++;
      fmid[t] = m;
    }
    A.force = upload_prog(10, fent);
    A.force_start = upload_prog(11, fstart);
    A.force_mid = upload_prog(12, fmid);
  }

  // Chunks per tile for the work-queue schedule: the K in 2..16 whose rounds
  // ceil(tiles * K / resident) / K come closest to the ideal tiles / resident
  // (1 when every tile is resident at once).
  static uint32_t choose_chunks(uint32_t tiles, uint32_t resident) {
    if (resident == 0 || tiles <= resident) return 1;
    uint32_t best = 1;
    double best_cost = std::ceil(double(tiles) / resident);
    for (uint32_t k = 2; k <= 16; k++) {
      const double cost = std::ceil(double(tiles) * k / resident) / k;
      if (cost < best_cost - 1e-9) { best = k; best_cost = cost; }
    }
    return best;
  }

  // A prepared run and its launch schedule.
  struct Launch {
    Prepared R;
    uint64_t cycles = 0;
    uint32_t chunks = 1, resident = 0;
  };
  Launch plan_launch(const RunSpec& spec, bool with_outputs) {
    Launch L;
    L.R = prepare(spec, with_outputs);
    L.cycles = spec.cycles;
    // one pass when the one-pass kernel holds every tile at once, else the
    // work queue with its own residency
    const uint32_t one_pass = resident_tiles(tile_fn(L.R.W, L.R.byte_addr, L.R.M), L.R.lds);
    L.resident = L.R.tiles <= one_pass ? one_pass : resident_tiles(queue_fn(L.R.W, L.R.byte_addr, L.R.M), L.R.lds);
    L.chunks = opt_.tile_chunks ? opt_.tile_chunks : (L.R.tiles <= one_pass ? 1u : choose_chunks(L.R.tiles, L.resident));
    if (L.chunks > 1) qctl_.alloc(size_t(L.R.tiles) + 2);
    if (L.R.lds > 65536)
      RRTL_HIP_CHECK(hipFuncSetAttribute(tile_fn(L.R.W, L.R.byte_addr, L.R.M), hipFuncAttributeMaxDynamicSharedMemorySize,
                                         int(L.R.lds)));
    return L;
  }
  // The run's kernels on stream s: kernels only (no copies, fills or host
  // waits), so the sequence can be captured into a graph.  Returns the launches.
  uint64_t enqueue(const Launch& L, hipStream_t s) {
    TileArgs A = L.R.A;
    const uint64_t chunk = opt_.ticks_per_launch ? opt_.ticks_per_launch : std::max<uint64_t>(L.cycles, 1);
    uint64_t launches = 0, queue_launches = 0;
    last_chunks_ = 1;
    for (uint64_t done = 0; done < L.cycles; done += chunk) {
      A.cycle_base = done;
      A.cycles = std::min<uint64_t>(chunk, L.cycles - done);
      if (A.trace) A.trace = trace_.p + done * A.n_obs * nwords_;
      if (L.chunks > 1 && A.cycles > 1) {
        QueueArgs Q{};
        Q.ctl = qctl_.p; Q.n_tiles = L.R.tiles;
        Q.chunk = (A.cycles + L.chunks - 1) / L.chunks;
        Q.n_chunks = uint32_t((A.cycles + Q.chunk - 1) / Q.chunk);
        Q.wait_limit = wall_khz_ * 1000u * 120u;  // two minutes
        const uint32_t err_mode = queue_launches == inject_queue_error_ ? 2u : queue_launches == 0 ? 1u : 0u;
        hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(256), 0, s, qctl_.p, L.R.tiles + 2, err_mode);
        queue_launches++;
        const uint32_t grid = std::min<uint32_t>(std::max<uint32_t>(L.resident, 1), L.R.tiles * Q.n_chunks);
        launch_queue(L.R, grid, A, Q, s);
        last_chunks_ = Q.n_chunks;
      } else {
        launch_tile(L.R, A, s);
      }
      launches++;
    }
    return launches;
  }
  // After the run's kernels finished: the work queue's error flag, stats.
  void complete(const Launch& L) {
    if (L.chunks > 1) {
      uint32_t err = 0;
      RRTL_HIP_CHECK(hipMemcpy(&err, qctl_.p + 1, sizeof(err), hipMemcpyDeviceToHost));
      if (err) throw Error("E_HIP_INTERNAL: tile work queue: a chunk waited past its limit for the previous chunk");
    }
    stats_.lane_cycles = uint64_t(lanes_) * L.cycles;
    wires_stale_ = wires_stale_ || L.cycles > 0;
  }

  void run_impl(const RunSpec& spec, bool with_outputs) {
    const Launch L = plan_launch(spec, with_outputs);
    stats_ = RunStats{};
    RRTL_HIP_CHECK(hipStreamSynchronize(stream_));
    RRTL_HIP_CHECK(hipEventRecord(ev0_, stream_));
    stats_.launches = enqueue(L, stream_);
    RRTL_HIP_CHECK(hipEventRecord(ev1_, stream_));
    RRTL_HIP_CHECK(hipEventSynchronize(ev1_));
    float ms = 0;
    RRTL_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
    stats_.kernel_ms = ms;
    complete(L);
  }

  void async_prepare(const RunSpec& spec) override {
    async_ = plan_launch(spec, true);
    batch_digest_ = async_.R.A.digest != nullptr;
    batch_detect_ = async_.R.A.detect != nullptr;
    RRTL_HIP_CHECK(hipStreamSynchronize(stream_));
  }
  uint64_t async_enqueue(hipStream_t s) override { return enqueue(async_, s); }
  void async_finish(const RunSpec& spec) override {
    (void)spec;
    stats_ = RunStats{};
    complete(async_);
  }

  BatchJob batch_prepare(const RunSpec& spec, const BatchRequest& req) override {
    (void)req;  // a tile needs its LDS whatever the batch; the batch sizes its budget from the tiles
    Prepared R = prepare(spec, true);
    batch_digest_ = R.A.digest != nullptr;
    batch_detect_ = R.A.detect != nullptr;
    BatchJob J;
    J.kind = ENGINE_GATE;
    J.call = "rrtl::gate::tile_body<" + std::to_string(R.W) + ", " + (R.byte_addr ? "true" : "false") + ", " +
             std::to_string(R.M) + ">(U.g, item, lds);";
    J.form = "gate tile W=" + std::to_string(R.W) + " M=" + std::to_string(R.M);
    J.n_items = R.tiles; J.lds_bytes = R.lds;
    // per tile and cycle: one barrier-separated pass per stage plus its gates
    J.item_cost = double(spec.cycles) * (double(R.A.n_cycle_stages + R.A.n_post_stages) * (48.0 + 8.0 * R.M) +
                                         double(last_gates_) * R.W / 16.0);
    J.args.resize(sizeof(TileArgs));
    std::memcpy(J.args.data(), &R.A, sizeof(TileArgs));
    RRTL_HIP_CHECK(hipStreamSynchronize(stream_));
    return J;
  }
  void batch_reset(hipStream_t s) override {
    if (batch_digest_) RRTL_HIP_CHECK(hipMemsetAsync(digest_.p, 0, digest_.n * 4, s));
    if (batch_detect_) fill_u32(detect_.p, nwords_, 0u, s);
  }
  void batch_finish(const RunSpec& spec) override {
    stats_ = RunStats{};
    stats_.lane_cycles = uint64_t(lanes_) * spec.cycles;
    wires_stale_ = wires_stale_ || spec.cycles > 0;
  }

  // Kernel entry points of the compiled variant (W, byte addressing, M).
  static const void* tile_fn(uint32_t W, bool byte_addr, uint32_t M) {
    const void* fn = nullptr;
    with_tile_variant(W, byte_addr, M, [&](auto w, auto ba, auto m) {
      fn = reinterpret_cast<const void*>(&tile_kernel<decltype(w)::value, decltype(ba)::value, decltype(m)::value>);
    });
    return fn;
  }
  static const void* queue_fn(uint32_t W, bool byte_addr, uint32_t M) {
    const void* fn = nullptr;
    with_tile_variant(W, byte_addr, M, [&](auto w, auto ba, auto m) {
      fn = reinterpret_cast<const void*>(&tile_queue_kernel<decltype(w)::value, decltype(ba)::value, decltype(m)::value>);
    });
    return fn;
  }
  // Workgroups of the queue kernel the device holds at once.
  // Tiles the device holds at once for a kernel (the one-pass and the
  // work-queue kernels have different register budgets).
  uint32_t resident_tiles(const void* fn, size_t lds) {
    if (lds > 65536) RRTL_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    int per_cu = 0;
    RRTL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TILE_THREADS, lds));
    return uint32_t(std::max(per_cu, 0)) * cus_;
  }
  void launch_queue(const Prepared& R, uint32_t grid, const TileArgs& A, const QueueArgs& Q, hipStream_t s) {
    with_tile_variant(R.W, R.byte_addr, R.M, [&](auto w, auto ba, auto m) {
      hipLaunchKernelGGL((tile_queue_kernel<decltype(w)::value, decltype(ba)::value, decltype(m)::value>), dim3(grid),
                         dim3(TILE_THREADS), R.lds, s, A, Q);
    });
    RRTL_HIP_CHECK(hipGetLastError());
  }
  void launch_tile(const Prepared& R, const TileArgs& A, hipStream_t s) {
    with_tile_variant(R.W, R.byte_addr, R.M, [&](auto w, auto ba, auto m) {
      hipLaunchKernelGGL((tile_kernel<decltype(w)::value, decltype(ba)::value, decltype(m)::value>), dim3(R.tiles),
                         dim3(TILE_THREADS), R.lds, s, A);
    });
    RRTL_HIP_CHECK(hipGetLastError());
  }

  GateNet N;
  uint64_t nwords_ = 0;
  uint32_t rows_ = 0, stage_row_ = 0;
  size_t lds_max_ = 65536;
  DevBuf<uint32_t> state_;
  DevBuf<uint4> settle_g_, next_g_, commit_g_;
  DevBuf<uint32_t> fstart_, fent_, tile_force_, tile_fstart_, digest_, trace_, detect_;
  DevBuf<uint64_t> group_key_;
  DevBuf<uint32_t> prog_bufs_[16];
  std::vector<Force> forces_;
  bool forces_dirty_ = true, wires_stale_ = false, has_digest_ = false;
  size_t trace_bytes_ = 0;
  Launch async_;  // the run of a concurrent batch
  uint32_t W_ = 0, M_ = 2, last_slots_ = 0, last_levels_ = 0, last_gates_ = 0, last_chunks_ = 1, cus_ = 1;
  uint64_t wall_khz_ = 100000;
  uint64_t inject_queue_error_ = ~uint64_t(0);
  DevBuf<uint32_t> qctl_;  // work-queue counter, error flag, per-tile chunk counts
  Prepared plan_;
  std::string plan_key_;
  bool have_plan_ = false, batch_digest_ = false, batch_detect_ = false, has_detect_ = false;
  uint64_t forces_version_ = 0;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

}  // namespace

Engine* make_gate_engine(Program prog, uint32_t lanes, const EngineOptions& opt) {
  return new GateEngine(std::move(prog), lanes, opt);
}

// gate-engine specific accessors used by the C ABI
void gate_engine_get_words(Engine* e, uint32_t sig, uint32_t* out) { static_cast<GateEngine*>(e)->get_signal_words(sig, out); }
void gate_engine_tile_info(Engine* e, uint32_t* info) {
  auto* g = static_cast<GateEngine*>(e);
  info[0] = g->tile_words(); info[1] = g->tile_slots(); info[2] = g->tile_levels(); info[3] = g->tile_gates();
}

}  // namespace rrtl
