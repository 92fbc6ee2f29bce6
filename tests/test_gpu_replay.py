"""Device-side trace replay against the reference's replay loop
(SimBackend::replay_trace, crates/rrtl-sim-ir/src/lib.rs:5270-5345) run over
the packed oracle: per step the inputs are set raw, comb is evaluated, every
active check compares each checked lane (ReplayCheckMode: lane 0 or all;
ReplayLaneMode::Independent: all, with per-lane values), then
tick_from_evaluated_no_post_eval.  The mismatch count, the first
max_mismatches in (step, check, lane) order (ReplayReport::record_mismatch,
:4388) and the final state must match."""
import numpy as np
import pytest

import oracle
from rrtl_amd import lower
from rrtl_amd.backend import (REPLAY_ALL_LANES, REPLAY_INDEPENDENT, REPLAY_LANE0, REPLAY_REPLICATED, HipError,
                              HipSimulator)
from rrtl_amd.program import KIND_INPUT
from conftest import flush_print
from designgen import random_design

pytestmark = pytest.mark.gpu


def u128(v):
    return sum(int(x) << (32 * k) for k, x in enumerate(list(v)[:4]))


def oracle_replay(ref, p, lanes, inputs, in_vals, checks, expected, lane_mode, check_mode, active, max_mm):
    indep = lane_mode == REPLAY_INDEPENDENT
    to_check = lanes if (indep or check_mode == REPLAY_ALL_LANES) else 1
    count, mism, actual = 0, [], []
    for s in range(in_vals.shape[0]):
        off = 0
        for sig in inputs:
            n = p.signals[sig].limbs
            v = in_vals[s, off:off + n] if indep else np.repeat(in_vals[s, off:off + n, None], lanes, axis=1)
            ref.set_signal_limbs(sig, np.ascontiguousarray(v), eval=False)
            off += n
        ref.eval_combinational()
        off, row = 0, []
        for j, sig in enumerate(checks):
            n = p.signals[sig].limbs
            act = ref.get_signal_limbs(sig)
            row.append(act)
            if active is None or (int(active[s, j // 32]) >> (j % 32)) & 1:
                for lane in range(to_check):
                    exp = expected[s, off:off + n, lane] if indep else expected[s, off:off + n]
                    if not np.array_equal(act[:, lane], exp):
                        count += 1
                        if len(mism) < max_mm:
                            mism.append((s, j, lane, u128(exp), u128(act[:, lane])))
            off += n
        actual.append(row)
        ref.tick_from_evaluated()
    return count, mism, actual


def setup(seed, lanes, codegen):
    p = lower.lower_module(random_design(seed))
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen)
    ref = oracle.PackedRef(p, lanes)
    rng = np.random.default_rng(seed)
    # lanes start from different register states, so replicated inputs still diverge
    for i, s in enumerate(p.signals):
        if s.kind != KIND_INPUT:
            v = rng.integers(0, 2**32, size=(s.limbs, lanes), dtype=np.uint32)
            sim.set_signal_limbs(i, v, settle=False)
            ref.set_signal_limbs(i, v, eval=False)
    inputs = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    checks = [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT and s.limbs <= 4]
    return p, sim, ref, rng, inputs, checks


def perturb(rng, p, checks, expected, per_lane, frac=0.03):
    """Flips one bit inside the width of about `frac` of the expected values."""
    exp = expected.copy()
    off = 0
    for sig in checks:
        s = p.signals[sig]
        top = s.limbs - 1
        shape = (exp.shape[0],) + ((exp.shape[2],) if per_lane else ())
        hit = rng.random(shape) < frac
        bit = rng.integers(0, s.width, size=shape)
        limb, pos = bit // 32, bit % 32
        for l in range(s.limbs):
            sel = hit & (limb == l)
            flip = np.where(sel, np.left_shift(np.uint32(1), pos.astype(np.uint32)), 0).astype(np.uint32)
            exp[:, off + l] ^= flip
        assert top >= 0
        off += s.limbs
    return exp


@pytest.mark.parametrize("codegen", ["on", "off"])
@pytest.mark.parametrize("lane_mode,check_mode", [(REPLAY_REPLICATED, REPLAY_LANE0), (REPLAY_REPLICATED, REPLAY_ALL_LANES),
                                                  (REPLAY_INDEPENDENT, REPLAY_ALL_LANES)])
def test_replay_matches_reference_loop(lane_mode, check_mode, codegen):
    flush_print(f"case: replay lane_mode={lane_mode} check_mode={check_mode} codegen={codegen}")
    lanes, steps, max_mm = 70, 30, 16
    for seed in (3, 8):
        p, sim, ref, rng, inputs, checks = setup(seed, lanes, codegen)
        indep = lane_mode == REPLAY_INDEPENDENT
        in_limbs = sum(p.signals[i].limbs for i in inputs)
        ck_limbs = sum(p.signals[i].limbs for i in checks)
        shape = (steps, in_limbs, lanes) if indep else (steps, in_limbs)
        in_vals = rng.integers(0, 2**32, size=shape, dtype=np.uint32)
        # the true values from a first oracle pass on a copy of the state, then perturbed
        probe = oracle.PackedRef(p, lanes)
        probe.restore(*ref.snapshot())
        _, _, actual = oracle_replay(probe, p, lanes, inputs, in_vals, checks,
                                     np.zeros((steps, ck_limbs, lanes) if indep else (steps, ck_limbs), np.uint32),
                                     lane_mode, check_mode, None, 0)
        true = np.concatenate([np.concatenate(row, axis=0)[None] for row in actual], axis=0)  # [steps][limbs][lanes]
        expected = true if indep else true[:, :, 0]
        expected = perturb(rng, p, checks, np.ascontiguousarray(expected), indep)
        active = rng.integers(0, 2**32, size=(steps, (len(checks) + 31) // 32), dtype=np.uint32) if seed == 8 else None
        cidx = [100 + 3 * j for j in range(len(checks))]
        got = sim.replay(inputs, in_vals, checks, expected, lane_mode=lane_mode, check_mode=check_mode,
                         check_active=active, check_index=cidx, max_mismatches=max_mm)
        count, mism, _ = oracle_replay(ref, p, lanes, inputs, in_vals, checks, expected, lane_mode, check_mode,
                                       active, max_mm)
        assert count > 0
        assert got["mismatch_count"] == count, (seed, got["mismatch_count"], count)
        assert got["mismatches"] == [(s, cidx[j], lane, e, a) for s, j, lane, e, a in mism], seed
        vs, ms = sim.snapshot()
        vr, mr = ref.snapshot()
        assert np.array_equal(vs, vr) and np.array_equal(ms, mr), seed
        sim.close()


def test_replay_clean_trace():
    """A plan whose expected values are the true ones reports nothing."""
    p, sim, ref, rng, inputs, checks = setup(5, 40, "auto")
    steps = 12
    in_vals = rng.integers(0, 2**32, size=(steps, sum(p.signals[i].limbs for i in inputs)), dtype=np.uint32)
    probe = oracle.PackedRef(p, 40)
    probe.restore(*ref.snapshot())
    _, _, actual = oracle_replay(probe, p, 40, inputs, in_vals, checks,
                                 np.zeros((steps, sum(p.signals[i].limbs for i in checks)), np.uint32),
                                 REPLAY_REPLICATED, REPLAY_ALL_LANES, None, 0)
    true = np.concatenate([np.concatenate(row, axis=0)[None] for row in actual], axis=0)
    # lane 0's values are right for lane 0 only; in Lane0Fast mode nothing differs
    got = sim.replay(inputs, in_vals, checks, np.ascontiguousarray(true[:, :, 0]), check_mode=REPLAY_LANE0)
    assert got == {"mismatch_count": 0, "mismatches": []}


@pytest.mark.parametrize("lane_mode,check_mode", [(REPLAY_REPLICATED, REPLAY_LANE0), (REPLAY_REPLICATED, REPLAY_ALL_LANES),
                                                  (REPLAY_INDEPENDENT, REPLAY_ALL_LANES)])
def test_gate_engine_replay_matches_reference_loop(lane_mode, check_mode):
    """The gate engine's replay (the reference's host loop over its API calls)
    on random 1-bit designs with resets: mismatch count, first mismatches and
    final state equal the reference loop over the packed oracle."""
    from designgen import random_gate_design
    flush_print(f"case: gate replay lane_mode={lane_mode} check_mode={check_mode}")
    lanes, steps, max_mm = 100, 20, 12
    for seed in (4, 9):
        p = lower.lower_module(random_gate_design(seed, observe_regs=True, async_resets=False))
        sim = HipSimulator(p, lanes, engine="gate")
        ref = oracle.PackedRef(p, lanes)
        rng = np.random.default_rng(seed)
        for i, sg in enumerate(p.signals):
            if sg.kind != KIND_INPUT:
                v = rng.integers(0, 2, size=(1, lanes), dtype=np.uint32)
                sim.set_signal_limbs(i, v, settle=False)
                ref.set_signal_limbs(i, v, eval=False)
        inputs = [i for i, sg in enumerate(p.signals) if sg.kind == KIND_INPUT]
        checks = [i for i, sg in enumerate(p.signals) if sg.kind != KIND_INPUT]
        indep = lane_mode == REPLAY_INDEPENDENT
        shape = (steps, len(inputs), lanes) if indep else (steps, len(inputs))
        in_vals = rng.integers(0, 2, size=shape, dtype=np.uint32)
        expected = rng.integers(0, 2, size=(steps, len(checks), lanes) if indep else (steps, len(checks)),
                                dtype=np.uint32)
        active = rng.integers(0, 2**32, size=(steps, (len(checks) + 31) // 32), dtype=np.uint32) if seed == 9 else None
        got = sim.replay(inputs, in_vals, checks, expected, lane_mode=lane_mode, check_mode=check_mode,
                         check_active=active, max_mismatches=max_mm)
        count, mism, _ = oracle_replay(ref, p, lanes, inputs, in_vals, checks, expected, lane_mode, check_mode,
                                       active, max_mm)
        assert count > 0
        assert got["mismatch_count"] == count, seed
        assert got["mismatches"] == mism, seed
        for i in range(len(p.signals)):
            assert np.array_equal(sim.get_signal_limbs(i), ref.get_signal_limbs(i)), p.signals[i].name
        sim.close()


def test_replay_errors():
    """A signal index past the program is rejected (by the binding's name
    check, or by the C ABI's replay check); a null report buffer is rejected
    by the C ABI."""
    import ctypes
    from rrtl_amd.backend import ReplaySpec
    p, sim, ref, rng, inputs, checks = setup(5, 40, "auto")
    with pytest.raises(HipError) as e:
        sim.replay([len(p.signals) + 3], np.zeros((2, 1), np.uint32), [], np.zeros((2, 0), np.uint32))
    assert e.value.code in ("E_SIM_IR_SIGNAL", "E_SIM_IR_REPLAY_SIGNAL")
    bad = np.array([len(p.signals) + 3], np.uint32)
    vals = np.zeros((2, 1), np.uint32)
    sp = ReplaySpec(ctypes.sizeof(ReplaySpec), 2, REPLAY_REPLICATED, REPLAY_LANE0, 1, bad.ctypes.data,
                    vals.ctypes.data, 0, None, None, None, None, 0)
    count, n = ctypes.c_uint64(), ctypes.c_uint32()
    st = sim.L.rrtl_hip_replay(sim.h, ctypes.byref(sp), ctypes.byref(count), None, ctypes.byref(n))
    assert st != 0 and sim.L.rrtl_hip_last_error(sim.h).decode().startswith("E_SIM_IR_REPLAY_SIGNAL")
