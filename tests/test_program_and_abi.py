"""Program format, lowering structure and the C-ABI library (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from rrtl_amd import backend, designs, ir, lower
from rrtl_amd.program import OP_ADD, OP_MUX, OP_SIGNAL, Program
from designgen import random_design

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_blob_roundtrip():
    for seed in range(10):
        p = lower.lower_module(random_design(seed))
        q = Program.from_blob(p.to_blob())
        assert q.to_blob() == p.to_blob()
        assert [s.name for s in q.signals] == [s.name for s in p.signals]


def test_counter_lowering_structure():
    """Packets follow lower_packets_to_block: effects on the last machine packet,
    tick_next captures the register with its reset (lib.rs:2997-3048, :12816)."""
    p = lower.lower_module(designs.counter())
    assert [s.name for s in p.signals] == ["clk", "rst", "en", "out", "count"]
    assert [s.offset for s in p.signals] == [0, 1, 2, 3, 4]
    assert p.total_signal_words == 5
    comb = p.comb
    assert len(comb.packets) == 1 and comb.packets[0].instrs[0].op == OP_SIGNAL
    nxt = p.tick_next
    last = nxt.packets[-1]
    assert len(last.effects) == 1 and last.effects[0].reset is not None
    ops = [i.op for pk in nxt.packets for i in pk.instrs]
    assert OP_ADD in ops and OP_MUX in ops


def test_levelization_with_chain():
    m = ir.Module("Chain")
    a = m.input("a", 4)
    w1 = m.wire("w1", 4)
    w2 = m.wire("w2", 4)
    w3 = m.wire("w3", 4)
    m.assign(w3, w2 + w1)   # declared first, depends on later wires
    m.assign(w2, w1 ^ a)
    m.assign(w1, ~a)
    p = lower.lower_module(m)
    stores = []
    for pk in p.comb.packets:
        for e in pk.effects:
            stores.append(p.signals[e.dst].name)
    assert stores == ["w1", "w2", "w3"]


def test_design_rules():
    m = ir.Module("Bad")
    a = m.input("a", 4)
    b = m.input("b", 5)
    with pytest.raises(ir.DesignError):
        _ = a + b
    w = m.wire("w", 4)
    m.assign(w, a)
    with pytest.raises(ir.DesignError):
        m.assign(w, a)
    with pytest.raises(ir.DesignError):
        ir.mux(a, a, a)


def test_library_exports_every_header_symbol():
    lib = backend.load_library()
    hdr = open(os.path.join(ROOT, "include", "rrtl_hip.h")).read()
    names = set(re.findall(r"\b(rrtl_hip_[a-z_0-9]+)\s*\(", hdr))
    assert names, "no declarations parsed"
    assert set(backend.EXPORTS) == names
    for n in names:
        assert hasattr(lib, n), n


def test_library_exports_every_surrogate_symbol():
    from rrtl_amd import surrogate
    lib = backend.load_library()
    hdr = open(os.path.join(ROOT, "include", "rrtl_mlp.h")).read()
    names = set(re.findall(r"\b(rrtl_mlp_[a-z_0-9]+)\s*\(", hdr))
    assert names, "no declarations parsed"
    assert set(surrogate.EXPORTS) == names
    for n in names:
        assert hasattr(lib, n), n


def test_surrogate_config_rejected_without_gpu_work():
    """Shape checks run before any device call."""
    from rrtl_amd import surrogate
    lib = surrogate._lib()
    c = surrogate.MlpConfig()
    c.struct_size = ctypes.sizeof(c)
    c.n_layers = 2
    c.dims[0], c.dims[1], c.dims[2] = 100, 128, 1  # input width not a multiple of 128
    c.max_batch = 128
    h = ctypes.c_void_p()
    assert lib.rrtl_mlp_create(ctypes.byref(c), ctypes.byref(h)) != 0
    assert lib.rrtl_mlp_create_error().decode().startswith("E_MLP_CONFIG")
    c.dims[0], c.dims[2] = 128, 2  # more than one output
    assert lib.rrtl_mlp_create(ctypes.byref(c), ctypes.byref(h)) != 0
    assert lib.rrtl_mlp_create_error().decode().startswith("E_MLP_CONFIG")


def test_create_rejects_bad_blob_without_gpu_work():
    lib = backend.load_library()
    h = ctypes.c_void_p()
    st = lib.rrtl_hip_create(b"not a program blob at all......." * 2, 64, 4, None, ctypes.byref(h))
    assert st == 2  # RRTL_E_PROGRAM
    assert lib.rrtl_hip_create_error().decode().startswith("E_PROGRAM_MAGIC")
    st = lib.rrtl_hip_create(b"x", 1, 0, None, ctypes.byref(h))
    assert lib.rrtl_hip_create_error().decode().startswith("E_SIM_IR_LANES")


def test_autotune_recommendation_lookup(tmp_path):
    """Exact and nearest matching like find_gpu_autotune_recommendation_with_mode
    (crates/rrtl-gpu-sim/src/lib.rs:108-172)."""
    from rrtl_amd.autotune import (HipAutotuneRecommendation, find_recommendation, load_recommendations,
                                   options_from_recommendations, save_recommendations)
    items = [HipAutotuneRecommendation("net", 4096, 1000, "gate", tile_words=2, autotune_metric_ns=1.0),
             HipAutotuneRecommendation("net", 65536, 10000, "gate", tile_words=4, autotune_metric_ns=0.5),
             HipAutotuneRecommendation("ctr", 4096, 1000, "word", block_threads=128, autotune_metric_ns=2.0)]
    p = tmp_path / "rec.json"
    save_recommendations(str(p), items)
    back = load_recommendations(str(p))
    assert back == items
    assert find_recommendation(back, "net", 65536, 10000).tile_words == 4
    assert find_recommendation(back, "net", 60000, 9000) is None
    assert find_recommendation(back, "net", 60000, 9000, mode="nearest").tile_words == 4
    assert options_from_recommendations(back, "ctr", 4096, 1000)["block_threads"] == 128
    with pytest.raises(KeyError):
        options_from_recommendations(back, "nope", 1, 1)
