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


def test_generated_kernels_compile_for_gfx950():
    """Per-design code generation (gpu_codegen::emit_kernel counterpart) and
    the run-time compiler, both without a GPU: RV32I is all one-limb records
    (straight-line code, no interpreted records); a 100-bit counter keeps its
    wide records in interpreted form; every storage mode compiles."""
    from rrtl_amd import riscv
    core = lower.lower_module(riscv.rv32i_core())
    rv = backend.emit_kernel(core)
    assert "exec_rec" not in rv and "rrtl_word_jit" in rv and "C.lmem" not in rv
    assert backend.compile_kernel(rv) > 0
    # at 64 lanes per workgroup the 32-word register file is staged in LDS
    # (after the core's signal words) and addressed at constant offsets from the
    # thread's column; the 1024-word memory stays in HBM
    staged = backend.emit_kernel(core, backend.STORAGE_ALL_LDS, 64)
    sig_words = sum(s.limbs for s in core.signals)
    assert "C.lmem(0u, w) = C.mem(1024u, w)" in staged
    assert f"L[{sig_words * 64}u + (" in staged.split("jit_eval")[1]
    assert "C.mem(0u, uint32_t(" in staged
    assert backend.compile_kernel(staged) > 0
    wide = lower.lower_module(designs.counter(100))
    for storage in (backend.STORAGE_HBM, backend.STORAGE_SCRATCH_LDS, backend.STORAGE_ALL_LDS):
        src = backend.emit_kernel(wide, storage)
        assert "exec_rec(C, R)" in src and f"word_body<{storage}," in src
        assert backend.compile_kernel(src) > 0
    for d in (designs.ready_valid_fifo(), random_design(3)):
        assert backend.compile_kernel(backend.emit_kernel(lower.lower_module(d))) > 0


def test_generated_kernel_abi_errors():
    lib = backend.load_library()
    p = lower.lower_module(designs.counter())
    blob = p.to_blob()
    n = ctypes.c_size_t()
    assert lib.rrtl_hip_emit_kernel(blob, len(blob), 3, 0, None, 0, ctypes.byref(n)) != 0
    assert lib.rrtl_hip_create_error().decode().startswith("E_HIP_OPTIONS")
    # a short buffer gets a NUL-terminated prefix and the full length
    full = backend.emit_kernel(p)
    buf = ctypes.create_string_buffer(17)
    assert lib.rrtl_hip_emit_kernel(blob, len(blob), 2, 0, buf, 17, ctypes.byref(n)) == 0
    assert n.value == len(full) and buf.value.decode() == full[:16]
    with pytest.raises(backend.HipError) as e:
        backend.compile_kernel('#include "word_device.hpp"\n__global__ void k() { undefined_name(); }\n')
    assert e.value.code == "E_HIP_RUNTIME" and "undefined_name" in str(e.value)


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


def test_fused_records_in_generated_source():
    """The record encoder folds single-reader inner operations (the reference
    interpreter's fusion menu): the generated kernel source holds the fused
    expressions and fewer value definitions than the design has operations."""
    import re
    from designgen import fusion_design
    from rrtl_amd import backend
    src = backend.emit_kernel(lower.lower_module(fusion_design(8)))
    assert re.search(r"~\(\(\w+ & \w+\) \| \w+\)", src), "AOI21"
    assert re.search(r"~\(\(\w+ \| \w+\) & \w+\)", src), "OAI21"
    assert re.search(r"\(\w+ \^ \w+ \^ \w+\)", src), "XOR3"
    assert re.search(r"\(\w+ \+ \w+ \+ \w+\)", src), "ADD3"
    assert re.search(r"\(\w+ \* \w+ \+ \w+\)", src), "MULADD"
    assert re.search(r"\(\w+ \* \w+ - \w+\)", src), "MULSUB"
    assert re.search(r"~\(\(\w+ & \w+\) \| \(\w+ & \w+\)\)", src), "AOI22"
    assert re.search(r"~\(\(\w+ \| \w+\) & \(\w+ \| \w+\)\)", src), "OAI22"
    assert re.search(r"~\(\w+ & \w+\)", src), "NAND"


def test_split_kernel_source_compiles():
    """The picorv32 kernel split between 2 and 4 waves per 64 lanes: one code
    block per wave, estimated to beat one wave, and hiprtc compiles it for
    gfx950 (no device needed); a wave count other than 2 or 4 is refused."""
    from rrtl_amd import backend, picorv32 as pv
    p = pv.program()
    for waves in (2, 4):
        src = backend.emit_split_kernel(p, waves)
        head = src.splitlines()[0]
        m = re.search(r"estimated speed-up ([0-9.]+), exchange rows (\d+)", head)
        assert m and float(m.group(1)) > 1.4, head
        assert all(f"if (C.wave == {g}u)" in src for g in range(waves))
        assert f"jit::JitExec, {waves}>" in src
        assert backend.compile_kernel(src) > 0
    with pytest.raises(backend.HipError) as e:
        backend.emit_split_kernel(p, 3)
    assert e.value.code == "E_HIP_OPTIONS"


def test_split_plans_verified_on_designs():
    """Every split the engine would build passes its own hazard check (a
    violation raises E_HIP_INTERNAL): picorv32, the single-cycle RV32I core
    and random one-limb designs, at 2 and 4 waves; designs with interpreted
    records are refused with E_HIP_OPTIONS."""
    import sys
    from rrtl_amd import backend, lower, picorv32 as pv, riscv
    sys.path.insert(0, os.path.dirname(__file__))
    from designgen import random_design
    progs = [pv.program(), lower.lower_module(riscv.rv32i_core())]
    progs += [lower.lower_module(random_design(s, async_resets=False, widths=(1, 4, 8, 13, 16, 31, 32)))
              for s in range(12)]
    split = 0
    for p in progs:
        for waves in (2, 4):
            try:
                backend.emit_split_kernel(p, waves)
                split += 1
            except backend.HipError as e:
                assert e.code == "E_HIP_OPTIONS", str(e)
    assert split >= 10


def test_macro_records_form_on_designs():
    """The interpreter's macro records (cones of single-reader one-limb
    operations with their intermediates in registers) form on picorv32, the
    FIFO and random designs: fewer records than the fusion-menu encoding,
    every macro holds 2..16 operations, and fewer scratch slots are live."""
    from rrtl_amd import backend, picorv32 as pv
    from designgen import random_design
    cases = [pv.program(), lower.lower_module(designs.ready_valid_fifo())]
    cases += [lower.lower_module(random_design(seed, async_resets=True)) for seed in range(6)]
    for p in cases:
        mac = backend.word_program_info(p, macros=True)
        menu = backend.word_program_info(p, macros=False)
        assert menu["macro_records"] == 0 and menu["macro_ops"] == 0
        assert mac["macro_records"] > 0
        assert 2 * mac["macro_records"] <= mac["macro_ops"] <= 16 * mac["macro_records"]
        assert mac["records"] < menu["records"] and mac["scratch_slots"] <= menu["scratch_slots"]
    pico = backend.word_program_info(pv.program(), macros=True)
    assert pico["records"] < 0.6 * backend.word_program_info(pv.program(), macros=False)["records"]
