"""Autotuner for the HIP engines' launch shape.

Mirrors the reference's GPU autotune recommendations
(``GpuAutotuneRecommendation`` / ``find_gpu_autotune_recommendation_with_mode`` /
``autotune_match_key``, crates/rrtl-gpu-sim/src/lib.rs:55-190): candidates are
measured on the device, the fastest becomes a recommendation keyed by
``(case, lanes, steps)``, recommendations are stored as JSON, and lookups match
exactly or by the nearest (lanes, steps) with ties broken by the measured
metric.

Knobs: the gate engine's tile width (u32 lane-words per thread, i.e. 32 to
256 lanes per thread sharing one LDS-resident copy of the live gate values,
which sets LDS per workgroup and so workgroups per CU), gate positions per
thread per stage and cycles per launch; the word engine's workgroup size,
active lanes per wave (below one wave's width), waves per 64 lanes (the
generated code split between several threads per lane) and, with them, lanes
per split workgroup, cycles per launch and, for runs that halt, lane
compaction between launches.  Shapes a design
cannot take (a split of code with interpreted records) are recorded with
their error and skipped.
"""
from __future__ import annotations

import json
from dataclasses import asdict, dataclass, field
from typing import Callable, Dict, Iterable, List, Optional, Sequence

from .backend import HipError, HipSimulator
from .program import Program


@dataclass
class HipAutotuneRecommendation:
    case: str
    lanes: int
    steps: int
    engine: str
    tile_words: int = 0
    block_threads: int = 0
    ticks_per_launch: int = 0
    wave_lanes: int = 0
    stage_gates: int = 0
    compact_lanes: bool = False
    lane_waves: int = 0
    autotune_metric: str = "kernel_ns_per_lane_cycle"
    autotune_metric_ns: Optional[float] = None
    candidates: List[dict] = field(default_factory=list)

    def options(self) -> dict:
        return {"engine": self.engine, "tile_words": self.tile_words, "block_threads": self.block_threads,
                "ticks_per_launch": self.ticks_per_launch, "wave_lanes": self.wave_lanes,
                "stage_gates": self.stage_gates, "compact_lanes": self.compact_lanes, "lane_waves": self.lane_waves}


def _candidates(engine: str, steps: int, probe: int, halting: bool = False) -> List[dict]:
    # a cycles-per-launch candidate only differs from one launch when the probe
    # spans at least two of its launches
    tpl = [0] + [t for t in (2048,) if t < steps and 2 * t <= probe]
    if engine == "gate":
        return [{"tile_words": w, "stage_gates": m, "ticks_per_launch": t}
                for w in (1, 2, 4, 8) for m in (2, 3, 4) for t in tpl]
    out = ([{"block_threads": b, "wave_lanes": 64, "lane_waves": 1, "ticks_per_launch": t}
            for b in (64, 128, 256) for t in tpl] +
           [{"block_threads": 64, "wave_lanes": w, "lane_waves": 1, "ticks_per_launch": t}
            for w in (8, 16, 32) for t in tpl] +
           [{"lane_waves": g, "wave_lanes": w, "ticks_per_launch": t} for g in (2, 4) for w in (0, 32) for t in tpl])
    if halting:
        ct = [t for t in (256, 1024) if 2 * t <= probe] or [max(1, probe // 2)]
        out += [{"block_threads": 0, "wave_lanes": 0, "ticks_per_launch": t, "compact_lanes": True} for t in ct]
        out += [{"lane_waves": g, "ticks_per_launch": t, "compact_lanes": True} for g in (2, 4) for t in ct]
    return out


def autotune(program: Program, lanes: int, steps: int, case: str, engine: str = "auto",
             run_kwargs: Optional[dict] = None, probe_steps: Optional[int] = None, repeats: int = 2,
             device: int = 0, setup: Optional[Callable[[HipSimulator], None]] = None) -> HipAutotuneRecommendation:
    """Measures every candidate on ``probe_steps`` cycles (default min(steps, 4096))
    of the same run (``setup`` prepares each fresh engine, e.g. loads memories)
    and returns the fastest per simulated lane-cycle as a recommendation."""
    if engine == "auto":
        engine = "gate" if program.is_gate_level() else "word"
    probe = probe_steps or min(steps, 4096)
    run_kwargs = dict(run_kwargs or {})
    results = []
    best = None
    for cand in _candidates(engine, steps, probe, halting=run_kwargs.get("halt") is not None):
        try:
            sim = HipSimulator(program, lanes, engine=engine, device=device, **cand)
            if setup:
                setup(sim)
        except HipError as e:
            results.append({**cand, "error": e.code})
            continue
        try:
            sim.run(probe, **run_kwargs)  # warm
            ns = min(_timed(sim, probe, run_kwargs) for _ in range(repeats))
        except HipError as e:
            results.append({**cand, "error": e.code})
            sim.close()
            continue
        row = {**cand, "ns_per_lane_cycle": ns}
        if engine == "gate":
            row.update(sim.tile_info())
        results.append(row)
        if best is None or ns < best[1]:
            best = (cand, ns)
        sim.close()
    if best is None:
        raise HipError(3, "E_HIP_AUTOTUNE: no candidate launch shape ran")
    return HipAutotuneRecommendation(case=case, lanes=lanes, steps=steps, engine=engine,
                                     tile_words=best[0].get("tile_words", 0),
                                     block_threads=best[0].get("block_threads", 0),
                                     ticks_per_launch=best[0].get("ticks_per_launch", 0),
                                     wave_lanes=best[0].get("wave_lanes", 0),
                                     stage_gates=best[0].get("stage_gates", 0),
                                     compact_lanes=bool(best[0].get("compact_lanes", False)),
                                     lane_waves=best[0].get("lane_waves", 0),
                                     autotune_metric_ns=best[1], candidates=results)


def _timed(sim: HipSimulator, steps: int, run_kwargs: dict) -> float:
    """Kernel ns per simulated lane-cycle of one run (halted lanes do not count)."""
    sim.run(steps, **run_kwargs)
    st = sim.stats()
    return st["kernel_ms"] * 1e6 / max(1, st["lane_cycles"])


def save_recommendations(path: str, items: Iterable[HipAutotuneRecommendation]) -> None:
    with open(path, "w") as f:
        json.dump([asdict(i) for i in items], f, indent=1)


def load_recommendations(path: str) -> List[HipAutotuneRecommendation]:
    try:
        with open(path) as f:
            raw = json.load(f)
        return [HipAutotuneRecommendation(**r) for r in raw]
    except (OSError, ValueError, TypeError) as e:
        raise ValueError(f"E_HIP_AUTOTUNE_JSON: failed to load autotune recommendations: {e}")


def _match_key(r: HipAutotuneRecommendation, lanes: int, steps: int):
    return (abs(r.lanes - lanes), abs(r.steps - steps),
            r.autotune_metric_ns if r.autotune_metric_ns is not None else float("inf"), r.lanes, r.steps)


def find_recommendation(items: Sequence[HipAutotuneRecommendation], case: str, lanes: int, steps: int,
                        mode: str = "exact") -> Optional[HipAutotuneRecommendation]:
    if mode == "exact":
        for r in items:
            if r.case == case and r.lanes == lanes and r.steps == steps:
                return r
        return None
    cands = [r for r in items if r.case == case]
    return min(cands, key=lambda r: _match_key(r, lanes, steps)) if cands else None


def options_from_recommendations(items, case: str, lanes: int, steps: int, mode: str = "exact") -> dict:
    r = find_recommendation(items, case, lanes, steps, mode)
    if r is None:
        raise KeyError(f"E_HIP_AUTOTUNE_RECOMMENDATION: no recommendation for case `{case}` lanes={lanes} steps={steps}")
    return r.options()
