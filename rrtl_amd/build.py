"""Builds the in-tree HIP library ``rrtl_amd/_lib/librrtl_hip.so`` for gfx950.

Every source is compiled with hipcc for ``--offload-arch=gfx950`` (cross-
compiles without a GPU).  Packed FP32 instruction generation is switched off
(``-packed-fp32-ops``): this library has no floating point on the simulation
path, and the flag keeps the compiler from emitting those forms anywhere.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_lib")
LIB = os.path.join(OUT_DIR, "librrtl_hip.so")
SOURCES = ["gate_engine.hip", "word_engine.hip", "batch.hip", "mlp.hip", "capi.cpp", "rtc.cpp"]
# device headers that kernels generated at run time include (embedded as text)
RTC_HEADERS = ["device_defs.hpp", "word_device.hpp", "gate_device.hpp", "batch_device.hpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RRTL_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"), "-I", OUT_DIR]


def _compile(src: str, obj: str) -> None:
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")


def _write_rtc_headers() -> None:
    """rtc_headers.inc: the device headers as C++ raw string literals."""
    names, texts = [], []
    for h in RTC_HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            text = f.read()
        if ")RRTL_RTC\"" in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter")
        names.append(f'"{h}"')
        texts.append(f'R"RRTL_RTC({text})RRTL_RTC"')
    body = (f"static const int kRtcHeaderCount = {len(RTC_HEADERS)};\n"
            f"static const char* const kRtcHeaderNames[] = {{{', '.join(names)}}};\n"
            f"static const char* const kRtcHeaderTexts[] = {{\n" + ",\n".join(texts) + "};\n")
    path = os.path.join(OUT_DIR, "rtc_headers.inc")
    with open(path + ".tmp", "w") as f:
        f.write(body)
    os.replace(path + ".tmp", path)


def build_library(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    _write_rtc_headers()
    objs = [os.path.join(OUT_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        list(ex.map(lambda p: _compile(*p), zip(SOURCES, objs)))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of librrtl_hip.so failed")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build_library(verbose=True)
