"""Writes picorv32_soc.rrtl, the packed program of the reference's picorv32
SoC (bench/sv/picorv32.v + bench/sv/picorv32_soc.v, top ``picorv32_soc``)
as imported by ``rrtl_amd.verilog`` and lowered by ``rrtl_amd.lower``, plus
picorv32_soc.json (source digests and program statistics).

The Verilog sources are not part of this repository; the blob is the derived
program, so the GPU runs and tests need no source.  Usage:
    python -m rrtl_amd.data.make_picorv32 [/root/reference/bench/sv]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ("picorv32.v", "picorv32_soc.v")
TOP = "picorv32_soc"


def build_program(sv_dir: str):
    from rrtl_amd import lower, verilog
    texts = [open(os.path.join(sv_dir, f)).read() for f in SOURCES]
    imp = verilog.import_verilog("\n".join(texts), TOP)
    return lower.lower_module(imp.module), imp, texts


def main(sv_dir: str = "/root/reference/bench/sv") -> None:
    p, imp, texts = build_program(sv_dir)
    blob = p.to_blob()
    with open(os.path.join(HERE, "picorv32_soc.rrtl"), "wb") as f:
        f.write(blob)
    info = {
        "top": TOP,
        "sources": {f: hashlib.sha256(t.encode()).hexdigest() for f, t in zip(SOURCES, texts)},
        "blob_sha256": hashlib.sha256(blob).hexdigest(),
        "importer_stats": imp.stats,
        "signals": len(p.signals),
        "memories": [m.name for m in p.memories],
        "instrs_per_block": [sum(len(k.instrs) for k in b.packets) for b in p.blocks],
    }
    with open(os.path.join(HERE, "picorv32_soc.json"), "w") as f:
        json.dump(info, f, indent=1)
        f.write("\n")
    print(json.dumps(info, indent=1))


if __name__ == "__main__":
    root = os.path.dirname(os.path.dirname(HERE))
    if root not in sys.path:
        sys.path.insert(0, root)
    main(*sys.argv[1:])
