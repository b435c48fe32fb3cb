"""Build the HIP library ``larndsim_amd/liblsim.so`` for gfx950 with hipcc (in-tree).

Usage: python build.py [--arch gfx950] [--debug]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "lsim.hip")
OUT = os.path.join(HERE, "larndsim_amd", "liblsim.so")
HEADERS = [os.path.join(HERE, "csrc", h) for h in ("device.h", "scan_sort.h", "forward.h", "backward.h", "loss.h", "chop.h", "probfee.h", "sdtw.h", "param.h")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm not installed?)")


def build(arch: str = "gfx950", debug: bool = False, verbose: bool = True) -> str:
    tmp = OUT + ".tmp"
    cmd = [
        hipcc(), f"--offload-arch={arch}", "-std=c++17", "-fPIC", "-shared",
        "-O1" if debug else "-O3",
        # packed fp32 instructions stay disabled (the compiler would otherwise form v_pk_* ops)
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result",
        "-o", tmp, SRC,
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


def up_to_date() -> bool:
    if not os.path.exists(OUT):
        return False
    t = os.path.getmtime(OUT)
    return all(os.path.getmtime(p) <= t for p in [SRC] + HEADERS)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args()
    print(build(a.arch, a.debug))
    sys.exit(0)
