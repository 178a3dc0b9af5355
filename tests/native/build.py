"""Builds the test-only probe library libguide_probe.so (gfx950) from
guide_probe.hip, with the product library's compile flags and headers, and
libqmc_host.so, the cached tracer's stratified sampler (dev_qmc.h) for the
CPU.

Test infrastructure: only tests/ load it.
"""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
LIB = HERE / "libguide_probe.so"
QMC_LIB = HERE / "libqmc_host.so"


def _product_build():
    spec = importlib.util.spec_from_file_location("_lambent_build", ROOT / "lambent_amd" / "build.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def build(verbose: bool = True) -> Path:
    pb = _product_build()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [pb._hipcc(), *pb.COMMON, *pb.HIP_FLAGS, f"-I{pb.CSRC}", "-shared", "-o", str(tmp),
           str(HERE / "guide_probe.hip")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"probe build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    tmp = QMC_LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{pb.CSRC}", "-o", str(tmp),
           str(HERE / "qmc_host.cpp")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"sampler build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, QMC_LIB)
    if verbose:
        print(f"built {QMC_LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
