"""Build the gfx950 tracer library in-tree with hipcc.

    python -m blackhorizon_amd.build [--resource-usage]

Produces ``blackhorizon_amd/lib/libbhtrace.so`` from the HIP sources in
``blackhorizon_amd/csrc``.  The precise build is compiled with
``-ffp-contract=off`` (one rounding per operation, NumPy order); the fast
build lets the compiler contract multiply-adds into FMAs.
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIBNAME = "libbhtrace.so"
ARCH = os.environ.get("BH_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

UNITS = {
    "bh_precise": ["-ffp-contract=off"],
    "bh_fast": ["-ffp-contract=fast"],
    "bh_abi": ["-ffp-contract=off"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit: str, extra: list[str], objdir: str, resource_usage: bool) -> str:
    src = os.path.join(CSRC, unit + ".hip")
    obj = os.path.join(objdir, unit + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", src, "-o", obj]
    if resource_usage:
        cmd.append("-Rpass-analysis=kernel-resource-usage")
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{proc.stderr}")
    if resource_usage and proc.stderr:
        with open(os.path.join(objdir, unit + ".resource.txt"), "w") as f:
            f.write(proc.stderr)
    return obj


def build(resource_usage: bool = False, verbose: bool = False) -> str:
    """Compile every unit and link the shared library; returns its path."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        futs = [ex.submit(_compile, u, e, objdir, resource_usage) for u, e in UNITS.items()]
        objs = [f.result() for f in futs]
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--resource-usage", action="store_true",
                    help="write per-kernel VGPR/SGPR/occupancy remarks under _build/")
    args = ap.parse_args(argv)
    build(resource_usage=args.resource_usage, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
