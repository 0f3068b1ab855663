"""Build of the engine's native library, ``zeroslam_amd/libzsba.so``.

Every HIP translation unit is compiled for gfx950 only with ``hipcc`` and linked
into one shared library with a plain C ABI (``include/zsba.h``).  The binary64
kernels are compiled without floating-point contraction so that per-observation
arithmetic is the reference's IEEE sequence.  The mixed path's binary32 kernels
are compiled the same way: they evaluate each observation's residual and robust
weight in binary64 in the reference's order.  Packed-FP32 instruction selection
is disabled for all device code.

Usage: ``python -m zeroslam_amd.build`` (also called by ``__graft_entry__.build``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OUT = PKG / "libzsba.so"
BUILD_DIR = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", f"-I{INCLUDE}", f"-I{CSRC}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]

# (source, extra flags).  FP64 parity code: no contraction anywhere.
UNITS = [
    ("synth.cpp", ["-ffp-contract=off", "-x", "c++"]),
    ("output.cpp", ["-ffp-contract=off", "-x", "c++"]),
    ("engine.hip", ["-ffp-contract=off"]),
    ("kernels_fp64.hip", ["-ffp-contract=off"]),
    ("kernels_fp32.hip", ["-ffp-contract=off"]),
    ("kernels_chol.hip", ["-ffp-contract=off"]),
    ("kernels_band.hip", ["-ffp-contract=off"]),
    ("kernels_batch.hip", ["-ffp-contract=off"]),
    ("kernels_adjoint.hip", ["-ffp-contract=off"]),
    ("kernels_match.hip", ["-ffp-contract=off"]),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    # hipcc's host pass warns that it does not know -packed-fp32-ops: harmless.
    for line in proc.stdout.splitlines():
        if "packed-fp32-ops" in line or not line.strip():
            continue
        if "warning" in line or "error" in line:
            sys.stderr.write(line + "\n")


def build(verbose: bool = False) -> Path:
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found; the engine needs ROCm to build")
    BUILD_DIR.mkdir(exist_ok=True)
    jobs = []
    for src, flags in UNITS:
        path = CSRC / src
        if not path.exists():
            continue
        obj = BUILD_DIR / (path.stem + ".o")
        cmd = [HIPCC, *COMMON, *flags, "-c", str(path), "-o", str(obj)]
        if src.endswith(".cpp"):
            # host-only unit: no device code, no offload flags
            cmd = ["g++", "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}", "-c", str(path),
                   "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        jobs.append((cmd, str(obj)))
    # the translation units are independent: compile them side by side
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(_run, cmd) for cmd, _ in jobs]:
            f.result()
    objects = [obj for _, obj in jobs]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objects, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, OUT)
    return OUT


def build_oracle(verbose: bool = False) -> None:
    """Compile the CPU oracle (test infrastructure) and, where the reference
    checkout exists, the reference driver and the map harnesses (the reference's
    map.hpp with its own factor graph and with the drop-in) into ``oracle/_ref``."""
    targets = ["oracle"]
    if Path("/root/reference/source/factor_graph.hpp").exists():
        targets += ["ref", "map", "output"]
    cmd = ["make", "-s", "-C", str(ROOT / "oracle"), *targets]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    build_oracle(verbose="-v" in sys.argv)
    print(OUT)
