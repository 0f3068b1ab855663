"""The header-only C++ drop-in (include/zsba_factor_graph.hpp) compiles and links
against libzsba.so as a map::optimise-shaped caller would use it; on a GPU it
solves bit for bit as the C ABI does."""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cxx" / "shim_example.cpp"
LIB = ROOT / "zeroslam_amd"


def build(out: Path, extra=()) -> Path:
    exe = out / "shim_example"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", *extra, str(SRC),
           f"-L{LIB}", "-l:libzsba.so", f"-Wl,-rpath,{LIB}", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    return exe


def test_shim_compiles_and_links(tmp_path):
    assert build(tmp_path).exists()


@pytest.mark.skipif(not Path("/root/reference/source/matrix.hpp").exists(), reason="reference checkout not present")
def test_shim_accepts_the_reference_matrix_type(tmp_path):
    """With ZSBA_SHIM_MATRIX set to the reference's matrix::matrix<double, 0, 0>,
    the caller's parameter matrices are the reference's own type (compiled in
    place from the reference checkout; nothing is copied)."""
    src = tmp_path / "with_reference_matrix.cpp"
    src.write_text('#include "matrix.hpp"\n#define ZSBA_SHIM_MATRIX ::matrix::matrix<double, 0, 0>\n'
                   f'#include "{SRC}"\n')
    exe = tmp_path / "shim_ref_matrix"
    cmd = ["g++", "-std=c++17", "-O1", "-fno-exceptions", "-Wno-invalid-constexpr", f"-I{ROOT / 'include'}",
           "-I/root/reference/source", str(src), f"-L{LIB}", "-l:libzsba.so", f"-Wl,-rpath,{LIB}",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-3000:]


@pytest.mark.gpu
@pytest.mark.parametrize("config,iters", [(0, 10), (1, 12)])
def test_shim_solves_like_the_c_abi(tmp_path, config, iters):
    exe = build(tmp_path)
    proc = subprocess.run([str(exe), str(config), str(iters)], capture_output=True, text=True, timeout=120,
                          env=dict(os.environ))
    print(proc.stdout, proc.stderr)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "identical 1" in proc.stdout
