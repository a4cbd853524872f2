"""The drop-in serves both of the reference's FLOTANTE builds (types.h:3-23
of the reference): a caller compiled with FLOTANTE = float links the float
entry points (float A->x and spL->x, factorization in FP32), one compiled
with double the double ones, and one compiled with any other FLOTANTE fails
to link rather than misreading the value arrays.  The caller here is the
mdgpu driver (mdgpu_amd/cli/mdgpu_main.cpp), written against the reference
headers."""
import os
import subprocess

import numpy as np
import pytest

from mdgpu_amd import gen
from tests.test_dropin_cli import write_triplets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mdgpu_amd")
SRC = os.path.join(PKG, "cli", "mdgpu_main.cpp")


def build_caller(out, flotante, src=SRC):
    return subprocess.run(["g++", "-O1", "-std=c++17", f"-DFLOTANTE={flotante}", f"-I{ROOT}/include", src, "-o",
                           str(out), f"-L{PKG}", "-lmdgpu_amd", f"-Wl,-rpath,{PKG}"],
                          capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("flotante", ["float", "double"])
def test_caller_links_its_flotante(tmp_path, flotante):
    r = build_caller(tmp_path / "mdgpu", flotante)
    assert r.returncode == 0, r.stderr
    syms = subprocess.run(["nm", "-C", str(tmp_path / "mdgpu")], capture_output=True, text=True).stdout
    assert f"mdgpu_flotante_{flotante}::Factor(" in syms
    other = "double" if flotante == "float" else "float"
    assert f"mdgpu_flotante_{other}::" not in syms


def test_other_flotante_fails_to_link(tmp_path):
    src = tmp_path / "caller.cpp"
    src.write_text('#include "cs.h"\n#include "etree.h"\n#include "mdgpu.h"\n'
                   "int main() { cs* T = cs_spalloc(2, 2, 4, 1, 1); Factor(T, T, 0, 0); return 0; }\n")
    r = build_caller(tmp_path / "caller", "long", src=str(src))
    assert r.returncode != 0
    assert "undefined reference" in r.stderr and "mdgpu_flotante_long" in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("flotante,bound", [("float", 1e-5), ("double", 1e-14)])
def test_float_and_double_callers_factor_and_solve(tmp_path, device, flotante, bound):
    """The driver built with each FLOTANTE loads the triplets in its own
    type, factors through Factor() (FP32 for float, FP64 for double) and
    solves with cs_lsolve/cs_ltsolve in FLOTANTE: backward error of float's
    or double's rounding."""
    exe = tmp_path / "mdgpu"
    r = build_caller(exe, flotante)
    assert r.returncode == 0, r.stderr
    p, i, x = gen.laplace2d(60)
    mat = tmp_path / "a.txt"
    write_triplets(mat, p, i, x)
    r = subprocess.run([str(exe), "-amd", "-rn", "1000", str(mat)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    resid = float([ln for ln in r.stdout.splitlines() if ln.startswith("resid:")][0].split()[1])
    assert resid < bound, resid
    if flotante == "float":
        assert resid > 1e-12      # the factor and solves really ran in float
