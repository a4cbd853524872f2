"""The C++ drop-in interface (include/cs.h, etree.h, mdgpu.h), exercised by
the mdgpu command-line driver built on it, which follows the reference
driver's steps (mdgpu.cu:1666-1973)."""
import os
import subprocess

import numpy as np
import pytest

from mdgpu_amd import gen
from tests.test_symbolic_golden import SMALL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "mdgpu_amd", "mdgpu")


def write_triplets(path, p, i, x):
    with open(path, "w") as f:
        for j in range(len(p) - 1):
            for t in range(p[j], p[j + 1]):
                f.write(f"{i[t] + 1} {j + 1} {x[t]:.17g}\n")


@pytest.mark.parametrize("flags,golden", [(["-amd"], "c1_100x100_plain0"),
                                          (["-amd", "-rn", "1000"], "c1_100x100_rn1000"),
                                          (["-amd", "-rn", "0"], "c1_100x100_rn0")])
def test_cli_symbolic_matches_reference(tmp_path, flags, golden):
    p, i, x = gen.laplace2d(100)
    mat = tmp_path / "a.txt"
    write_triplets(mat, p, i, x)
    r = subprocess.run([CLI, *flags, "-symbolic", "-dump", str(tmp_path), str(mat)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for name in ["amd_perm", "perm_final", "count_final", "lp", "alow_p", "alow_i"]:
        mine = np.fromfile(tmp_path / f"{name}.i32", dtype=np.int32)
        np.testing.assert_array_equal(mine, SMALL[f"{golden}__{name}"], err_msg=name)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [["-amd"], ["-amd", "-sn"], ["-amd", "-rn", "1000"]])
def test_cli_factor_and_residual(tmp_path, flags, device):
    p, i, x = gen.laplace2d(60)
    mat = tmp_path / "a.txt"
    write_triplets(mat, p, i, x)
    r = subprocess.run([CLI, *flags, str(mat)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    resid = float([ln for ln in r.stdout.splitlines() if ln.startswith("resid:")][0].split()[1])
    assert resid < 1e-14


def write_matrix_market(path, p, i, x, general=False):
    """Matrix Market coordinate file of the symmetric matrix: its lower
    triangle with the symmetric qualifier, or (general) both triangles."""
    ents = []
    for j in range(len(p) - 1):
        for t in range(p[j], p[j + 1]):
            ents.append((i[t] + 1, j + 1, x[t]))
            if general and i[t] != j:
                ents.append((j + 1, i[t] + 1, x[t]))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real " + ("general" if general else "symmetric") + "\n")
        f.write("% written by the test\n")
        n = len(p) - 1
        f.write(f"{n} {n} {len(ents)}\n")
        for a, b, v in ents:
            f.write(f"{a} {b} {v:.17g}\n")


@pytest.mark.parametrize("general", [False, True])
def test_cli_reads_matrix_market(tmp_path, general):
    """A Matrix Market file (symmetric or general) gives the same analysis
    as the reference's triplet file of its lower triangle."""
    p, i, x = gen.laplace2d(100)
    mat = tmp_path / "a.mtx"
    write_matrix_market(mat, p, i, x, general)
    r = subprocess.run([CLI, "-amd", "-rn", "1000", "-symbolic", "-dump", str(tmp_path), str(mat)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for name in ["amd_perm", "perm_final", "count_final", "lp", "alow_p", "alow_i"]:
        mine = np.fromfile(tmp_path / f"{name}.i32", dtype=np.int32)
        np.testing.assert_array_equal(mine, SMALL[f"c1_100x100_rn1000__{name}"], err_msg=name)


def test_cli_rejects_cpu_mode(tmp_path):
    p, i, x = gen.laplace2d(10)
    mat = tmp_path / "a.txt"
    write_triplets(mat, p, i, x)
    r = subprocess.run([CLI, "-cpu", "-amd", str(mat)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "-cpu is not supported" in r.stderr


@pytest.mark.gpu
def test_cli_report_line_items(tmp_path, device):
    """The report carries the reference's line items (mdgpu.cu:1901-1930);
    MFlops is the reference's padded sum (c_j^2 + 1), and FactorAux is the
    sum of its GPU parts."""
    from mdgpu_amd import symbolic
    p, i, x = gen.laplace2d(80)
    mat = tmp_path / "a.mtx"
    write_matrix_market(mat, p, i, x)
    r = subprocess.run([CLI, "-amd", "-rn", "1000", str(mat)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    items = {}
    for ln in r.stdout.splitlines():
        if "=" in ln:
            k, v = ln.split("=", 1)
            items[k.strip()] = float(v.split()[0])
    for k in ["MFlops", "MFlops/s", "Factor", "FactorAux", "BLAS GPU", "TRSM GPU", "GEMM GPU", "Memcpy",
              "Memcpy2", "MemcpyX", "Merge", "GetFk", "ExtendAdd", "Malloc GPU", "Free GPU", "Load", "Symbolic"]:
        assert k in items, k
    S = symbolic.analyze(p, i, True, "rn", 1000)
    assert abs(items["MFlops"] - S.flops_padded / 1e6) <= 0.01
    parts = items["GetFk"] + items["ExtendAdd"] + items["BLAS GPU"] + items["Otros"]
    assert abs(items["FactorAux"] - parts) <= 1e-5 and 0 < items["FactorAux"] <= items["Factor"]
    assert items["MFlops/s"] > 0
