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
