"""The numeric schedule (extend-add rounds and splits, update-arena offsets,
panel / blocked potrf+TRSM / GEMM jobs at several outer block widths, with
and without the update-matrix look-ahead,
diagonal-block inverses),
executed step by step on the CPU by tests/emulate/plan_emulator.cpp with the
device kernels' loops, reproduces the oracle's factor.  This checks the plan
without a GPU; the kernels themselves are checked by the gpu tests."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from mdgpu_amd import gen, symbolic
from tests.helpers import oracle_factor, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SO = os.path.join(HERE, "emulate", "libemulator.so")


@pytest.fixture(scope="module")
def emulator():
    src = [os.path.join(HERE, "emulate", "plan_emulator.cpp"), os.path.join(ROOT, "mdgpu_amd", "csrc", "plan.cpp"),
           os.path.join(ROOT, "mdgpu_amd", "csrc", "symbolic.cpp"),
           os.path.join(ROOT, "mdgpu_amd", "csrc", "selinv_plan.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", SO, *src], check=True)
    return C.CDLL(SO)


CASES = [
    ("grid2d_7_plain", lambda: gen.laplace2d(7), 0, 0),
    ("st27_12_rn100000", lambda: gen.stencil27(12, seed=1), 2, 100000),
    ("c1_rn1000", lambda: gen.laplace2d(100), 2, 1000),
    ("grid3d_16_rn4000", lambda: gen.laplace3d_7pt(16), 2, 4000),
    ("st27_10_rn100000", lambda: gen.stencil27(10, seed=2), 2, 100000),
    ("grid3d_12_sn", lambda: gen.laplace3d_7pt(12), 1, 0),
    ("bbt_400_rn500", lambda: gen.bbt_plus_ni(400, seed=9), 2, 500),
]


@pytest.mark.parametrize("nb_outer", [64, 128, 256, -256])   # -256: no update-matrix look-ahead
@pytest.mark.parametrize("name,make,mode,mz", CASES, ids=[c[0] for c in CASES])
def test_schedule_reproduces_oracle(emulator, name, make, mode, mz, nb_outer):
    p, i, x = make()
    S = symbolic.analyze(p, i, True, ["plain", "sn", "rn"][mode], mz)
    lp, li, lx, _, _ = oracle_factor(p, i, x, S)
    I, D = C.POINTER(C.c_int), C.POINTER(C.c_double)
    Lx = np.zeros(S.lnz)
    Li = np.zeros(S.lnz, np.int32)
    st = emulator.emulate_factor(len(p) - 1, p.ctypes.data_as(I), i.ctypes.data_as(I), x.ctypes.data_as(D), 1,
                                 mode, mz, nb_outer, Lx.ctypes.data_as(D), Li.ctypes.data_as(I))
    assert st == 0
    assert np.array_equal(Li, li)
    assert rel_err(Lx, lx) < 1e-12


@pytest.mark.parametrize("nb_outer", [64, 256])
@pytest.mark.parametrize("name,make,mode,mz", CASES, ids=[c[0] for c in CASES])
def test_presplit_schedule_reproduces_oracle(emulator, name, make, mode, mz, nb_outer):
    """The bf16 variants' schedule: block steps that also write the split
    planes, GEMMs that read them (kept unsplit here, in FP64, NaN until a
    block step writes them and past a front's last column) reproduce the
    oracle's factor."""
    p, i, x = make()
    S = symbolic.analyze(p, i, True, ["plain", "sn", "rn"][mode], mz)
    lp, li, lx, _, _ = oracle_factor(p, i, x, S)
    I, D = C.POINTER(C.c_int), C.POINTER(C.c_double)
    Lx = np.zeros(S.lnz)
    Li = np.zeros(S.lnz, np.int32)
    st = emulator.emulate_factor_presplit(len(p) - 1, p.ctypes.data_as(I), i.ctypes.data_as(I),
                                          x.ctypes.data_as(D), 1, mode, mz, nb_outer, Lx.ctypes.data_as(D),
                                          Li.ctypes.data_as(I))
    if st == 4:
        pytest.skip("no blocked front (no split planes)")
    assert st == 0
    assert np.array_equal(Li, li)
    assert rel_err(Lx, lx) < 1e-12


@pytest.mark.parametrize("u_stride,u_flush,presplit", [(1, 1, 0), (3, 1, 1), (2, 2, 0), (3, 3, 1), (2, 0, 0)])
@pytest.mark.parametrize("name,make,mode,mz", CASES[1:5], ids=[c[0] for c in CASES[1:5]])
def test_update_matrix_stride_reproduces_oracle(emulator, name, make, mode, mz, u_stride, u_flush, presplit):
    """Update-matrix look-ahead applied every u_stride outer blocks (64
    columns) and at the last u_flush blocks before a front's last, with and
    without the split planes."""
    p, i, x = make()
    S = symbolic.analyze(p, i, True, ["plain", "sn", "rn"][mode], mz)
    lp, li, lx, _, _ = oracle_factor(p, i, x, S)
    I, D = C.POINTER(C.c_int), C.POINTER(C.c_double)
    Lx = np.zeros(S.lnz)
    Li = np.zeros(S.lnz, np.int32)
    st = emulator.emulate_factor_ustride(len(p) - 1, p.ctypes.data_as(I), i.ctypes.data_as(I),
                                         x.ctypes.data_as(D), 1, mode, mz, 64, u_stride, u_flush, presplit,
                                         Lx.ctypes.data_as(D), Li.ctypes.data_as(I))
    if st == 4:
        pytest.skip("no blocked front (no split planes)")
    assert st == 0
    assert np.array_equal(Li, li)
    assert rel_err(Lx, lx) < 1e-12


SELINV_CASES = [
    ("grid2d_7_plain", lambda: gen.laplace2d(7), 0, 0),
    ("st27_12_rn100000", lambda: gen.stencil27(12, seed=1), 2, 100000),
    ("grid2d_30_rn500", lambda: gen.laplace2d(30), 2, 500),
    ("grid3d_9_sn", lambda: gen.laplace3d_7pt(9), 1, 0),
    ("st27_9_rn100000", lambda: gen.stencil27(9, seed=2), 2, 100000),
    ("bbt_300_rn500", lambda: gen.bbt_plus_ni(300, seed=9), 2, 500),
]


@pytest.mark.parametrize("name,make,mode,mz", SELINV_CASES, ids=[c[0] for c in SELINV_CASES])
def test_selinv_schedule_matches_dense_inverse(emulator, name, make, mode, mz):
    """The selected-inverse schedule, run on the CPU with the device kernels'
    loops, equals the dense inverse of P A P^T on the pattern of L (numpy)."""
    p, i, x = make()
    S = symbolic.analyze(p, i, True, ["plain", "sn", "rn"][mode], mz)
    I, D = C.POINTER(C.c_int), C.POINTER(C.c_double)
    Lx = np.zeros(S.lnz)
    Li = np.zeros(S.lnz, np.int32)
    Zx = np.zeros(S.lnz)
    st = emulator.emulate_factor_selinv(len(p) - 1, p.ctypes.data_as(I), i.ctypes.data_as(I), x.ctypes.data_as(D), 1,
                                        mode, mz, 128, Lx.ctypes.data_as(D), Li.ctypes.data_as(I),
                                        Zx.ctypes.data_as(D))
    assert st == 0
    from mdgpu_amd import _lib as L
    perm = S.array(L.ARR_PERM)
    A = gen.full_from_lower(p, i, x).toarray()
    Zd = np.linalg.inv(A[np.ix_(perm, perm)])
    lp = S.array(L.ARR_LP)
    cols = np.repeat(np.arange(S.n), np.diff(lp))
    ref = Zd[Li, cols]
    assert rel_err(Zx, ref) < 1e-11


@pytest.mark.parametrize("kind,want", [(0, "forward chain order"), (1, "backward chain order"),
                                       (2, "chain counter"), (3, "left-looking range"),
                                       (4, "chain block inverse"), (5, "side stream never joined")])
def test_validator_rejects_planted_mistakes(emulator, kind, want):
    """A mistake planted in a valid plan (chained-solve order, counters,
    inverses; left-looking block range; a side-stream launch never joined)
    is caught by validate_plan with the matching message, before anything
    would run on a GPU."""
    p, i, x = gen.stencil27(16, seed=1)
    I = C.POINTER(C.c_int)
    st = emulator.emulate_validate_mistake(len(p) - 1, p.ctypes.data_as(I), i.ctypes.data_as(I), 2, 100000, kind)
    assert st == 0, (want, st)
