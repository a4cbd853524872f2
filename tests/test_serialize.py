"""The on-disk symbolic analysis (mdgpu_symbolic_save / _load) round-trips
every array and plans identically; damaged files are rejected."""
import ctypes as C

import numpy as np
import pytest

from mdgpu_amd import _lib as L
from mdgpu_amd import gen, symbolic

ARRS = [L.ARR_AMD_PERM, L.ARR_PARENT, L.ARR_POST, L.ARR_COUNT, L.ARR_PERM, L.ARR_COUNT_FINAL, L.ARR_LP,
        L.ARR_ALOW_P, L.ARR_ALOW_I, L.ARR_NODES, L.ARR_FRONT_PARENT, L.ARR_Q_OFF, L.ARR_Q, L.ARR_LEVEL,
        L.ARR_SEG_COL, L.ARR_SEG_NNZ]


def plan_info(S, prec):
    out = (C.c_int64 * L.NUM_NINFO)()
    L.check(L.lib().mdgpu_plan_check(S.handle, prec, out))
    return list(out)


@pytest.mark.parametrize("mode,mz", [("rn", 500), ("sn", 0)])
def test_round_trip(tmp_path, mode, mz):
    p, i, x = gen.laplace3d_7pt(10)
    S = symbolic.analyze(p, i, True, mode, mz)
    path = tmp_path / "a.sym"
    S.save(path)
    T = symbolic.load(path)
    assert T.sizes == S.sizes and (T.flops, T.flops_padded) == (S.flops, S.flops_padded)
    for w in ARRS:
        assert np.array_equal(T.array(w), S.array(w)), w
    if mode == "rn":
        assert np.array_equal(T.array(L.ARR_RELAXED_PERM), S.array(L.ARR_RELAXED_PERM))
    for prec in (L.FP64, L.FP32):
        assert plan_info(T, prec) == plan_info(S, prec)


def test_batch_round_trip(tmp_path):
    mats = [gen.laplace2d(8), gen.bbt_plus_ni(90, seed=1)]
    S = symbolic.analyze_batch([(p, i) for p, i, _ in mats], True, "rn", 100)
    S.save(tmp_path / "b.sym")
    T = symbolic.load(tmp_path / "b.sym")
    assert T.nseg == 2 and np.array_equal(T.array(L.ARR_SEG_COL), S.array(L.ARR_SEG_COL))


def test_damaged_files_are_rejected(tmp_path):
    p, i, x = gen.laplace2d(12)
    S = symbolic.analyze(p, i, True, "rn", 100)
    path = tmp_path / "c.sym"
    S.save(path)
    data = bytearray(path.read_bytes())
    bad = tmp_path / "bad.sym"
    for mutate in (lambda d: d.__setitem__(40, d[40] ^ 1), lambda d: d.__delitem__(slice(len(d) // 2, None)),
                   lambda d: d.__setitem__(0, ord("X"))):
        d = bytearray(data)
        mutate(d)
        bad.write_bytes(bytes(d))
        with pytest.raises(L.MdgpuError):
            symbolic.load(bad)
    with pytest.raises(L.MdgpuError):
        symbolic.load(tmp_path / "missing.sym")
