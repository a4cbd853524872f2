"""The C ABI library loads, exports every declared symbol, and reports
errors through status codes (no compute on a GPU here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mdgpu_amd import _lib as L
from mdgpu_amd import gen, symbolic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mdgpu_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mdgpu_[a-z_0-9]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(L.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _analyze_raw(p, i):
    p = np.ascontiguousarray(p, np.int32)
    i = np.ascontiguousarray(i, np.int32)
    h = C.c_void_p()
    opts = L.AnalyzeOpts(1, L.PART_PLAIN, 0)
    st = L.lib().mdgpu_analyze(len(p) - 1, p.ctypes.data_as(L._I32P), i.ctypes.data_as(L._I32P),
                               C.byref(opts), C.byref(h))
    msg = L.lib().mdgpu_last_error().decode()
    if h:
        L.lib().mdgpu_symbolic_destroy(h)
    return st, msg


def test_invalid_inputs_are_rejected_with_status():
    p, i, x = gen.laplace2d(4)
    assert _analyze_raw(p, i)[0] == L.MDGPU_OK
    bad = i.copy()
    bad[0] = 3                                  # breaks row order in column 0
    st, msg = _analyze_raw(p, bad)
    assert st == L.MDGPU_ERR_ARG and "sorted" in msg
    nodiag = i.copy()
    nodiag[p[2]] = 3                            # column 2 loses its diagonal
    st, msg = _analyze_raw(p, nodiag)
    assert st == L.MDGPU_ERR_ARG
    upper = i.copy()
    upper[p[5]] = 1                             # an entry above the diagonal
    st, msg = _analyze_raw(p, upper)
    assert st == L.MDGPU_ERR_ARG and "lower" in msg


def test_plan_check_on_host():
    p, i, x = gen.laplace3d_7pt(10)
    S = symbolic.analyze(p, i, True, "rn", 500)
    for prec in (L.FP64, L.FP32, L.BF16, L.BF16X3):
        out = (C.c_int64 * L.NUM_NINFO)()
        L.check(L.lib().mdgpu_plan_check(S.handle, prec, out))
        assert out[L.NUM_LSTORE] >= S.lnz and out[L.NUM_STEPS] > 0


def test_unknown_precision_rejected():
    p, i, x = gen.laplace2d(5)
    S = symbolic.analyze(p, i, True, "plain", 0)
    out = (C.c_int64 * L.NUM_NINFO)()
    with pytest.raises(L.MdgpuError):
        L.check(L.lib().mdgpu_plan_check(S.handle, 7, out))
