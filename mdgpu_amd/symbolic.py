"""Host mirror of the reference's symbolic analysis, over the C ABI.

``analyze(p, i, amd=True, mode="rn", max_zeros=c)`` corresponds to the
reference driver's flags ``-amd`` / ``-sn`` / ``-rn c`` (mdgpu.cu:1697-1722)
and returns a :class:`Symbolic` whose arrays are the reference's
intermediates (``cs_amd`` permutation, ``cs_etree`` parent, ``cs_post``,
``cs_counts``, node lists of GetETree / GetSuperETree / GetRelaxedETree,
``spL->p``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

_MODES = {"plain": L.PART_PLAIN, "sn": L.PART_SUPERNODAL, "rn": L.PART_RELAXED}


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Symbolic:
    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep
        sizes = (C.c_int64 * L.SYM_NSIZES)()
        L.check(L.lib().mdgpu_symbolic_sizes(self._h, sizes))
        self.sizes = list(sizes)
        self.n = sizes[L.SYM_N]
        self.nnodes = sizes[L.SYM_NNODES]
        self.lnz = sizes[L.SYM_LNZ]
        self.nnz_a = sizes[L.SYM_NNZ_A]
        self.max_front = sizes[L.SYM_MAXFRONT]
        self.nlevels = sizes[L.SYM_NLEVELS]
        self.nseg = sizes[L.SYM_NSEG]
        fl = (C.c_double * 2)()
        L.check(L.lib().mdgpu_symbolic_flops(self._h, fl))
        self.flops, self.flops_padded = fl[0], fl[1]

    @property
    def handle(self):
        return self._h

    def array(self, which):
        n, nn = self.n, self.nnodes
        length = {
            L.ARR_LP: n + 1, L.ARR_ALOW_P: n + 1, L.ARR_ALOW_I: self.nnz_a,
            L.ARR_NODES: self.sizes[L.SYM_NODES_FLAT], L.ARR_FRONT_PARENT: nn,
            L.ARR_Q_OFF: nn + 1, L.ARR_Q: self.sizes[L.SYM_QLEN], L.ARR_LEVEL: nn,
            L.ARR_SEG_COL: self.nseg + 1, L.ARR_SEG_NNZ: self.nseg + 1,
        }.get(which, n)
        out = np.empty(max(length, 1), dtype=np.int32)
        L.check(L.lib().mdgpu_symbolic_array(self._h, which, out.ctypes.data_as(L._I32P), length))
        return out[:length]

    def nodes(self):
        """List of (first, ncols, rows, zeros, kids) in processing order."""
        flat = self.array(L.ARR_NODES)
        out, t = [], 0
        while t < len(flat):
            first, ncols, rows, zeros, nk = flat[t:t + 5]
            out.append((int(first), int(ncols), int(rows), int(zeros), [int(k) for k in flat[t + 5:t + 5 + nk]]))
            t += 5 + nk
        return out

    def save(self, path):
        """Write the analysis to `path` (see :func:`load`)."""
        L.check(L.lib().mdgpu_symbolic_save(self._h, str(path).encode()))

    def close(self):
        if self._h:
            L.lib().mdgpu_symbolic_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def analyze(p, i, amd=True, mode="plain", max_zeros=0):
    p, i = _i32(p), _i32(i)
    n = len(p) - 1
    opts = L.AnalyzeOpts(int(bool(amd)), _MODES[mode], int(max_zeros))
    h = C.c_void_p()
    L.check(L.lib().mdgpu_analyze(n, p.ctypes.data_as(L._I32P), i.ctypes.data_as(L._I32P),
                                  C.byref(opts), C.byref(h)))
    return Symbolic(h, (p, i))


def analyze_budgets(p, i, budgets, amd=True):
    """Relaxed analyses (mode "rn") of one matrix for each zero budget in
    ``budgets``, sharing the ordering, elimination tree and column counts;
    element k equals ``analyze(p, i, amd, "rn", budgets[k])``."""
    p, i = _i32(p), _i32(i)
    b = _i32(budgets)
    hs = (C.c_void_p * len(b))()
    L.check(L.lib().mdgpu_analyze_budgets(len(p) - 1, p.ctypes.data_as(L._I32P), i.ctypes.data_as(L._I32P),
                                          int(bool(amd)), len(b), b.ctypes.data_as(L._I32P), hs))
    return [Symbolic(C.c_void_p(h), (p, i)) for h in hs]


def analyze_batch(mats, amd=True, mode="plain", max_zeros=0, nthreads=0):
    """One analysis for a batch of independent matrices ``mats`` = [(p, i), ...]
    (lower-triangular CSC patterns): each matrix is analysed exactly as by
    :func:`analyze`, and the results are joined block-diagonally.  Values,
    right-hand sides and solutions of the batch are the concatenations of
    the matrices' own; ``array(ARR_SEG_COL)`` / ``array(ARR_SEG_NNZ)`` give the
    offsets."""
    ps = [_i32(p) for p, _ in mats]
    iis = [_i32(i) for _, i in mats]
    k = len(mats)
    ns = _i32([len(p) - 1 for p in ps])
    PP = (L._I32P * k)(*[p.ctypes.data_as(L._I32P) for p in ps])
    II = (L._I32P * k)(*[i.ctypes.data_as(L._I32P) for i in iis])
    o = L.AnalyzeOpts(1 if amd else 0, _MODES[mode], int(max_zeros))
    h = C.c_void_p()
    L.check(L.lib().mdgpu_analyze_batch(k, ns.ctypes.data_as(L._I32P), PP, II, C.byref(o), int(nthreads),
                                        C.byref(h)))
    return Symbolic(h, (ps, iis, ns))


def load(path):
    """Read an analysis written by :meth:`Symbolic.save`."""
    h = C.c_void_p()
    L.check(L.lib().mdgpu_symbolic_load(str(path).encode(), C.byref(h)))
    return Symbolic(h, None)
