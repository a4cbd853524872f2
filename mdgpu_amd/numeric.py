"""Host mirror of the numeric phase over the C ABI (include/mdgpu_c.h).

``Factorization(sym, precision)`` owns a device arena; ``factorize(ax)``
replaces the reference's ``Factor(A, spL, listaNodos, cantNodos)``
(mdgpu.cu:1394-1608), ``solve`` its ``cs_lsolve``/``cs_ltsolve`` pair
(cs.cpp:1086-1119), ``refine`` adds FP64 iterative refinement to the
reference's backward-error check (mdgpu.cu:1972), and ``export_L`` returns L
in the reference's ``spL`` CSC layout.  Device vectors are torch tensors on
the current device; torch is only used to hold memory and streams.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

PRECISIONS = {"fp64": L.FP64, "fp32": L.FP32, "bf16": L.BF16, "bf16x3": L.BF16X3}


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    return C.c_void_p(t.data_ptr())


class Factorization:
    def __init__(self, sym, precision="fp64"):
        self.sym = sym
        self.precision = precision
        h = C.c_void_p()
        L.check(L.lib().mdgpu_numeric_create(sym.handle, PRECISIONS[precision], C.byref(h)))
        self._h = h
        self._ax = None

    def set_graph(self, enable):
        L.check(L.lib().mdgpu_numeric_set_graph(self._h, int(bool(enable))))
        return self

    def info(self):
        out = (C.c_int64 * L.NUM_NINFO)()
        L.check(L.lib().mdgpu_numeric_info(self._h, out))
        names = ["device_bytes", "lstore", "arena", "steps", "levels", "panel_fronts", "diag_jobs", "precision",
                 "fp64_fallbacks"]
        return dict(zip(names, list(out)))

    def factorize(self, ax):
        """ax: float64 device tensor with the caller's lower-triangle values."""
        torch = _torch()
        ax = ax.to(device="cuda", dtype=torch.float64).contiguous()
        self._ax = ax
        L.check(L.lib().mdgpu_factorize(self._h, _dptr(ax), _stream()))
        return self

    def factorize_profile(self, ax):
        """Factorize with per-kernel-class HIP-event timing.  Returns
        {class: {"ms", "launches", "flops", "bytes"}}."""
        torch = _torch()
        ax = ax.to(device="cuda", dtype=torch.float64).contiguous()
        self._ax = ax
        out = (C.c_double * (4 * len(L.KIND_NAMES)))()
        L.check(L.lib().mdgpu_factorize_profile(self._h, _dptr(ax), _stream(), out))
        return {k: {"ms": out[4 * t], "launches": int(out[4 * t + 1]), "flops": out[4 * t + 2],
                    "bytes": out[4 * t + 3]} for t, k in enumerate(L.KIND_NAMES)}

    def autotune(self, ax):
        """Pick the GEMM kernel per launch (64- or 128-row 16x16x4 MFMA tiles,
        and in FP32 the 32x32x2 MFMA kernel's 128- and 64-row tiles) by timing
        each on `ax`; returns
        the number of launches on 64-row tiles.  Leaves `ax` factored."""
        torch = _torch()
        ax = ax.to(device="cuda", dtype=torch.float64).contiguous()
        self._ax = ax
        out = C.c_int64()
        L.check(L.lib().mdgpu_numeric_autotune(self._h, _dptr(ax), _stream(), C.byref(out)))
        return out.value

    def factorize_steps(self, ax):
        """Factorize with a HIP-event time for every launch of the schedule
        (one launch after the other): array of rows (kernel class, ms,
        workgroups, flops, bytes, level, look-ahead stream, GEMM kernel code)."""
        torch = _torch()
        ax = ax.to(device="cuda", dtype=torch.float64).contiguous()
        self._ax = ax
        nsteps = self.info()["steps"]
        out = np.zeros(8 * max(nsteps, 1))
        L.check(L.lib().mdgpu_factorize_steps(self._h, _dptr(ax), _stream(), out.ctypes.data_as(L._DP), len(out)))
        return out[:8 * nsteps].reshape(nsteps, 8)

    def solve(self, b, out=None):
        """x = A^{-1} b; b of shape (n,) or (nrhs, n) (one right-hand side per row)."""
        torch = _torch()
        b = b.to(device="cuda", dtype=torch.float64).contiguous()
        x = out if out is not None else torch.empty_like(b)
        if b.dim() == 2:
            L.check(L.lib().mdgpu_solve_nrhs(self._h, _dptr(b), _dptr(x), b.shape[0], b.shape[1], _stream()))
        else:
            L.check(L.lib().mdgpu_solve(self._h, _dptr(b), _dptr(x), _stream()))
        return x

    def refine(self, b, tol=1e-12, maxit=30, out=None, raise_on_fail=True):
        torch = _torch()
        b = b.to(device="cuda", dtype=torch.float64).contiguous()
        x = out if out is not None else torch.empty_like(b)
        eta = C.c_double()
        it = C.c_int()
        st = L.lib().mdgpu_refine(self._h, _dptr(b), _dptr(x), tol, maxit, C.byref(eta), C.byref(it), _stream())
        if st != L.MDGPU_OK and (raise_on_fail or st != L.MDGPU_ERR_NOT_CONVERGED):
            L.check(st)
        return x, eta.value, it.value

    def solve_status(self):
        """Status bits of the solves since the last call (1: a chained
        solve's wait gave up); reading clears them."""
        v = C.c_int()
        L.check(L.lib().mdgpu_solve_status(self._h, C.byref(v), _stream()))
        return v.value

    def segment_eta(self):
        """Backward error of each matrix of a batch after the last refine."""
        out = np.empty(max(self.sym.nseg, 1), dtype=np.float64)
        L.check(L.lib().mdgpu_refine_segment_eta(self._h, out.ctypes.data_as(L._DP), len(out)))
        return out

    def logdet(self):
        out = C.c_double()
        L.check(L.lib().mdgpu_logdet(self._h, C.byref(out), _stream()))
        return out.value

    # ------------------------------------------------------- reverse mode
    def selinv(self):
        """Selected inverse Z = A^{-1} on the pattern of L (FP64 factor)."""
        L.check(L.lib().mdgpu_selinv(self._h, _stream()))
        return self

    def selinv_export(self):
        """Z on the pattern of L, in export_L's layout (host array)."""
        _torch().cuda.current_stream().synchronize()
        Zx = np.empty(self.sym.lnz, dtype=np.float64)
        L.check(L.lib().mdgpu_selinv_export(self._h, Zx.ctypes.data_as(L._DP)))
        return Zx

    def logdet_grad(self):
        """d log det A / d a_e for the caller's lower-triangle entries (device)."""
        torch = _torch()
        g = torch.empty(self.sym.nnz_a, dtype=torch.float64, device="cuda")
        L.check(L.lib().mdgpu_logdet_grad(self._h, _dptr(g), _stream()))
        return g

    def solve_adjoint(self, x, xbar):
        """Adjoint of x = A^{-1} b: returns (bbar = A^{-1} xbar, dL/da_e)."""
        torch = _torch()
        x = x.to(device="cuda", dtype=torch.float64).contiguous()
        xbar = xbar.to(device="cuda", dtype=torch.float64).contiguous()
        bbar = torch.empty_like(xbar)
        g = torch.empty(self.sym.nnz_a, dtype=torch.float64, device="cuda")
        L.check(L.lib().mdgpu_solve_adjoint(self._h, _dptr(x), _dptr(xbar), _dptr(bbar), _dptr(g), _stream()))
        return bbar, g

    def export_L(self):
        n, lnz = self.sym.n, self.sym.lnz
        Lp = np.empty(n + 1, dtype=np.int32)
        Li = np.empty(lnz, dtype=np.int32)
        Lx = np.empty(lnz, dtype=np.float64)
        L.check(L.lib().mdgpu_export_L(self._h, Lp.ctypes.data_as(L._I32P), Li.ctypes.data_as(L._I32P),
                                       Lx.ctypes.data_as(L._DP)))
        return Lp, Li, Lx

    def export_L_segment(self, k):
        """L of matrix k of a batch in its own numbering (Lp, Li, Lx)."""
        from . import _lib
        seg = self.sym.array(_lib.ARR_SEG_COL)
        n = int(seg[k + 1] - seg[k])
        Lp = np.empty(n + 1, dtype=np.int32)
        L.check(L.lib().mdgpu_export_L_segment(self._h, k, Lp.ctypes.data_as(L._I32P), None, None))
        Li = np.empty(int(Lp[-1]), dtype=np.int32)
        Lx = np.empty(int(Lp[-1]), dtype=np.float64)
        L.check(L.lib().mdgpu_export_L_segment(self._h, k, Lp.ctypes.data_as(L._I32P), Li.ctypes.data_as(L._I32P),
                                               Lx.ctypes.data_as(L._DP)))
        return Lp, Li, Lx

    def close(self):
        if self._h:
            L.lib().mdgpu_numeric_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_partition(syms, ax, precision="fp32", reps=2, solves=1):
    """Partition tuner: for each analysis in ``syms`` (one matrix, e.g.
    ``symbolic.analyze_budgets``) time a factorization plus ``solves``
    solves on the GPU (GEMM tiles autotuned first).  Returns (index of the
    fastest, [ms per analysis])."""
    torch = _torch()
    ax = ax.to(device="cuda", dtype=torch.float64).contiguous()
    hs = (C.c_void_p * len(syms))(*[S.handle.value if isinstance(S.handle, C.c_void_p) else S.handle
                                     for S in syms])
    ms = np.zeros(len(syms))
    best = C.c_int()
    L.check(L.lib().mdgpu_tune_partition(hs, len(syms), _dptr(ax), PRECISIONS[precision], int(reps), int(solves),
                                         _stream(), ms.ctypes.data_as(L._DP), C.byref(best)))
    return best.value, ms.tolist()


def backward_error(p, i, x, xs, b):
    """eta = ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf) on the GPU for
    the symmetric matrix with lower-triangle CSC (p, i, x) (host arrays),
    xs and b device or host vectors (mdgpu_backward_error)."""
    torch = _torch()
    import scipy.sparse as sp
    n = len(p) - 1
    Lo = sp.csc_matrix((np.asarray(x, np.float64), np.asarray(i), np.asarray(p)), shape=(n, n))
    A = (Lo + sp.tril(Lo, -1).T).tocsr()
    A.sort_indices()
    dev = dict(device="cuda")
    rp = torch.from_numpy(A.indptr.astype(np.int32)).to(**dev)
    ci = torch.from_numpy(A.indices.astype(np.int32)).to(**dev)
    av = torch.from_numpy(A.data).to(**dev)
    xs = torch.as_tensor(xs, dtype=torch.float64).to(**dev).contiguous()
    b = torch.as_tensor(b, dtype=torch.float64).to(**dev).contiguous()
    eta = C.c_double()
    L.check(L.lib().mdgpu_backward_error(n, _dptr(rp), _dptr(ci), _dptr(av), _dptr(xs), _dptr(b), C.byref(eta),
                                         _stream()))
    return eta.value
