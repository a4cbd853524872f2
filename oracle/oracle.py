"""ctypes front end of the CPU oracle (oracle/mf_oracle.c) and of the
reference-symbolic harness (oracle/_ref/ref_symbolic).  TEST INFRASTRUCTURE:
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_SYMBOLIC = os.path.join(HERE, "_ref", "ref_symbolic")

_lib = None
_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_factor.restype = C.c_int
        L.oracle_factor.argtypes = [C.c_int, _I, _I, _D, C.c_int, _I, _I, _I, _I, _I, _I, _D, _D]
        L.oracle_lsolve.argtypes = [C.c_int, _I, _I, _D, _D]
        L.oracle_ltsolve.argtypes = [C.c_int, _I, _I, _D, _D]
        L.oracle_backward_error.restype = C.c_double
        L.oracle_backward_error.argtypes = [C.c_int, _I, _I, _D, _D, _D]
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(_I)


def _dp(a):
    return a.ctypes.data_as(_D)


def factor(ap, ai, ax, first, ncols, kids_lists, lp):
    """Multifrontal FP64 factorization of the permuted lower triangle over a
    postordered node list.  Returns (Li, Lx, flops, bad_col)."""
    n = len(ap) - 1
    ap = np.ascontiguousarray(ap, dtype=np.int32)
    ai = np.ascontiguousarray(ai, dtype=np.int32)
    ax = np.ascontiguousarray(ax, dtype=np.float64)
    first = np.ascontiguousarray(first, dtype=np.int32)
    ncols = np.ascontiguousarray(ncols, dtype=np.int32)
    kid_ptr = np.zeros(len(kids_lists) + 1, dtype=np.int32)
    kid_ptr[1:] = np.cumsum([len(k) for k in kids_lists])
    kids = np.ascontiguousarray(np.concatenate([np.asarray(k, dtype=np.int32) for k in kids_lists])
                                if kid_ptr[-1] else np.zeros(1, np.int32), dtype=np.int32)
    lp = np.ascontiguousarray(lp, dtype=np.int32)
    lnz = int(lp[-1])
    Li = np.zeros(max(lnz, 1), dtype=np.int32)
    Lx = np.zeros(max(lnz, 1), dtype=np.float64)
    flops = C.c_double()
    bad = lib().oracle_factor(n, _ip(ap), _ip(ai), _dp(ax), len(first), _ip(first), _ip(ncols), _ip(kid_ptr),
                              _ip(kids), _ip(lp), _ip(Li), _dp(Lx), C.byref(flops))
    return Li[:lnz], Lx[:lnz], flops.value, bad


def lsolve(lp, li, lx, b):
    x = np.array(b, dtype=np.float64, copy=True)
    lib().oracle_lsolve(len(lp) - 1, _ip(np.ascontiguousarray(lp, np.int32)), _ip(np.ascontiguousarray(li, np.int32)),
                        _dp(np.ascontiguousarray(lx, np.float64)), _dp(x))
    return x


def ltsolve(lp, li, lx, b):
    x = np.array(b, dtype=np.float64, copy=True)
    lib().oracle_ltsolve(len(lp) - 1, _ip(np.ascontiguousarray(lp, np.int32)), _ip(np.ascontiguousarray(li, np.int32)),
                         _dp(np.ascontiguousarray(lx, np.float64)), _dp(x))
    return x


def backward_error(ap, ai, ax, x, b):
    return lib().oracle_backward_error(len(ap) - 1, _ip(np.ascontiguousarray(ap, np.int32)),
                                       _ip(np.ascontiguousarray(ai, np.int32)), _dp(np.ascontiguousarray(ax, np.float64)),
                                       _dp(np.ascontiguousarray(x, np.float64)), _dp(np.ascontiguousarray(b, np.float64)))


def solve_with_symbolic(sym_arrays, ax_in, b):
    """Full oracle pipeline on the caller's values: permute A with the
    analysis' permutation, factor, solve, unpermute.  sym_arrays: dict with
    perm, alow_p, alow_i, a_src, nodes (list of (first, ncols, kids)), lp."""
    perm = sym_arrays["perm"]
    ax = np.asarray(ax_in, dtype=np.float64)[sym_arrays["a_src"]]
    nodes = sym_arrays["nodes"]
    Li, Lx, flops, bad = factor(sym_arrays["alow_p"], sym_arrays["alow_i"], ax, [nd[0] for nd in nodes],
                                [nd[1] for nd in nodes], [nd[2] for nd in nodes], sym_arrays["lp"])
    xp = np.asarray(b, dtype=np.float64)[perm]
    y = lsolve(sym_arrays["lp"], Li, Lx, xp)
    z = ltsolve(sym_arrays["lp"], Li, Lx, y)
    x = np.empty_like(z)
    x[perm] = z
    return x, Li, Lx, flops, bad


def run_ref_symbolic(p, i, workdir, amd=True, mode="plain", max_zeros=0):
    """Run the reference-symbolic harness; returns dict name -> int32 array."""
    os.makedirs(workdir, exist_ok=True)
    path = os.path.join(workdir, "in.bin")
    with open(path, "wb") as f:
        np.array([len(p) - 1, len(i)], dtype=np.int32).tofile(f)
        np.asarray(p, dtype=np.int32).tofile(f)
        np.asarray(i, dtype=np.int32).tofile(f)
    subprocess.run([REF_SYMBOLIC, path, workdir, str(int(amd)), mode, str(max_zeros)], check=True,
                   capture_output=True)
    out = {}
    for name in os.listdir(workdir):
        if name.endswith(".i32"):
            out[name[:-4]] = np.fromfile(os.path.join(workdir, name), dtype=np.int32)
    return out
