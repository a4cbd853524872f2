"""Shared test helpers: permuted matrices, oracle runs, comparisons."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from mdgpu_amd import _lib as L
from mdgpu_amd import gen, symbolic


def permuted_lower(p, i, x, perm):
    """Lower triangle of P A P^T (new k = old perm[k]) as sorted scipy CSC."""
    A = gen.full_from_lower(p, i, x)
    B = sp.tril(A[perm][:, perm]).tocsc()
    B.sort_indices()
    return B


def oracle_factor(p, i, x, S):
    """Run the CPU oracle on the analysis S; returns (Lp, Li, Lx, flops, bad)."""
    from oracle import oracle as O
    perm = S.array(L.ARR_PERM)
    B = permuted_lower(p, i, x, perm)
    nodes = S.nodes()
    lp = S.array(L.ARR_LP)
    Li, Lx, flops, bad = O.factor(B.indptr, B.indices, B.data, [nd[0] for nd in nodes],
                                  [nd[1] for nd in nodes], [nd[4] for nd in nodes], lp)
    return lp, Li, Lx, flops, bad


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def cases():
    """(name, generator, mode, max_zeros) used by the parity tests."""
    return [
        ("grid2d_7_plain", lambda: gen.laplace2d(7), "plain", 0),
        ("grid3d_543_sn", lambda: gen.stencil_lower((5, 4, 3), 7, 6.1), "sn", 0),
        ("c1_100x100_plain", lambda: gen.laplace2d(100), "plain", 0),
        ("c1_100x100_rn1000", lambda: gen.laplace2d(100), "rn", 1000),
        ("grid3d_16_rn4000", lambda: gen.laplace3d_7pt(16), "rn", 4000),
        ("st27_12_rn100000", lambda: gen.stencil27(12, seed=3), "rn", 100000),
        ("st27_16_aniso_rn1000", lambda: gen.stencil27(16, seed=4, anisotropic=True), "rn", 1000),
        ("bbt_600_rn1000", lambda: gen.bbt_plus_ni(600, seed=5), "rn", 1000),
    ]
