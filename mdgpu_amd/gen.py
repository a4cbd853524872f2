"""Synthetic SPD test matrices for the five benchmark workloads.

Every generator returns the LOWER triangle (diagonal included) of a symmetric
positive definite matrix as compressed-sparse-column arrays ``(p, i, x)`` with
int32 indices and float64 values, rows sorted inside each column -- the input
layout the reference driver builds with ``cs_compress`` from a lower-triangular
triplet file (mdgpu.cu:1727-1728, cs.cpp:251-270).  Grid points are numbered
x-fastest: node(ix, iy, iz) = ix + kx*(iy + ky*iz).

Workloads:
  * 2D 5-point Laplacian, 100x100, diagonal 4+1e-2, off-diagonals -1.
  * 3D 7-point Laplacian, 48^3, diagonal 6+1e-3, off-diagonals -1.
  * 3D 27-point stencil, 64^3, off-diagonal -w with w ~ U(0.5, 1.5) drawn once
    per undirected edge, diagonal = (1+1e-4) * sum_j |a_ij|; the anisotropic
    variant multiplies every edge whose z-offset is nonzero by 1e4.
  * 512 uneven matrices (half stencils with a random diagonal shift, half
    B*B^T + n*I with B having 5 N(0,1) entries per column).
  * 2D GMRF precision Q = tau*(kappa^2*I + L), L the 5-point Laplacian (256^2).
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "lower_csc_from_coo", "stencil_lower", "laplace2d", "laplace3d_7pt",
    "stencil27", "gmrf_precision", "bbt_plus_ni", "batch_matrices",
    "full_from_lower", "spmv_full", "rhs_from_xtrue",
]


def lower_csc_from_coo(n, rows, cols, vals):
    """Strictly-lower (row > col) COO + diagonal -> sorted lower CSC."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    p = np.zeros(n + 1, dtype=np.int64)
    np.add.at(p, cols + 1, 1)
    p = np.cumsum(p)
    return p.astype(np.int32), rows.astype(np.int32), vals


def _grid_edges(dims, offsets):
    """Undirected edges (u < v) of a regular grid for the given forward offsets.

    Offsets are (dx, dy, dz) with a positive linear-index step; edges are
    emitted offset by offset, each in increasing u, which fixes the order in
    which random edge weights are drawn."""
    kx, ky, kz = dims
    node = np.arange(kx * ky * kz, dtype=np.int64)
    ix, iy, iz = node % kx, (node // kx) % ky, node // (kx * ky)
    us, vs, dzs = [], [], []
    for dx, dy, dz in offsets:
        jx, jy, jz = ix + dx, iy + dy, iz + dz
        ok = (jx >= 0) & (jx < kx) & (jy >= 0) & (jy < ky) & (jz >= 0) & (jz < kz)
        u = node[ok]
        v = jx[ok] + kx * (jy[ok] + ky * jz[ok])
        us.append(u)
        vs.append(v)
        dzs.append(np.full(u.shape, dz))
    return np.concatenate(us), np.concatenate(vs), np.concatenate(dzs)


def _forward_offsets(npts):
    if npts == 5:
        return [(1, 0, 0), (0, 1, 0)]
    if npts == 7:
        return [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    if npts == 27:
        offs = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    lin = dx + 3 * dy + 9 * dz  # sign of the linear-index step
                    if lin > 0:
                        offs.append((dx, dy, dz))
        return offs
    raise ValueError(npts)


def stencil_lower(dims, npts, diag, off=-1.0):
    """Constant-coefficient stencil: lower CSC with `diag` and `off` values."""
    n = int(np.prod(dims))
    u, v, _ = _grid_edges(dims, _forward_offsets(npts))
    rows = np.concatenate([v, np.arange(n)])
    cols = np.concatenate([u, np.arange(n)])
    vals = np.concatenate([np.full(u.shape, off, dtype=np.float64), np.full(n, diag, dtype=np.float64)])
    return lower_csc_from_coo(n, rows, cols, vals)


def laplace2d(k=100, diag=4.0 + 1e-2):
    return stencil_lower((k, k, 1), 5, diag)


def laplace3d_7pt(k=48, diag=6.0 + 1e-3):
    return stencil_lower((k, k, k), 7, diag)


def stencil27(k=64, seed=0, anisotropic=False, margin=1e-4):
    """27-point stencil with random symmetric weights (one draw per edge)."""
    dims = (k, k, k)
    n = k ** 3
    u, v, dz = _grid_edges(dims, _forward_offsets(27))
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, size=u.shape[0])
    if anisotropic:
        w = np.where(dz != 0, w * 1e4, w)
    rowsum = np.zeros(n)
    np.add.at(rowsum, u, w)
    np.add.at(rowsum, v, w)
    rows = np.concatenate([v, np.arange(n)])
    cols = np.concatenate([u, np.arange(n)])
    vals = np.concatenate([-w, rowsum * (1.0 + margin)])
    return lower_csc_from_coo(n, rows, cols, vals)


def gmrf_precision(k=256, log_tau=0.0, log_kappa=-1.0):
    """Q(theta) = tau*(kappa^2 I + L), L = 5-point Dirichlet Laplacian."""
    tau, kappa = np.exp(log_tau), np.exp(log_kappa)
    return stencil_lower((k, k, 1), 5, tau * (kappa * kappa + 4.0), off=-tau)


def bbt_plus_ni(n, seed, nnz_per_col=5):
    """A = B B^T + n I with B n-by-n, `nnz_per_col` N(0,1) entries per column."""
    rng = np.random.default_rng(seed)
    cols_b = np.repeat(np.arange(n), nnz_per_col)
    rows_b = np.concatenate([rng.choice(n, size=nnz_per_col, replace=False) for _ in range(n)])
    vals_b = rng.standard_normal(n * nnz_per_col)
    import scipy.sparse as sp
    B = sp.csc_matrix((vals_b, (rows_b, cols_b)), shape=(n, n))
    A = (B @ B.T).tocoo()
    A = A + sp.identity(n, format="coo") * float(n)
    A = sp.tril(A.tocsc(), format="coo")
    return lower_csc_from_coo(n, A.row, A.col, A.data)


def _last_side(n_target, area, nmin, nmax):
    """Last grid side k for a grid of `area` points per layer: the nearest
    to n_target / area with nmin <= area k <= nmax."""
    k = max(1, int(round(n_target / area)))
    while area * k < nmin:
        k += 1
    while area * k > nmax and k > 1:
        k -= 1
    return k


def batch_matrices(count=512, seed=0, nmin=500, nmax=20000):
    """Uneven batch: n log-uniform in [nmin, nmax]; even entries are 2D/3D
    Laplacian stencils with diagonal shift U(1e-3, 1e-1), odd ones B B^T + nI."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n_target = int(np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        sub_seed = int(rng.integers(0, 2**31 - 1))
        if t % 2 == 0:
            shift = float(rng.uniform(1e-3, 1e-1))
            # grid sides near the target; the last side is rounded so that
            # n = kx ky kz stays inside [nmin, nmax]
            if rng.uniform() < 0.5:
                kx = int(np.floor(np.sqrt(n_target)))
                ky = _last_side(n_target, kx, nmin, nmax)
                out.append(stencil_lower((kx, ky, 1), 5, 4.0 + shift))
            else:
                kx = max(2, int(np.floor(n_target ** (1.0 / 3.0))))
                ky = kx
                kz = _last_side(n_target, kx * ky, nmin, nmax)
                out.append(stencil_lower((kx, ky, kz), 7, 6.0 + shift))
        else:
            out.append(bbt_plus_ni(n_target, sub_seed))
    return out


def full_from_lower(p, i, x):
    """Full symmetric scipy CSC matrix from lower CSC arrays."""
    import scipy.sparse as sp
    n = len(p) - 1
    L = sp.csc_matrix((x, i, p), shape=(n, n))
    return (L + sp.tril(L, -1).T).tocsc()


def spmv_full(p, i, x, v):
    return full_from_lower(p, i, x) @ v


def rhs_from_xtrue(p, i, x, seed=0):
    """x_true ~ U(-1, 1) from `seed`, b = A x_true."""
    n = len(p) - 1
    rng = np.random.default_rng(seed)
    xt = rng.uniform(-1.0, 1.0, size=n)
    return xt, spmv_full(p, i, x, xt)
