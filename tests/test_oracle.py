"""The CPU oracle is pinned before it is trusted: dense LAPACK Cholesky of
P A P^T, the closed-form log-determinant of the Dirichlet Laplacian, the
reference's own flop formula, and triangular-solve round trips."""
import numpy as np
import pytest
import scipy.sparse as sp

from mdgpu_amd import _lib as L
from mdgpu_amd import gen, symbolic
from oracle import oracle as O
from tests.helpers import oracle_factor, permuted_lower

CASES = [
    ("grid2d_9", lambda: gen.laplace2d(9)),
    ("grid3d_6", lambda: gen.laplace3d_7pt(6)),
    ("st27_5", lambda: gen.stencil27(5, seed=7)),
    ("st27_5_aniso", lambda: gen.stencil27(5, seed=7, anisotropic=True)),
    ("bbt_150", lambda: gen.bbt_plus_ni(150, seed=1)),
]
MODES = [("plain", 0), ("sn", 0), ("rn", 0), ("rn", 60), ("rn", 100000)]


@pytest.mark.parametrize("mode,mz", MODES)
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_dense_cholesky(name, make, mode, mz):
    p, i, x = make()
    S = symbolic.analyze(p, i, True, mode, mz)
    lp, li, lx, flops, bad = oracle_factor(p, i, x, S)
    assert bad == -1
    n = S.n
    Lm = sp.csc_matrix((lx, li, lp), shape=(n, n)).toarray()
    B = permuted_lower(p, i, x, S.array(L.ARR_PERM))
    D = np.linalg.cholesky((B + sp.tril(B, -1).T).toarray())
    assert np.abs(Lm - D).max() <= 1e-13 * np.abs(D).max()
    # L's pattern is the analysis' row lists (explicit zeros included)
    assert np.all(np.abs(D[Lm == 0]) <= 1e-13 * np.abs(D).max())
    # the reference's flop formula: sum over padded columns of c_j^2 + 1
    assert flops == pytest.approx(S.flops_padded, rel=0, abs=0.5)


def test_logdet_closed_form_dirichlet_laplacian():
    k, shift = 24, 0.0137
    p, i, x = gen.stencil_lower((k, k, 1), 5, 4.0 + shift)
    S = symbolic.analyze(p, i, True, "rn", 300)
    lp, li, lx, _, bad = oracle_factor(p, i, x, S)
    logdet = 2.0 * np.sum(np.log(lx[lp[:-1]]))
    t = np.arange(1, k + 1) * np.pi / (k + 1)
    lam = shift + (2 - 2 * np.cos(t))[:, None] + (2 - 2 * np.cos(t))[None, :]
    assert logdet == pytest.approx(np.sum(np.log(lam)), rel=1e-13)


def test_solve_round_trip_and_backward_error():
    p, i, x = gen.laplace2d(30)
    S = symbolic.analyze(p, i, True, "rn", 200)
    lp, li, lx, _, _ = oracle_factor(p, i, x, S)
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=3)
    perm = S.array(L.ARR_PERM)
    y = O.ltsolve(lp, li, lx, O.lsolve(lp, li, lx, b[perm]))
    xs = np.empty_like(y)
    xs[perm] = y
    assert np.abs(xs - xt).max() < 1e-12
    eta = O.backward_error(p, i, x, xs, b)
    assert eta < 1e-15
    # the formula itself, restated with scipy
    A = gen.full_from_lower(p, i, x)
    r = b - A @ xs
    ref = np.abs(r).max() / (abs(A).sum(axis=0).max() * np.abs(xs).max() + np.abs(b).max())
    assert eta == pytest.approx(ref, rel=1e-12)


def test_planted_error_in_oracle_is_detected():
    """A plausible mistake (dropping one child's update) must fail the dense check."""
    p, i, x = gen.laplace2d(9)
    S = symbolic.analyze(p, i, True, "plain", 0)
    nodes = S.nodes()
    B = permuted_lower(p, i, x, S.array(L.ARR_PERM))
    kids = [list(nd[4]) for nd in nodes]
    victim = next(t for t, k in enumerate(kids) if len(k) >= 2)
    kids[victim] = kids[victim][1:]
    lp = S.array(L.ARR_LP)
    try:
        li, lx, _, _ = O.factor(B.indptr, B.indices, B.data, [nd[0] for nd in nodes],
                                [nd[1] for nd in nodes], kids, lp)
    except Exception:
        return
    D = np.linalg.cholesky((B + sp.tril(B, -1).T).toarray())
    Lm = sp.csc_matrix((lx, li, lp), shape=(S.n, S.n)).toarray()
    assert np.abs(Lm - D).max() > 1e-6


def test_forest_factorization():
    p, i, x = gen.bbt_plus_ni(600, seed=5)
    S = symbolic.analyze(p, i, True, "rn", 1000)
    assert (S.array(L.ARR_PARENT) == -1).sum() > 1
    lp, li, lx, _, bad = oracle_factor(p, i, x, S)
    B = permuted_lower(p, i, x, S.array(L.ARR_PERM))
    D = np.linalg.cholesky((B + sp.tril(B, -1).T).toarray())
    Lm = sp.csc_matrix((lx, li, lp), shape=(S.n, S.n)).toarray()
    assert np.abs(Lm - D).max() <= 1e-12 * np.abs(D).max()


def test_backward_error_against_scipy_norms():
    """oracle.backward_error is the reference's eta = ||b - A x||_inf /
    (||A||_1 ||x||_inf + ||b||_inf) (mdgpu.cu:1972, cs_norm cs.cpp:236-248)
    on the full symmetric A rebuilt from the lower triangle: checked against
    scipy's sparse norms at a perturbed x (eta of order 1e-3).  The checks
    below fail for a dropped ||b|| term, the infinity norm of A in place of
    the 1-norm on a non-symmetric-looking weighting, or the lower triangle
    alone in the product."""
    import scipy.sparse.linalg as sla
    p, i, x = gen.stencil27(6, seed=2)
    A = gen.full_from_lower(p, i, x)
    n = A.shape[0]
    rng = np.random.default_rng(3)
    xv = rng.uniform(-1, 1, n)
    b = A @ xv + 1e-3 * rng.standard_normal(n)
    eta = O.backward_error(p, i, x, xv, b)
    r = b - A @ xv
    ref = np.abs(r).max() / (sla.norm(A, 1) * np.abs(xv).max() + np.abs(b).max())
    assert abs(eta - ref) <= 1e-12 * ref
    assert abs(eta - np.abs(r).max() / (sla.norm(A, 1) * np.abs(xv).max())) > 1e-3 * ref
