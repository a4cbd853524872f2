"""Reverse mode on the GPU: the selected inverse against a dense inverse,
the log-determinant gradient and the solve adjoint against dense FP64
references, the 2D Gaussian-Markov-field closed forms (log det, d/d log tau,
d/d log kappa from the Dirichlet Laplacian's eigenvalues), and the
torch.autograd wrapper against central finite differences and a dense
torch reference."""
import numpy as np
import pytest

from mdgpu_amd import _lib as L
from mdgpu_amd import gen, symbolic
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

CASES = [
    ("grid2d_30_rn500", lambda: gen.laplace2d(30), "rn", 500),
    ("grid3d_9_sn", lambda: gen.laplace3d_7pt(9), "sn", 0),
    ("st27_9_rn100000", lambda: gen.stencil27(9, seed=2), "rn", 100000),
    ("bbt_300_rn500", lambda: gen.bbt_plus_ni(300, seed=9), "rn", 500),
]


def caller_rows_cols(p, i):
    cols = np.repeat(np.arange(len(p) - 1), np.diff(p))
    return np.asarray(i), cols


@pytest.mark.parametrize("name,make,mode,mz", CASES, ids=[c[0] for c in CASES])
def test_selinv_matches_dense_inverse(device, name, make, mode, mz):
    import torch
    from mdgpu_amd.numeric import Factorization
    p, i, x = make()
    S = symbolic.analyze(p, i, True, mode, mz)
    F = Factorization(S, "fp64").factorize(torch.from_numpy(x).to(device)).selinv()
    Zx = F.selinv_export()
    _, Li, _ = F.export_L()
    perm = S.array(L.ARR_PERM)
    A = gen.full_from_lower(p, i, x).toarray()
    Zd = np.linalg.inv(A[np.ix_(perm, perm)])
    lp = S.array(L.ARR_LP)
    cols = np.repeat(np.arange(S.n), np.diff(lp))
    assert rel_err(Zx, Zd[Li, cols]) < 1e-11


@pytest.mark.parametrize("name,make,mode,mz", CASES[:2] + CASES[3:], ids=[c[0] for c in CASES[:2] + CASES[3:]])
def test_logdet_grad_and_solve_adjoint_match_dense(device, name, make, mode, mz):
    import torch
    from mdgpu_amd.numeric import Factorization
    p, i, x = make()
    S = symbolic.analyze(p, i, True, mode, mz)
    F = Factorization(S, "fp64").factorize(torch.from_numpy(x).to(device))
    A = gen.full_from_lower(p, i, x).toarray()
    Ainv = np.linalg.inv(A)
    r, c = caller_rows_cols(p, i)
    g = F.logdet_grad().cpu().numpy()
    ref = np.where(r == c, 1.0, 2.0) * Ainv[r, c]
    assert rel_err(g, ref) < 1e-11
    rng = np.random.default_rng(3)
    b = rng.standard_normal(S.n)
    xbar = rng.standard_normal(S.n)
    xs = F.solve(torch.from_numpy(b).to(device))
    bbar, ga = F.solve_adjoint(xs, torch.from_numpy(xbar).to(device))
    xd = Ainv @ b
    bd = Ainv @ xbar
    assert rel_err(bbar.cpu().numpy(), bd) < 1e-11
    refa = np.where(r == c, -bd[r] * xd[r], -(bd[r] * xd[c] + bd[c] * xd[r]))
    assert rel_err(ga.cpu().numpy(), refa) < 1e-11


def gmrf_parts(k):
    """Lower-triangle pattern of the 5-point grid and the values of the
    Dirichlet Laplacian (diag 4, off -1) and of the identity on it."""
    p, i, lap = gen.stencil_lower((k, k, 1), 5, 4.0, off=-1.0)
    r, c = caller_rows_cols(p, i)
    return p, i, lap, (r == c).astype(np.float64)


def gmrf_vals(theta, lap, eye):
    import torch
    tau, kappa = torch.exp(theta[0]), torch.exp(theta[1])
    return tau * (kappa * kappa * eye + lap)


def test_gmrf_closed_forms_256(device):
    """Q = tau (kappa^2 I + L) on a 256 x 256 grid at theta = (0, -1):
    log det Q = sum log tau (kappa^2 + lam_ij), d/d log tau = n,
    d/d log kappa = sum 2 kappa^2 / (kappa^2 + lam_ij), with
    lam_ij = 4 - 2 cos(i pi / 257) - 2 cos(j pi / 257)."""
    import torch
    from mdgpu_amd.autograd import logdet_solve
    k = 256
    p, i, lap, eye = gmrf_parts(k)
    S = symbolic.analyze(p, i, True, "rn", 1000)
    theta = torch.tensor([0.0, -1.0], dtype=torch.float64, device=device, requires_grad=True)
    vals = gmrf_vals(theta, torch.from_numpy(lap).to(device), torch.from_numpy(eye).to(device))
    b = torch.from_numpy(np.random.default_rng(0).standard_normal(k * k)).to(device)
    ld, x = logdet_solve(vals, b, S)
    ld.backward()
    t = np.arange(1, k + 1) * np.pi / (k + 1)
    lam = (2 - 2 * np.cos(t))[:, None] + (2 - 2 * np.cos(t))[None, :]
    kap2 = np.exp(-2.0)
    ld_ref = np.sum(np.log(kap2 + lam))
    assert abs(ld.item() - ld_ref) <= 1e-10 * abs(ld_ref)
    g = theta.grad.cpu().numpy()
    assert abs(g[0] - k * k) <= 1e-9 * k * k
    gk = np.sum(2 * kap2 / (kap2 + lam))
    assert abs(g[1] - gk) <= 1e-9 * abs(gk)


def test_autograd_matches_fd_and_dense_32(device):
    """l(theta) = log det Q + w . Q^{-1} b on a 32 x 32 grid: the gradient
    over theta and over b against central differences (h = 1e-6, 1e-6
    relative) and against a dense FP64 torch reference (1e-10)."""
    import torch
    from mdgpu_amd.autograd import logdet_solve
    k = 32
    p, i, lap, eye = gmrf_parts(k)
    S = symbolic.analyze(p, i, True, "rn", 200)
    rng = np.random.default_rng(1)
    bn = rng.standard_normal(k * k)
    wn = rng.standard_normal(k * k)
    lap_t, eye_t = torch.from_numpy(lap).to(device), torch.from_numpy(eye).to(device)
    w = torch.from_numpy(wn).to(device)

    def loss(th, bb):
        ld, x = logdet_solve(gmrf_vals(th, lap_t, eye_t), bb, S)
        return ld + torch.dot(w, x)

    theta = torch.tensor([0.3, -0.7], dtype=torch.float64, device=device, requires_grad=True)
    b = torch.from_numpy(bn).to(device).requires_grad_(True)
    loss(theta, b).backward()
    g = theta.grad.cpu().numpy()
    gb = b.grad.cpu().numpy()
    h = 1e-6
    for t in range(2):
        e = torch.zeros(2, dtype=torch.float64, device=device)
        e[t] = h
        with torch.no_grad():
            fd = (loss(theta + e, b).item() - loss(theta - e, b).item()) / (2 * h)
        assert abs(g[t] - fd) <= 1e-6 * max(1.0, abs(fd)), (t, g[t], fd)
    # dense FP64 reference on the CPU
    r, c = caller_rows_cols(p, i)
    th = theta.detach().cpu().clone().requires_grad_(True)
    bc = torch.from_numpy(bn).clone().requires_grad_(True)
    v = gmrf_vals(th, torch.from_numpy(lap), torch.from_numpy(eye))
    Q = torch.zeros(k * k, k * k, dtype=torch.float64)
    Q = Q.index_put((torch.from_numpy(r), torch.from_numpy(c)), v)
    Q = Q + Q.T - torch.diag(torch.diagonal(Q))
    ld = torch.linalg.slogdet(Q)[1]
    xd = torch.linalg.solve(Q, bc)
    (ld + torch.dot(torch.from_numpy(wn), xd)).backward()
    assert rel_err(g, th.grad.numpy()) < 1e-10
    assert rel_err(gb, bc.grad.numpy()) < 1e-10
