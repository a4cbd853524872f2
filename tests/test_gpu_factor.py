"""GPU parity: the HIP factorization and solves against the CPU oracle."""
import numpy as np
import pytest

from mdgpu_amd import gen, symbolic
from mdgpu_amd.numeric import Factorization
from tests.helpers import cases, oracle_factor, rel_err

pytestmark = pytest.mark.gpu

CASES = cases()


@pytest.mark.parametrize("name,make,mode,mz", CASES, ids=[c[0] for c in CASES])
def test_fp64_factor_matches_oracle(device, name, make, mode, mz):
    import torch
    p, i, x = make()
    S = symbolic.analyze(p, i, True, mode, mz)
    lp, Li_ref, Lx_ref, _, bad = oracle_factor(p, i, x, S)
    assert bad == -1
    F = Factorization(S, "fp64").factorize(torch.from_numpy(x).to(device))
    Lp, Li, Lx = F.export_L()
    assert np.array_equal(Lp, lp)
    assert np.array_equal(Li, Li_ref)            # pattern of L bit-exact
    assert rel_err(Lx, Lx_ref) <= 1e-10          # FP64 factor within 1e-10 of the oracle


@pytest.mark.parametrize("name,make,mode,mz", CASES, ids=[c[0] for c in CASES])
def test_fp64_solve_backward_error(device, name, make, mode, mz):
    import torch
    p, i, x = make()
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=1)
    S = symbolic.analyze(p, i, True, mode, mz)
    F = Factorization(S, "fp64").factorize(torch.from_numpy(x).to(device))
    xs = F.solve(torch.from_numpy(b).to(device)).cpu().numpy()
    from oracle import oracle as O
    eta = O.backward_error(p, i, x, xs, b)
    assert eta <= 1e-14, eta


@pytest.mark.parametrize("name,make,mode,mz", CASES, ids=[c[0] for c in CASES])
def test_fp32_refined_to_1e12(device, name, make, mode, mz):
    import torch
    p, i, x = make()
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=2)
    S = symbolic.analyze(p, i, True, mode, mz)
    F = Factorization(S, "fp32").factorize(torch.from_numpy(x).to(device))
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=30)
    from oracle import oracle as O
    eta_cpu = O.backward_error(p, i, x, xs.cpu().numpy(), b)
    assert eta <= 1e-12 and eta_cpu <= 1e-12, (eta, eta_cpu, iters)


@pytest.mark.parametrize("name,make,mode,mz", CASES, ids=[c[0] for c in CASES])
def test_bf16x3_refined_to_1e12(device, name, make, mode, mz):
    """Split-bf16 Schur updates (hi*hi + hi*lo + lo*hi on bf16 MFMA) + FP64 IR."""
    import torch
    p, i, x = make()
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=5)
    S = symbolic.analyze(p, i, True, mode, mz)
    F = Factorization(S, "bf16x3").factorize(torch.from_numpy(x).to(device))
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=50)
    from oracle import oracle as O
    eta_cpu = O.backward_error(p, i, x, xs.cpu().numpy(), b)
    assert eta <= 1e-12 and eta_cpu <= 1e-12, (eta, eta_cpu, iters)


@pytest.mark.parametrize("prec", ["bf16x3", "bf16"])
@pytest.mark.parametrize("nb_outer", ["64", "256"])
def test_split_planes_equal_staged_conversion(device, monkeypatch, prec, nb_outer):
    """The bf16 GEMMs that read the split planes the block steps wrote give
    the factor the staged kernels give (they convert the FP32 tiles as they
    load them), bit for bit: same hi/lo values, same k order and products.
    Fronts of several outer blocks with pivot counts that are not multiples
    of 32 (the planes kernel's last k-tile is masked)."""
    import torch
    p, i, x = gen.stencil27(14, seed=3)
    S = symbolic.analyze(p, i, True, "rn", 100000)
    ax = torch.from_numpy(x).to(device)
    monkeypatch.setenv("MDGPU_NB_OUTER", nb_outer)
    out = {}
    for presplit in ("0", "1"):
        monkeypatch.setenv("MDGPU_PRESPLIT", presplit)
        F = Factorization(S, prec).factorize(ax)
        out[presplit] = F.export_L()[2]
    cs = [nd[1] for nd in S.nodes()]
    assert max(cs) > 2 * int(nb_outer) and any(c % 32 for c in cs if c > int(nb_outer))
    assert np.isfinite(out["1"]).all()
    assert np.array_equal(out["0"], out["1"])


BF16_CASES = [c for c in CASES if c[0] in ("grid2d_7_plain", "bbt_600_rn1000")]


@pytest.mark.parametrize("name,make,mode,mz", BF16_CASES, ids=[c[0] for c in BF16_CASES])
def test_bf16_refined_to_1e12(device, name, make, mode, mz):
    """Single bf16 products converge under FP64 IR on well-conditioned cases
    (kappa ~ 25 and ~ 1); the factor itself is close to the FP32 one."""
    import torch
    p, i, x = make()
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=6)
    S = symbolic.analyze(p, i, True, mode, mz)
    ax = torch.from_numpy(x).to(device)
    F = Factorization(S, "bf16").factorize(ax)
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=60)
    from oracle import oracle as O
    eta_cpu = O.backward_error(p, i, x, xs.cpu().numpy(), b)
    assert eta <= 1e-12 and eta_cpu <= 1e-12, (eta, eta_cpu, iters)
    _, _, l16 = F.export_L()
    _, _, l32 = Factorization(S, "fp32").factorize(ax).export_L()
    assert rel_err(l16, l32) < 5e-2


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_replays_match_eager(device, prec):
    """The captured factorization/solve graphs are replayed several times with
    eager work and synchronizations in between; every replay must equal the
    eager launch sequence bit for bit."""
    import torch
    p, i, x = gen.laplace3d_7pt(14)
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=4)
    S = symbolic.analyze(p, i, True, "rn", 2000)
    ax = torch.from_numpy(x).to(device)
    bd = torch.from_numpy(b).to(device)
    E = Factorization(S, prec).set_graph(False).factorize(ax)
    _, _, lx_eager = E.export_L()
    xe = E.solve(bd).cpu().numpy()
    G = Factorization(S, prec).set_graph(True)
    for rep in range(3):
        G.factorize(ax)
        _, _, lx = G.export_L()
        assert np.array_equal(lx, lx_eager), rep
        xs = G.solve(bd).cpu().numpy()
        assert np.array_equal(xs, xe), rep
        E.solve(bd)                      # eager work between replays
        torch.cuda.synchronize()


def _dense_lower(A):
    n = A.shape[0]
    r, c = np.tril_indices(n)
    return gen.lower_csc_from_coo(n, r, c, A[r, c])


def test_pivot_growth_fallback_to_fp64(device):
    """A diagonal block with heavy cancellation (ones + 1e-4 I: pivots ~1e-4
    of the assembled diagonal) fails the FP32 pivot-growth check and is
    refactored in FP64; refinement still reaches 1e-12.  A well-conditioned
    matrix never falls back."""
    import torch
    from oracle import oracle as O
    n = 130
    A = np.ones((n, n)) + 1e-4 * np.eye(n)
    p, i, x = _dense_lower(A)
    S = symbolic.analyze(p, i, False, "sn", 0)
    F = Factorization(S, "fp32").factorize(torch.from_numpy(x).to(device))
    assert F.info()["fp64_fallbacks"] > 0
    b = A @ np.random.default_rng(0).uniform(-1, 1, n)
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=80)
    assert eta <= 1e-12 and O.backward_error(p, i, x, xs.cpu().numpy(), b) <= 1e-12, (eta, iters)
    p, i, x = gen.laplace2d(100)
    S = symbolic.analyze(p, i, True, "rn", 1000)
    G = Factorization(S, "fp32").factorize(torch.from_numpy(x).to(device))
    assert G.info()["fp64_fallbacks"] == 0


# plus a case whose large fronts have chains of up to ten 128-column blocks
SOLVE_CASES = CASES + [("st27_20_rn100000", lambda: gen.stencil27(20, seed=6), "rn", 100000)]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name,make,mode,mz", SOLVE_CASES, ids=[c[0] for c in SOLVE_CASES])
def test_solve_matches_exported_factor(device, name, make, mode, mz, prec):
    """The solve (small fronts one workgroup each, large fronts by the chained
    128-column block launches) against triangular solves in FP64 on the CPU
    with the factor the GPU computed.  FP64: same L, only the summation order
    differs (1e-10).  FP32: the solves apply the diagonal-block inverses
    rounded once to FP32, a perturbation of relative size ~1e-7 (bound 1e-5;
    a wrong index or a missed block gives O(1))."""
    import torch
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve_triangular
    from mdgpu_amd import _lib as L
    p, i, x = make()
    S = symbolic.analyze(p, i, True, mode, mz)
    F = Factorization(S, prec).factorize(torch.from_numpy(x).to(device))
    Lp, Li, Lx = F.export_L()
    Lc = sp.csc_matrix((Lx, Li, Lp), shape=(S.n, S.n)).tocsr()
    perm = S.array(L.ARR_PERM)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(S.n)
    y = spsolve_triangular(Lc, b[perm], lower=True)
    z = spsolve_triangular(Lc.T.tocsr(), y, lower=False)
    ref = np.empty(S.n)
    ref[perm] = z
    for rep in range(2):                     # the chain's counters and markers are reset per solve
        xs = F.solve(torch.from_numpy(b).to(device)).cpu().numpy()
        assert np.all(np.isfinite(xs)), (name, prec, rep)
        assert rel_err(xs, ref) <= (1e-10 if prec == "fp64" else 1e-5), (name, prec, rep, rel_err(xs, ref))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_solve_nrhs_matches_single_solves(device, prec):
    """mdgpu_solve_nrhs: each right-hand side of the block gives exactly the
    single solve's result, and the FP64 block solves to eta <= 1e-14."""
    import torch
    p, i, x = gen.stencil27(12, seed=3)
    S = symbolic.analyze(p, i, True, "rn", 1000)
    F = Factorization(S, prec).factorize(torch.from_numpy(x).to(device))
    rng = np.random.default_rng(7)
    B = torch.from_numpy(rng.standard_normal((3, S.n))).to(device)
    X = F.solve(B).cpu().numpy()
    for r in range(3):
        assert np.array_equal(X[r], F.solve(B[r]).cpu().numpy()), r
    if prec == "fp64":
        from oracle import oracle as O
        for r in range(3):
            assert O.backward_error(p, i, x, X[r], B[r].cpu().numpy()) <= 1e-14


@pytest.mark.parametrize("nb_outer", ["64", "128"])
def test_lookahead_outer_blocks(device, monkeypatch, nb_outer):
    """Narrow outer blocks put many trailing updates on the look-ahead side
    stream; the FP64 factor still matches the oracle and graph replays equal
    the eager run bit for bit."""
    import torch
    monkeypatch.setenv("MDGPU_NB_OUTER", nb_outer)
    for name, make, mode, mz in [c for c in CASES if c[0] in ("grid3d_16_rn4000", "st27_12_rn100000")]:
        p, i, x = make()
        S = symbolic.analyze(p, i, True, mode, mz)
        lp, Li_ref, Lx_ref, _, bad = oracle_factor(p, i, x, S)
        ax = torch.from_numpy(x).to(device)
        E = Factorization(S, "fp64").set_graph(False).factorize(ax)
        _, Li, Lx = E.export_L()
        assert np.array_equal(Li, Li_ref) and rel_err(Lx, Lx_ref) <= 1e-10, name
        G = Factorization(S, "fp64").set_graph(True)
        for rep in range(2):
            G.factorize(ax)
            assert np.array_equal(G.export_L()[2], Lx), (name, rep)


def test_fp32_gemm_kernels_agree(device, monkeypatch):
    """Every FP32 GEMM kernel, forced on every launch (MDGPU_GEMM_KERNEL):
    the factor stays within FP32 rounding of the oracle's FP64 factor, and the
    32x32x2 kernel's 128- and 64-row tiles give the same factor bit for bit
    (the same products in the same order per entry; their k-tiles are 16 and
    32 deep, so the partial last k-tile of the fronts whose pivot counts are
    not multiples of 32 takes both register-staged paths), with and without
    two-level accumulation."""
    import torch
    p, i, x = gen.stencil27(12, seed=1)
    S = symbolic.analyze(p, i, True, "rn", 100000)
    lp, Li_ref, Lx_ref, _, bad = oracle_factor(p, i, x, S)
    ax = torch.from_numpy(x).to(device)
    out = {}
    for code in (129, 65, 129 | 256, 65 | 256, 128, 64, 64 | 256):
        monkeypatch.setenv("MDGPU_GEMM_KERNEL", str(code))
        F = Factorization(S, "fp32").factorize(ax)
        _, Li, Lx = F.export_L()
        assert np.array_equal(Li, Li_ref), code
        assert rel_err(Lx, Lx_ref) <= 2e-5, (code, rel_err(Lx, Lx_ref))
        out[code] = Lx
    assert np.array_equal(out[129], out[65])
    assert np.array_equal(out[129 | 256], out[65 | 256])
    assert not np.array_equal(out[129], out[128])   # the 16x16x4 kernel rounds differently


def test_autotune_keeps_the_factor(device):
    """The autotuner only changes GEMM tile sizes: the FP64 factor still
    matches the oracle, and graph replays after tuning reproduce it."""
    import torch
    name, make, mode, mz = [c for c in CASES if c[0] == "st27_12_rn100000"][0]
    p, i, x = make()
    S = symbolic.analyze(p, i, True, mode, mz)
    lp, Li_ref, Lx_ref, _, bad = oracle_factor(p, i, x, S)
    ax = torch.from_numpy(x).to(device)
    F = Factorization(S, "fp64")
    n64 = F.autotune(ax)
    assert 0 <= n64 <= F.info()["steps"]
    _, Li, Lx = F.export_L()
    assert np.array_equal(Li, Li_ref) and rel_err(Lx, Lx_ref) <= 1e-10
    F.factorize(ax)
    assert np.array_equal(F.export_L()[2], Lx)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_indefinite_matrix_is_reported(device, prec):
    """A - sigma I with sigma above lambda_min: the factorization meets a
    non-positive pivot late (in the large fronts of the blocked path) and
    returns MDGPU_ERR_NOT_SPD naming the column; a following factorization
    of the SPD matrix on the same handle succeeds."""
    import torch
    from mdgpu_amd import _lib as L
    p, i, x = gen.laplace3d_7pt(16)          # lambda_min = 6.001 - 6 cos(pi/17) ~ 0.103
    n = len(p) - 1
    cols = np.repeat(np.arange(n), np.diff(p))
    xs = x.copy()
    xs[i == cols] -= 0.2
    S = symbolic.analyze(p, i, True, "rn", 4000)
    F = Factorization(S, prec)
    with pytest.raises(L.MdgpuError) as e:
        F.factorize(torch.from_numpy(xs).to(device))
    assert e.value.code == L.MDGPU_ERR_NOT_SPD and "non-positive pivot" in str(e.value)
    F.factorize(torch.from_numpy(x).to(device))
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=9)
    _, eta, _ = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=30)
    assert eta <= 1e-12


def test_pivot_growth_fallback_lds_front(device):
    """A front factored wholly in LDS (one 16 x 16 dense supernode, 1 KB in
    FP32) with heavy cancellation (ones + 1e-4 I: every pivot after the first
    is ~1e-4 of its assembled diagonal, below the 2^-12 threshold) fails the
    FP32 pivot-growth check and is refactored in FP64 by the fallback launch.
    The FP32 factor then equals the FP64 factor rounded to FP32, and
    refinement reaches 1e-12; the FP64 path never falls back."""
    import torch
    from oracle import oracle as O
    n = 16
    A = np.ones((n, n)) + 1e-4 * np.eye(n)
    p, i, x = _dense_lower(A)
    S = symbolic.analyze(p, i, False, "sn", 0)
    assert S.nnodes == 1
    ax = torch.from_numpy(x).to(device)
    F = Factorization(S, "fp32").factorize(ax)
    info = F.info()
    assert info["panel_fronts"] == 1 and info["diag_jobs"] == 0, info
    assert info["fp64_fallbacks"] == 1, info
    D = Factorization(S, "fp64").factorize(ax)
    assert D.info()["fp64_fallbacks"] == 0
    _, Li, Lx32 = F.export_L()
    _, Li64, Lx64 = D.export_L()
    assert np.array_equal(Li, Li64)
    # the fallback factors the FP32-rounded assembled values in FP64: within
    # FP32 rounding (relative to the largest entry) of the FP64 factor
    assert rel_err(Lx32, Lx64) <= 1e-5, rel_err(Lx32, Lx64)
    b = A @ np.random.default_rng(0).uniform(-1, 1, n)
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=80)
    assert eta <= 1e-12 and O.backward_error(p, i, x, xs.cpu().numpy(), b) <= 1e-12, (eta, iters)
    # a second factorization on the same handle: the fallback list starts empty again
    F.factorize(ax)
    assert F.info()["fp64_fallbacks"] == 1
    assert np.array_equal(F.export_L()[2], Lx32)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_refine_eta_equals_oracle_backward_error(device, prec):
    """The backward error mdgpu_refine reports is the reference's
    ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf) (mdgpu.cu:1972) of the
    x it returns: with maxit = 0 the x is the unrefined reduced-precision
    solve (eta ~1e-8, far above rounding), and the oracle's backward_error of
    that same x must agree to 1e-6 relative.  A dropped ||b|| term or a wrong
    norm changes eta by O(1)."""
    import torch
    from oracle import oracle as O
    p, i, x = gen.stencil27(14, seed=8)
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=8)
    S = symbolic.analyze(p, i, True, "rn", 1000)
    F = Factorization(S, prec).factorize(torch.from_numpy(x).to(device))
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=0, raise_on_fail=False)
    assert iters == 0
    eta_cpu = O.backward_error(p, i, x, xs.cpu().numpy(), b)
    assert eta_cpu > 1e-11, eta_cpu             # far above FP64 rounding
    assert abs(eta - eta_cpu) <= 1e-6 * eta_cpu, (eta, eta_cpu)
    # after refinement the residual is rounding (summed in another order on
    # the GPU): agreement to 5%
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=30)
    eta_cpu = O.backward_error(p, i, x, xs.cpu().numpy(), b)
    assert eta <= 1e-12 and eta_cpu <= 1e-12
    assert abs(eta - eta_cpu) <= 5e-2 * eta_cpu + 1e-16, (eta, eta_cpu)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_solve_in_place_vector_path_equals_lds_path(device, monkeypatch, prec):
    """With MDGPU_SOLVE_LDS_ROWS=0 every front vector of the forward solve is
    accumulated in place in global memory instead of in LDS (the path fronts
    longer than the LDS capacity take).  Children are added in the same
    order either way, so the solution is bit-identical, and it matches
    triangular solves with the exported factor."""
    import torch
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve_triangular
    p, i, x = gen.stencil27(20, seed=6)
    S = symbolic.analyze(p, i, True, "rn", 100000)
    ax = torch.from_numpy(x).to(device)
    b = torch.from_numpy(np.random.default_rng(12).standard_normal(S.n)).to(device)
    F = Factorization(S, prec).factorize(ax)
    x_lds = F.solve(b).cpu().numpy()
    monkeypatch.setenv("MDGPU_SOLVE_LDS_ROWS", "0")
    G = Factorization(S, prec).factorize(ax)
    x_glob = G.solve(b).cpu().numpy()
    assert np.array_equal(x_lds, x_glob)
    from mdgpu_amd import _lib as L
    Lp, Li, Lx = G.export_L()
    Lc = sp.csc_matrix((Lx, Li, Lp), shape=(S.n, S.n)).tocsr()
    perm = S.array(L.ARR_PERM)
    bh = b.cpu().numpy()
    z = spsolve_triangular(Lc.T.tocsr(), spsolve_triangular(Lc, bh[perm], lower=True), lower=False)
    ref = np.empty(S.n)
    ref[perm] = z
    assert rel_err(x_glob, ref) <= (1e-10 if prec == "fp64" else 1e-5)


def test_calls_on_two_streams_run_in_call_order(device):
    """Solves on one handle issued from two streams without synchronizing
    (they share the handle's solve buffers) run in call order on the GPU:
    each result equals the solve of its own right-hand side."""
    import torch
    p, i, x = gen.stencil27(16, seed=2)
    S = symbolic.analyze(p, i, True, "rn", 1000)
    F = Factorization(S, "fp64").factorize(torch.from_numpy(x).to(device))
    rng = np.random.default_rng(5)
    bs = [torch.from_numpy(rng.standard_normal(S.n)).to(device) for _ in range(6)]
    ref = [F.solve(b).cpu().numpy() for b in bs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = []
    for k, b in enumerate(bs):
        with torch.cuda.stream(streams[k % 2]):
            outs.append(F.solve(b))
    torch.cuda.synchronize()
    for k in range(len(bs)):
        assert np.array_equal(outs[k].cpu().numpy(), ref[k]), k


def test_chained_wait_give_up_is_reported(device, monkeypatch):
    """MDGPU_CHAIN_SPIN=0 makes every wait of the chained block solves give
    up at its first poll of a value not yet stored (every kernel still
    exits): the solve status reports it once, mdgpu_refine returns
    MDGPU_ERR_DEVICE, and a handle with the normal limit solves correctly."""
    import torch
    from mdgpu_amd import _lib as L
    p, i, x = gen.stencil27(20, seed=6)
    S = symbolic.analyze(p, i, True, "rn", 100000)
    ax = torch.from_numpy(x).to(device)
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=3)
    bd = torch.from_numpy(b).to(device)
    monkeypatch.setenv("MDGPU_CHAIN_SPIN", "0")
    F = Factorization(S, "fp64").factorize(ax)
    assert F.solve_status() == 0
    xs = F.solve(bd).cpu().numpy()
    assert F.solve_status() == L.SOLVE_TIMED_OUT
    assert F.solve_status() == 0                 # reading clears it
    assert not np.all(np.isfinite(xs))
    with pytest.raises(L.MdgpuError) as e:
        F.refine(bd, tol=1e-12, maxit=5)
    assert e.value.code == L.MDGPU_ERR_DEVICE
    monkeypatch.delenv("MDGPU_CHAIN_SPIN")
    G = Factorization(S, "fp64").factorize(ax)
    xs = G.solve(bd).cpu().numpy()
    assert G.solve_status() == 0 and np.all(np.isfinite(xs))


def test_partition_tuner_picks_a_budget_and_keeps_parity(device):
    """mdgpu_tune_partition times each relaxed analysis of the budget sweep
    (factorization + solves) and returns the fastest; the analysis it picks
    factors to the oracle's L like any other."""
    import torch
    from mdgpu_amd.numeric import tune_partition
    p, i, x = gen.stencil27(16, seed=1)
    budgets = [250, 1000, 4000, 16000]
    Ss = symbolic.analyze_budgets(p, i, budgets)
    ax = torch.from_numpy(x).to(device)
    best, ms = tune_partition(Ss, ax, "fp32", reps=2, solves=3)
    assert 0 <= best < len(budgets) and len(ms) == len(budgets)
    assert all(m > 0 for m in ms) and ms[best] == min(ms)
    S = Ss[best]
    lp, Li_ref, Lx_ref, _, bad = oracle_factor(p, i, x, S)
    _, Li, Lx = Factorization(S, "fp64").factorize(ax).export_L()
    assert bad == -1 and np.array_equal(Li, Li_ref) and rel_err(Lx, Lx_ref) <= 1e-10


def test_backward_error_entry_matches_oracle(device):
    """mdgpu_backward_error (the reference's eta check on the GPU) against
    the oracle's backward_error of the same x, at perturbed solutions where
    eta is far above rounding; matrix orders below, at and above one
    workgroup's 1024 rows."""
    from mdgpu_amd.numeric import backward_error
    from oracle import oracle as O
    rng = np.random.default_rng(21)
    for p, i, x in [gen.laplace2d(32), gen.stencil27(12, seed=1), gen.laplace2d(100), gen.bbt_plus_ni(600, seed=2)]:
        n = len(p) - 1
        xt, b = gen.rhs_from_xtrue(p, i, x, seed=3)
        xs = xt + 1e-6 * rng.standard_normal(n)
        e_gpu = backward_error(p, i, x, xs, b)
        e_cpu = O.backward_error(p, i, x, xs, b)
        # the residual at eta ~1e-6 carries rounding of ~1e-16 ||A|| ||x||
        # (summation orders differ): agreement to 1e-8 relative
        assert e_cpu > 1e-12 and abs(e_gpu - e_cpu) <= 1e-8 * e_cpu, (n, e_gpu, e_cpu)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_cg_refinement_converges_and_needs_no_more_solves(device, monkeypatch, prec):
    """MDGPU_REFINE=cg: conjugate gradients in FP64 preconditioned by the
    reduced-precision factor reach the same 1e-12 backward error (checked
    on the CPU) in at most one solve more than classical refinement (their
    first steps differ only by CG's step length), here on a matrix whose
    factor is poor (z-anisotropic 27-point stencil)."""
    import torch
    from oracle import oracle as O
    p, i, x = gen.stencil27(16, seed=4, anisotropic=True)
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=12)
    S = symbolic.analyze(p, i, True, "rn", 1000)
    ax = torch.from_numpy(x).to(device)
    bd = torch.from_numpy(b).to(device)
    its = {}
    for method in ("ir", "cg"):
        monkeypatch.setenv("MDGPU_REFINE", method)
        F = Factorization(S, prec).factorize(ax)
        xs, eta, iters = F.refine(bd, tol=1e-12, maxit=60)
        assert eta <= 1e-12 and O.backward_error(p, i, x, xs.cpu().numpy(), b) <= 1e-12, (method, eta, iters)
        its[method] = iters
    assert its["cg"] <= its["ir"] + 1, its
