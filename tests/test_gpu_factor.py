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
