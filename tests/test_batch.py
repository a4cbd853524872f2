"""Batches of independent matrices (mdgpu_analyze_batch): each matrix keeps
exactly its own analysis inside the block-diagonal union, and on the GPU the
level-batched factorization gives every matrix the same factor as a
single-matrix run (bit for bit in FP64), with the refinement stopping only
when every matrix reaches the tolerance."""
import ctypes as C

import numpy as np
import pytest

from mdgpu_amd import _lib as L
from mdgpu_amd import gen, symbolic


def small_batch():
    mats = [gen.laplace2d(9), gen.bbt_plus_ni(120, seed=3), gen.laplace3d_7pt(6),
            gen.stencil27(5, seed=1), gen.bbt_plus_ni(300, seed=4), gen.laplace2d(20)]
    return mats


ARRS = [L.ARR_AMD_PERM, L.ARR_PARENT, L.ARR_POST, L.ARR_COUNT, L.ARR_RELAXED_PERM, L.ARR_PERM,
        L.ARR_COUNT_FINAL]


@pytest.mark.parametrize("mode,mz", [("rn", 200), ("sn", 0), ("plain", 0)])
def test_batch_analysis_is_union_of_single_analyses(mode, mz):
    mats = small_batch()
    B = symbolic.analyze_batch([(p, i) for p, i, _ in mats], True, mode, mz, nthreads=3)
    singles = [symbolic.analyze(p, i, True, mode, mz) for p, i, _ in mats]
    seg = B.array(L.ARR_SEG_COL)
    segnz = B.array(L.ARR_SEG_NNZ)
    assert B.nseg == len(mats)
    assert list(seg) == list(np.concatenate([[0], np.cumsum([S.n for S in singles])]))
    assert list(segnz) == list(np.concatenate([[0], np.cumsum([len(i) for _, i, _ in mats])]))
    assert B.lnz == sum(S.lnz for S in singles)
    for which in ARRS:
        if which == L.ARR_RELAXED_PERM and mode != "rn":
            continue
        got = B.array(which)
        for k, S in enumerate(singles):
            ref = S.array(which).astype(np.int64)
            part = got[seg[k]:seg[k + 1]].astype(np.int64)
            if which in (L.ARR_COUNT, L.ARR_COUNT_FINAL):
                assert np.array_equal(part, ref)
            elif which == L.ARR_PARENT:
                assert np.array_equal(part, np.where(ref >= 0, ref + seg[k], -1))
            else:
                assert np.array_equal(part, ref + seg[k])
    # node lists: shifted columns and child indices
    bn = B.nodes()
    off_node = 0
    for k, S in enumerate(singles):
        sn = S.nodes()
        for t, (first, nc, rows, zeros, kids) in enumerate(sn):
            f2, nc2, r2, z2, k2 = bn[off_node + t]
            assert (f2, nc2, r2, z2) == (first + seg[k], nc, rows, zeros)
            assert k2 == [c + off_node for c in kids]
        off_node += len(sn)
    assert off_node == len(bn)
    for prec in (L.FP64, L.FP32):
        out = (C.c_int64 * L.NUM_NINFO)()
        L.check(L.lib().mdgpu_plan_check(B.handle, prec, out))


@pytest.mark.gpu
def test_batch_fp64_factor_equals_single_runs(device):
    import torch
    from mdgpu_amd.numeric import Factorization
    mats = small_batch()
    B = symbolic.analyze_batch([(p, i) for p, i, _ in mats], True, "rn", 200)
    ax = np.concatenate([x for _, _, x in mats])
    _, _, lx_b = Factorization(B, "fp64").factorize(torch.from_numpy(ax).to(device)).export_L()
    parts = []
    for p, i, x in mats:
        S = symbolic.analyze(p, i, True, "rn", 200)
        parts.append(Factorization(S, "fp64").factorize(torch.from_numpy(x).to(device)).export_L()[2])
    assert np.array_equal(lx_b, np.concatenate(parts))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_batch_refine_every_matrix_to_1e12(device, prec):
    import torch
    from mdgpu_amd.numeric import Factorization
    from oracle import oracle as O
    mats = small_batch()
    B = symbolic.analyze_batch([(p, i) for p, i, _ in mats], True, "rn", 200)
    seg = B.array(L.ARR_SEG_COL)
    ax = np.concatenate([x for _, _, x in mats])
    bs = [gen.rhs_from_xtrue(p, i, x, seed=7 + k)[1] for k, (p, i, x) in enumerate(mats)]
    b = np.concatenate(bs)
    F = Factorization(B, prec).factorize(torch.from_numpy(ax).to(device))
    xs, eta, iters = F.refine(torch.from_numpy(b).to(device), tol=1e-12, maxit=50)
    xs = xs.cpu().numpy()
    etas = F.segment_eta()
    assert len(etas) == len(mats) and eta == etas.max() and eta <= 1e-12
    for k, (p, i, x) in enumerate(mats):
        e = O.backward_error(p, i, x, xs[seg[k]:seg[k + 1]], bs[k])
        assert e <= 1e-12, (k, e, etas[k])
