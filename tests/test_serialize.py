"""The on-disk symbolic analysis (mdgpu_symbolic_save / _load) round-trips
every array and plans identically; damaged files are rejected."""
import ctypes as C

import numpy as np
import pytest

from mdgpu_amd import _lib as L
from mdgpu_amd import gen, symbolic

ARRS = [L.ARR_AMD_PERM, L.ARR_PARENT, L.ARR_POST, L.ARR_COUNT, L.ARR_PERM, L.ARR_COUNT_FINAL, L.ARR_LP,
        L.ARR_ALOW_P, L.ARR_ALOW_I, L.ARR_NODES, L.ARR_FRONT_PARENT, L.ARR_Q_OFF, L.ARR_Q, L.ARR_LEVEL,
        L.ARR_SEG_COL, L.ARR_SEG_NNZ]


def plan_info(S, prec):
    out = (C.c_int64 * L.NUM_NINFO)()
    L.check(L.lib().mdgpu_plan_check(S.handle, prec, out))
    return list(out)


@pytest.mark.parametrize("mode,mz", [("rn", 500), ("sn", 0)])
def test_round_trip(tmp_path, mode, mz):
    p, i, x = gen.laplace3d_7pt(10)
    S = symbolic.analyze(p, i, True, mode, mz)
    path = tmp_path / "a.sym"
    S.save(path)
    T = symbolic.load(path)
    assert T.sizes == S.sizes and (T.flops, T.flops_padded) == (S.flops, S.flops_padded)
    for w in ARRS:
        assert np.array_equal(T.array(w), S.array(w)), w
    if mode == "rn":
        assert np.array_equal(T.array(L.ARR_RELAXED_PERM), S.array(L.ARR_RELAXED_PERM))
    for prec in (L.FP64, L.FP32):
        assert plan_info(T, prec) == plan_info(S, prec)


def test_batch_round_trip(tmp_path):
    mats = [gen.laplace2d(8), gen.bbt_plus_ni(90, seed=1)]
    S = symbolic.analyze_batch([(p, i) for p, i, _ in mats], True, "rn", 100)
    S.save(tmp_path / "b.sym")
    T = symbolic.load(tmp_path / "b.sym")
    assert T.nseg == 2 and np.array_equal(T.array(L.ARR_SEG_COL), S.array(L.ARR_SEG_COL))


def test_damaged_files_are_rejected(tmp_path):
    p, i, x = gen.laplace2d(12)
    S = symbolic.analyze(p, i, True, "rn", 100)
    path = tmp_path / "c.sym"
    S.save(path)
    data = bytearray(path.read_bytes())
    bad = tmp_path / "bad.sym"
    for mutate in (lambda d: d.__setitem__(40, d[40] ^ 1), lambda d: d.__delitem__(slice(len(d) // 2, None)),
                   lambda d: d.__setitem__(0, ord("X"))):
        d = bytearray(data)
        mutate(d)
        bad.write_bytes(bytes(d))
        with pytest.raises(L.MdgpuError):
            symbolic.load(bad)
    with pytest.raises(L.MdgpuError):
        symbolic.load(tmp_path / "missing.sym")


def _fnv1a(b):
    h = 1469598103934665603
    for c in b:
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _vector_offsets(d):
    """Byte offsets of the int32 vectors at the head of the file: magic (8),
    version (4), n, nnz_in, use_amd, partition, max_zeros (4 each), then
    amd_perm, parent, post, count, relaxed_perm, perm, pinv, count_final,
    lp, ... each as an int64 length and its entries."""
    off, out = 32, {}
    for name in ["amd_perm", "parent", "post", "count", "relaxed_perm", "perm", "pinv", "count_final", "lp"]:
        ln = int(np.frombuffer(bytes(d[off:off + 8]), dtype="<i8")[0])
        out[name] = (off + 8, ln)
        off += 8 + 4 * ln
    return out


@pytest.mark.parametrize("field,delta", [("lp", 1), ("count_final", 1)])
def test_inconsistent_layout_with_valid_checksum_is_rejected(tmp_path, field, delta):
    """A file that passes the checksum but whose L layout is inconsistent
    (spL->p not the prefix sum of the counts, or a front column count that
    does not match its row list) is rejected: mdgpu_export_L would write
    outside the caller's buffers with it."""
    p, i, x = gen.laplace2d(12)
    S = symbolic.analyze(p, i, True, "rn", 100)
    path = tmp_path / "c.sym"
    S.save(path)
    d = bytearray(path.read_bytes())
    start, ln = _vector_offsets(d)[field]
    t = start + 4 * (ln // 2)
    v = int(np.frombuffer(bytes(d[t:t + 4]), dtype="<i4")[0]) + delta
    d[t:t + 4] = np.array([v], dtype="<i4").tobytes()
    d[-8:] = np.array([_fnv1a(d[:-8])], dtype="<u8").tobytes()
    bad = tmp_path / "bad.sym"
    bad.write_bytes(bytes(d))
    with pytest.raises(L.MdgpuError) as e:
        symbolic.load(bad)
    assert "inconsistent" in str(e.value)
    # the untouched file (checksum recomputed the same way) still loads
    d0 = bytearray(path.read_bytes())
    d0[-8:] = np.array([_fnv1a(d0[:-8])], dtype="<u8").tobytes()
    assert bytes(d0) == path.read_bytes()
    symbolic.load(path)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,mz", [("rn", 1000), ("sn", 0)])
def test_factor_from_loaded_analysis_matches_oracle(tmp_path, device, mode, mz):
    """Refactorization without analysis: a numeric handle built on an
    analysis read back from disk factors the matrix exactly as one built on
    the original (bit for bit in FP64), and both match the oracle."""
    import torch
    from mdgpu_amd.numeric import Factorization
    from oracle import oracle as O
    from tests.helpers import oracle_factor, rel_err
    p, i, x = gen.stencil27(14, seed=9)
    S = symbolic.analyze(p, i, True, mode, mz)
    S.save(tmp_path / "s.sym")
    T = symbolic.load(tmp_path / "s.sym")
    ax = torch.from_numpy(x).to(device)
    Lp, Li, Lx = Factorization(T, "fp64").factorize(ax).export_L()
    Lp0, Li0, Lx0 = Factorization(S, "fp64").factorize(ax).export_L()
    assert np.array_equal(Lp, Lp0) and np.array_equal(Li, Li0) and np.array_equal(Lx, Lx0)
    lp, Li_ref, Lx_ref, _, bad = oracle_factor(p, i, x, T)
    assert bad == -1 and np.array_equal(Li, Li_ref) and rel_err(Lx, Lx_ref) <= 1e-10
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=4)
    xs, eta, _ = Factorization(T, "fp32").factorize(ax).refine(torch.from_numpy(b).to(device), tol=1e-12)
    assert eta <= 1e-12 and O.backward_error(p, i, x, xs.cpu().numpy(), b) <= 1e-12
