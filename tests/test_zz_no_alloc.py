"""mdgpu_backward_error_noalloc (the backward-error check with its workspace
in the library's static device array, for callers that cannot allocate)
against mdgpu_backward_error on the same inputs: bit-identical, on two calls
in a row, the second with a smaller matrix (the shared workspace keeps the
first call's contents), and the size limit rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest

from mdgpu_amd import _lib as L
from mdgpu_amd import gen

pytestmark = pytest.mark.gpu


def _device_csr(p, i, x, dev):
    import scipy.sparse as sp
    import torch
    n = len(p) - 1
    Lo = sp.csc_matrix((x, i, p), shape=(n, n))
    A = (Lo + sp.tril(Lo, -1).T).tocsr()
    A.sort_indices()
    return [torch.from_numpy(a).to(dev) for a in (A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)]


def _eta(fn, n, rp, ci, av, xs, b):
    import torch
    eta = C.c_double()
    st = fn(n, C.c_void_p(rp.data_ptr()), C.c_void_p(ci.data_ptr()), C.c_void_p(av.data_ptr()),
            C.c_void_p(xs.data_ptr()), C.c_void_p(b.data_ptr()), C.byref(eta),
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return st, eta.value


def test_noalloc_equals_allocating_entry_point(device):
    import torch
    from oracle import oracle as O
    lib = L.lib()
    rng = np.random.default_rng(4)
    # a larger matrix first (5000+ rows: several workgroups), then a smaller one
    for p, i, x in [gen.stencil27(18, seed=2), gen.laplace2d(30)]:
        n = len(p) - 1
        xt, b = gen.rhs_from_xtrue(p, i, x, seed=5)
        xs = xt + 1e-7 * rng.standard_normal(n)
        rp, ci, av = _device_csr(p, i, x, device)
        xd = torch.from_numpy(xs).to(device)
        bd = torch.from_numpy(b).to(device)
        st0, e_alloc = _eta(lib.mdgpu_backward_error, n, rp, ci, av, xd, bd)
        st1, e_noalloc = _eta(lib.mdgpu_backward_error_noalloc, n, rp, ci, av, xd, bd)
        assert st0 == L.MDGPU_OK and st1 == L.MDGPU_OK
        assert e_noalloc == e_alloc, (n, e_noalloc, e_alloc)
        e_cpu = O.backward_error(p, i, x, xs, b)
        # the residual at eta ~1e-7 carries rounding of ~1e-16 ||A|| ||x||: ~1e-9 relative
        assert abs(e_noalloc - e_cpu) <= 1e-8 * e_cpu


def test_noalloc_rejects_sizes_beyond_its_workspace(device):
    import torch
    lib = L.lib()
    d = torch.zeros(8, dtype=torch.float64, device=device)
    k = torch.zeros(8, dtype=torch.int32, device=device)
    st, _ = _eta(lib.mdgpu_backward_error_noalloc, 4096 * 1024 + 1, k, k, d, d, d)
    assert st == L.MDGPU_ERR_ARG
    assert b"exceeds the limit" in lib.mdgpu_last_error()
