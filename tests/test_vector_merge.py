"""Row-list union and relative map (VectorMerge / VectorMergeIndices,
mdgpu.cu:535-595) -- the operations the front row lists and extend-add maps
are built with -- against the reference's own known-answer test
(testVectorMerge, mdgpu.cu:1610-1624) and against numpy set operations on
random sorted lists."""
import ctypes as C

import numpy as np
import pytest

from mdgpu_amd import _lib as L


def union(a, b):
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    out = np.empty(len(a) + len(b), np.int32)
    n = C.c_int()
    L.check(L.lib().mdgpu_row_union(a.ctypes.data_as(L._I32P), len(a), b.ctypes.data_as(L._I32P), len(b),
                                    out.ctypes.data_as(L._I32P), C.byref(n)))
    return out[:n.value]


def rmap(sup, sub):
    sup = np.ascontiguousarray(sup, np.int32)
    sub = np.ascontiguousarray(sub, np.int32)
    out = np.empty(max(len(sub), 1), np.int32)
    L.check(L.lib().mdgpu_row_map(sup.ctypes.data_as(L._I32P), len(sup), sub.ctypes.data_as(L._I32P), len(sub),
                                  out.ctypes.data_as(L._I32P)))
    return out[:len(sub)]


def test_reference_known_answer():
    """testVectorMerge: {2, 5} merged with {0, ..., 5} is {0, ..., 5}, and the
    map of {2, 5} into it is {2, 5}."""
    q = union([2, 5], [0, 1, 2, 3, 4, 5])
    assert q.tolist() == [0, 1, 2, 3, 4, 5]
    assert rmap(q, [2, 5]).tolist() == [2, 5]


@pytest.mark.parametrize("seed", range(5))
def test_random_lists_against_numpy(seed):
    rng = np.random.default_rng(seed)
    a = np.unique(rng.integers(0, 300, 40))
    b = np.unique(rng.integers(0, 300, 70))
    q = union(a, b)
    assert np.array_equal(q, np.union1d(a, b))
    assert np.array_equal(rmap(q, b), np.searchsorted(q, b))
    assert np.array_equal(q[rmap(q, a)], a)


def test_missing_row_and_unsorted_input_are_rejected():
    with pytest.raises(L.MdgpuError):
        rmap([0, 1, 3], [2])
    with pytest.raises(L.MdgpuError):
        union([3, 1], [0])
