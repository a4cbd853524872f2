"""Inputs of the symbolic golden vectors (shared by the generator and the tests)."""
from __future__ import annotations

import hashlib

import numpy as np

from mdgpu_amd import gen


def input_hash(p, i):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(p, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(i, dtype="<i4").tobytes())
    return h.hexdigest()


def random_connected(n, deg, seed, dense_rows=0):
    """Random symmetric pattern: a path (connectivity), `deg` random
    neighbours per node, and `dense_rows` rows coupled to most nodes (these
    exceed the minimum-degree dense threshold max(16, 10 sqrt(n)))."""
    rng = np.random.default_rng(seed)
    rows = [np.arange(1, n)]
    cols = [np.arange(0, n - 1)]
    a = rng.integers(0, n, size=n * deg)
    b = np.repeat(np.arange(n), deg)
    rows.append(np.maximum(a, b))
    cols.append(np.minimum(a, b))
    for t in range(dense_rows):
        d = int(rng.integers(0, n))
        others = rng.choice(n, size=int(0.8 * n), replace=False)
        rows.append(np.maximum(others, d))
        cols.append(np.minimum(others, d))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    keep = r != c
    r, c = r[keep], c[keep]
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = key // n, key % n
    rr = np.concatenate([r, np.arange(n)])
    cc = np.concatenate([c, np.arange(n)])
    vals = np.concatenate([-np.ones(len(r)), np.full(n, 1.0 + 2.0 * n)])
    return gen.lower_csc_from_coo(n, rr, cc, vals)


SMALL_CASES = []
for _mode, _mz in [("plain", 0), ("sn", 0), ("rn", 0), ("rn", 50)]:
    SMALL_CASES.append((f"grid2d_7_{_mode}{_mz}", lambda: gen.laplace2d(7), _mode, _mz))
    SMALL_CASES.append((f"grid3d_543_{_mode}{_mz}", lambda: gen.stencil_lower((5, 4, 3), 7, 6.1), _mode, _mz))
for _mode, _mz in [("plain", 0), ("sn", 0), ("rn", 0), ("rn", 1000), ("rn", 100000)]:
    SMALL_CASES.append((f"c1_100x100_{_mode}{_mz}", lambda: gen.laplace2d(100), _mode, _mz))
for _mode, _mz in [("plain", 0), ("rn", 200)]:
    SMALL_CASES.append((f"st27_6_{_mode}{_mz}", lambda: gen.stencil27(6, seed=11), _mode, _mz))
for _seed in range(4):
    SMALL_CASES.append((f"rand_400_s{_seed}_rn100", lambda s=_seed: random_connected(400, 3, s, dense_rows=2),
                        "rn", 100))
    SMALL_CASES.append((f"rand_900_s{_seed}_plain", lambda s=_seed: random_connected(900, 2, 100 + s), "plain", 0))
SMALL_CASES.append(("bbt_600_s5_plain", lambda: gen.bbt_plus_ni(600, seed=5), "plain", 0))
SMALL_CASES.append(("gmrf_32_rn1000", lambda: gen.gmrf_precision(32), "rn", 1000))

LARGE_CASES = [
    ("c2_48_plain", gen.laplace3d_7pt, "plain", 0),
    ("c2_48_sn", gen.laplace3d_7pt, "sn", 0),
    ("c2_48_rn1000", gen.laplace3d_7pt, "rn", 1000),
    ("c3_64_plain", gen.stencil27, "plain", 0),
    ("c3_64_sn", gen.stencil27, "sn", 0),
    ("c3_64_rn1000", gen.stencil27, "rn", 1000),
    ("c5_256_plain", gen.gmrf_precision, "plain", 0),
    ("c5_256_rn1000", gen.gmrf_precision, "rn", 1000),
]
# zero budgets of the amalgamation tuner's search (max_zeros candidates)
TUNER_BUDGETS = (250, 1000, 2000, 4000, 8000, 16000, 64000)
SWEEP_GRIDS = [("c2_48", gen.laplace3d_7pt), ("c3_64", gen.stencil27), ("c1_100", lambda: gen.laplace2d(100))]
for _mz in TUNER_BUDGETS:
    if _mz != 1000:
        LARGE_CASES.append((f"c2_48_rn{_mz}", gen.laplace3d_7pt, "rn", _mz))
        LARGE_CASES.append((f"c3_64_rn{_mz}", gen.stencil27, "rn", _mz))
    LARGE_CASES.append((f"c1_100_rn{_mz}", lambda: gen.laplace2d(100), "rn", _mz))
