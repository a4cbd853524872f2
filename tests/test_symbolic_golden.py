"""Symbolic phase vs the reference's own output (golden vectors).

The fixtures were produced by tests/golden/make_symbolic_golden.py, which runs
the reference's cs.cpp + etree.cpp (compiled from source) on the same inputs.
Every array must match bit for bit: the cs_amd permutation, cs_etree parent,
cs_post, cs_counts, the relaxed permutation and padded counts, the permuted
lower triangle, spL->p and the node list.  Two fields are exempt for the
fundamental-supernode mode: `ceros`, which GetSuperETree never initialises
(etree.cpp:69-87 leave it as malloc garbage), and the child order, which the
reference stores inconsistently with its own Factor (etree.cpp:89-103 vs
mdgpu.cu:1434-1440); the partition, row counts and parent relation are
compared instead.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from mdgpu_amd import _lib as L
from mdgpu_amd import symbolic
from tests.golden.golden_cases import LARGE_CASES, SMALL_CASES, SWEEP_GRIDS, TUNER_BUDGETS, input_hash

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = np.load(os.path.join(HERE, "symbolic_small.npz"))
with open(os.path.join(HERE, "symbolic_small_meta.json")) as _f:
    META = json.load(_f)
HASH_PATH = os.path.join(HERE, "symbolic_hashes.json")
HASHES = json.load(open(HASH_PATH)) if os.path.exists(HASH_PATH) else {}

ARR = {"amd_perm": L.ARR_AMD_PERM, "parent": L.ARR_PARENT, "post": L.ARR_POST, "count": L.ARR_COUNT,
       "relaxed_perm": L.ARR_RELAXED_PERM, "perm_final": L.ARR_PERM, "count_final": L.ARR_COUNT_FINAL,
       "lp": L.ARR_LP, "alow_p": L.ARR_ALOW_P, "alow_i": L.ARR_ALOW_I}


def parse_ref_nodes(flat):
    """Harness layout: count, then per node nodo, cols, rows, ceros, nHijos, child nodos."""
    out, t = [], 1
    for _ in range(int(flat[0])):
        nodo, cols, rows, ceros, nh = (int(v) for v in flat[t:t + 5])
        out.append((nodo, cols, rows, ceros, [int(v) for v in flat[t + 5:t + 5 + nh]]))
        t += 5 + nh
    return out


def check_nodes(S, ref_nodes, mode):
    mine = S.nodes()
    assert len(mine) == len(ref_nodes)
    for (rf, rc, rr, rz, rk), (mf, mc, mr, mz, mk) in zip(ref_nodes, mine):
        assert (rf, rc, rr) == (mf, mc, mr)
        if mode != "sn":
            assert rz == mz
            assert rk == [mine[k][0] for k in mk]
    # parent relation by first column
    ref_par = {k: nd[0] for nd in ref_nodes for k in nd[4]}
    my_par = {mine[k][0]: nd[0] for nd in mine for k in nd[4]}
    assert ref_par == my_par


@pytest.mark.parametrize("name,make,mode,mz", SMALL_CASES, ids=[c[0] for c in SMALL_CASES])
def test_small_cases_bit_exact(name, make, mode, mz):
    p, i, _ = make()
    assert input_hash(p, i) == META[name]["input"]
    S = symbolic.analyze(p, i, True, mode, mz)
    keys = [k[len(name) + 2:] for k in SMALL.files if k.startswith(name + "__")]
    assert keys
    for k in keys:
        ref = SMALL[f"{name}__{k}"]
        if k == "nodes":
            check_nodes(S, parse_ref_nodes(ref), mode)
        else:
            np.testing.assert_array_equal(S.array(ARR[k]), ref, err_msg=k)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


# the relaxed cases of the tuner's budgets other than 1000 are checked by the
# budget sweep below (one ordering for all budgets of a grid)
SINGLE_CASES = [c for c in LARGE_CASES if not (c[2] == "rn" and c[3] in TUNER_BUDGETS and c[3] != 1000)]


def check_hashes(S, ref, mode):
    for k, (length, digest) in ((k, v) for k, v in ref.items() if k != "input"):
        if k == "nodes":
            if mode == "sn":
                continue   # ceros / child order are not defined outputs in this mode
            nodes = S.nodes()
            flat = [len(nodes)]
            for f, c, r, z, kids in nodes:
                flat += [f, c, r, z, len(kids)] + [nodes[q][0] for q in kids]
            mine = np.asarray(flat, dtype=np.int32)
        else:
            mine = S.array(ARR[k])
        assert len(mine) == length, k
        assert _sha(mine) == digest, k


@pytest.mark.skipif(not HASHES, reason="large golden hashes not generated")
@pytest.mark.parametrize("name,make,mode,mz", SINGLE_CASES, ids=[c[0] for c in SINGLE_CASES])
def test_large_cases_hashes(name, make, mode, mz):
    if name not in HASHES:
        pytest.skip("no golden hash for this case")
    p, i, _ = make()
    ref = HASHES[name]
    assert input_hash(p, i) == ref["input"]
    check_hashes(symbolic.analyze(p, i, True, mode, mz), ref, mode)


@pytest.mark.skipif(not HASHES, reason="large golden hashes not generated")
@pytest.mark.parametrize("grid,make", SWEEP_GRIDS, ids=[g for g, _ in SWEEP_GRIDS])
def test_budget_sweep_matches_reference(grid, make):
    """mdgpu_analyze_budgets (one ordering, one relaxed partition per zero
    budget: the amalgamation tuner's candidates) gives, for every budget,
    the reference's relaxed analysis of the full-size grid bit for bit."""
    p, i, _ = make()
    Ss = symbolic.analyze_budgets(p, i, TUNER_BUDGETS)
    for mz, S in zip(TUNER_BUDGETS, Ss):
        ref = HASHES[f"{grid}_rn{mz}"]
        assert input_hash(p, i) == ref["input"]
        check_hashes(S, ref, "rn")


def test_forest_is_accepted_and_matches_reference_prefix():
    """B B^T + nI can be reducible: the reference stops at GetETree
    (etree.cpp:26-33) after computing the ordering, etree, postorder and
    counts, which must still match; the analysis here goes on."""
    name = "bbt_600_s5_plain"
    assert META[name]["forest"]
    p, i, _ = [c for c in SMALL_CASES if c[0] == name][0][1]()
    S = symbolic.analyze(p, i, True, "plain", 0)
    for k in ["amd_perm", "parent", "post", "count"]:
        np.testing.assert_array_equal(S.array(ARR[k]), SMALL[f"{name}__{k}"])
    assert (S.array(L.ARR_PARENT) == -1).sum() > 1
