"""Generate the symbolic golden vectors from the reference's own code.

Runs oracle/_ref/ref_symbolic (the reference's cs.cpp + etree.cpp, compiled
from /root/reference by `make -C oracle ref`) on deterministic inputs and
writes
  tests/golden/symbolic_small.npz   full arrays for small cases
  tests/golden/symbolic_hashes.json sha256 of every array for large cases
Only data is committed (inputs are regenerated from mdgpu_amd.gen and
checked against a stored hash of the input pattern).

usage: python tests/golden/make_symbolic_golden.py [--large | --large-only | --cases a,b,...]
(--large-only rewrites only the large-case hashes; the reference's relaxed
amalgamation on the 64^3 stencil takes about 20 minutes; --cases computes the
named large cases only and merges them into the existing hash file, so they
can run as separate processes)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.golden.golden_cases import LARGE_CASES, SMALL_CASES, input_hash  # noqa: E402
from oracle import oracle as O  # noqa: E402

ARRAYS = ["amd_perm", "parent", "post", "count", "relaxed_perm", "perm_final", "count_final",
          "lp", "alow_p", "alow_i", "nodes"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def run_case(name, make, mode, mz):
    p, i, _ = make()
    with tempfile.TemporaryDirectory() as d:
        try:
            out = O.run_ref_symbolic(p, i, d, True, mode, mz)
        except Exception as e:  # the reference exits(1) on an elimination forest
            out = {}
            for fn in os.listdir(d):
                if fn.endswith(".i32"):
                    out[fn[:-4]] = np.fromfile(os.path.join(d, fn), dtype=np.int32)
            out["_error"] = str(e)
    return p, i, out


def large_hashes(cases):
    hashes = {}
    for name, make, mode, mz in cases:
        p, i, out = run_case(name, make, mode, mz)
        hashes[name] = {"input": input_hash(p, i)}
        for k in ARRAYS:
            if k in out:
                hashes[name][k] = [int(len(out[k])), sha(out[k])]
        print(name, "done", flush=True)
    return hashes


def main():
    path = os.path.join(HERE, "symbolic_hashes.json")
    if "--cases" in sys.argv:
        names = sys.argv[sys.argv.index("--cases") + 1].split(",")
        new = large_hashes([c for c in LARGE_CASES if c[0] in names])
        # merge under an exclusive lock: several of these may run at once
        import fcntl
        with open(path + ".lock", "w") as lk:
            fcntl.flock(lk, fcntl.LOCK_EX)
            hashes = json.load(open(path)) if os.path.exists(path) else {}
            hashes.update(new)
            with open(path, "w") as f:
                json.dump(hashes, f, indent=1, sort_keys=True)
        return
    large_only = "--large-only" in sys.argv
    large = large_only or "--large" in sys.argv
    small = {}
    meta = {}
    for name, make, mode, mz in ([] if large_only else SMALL_CASES):
        p, i, out = run_case(name, make, mode, mz)
        meta[name] = {"input": input_hash(p, i), "forest": "_error" in out}
        for k in ARRAYS:
            if k in out:
                small[f"{name}__{k}"] = out[k]
        print(name, {k: len(v) for k, v in out.items() if k in ARRAYS}, "forest" if "_error" in out else "")
    if not large_only:
        np.savez_compressed(os.path.join(HERE, "symbolic_small.npz"), **small)
        with open(os.path.join(HERE, "symbolic_small_meta.json"), "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)
    if large:
        with open(path, "w") as f:
            json.dump(large_hashes(LARGE_CASES), f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
