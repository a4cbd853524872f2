"""Workload runner for rocprofv3 (kernel trace and hardware counters).

    python -m mdgpu_amd.profile_run --workload c3 --precision fp32 --factorizations 2 --solves 1

Analyses the workload (optionally autotunes the GEMM tiles), factorizes it
``--factorizations`` times and solves ``--solves`` times (eager launches, so
every kernel is its own dispatch),
then copies a buffer of ``--copy-mib`` MiB with a plain device copy.  The
copy moves a known number of bytes, so hardware byte counters
(FETCH_SIZE / WRITE_SIZE) read on it give the scale for the counters read
on the factorization kernels.  Prints one line per phase.
"""
from __future__ import annotations

import argparse
import sys
import time

from . import gen, symbolic
from .numeric import Factorization

WORKLOADS = {
    "c3": lambda: gen.stencil27(64, seed=0),
    "c2": lambda: gen.laplace3d_7pt(48),
    "c1": lambda: gen.laplace2d(100),
}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "bf16", "bf16x3"])
    ap.add_argument("--factorizations", type=int, default=2)
    ap.add_argument("--solves", type=int, default=1)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--mode", default="rn")
    ap.add_argument("--max-zeros", type=int, default=1000)
    ap.add_argument("--autotune", action="store_true",
                    help="pick the per-launch GEMM tiles first (as bench.py does); its own factorizations come first")
    args = ap.parse_args(argv)
    import torch
    p, i, x = WORKLOADS[args.workload]()
    t0 = time.perf_counter()
    S = symbolic.analyze(p, i, True, args.mode, args.max_zeros)
    print(f"analyzed n={S.n} fronts={S.nnodes} flops={S.flops:.3e} in {time.perf_counter() - t0:.1f}s", flush=True)
    ax = torch.from_numpy(x).cuda()
    F = Factorization(S, args.precision).set_graph(False)
    if args.autotune:
        print(f"autotune: {F.autotune(ax)} launches on 64-row tiles", flush=True)
    for k in range(args.factorizations):
        F.factorize(ax)
        torch.cuda.synchronize()
        print(f"factorization {k} done", flush=True)
    b = torch.ones(S.n, dtype=torch.float64, device="cuda")
    for k in range(args.solves):
        F.solve(b)
        torch.cuda.synchronize()
        print(f"solve {k} done", flush=True)
    src = torch.empty(args.copy_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"calibration copy of {args.copy_mib} MiB done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
