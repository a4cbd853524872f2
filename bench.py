"""Benchmark: multifrontal sparse Cholesky on one MI355X.

Workload (default): 3D 27-point stencil on a 64^3 grid (n = 262,144), edge
weights w ~ U(0.5, 1.5) (seed 0), diagonal = 1.0001 * row sum; b = A x_true
with x_true ~ U(-1, 1) (seed 0).  One step = numeric factorization in FP32 on
the GPU + FP64 iterative refinement to a backward error <= 1e-12.  The plain
FP64 path (factorization + refinement check) is timed the same way beside it.

value  = algorithmic factorization GFLOP/s of the FP32 path, flops =
         sum_j c_j^2 over the unpadded cs_counts column counts.
Inputs (matrix values, right-hand side) are resident in HBM before timing;
the symbolic analysis runs once, untimed (it is reused for refactorization).
The C3 run also times the z-anisotropic variant (z-edge weights x 1e4, same
pattern and analysis) and reports, for every workload, the fronts per level,
the extend-add GB/s and the MFMA utilisation of the large GEMM launches.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1|c4|c5]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "c3": ("3D 27-point stencil 64^3, random symmetric weights U(0.5,1.5), diag 1.0001*rowsum",
           lambda gen: gen.stencil27(64, seed=0)),
    "c2": ("3D 7-point Laplacian 48^3, diag 6+1e-3", lambda gen: gen.laplace3d_7pt(48)),
    "c1": ("2D 5-point Laplacian 100^2, diag 4+1e-2", lambda gen: gen.laplace2d(100)),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_path(F, ax, b, steps, warmup, tol, maxit=50):
    import torch
    for _ in range(warmup):
        F.factorize(ax)
        F.refine(b, tol=tol, maxit=maxit, raise_on_fail=False)
    torch.cuda.synchronize()
    tf, tr, etas, its = [], [], [], []
    x = None
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        F.factorize(ax)                       # ends with a stream sync (pivot check)
        t1 = time.perf_counter()
        x, eta, it = F.refine(b, tol=tol, maxit=maxit, raise_on_fail=False)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tr.append(t2 - t1)
        etas.append(eta)
        its.append(it)
    total = time.perf_counter() - t_all
    return {"factor_s": float(np.mean(tf)), "refine_s": float(np.mean(tr)), "total_s": total,
            "eta": float(max(etas)), "iters": int(max(its)), "x": x}


def level_report(S):
    """Fronts per level of the front tree (level = height above the leaves,
    the schedule's level): [level, fronts, max front rows, median front rows]."""
    nodes = S.nodes()
    lev = np.zeros(len(nodes), np.int64)
    for f, nd in enumerate(nodes):          # postorder: children first
        for k in nd[4]:
            lev[f] = max(lev[f], lev[k] + 1)
    rows = np.array([nd[2] for nd in nodes], np.int64)
    out = []
    for L in range(int(lev.max()) + 1):
        r = rows[lev == L]
        out.append([L, int(r.size), int(r.max()), int(np.median(r))])
    return out


def kernel_report(prof, steps, prec):
    """Extend-add HBM GB/s (algorithmic bytes / HIP-event time) and the MFMA
    utilisation of the large Schur-complement launches (>= 1024 tiles)."""
    ea = prof["extend_add"]
    g = steps[steps[:, 0] == 5]
    big = g[g[:, 2] >= 1024]
    util = None
    if len(big):
        util = float(big[:, 3].sum() / (big[:, 1].sum() * 1e-3) / 1e12 / PEAK_TFLOPS[prec])
    return {"extend_add_gbs": ea["bytes"] / max(ea["ms"], 1e-9) / 1e6,
            "extend_add_frac_hbm": ea["bytes"] / max(ea["ms"], 1e-9) / 1e6 / PEAK_HBM_GBS,
            "gemm_large_launches": int(len(big)), "gemm_large_mfma_util": util}


def permuted_lower(p, i, x, perm):
    """Lower triangle of P A P^T (new k = old perm[k]), sorted scipy CSC."""
    import scipy.sparse as sp
    from mdgpu_amd import gen
    A = gen.full_from_lower(p, i, x)
    B = sp.tril(A[perm][:, perm]).tocsc()
    B.sort_indices()
    return B


def cpu_sample(S, p, i, x, budget_flops):
    """Time the CPU oracle (single thread) on complete subtrees of the front
    tree, about budget_flops of factorization work: the largest subtrees
    that fit, so that the sample holds fronts as large as the budget allows
    (the whole tree when it fits)."""
    from mdgpu_amd import _lib as L
    from oracle import oracle as O
    nodes = S.nodes()
    count = S.array(L.ARR_COUNT_FINAL).astype(np.int64)
    nn = len(nodes)
    parent = np.full(nn, -1)
    for f, nd in enumerate(nodes):
        for k in nd[4]:
            parent[k] = f
    sub = np.zeros(nn)
    for f, nd in enumerate(nodes):        # postorder: children first
        c0, c = nd[0], nd[1]
        sub[f] += float(np.sum(count[c0:c0 + c] ** 2))
        if parent[f] >= 0:
            sub[parent[f]] += sub[f]
    # disjoint subtrees, largest first, while they fit in the budget
    inside = np.zeros(nn, bool)             # in a chosen subtree
    chosen, total = [], 0.0
    for f in sorted(range(nn), key=lambda t: -sub[t]):
        if sub[f] <= 0 or sub[f] > budget_flops - total:
            continue
        a, up = f, False
        while a >= 0:                       # skip subtrees of chosen ones
            if inside[a]:
                up = True
                break
            a = parent[a]
        if up:
            continue
        chosen.append(f)
        total += sub[f]
        stack = [f]
        while stack:
            t = stack.pop()
            inside[t] = True
            stack.extend(nodes[t][4])
    keep_idx = np.nonzero(inside)[0]
    remap = {int(f): t for t, f in enumerate(keep_idx)}
    first = [nodes[f][0] for f in keep_idx]
    ncols = [nodes[f][1] for f in keep_idx]
    kids = [[remap[k] for k in nodes[f][4]] for f in keep_idx]
    lens = np.zeros(S.n, np.int64)
    for f in keep_idx:
        c0, c = nodes[f][0], nodes[f][1]
        lens[c0:c0 + c] = count[c0:c0 + c]
    lp = np.zeros(S.n + 1, np.int64)
    lp[1:] = np.cumsum(lens)
    B = permuted_lower(p, i, x, S.array(L.ARR_PERM))
    t0 = time.perf_counter()
    _, _, flops_ref, bad = O.factor(B.indptr, B.indices, B.data, first, ncols, kids, lp.astype(np.int32))
    el = time.perf_counter() - t0
    alg = float(np.sum(lens.astype(np.float64) ** 2))
    return {"value": alg / el / 1e9, "seconds": el, "flops": alg, "fronts": int(len(keep_idx)),
            "subtrees": len(chosen)}


def bench_batch(args):
    """C4: 512 uneven independent matrices (n log-uniform in [500, 20000];
    half 2D/3D Laplacian stencils with a random diagonal shift, half
    B B^T + n I) factored in FP32 and refined in FP64 as ONE batch: one
    analysis (a forest), one level-batched launch sequence replayed from a
    hipGraph.  Every matrix finishes when the batch does, so its latency is
    the batch's time to solution."""
    import torch
    from mdgpu_amd import gen, symbolic
    from mdgpu_amd.numeric import Factorization
    t0 = time.perf_counter()
    mats = gen.batch_matrices(512, seed=0)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    S = symbolic.analyze_batch([(p, i) for p, i, _ in mats], True, args.mode, args.max_zeros, nthreads=0)
    t_an = time.perf_counter() - t0
    log(f"batch: {len(mats)} matrices, n={S.n} fronts={S.nnodes} lnz={S.lnz} flops={S.flops:.3e} "
        f"gen {t_gen:.1f}s analyze {t_an:.1f}s")
    ax = torch.from_numpy(np.concatenate([x for _, _, x in mats])).cuda()
    bs = [gen.rhs_from_xtrue(p, i, x, seed=k)[1] for k, (p, i, x) in enumerate(mats)]
    bd = torch.from_numpy(np.concatenate(bs)).cuda()
    F = Factorization(S, "fp32")
    tuned = F.autotune(ax) if not args.no_autotune else None
    r = time_path(F, ax, bd, args.steps, args.warmup, args.tol)
    etas = F.segment_eta()
    tts = r["factor_s"] + r["refine_s"]
    ns = [len(p) - 1 for p, _, _ in mats]
    F.close()
    del F
    torch.cuda.empty_cache()
    # per-matrix latency when a matrix is factored and refined alone (its own
    # analysis, handle and graph): a sample of 32 matrices spread over the
    # sizes; in the batch every matrix finishes with the batch
    order = np.argsort(ns)
    sample = sorted(set(int(order[k]) for k in np.linspace(0, len(mats) - 1, 32).round().astype(int)))
    solo = []
    for k in sample:
        p_, i_, x_ = mats[k]
        Sk = symbolic.analyze(p_, i_, True, args.mode, args.max_zeros)
        Fk = Factorization(Sk, "fp32")
        axk = torch.from_numpy(x_).cuda()
        bk = torch.from_numpy(bs[k]).cuda()
        rk = time_path(Fk, axk, bk, 3, 1, args.tol)
        solo.append((rk["factor_s"] + rk["refine_s"]) * 1e3)
        Fk.close()
    solo = np.array(solo)
    out = {
        "metric": "batched factor+refine throughput (matrices/s) at 1e-12 backward error, FP32+IR",
        "value": len(mats) / tts, "unit": "matrices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tts * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "512 matrices, n log-uniform [500, 20000], half stencils with shift U(1e-3,1e-1), "
                               "half B B^T + nI (5 N(0,1) per column)", "n_total": S.n, "fronts": S.nnodes,
                   "lnz": S.lnz, "flops_counts": S.flops, "n_min": min(ns), "n_max": max(ns),
                   "partition": f"{args.mode} {args.max_zeros}"},
        "factor_ms": r["factor_s"] * 1e3, "refine_ms": r["refine_s"] * 1e3, "ir_iters": r["iters"],
        "eta_max": float(etas.max()), "eta_p50": float(np.median(etas)),
        "batch_latency_ms": tts * 1e3,
        "solo_latency_ms_p50": float(np.percentile(solo, 50)), "solo_latency_ms_p99": float(np.percentile(solo, 99)),
        "solo_latency_sample": f"{len(sample)} matrices spread over the sizes (n {min(ns[k] for k in sample)}-"
                               f"{max(ns[k] for k in sample)}), each factored + refined alone: own analysis, "
                               "handle and graph; in the batch every matrix's latency is batch_latency_ms",
        "factor_gflops": S.flops / r["factor_s"] / 1e9, "analyze_s": t_an, "autotune_64_tile_launches": tuned,
    }
    print(json.dumps(out), flush=True)


def bench_reverse(args):
    """C5: Gaussian-Markov-field precision Q(theta) = tau (kappa^2 I + L) on a
    256^2 grid (L the 5-point Dirichlet Laplacian), theta = (log tau, log
    kappa) = (0, -1), b ~ N(0, 1).  One step = forward (FP64 factorization,
    log det Q, x = Q^{-1} b) + backward (selected inverse: d log det / dQ on
    the pattern; solve adjoint with xbar = 1: bbar and -sym(bbar x^T)), with
    d/d theta by the chain rule (dQ/dlog tau = Q, dQ/dlog kappa = 2 tau
    kappa^2 I).  Checked against the closed forms d log det / d log tau = n
    and d log det / d log kappa = sum 2 kappa^2 / (kappa^2 + lambda_ij)."""
    import torch
    from mdgpu_amd import gen, symbolic
    from mdgpu_amd.numeric import Factorization
    k, log_tau, log_kappa = 256, 0.0, -1.0
    p, i, x = gen.gmrf_precision(k, log_tau, log_kappa)
    S = symbolic.analyze(p, i, True, args.mode, args.max_zeros)
    n = S.n
    vals = torch.from_numpy(x).cuda()
    rng = np.random.default_rng(0)
    b = torch.from_numpy(rng.standard_normal(n)).cuda()
    xbar = torch.ones(n, dtype=torch.float64, device="cuda")
    cols = np.repeat(np.arange(n), np.diff(np.asarray(p)))
    diag = torch.from_numpy((np.asarray(i) == cols).astype(np.float64)).cuda()   # diagonal entries
    F = Factorization(S, "fp64")
    tuned = F.autotune(vals) if not args.no_autotune else None

    def step():
        F.factorize(vals)
        ld = F.logdet()
        xs = F.solve(b)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g = F.logdet_grad()
        bbar, ga = F.solve_adjoint(xs, xbar)
        torch.cuda.synchronize()
        return ld, xs, g, bbar, ga, t1

    for _ in range(args.warmup):
        step()
    tf, tb = [], []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ld, xs, g, bbar, ga, t1 = step()
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
    tau, kappa = np.exp(log_tau), np.exp(log_kappa)
    # d/dlog tau: dQ = Q (all entries); d/dlog kappa: dQ = 2 tau kappa^2 on the diagonal
    dld_dlogtau = float((g * vals).sum())
    dld_dlogkappa = float((g * diag).sum()) * 2.0 * tau * kappa * kappa
    ii = np.arange(1, k + 1)
    lam = (4.0 - 2.0 * np.cos(ii * np.pi / (k + 1)))[:, None] - 2.0 * np.cos(ii * np.pi / (k + 1))[None, :]
    ld_ref = float(np.sum(np.log(tau * (kappa * kappa + lam))))
    dk_ref = float(np.sum(2.0 * kappa * kappa / (kappa * kappa + lam)))
    out = {
        "metric": "reverse-mode step ms: FP64 factor + logdet + solve, then selected-inverse gradient + solve adjoint",
        "value": (np.mean(tf) + np.mean(tb)) * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": (np.mean(tf) + np.mean(tb)) * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
        "config": {"workload": "GMRF precision tau (kappa^2 I + L), 5-point Laplacian 256^2, "
                               "theta = (log tau, log kappa) = (0, -1), b ~ N(0,1), xbar = 1",
                   "n": n, "fronts": S.nnodes, "lnz": S.lnz, "partition": f"{args.mode} {args.max_zeros}"},
        "forward_ms": np.mean(tf) * 1e3, "backward_ms": np.mean(tb) * 1e3,
        "logdet": ld, "logdet_rel_err_closed_form": abs(ld - ld_ref) / abs(ld_ref),
        "dlogdet_dlogtau": dld_dlogtau, "dlogdet_dlogtau_rel_err": abs(dld_dlogtau - n) / n,
        "dlogdet_dlogkappa": dld_dlogkappa, "dlogdet_dlogkappa_rel_err": abs(dld_dlogkappa - dk_ref) / abs(dk_ref),
        "autotune_64_tile_launches": tuned,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c4", "c5"])
    ap.add_argument("--mode", default="rn")
    ap.add_argument("--max-zeros", type=int, default=1000)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--cpu-budget-gflop", type=float, default=75.0,
                    help="flops of the CPU baseline sample (about 11-15 s of one core at 5-7 GFLOP/s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-aniso", action="store_true", help="skip the z-anisotropic C3 variant")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default GEMM tile sizes")
    ap.add_argument("--budgets", default="250,1000,2000,4000,8000,16000,64000",
                    help="relaxed-amalgamation zero budgets the partition tuner tries per precision "
                         "(empty: --mode/--max-zeros for every precision)")
    ap.add_argument("--only", default="fp32,fp64,bf16x3,bf16",
                    help="comma-separated precisions to time (fp32 and fp64 are needed for the JSON line)")
    ap.add_argument("--steps-json", default="",
                    help="write per-launch FP32 timings (class, ms, wgs, flops, bytes, level, stream, GEMM kernel)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the path shards by subtree across GPUs by design, but only the single-GPU "
              "kernel path is built and measured)", file=sys.stderr)
        sys.exit(2)

    import torch
    from mdgpu_amd import gen, symbolic
    from mdgpu_amd.numeric import Factorization

    if args.workload == "c4":
        return bench_batch(args)
    if args.workload == "c5":
        return bench_reverse(args)
    desc, make = WORKLOADS[args.workload]
    p, i, x = make(gen)
    xt, b = gen.rhs_from_xtrue(p, i, x, seed=0)
    ax = torch.from_numpy(x).cuda()
    bd = torch.from_numpy(b).cuda()
    torch.cuda.synchronize()
    precs = [q for q in ("fp32", "fp64", "bf16x3", "bf16") if q in args.only.split(",")]
    # the analysis (and, with --budgets, the partition tuner: the relaxed
    # zero budget with the fastest factorization + solves, per precision)
    # is setup that depends on the pattern only: untimed, like the analysis
    budgets = [int(v) for v in args.budgets.split(",") if v.strip()]
    t0 = time.perf_counter()
    tuning = {}
    if budgets:
        cands = symbolic.analyze_budgets(p, i, budgets)
        t_an = time.perf_counter() - t0
        from mdgpu_amd.numeric import tune_partition
        sym_for = {}
        for prec in precs:
            # solves per time to solution: FP64 one, the refined paths ~3 (bf16: ~7)
            nsolve = {"fp64": 1, "bf16": 7}.get(prec, 3)
            best, ms = tune_partition(cands, ax, prec, reps=2, solves=nsolve)
            sym_for[prec] = cands[best]
            tuning[prec] = {"budget": budgets[best], "ms_factor_plus_solves": dict(zip(map(str, budgets), ms))}
            log(f"partition tuner {prec}: rn {budgets[best]}  " +
                " ".join(f"{bb}:{m:.2f}ms" for bb, m in zip(budgets, ms)))
    else:
        S0 = symbolic.analyze(p, i, True, args.mode, args.max_zeros)
        t_an = time.perf_counter() - t0
        sym_for = {prec: S0 for prec in precs}
    S = sym_for.get("fp32", next(iter(sym_for.values())))
    log(f"analyze: n={S.n} fronts={S.nnodes} lnz={S.lnz} max_front={S.max_front} "
        f"flops={S.flops:.3e} ({t_an:.1f}s)")

    res = {}
    for prec in precs:
        F = Factorization(sym_for[prec], prec)
        # per-launch GEMM tile sizes from timed factorizations (untimed setup,
        # like the analysis: it depends on the pattern only)
        tuned = F.autotune(ax) if not args.no_autotune else None
        r = time_path(F, ax, bd, args.steps, args.warmup, args.tol, maxit=50 if prec != "bf16" else 30)
        r["autotune_64_tile_launches"] = tuned
        xs = r.pop("x").cpu().numpy()
        r["rel_err_xtrue"] = float(np.abs(xs - xt).max() / np.abs(xt).max())
        r["x"] = xs
        if prec == "fp32":
            # same factorization with the values handed over in host memory (the
            # C++ drop-in's case): includes the PCIe upload of A's values
            ax_host = torch.from_numpy(x)
            tph = []
            for _ in range(max(args.steps, 1)):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                F.factorize(ax_host)
                tph.append(time.perf_counter() - t0)
            r["factor_s_host_values"] = float(np.mean(tph))
        prof = F.factorize_profile(ax)
        r["profile"] = prof
        if prec in ("fp32", "fp64"):
            steps = F.factorize_steps(ax)
            r["report"] = kernel_report(prof, steps, prec)
            if prec == "fp32" and args.steps_json:
                with open(args.steps_json, "w") as fh:
                    json.dump(steps.tolist(), fh)
        r["info"] = F.info()
        res[prec] = r
        log(f"{prec}: factor {r['factor_s']*1e3:.2f} ms  refine {r['refine_s']*1e3:.2f} ms "
            f"(it={r['iters']}, eta={r['eta']:.2e})  "
            + " ".join(f"{k}={v['ms']:.1f}ms" for k, v in prof.items()))
        F.close()
        del F
        torch.cuda.empty_cache()

    if "fp32" not in res or "fp64" not in res:
        log("timed " + ",".join(res) + " only (the JSON line needs fp32 and fp64)")
        return
    aniso = None
    if args.workload == "c3" and not args.no_aniso:
        # z-edges scaled by 1e4: same pattern, so the same analysis
        pa, ia, xa = gen.stencil27(64, seed=0, anisotropic=True)
        assert np.array_equal(pa, p) and np.array_equal(ia, i)
        xta, ba = gen.rhs_from_xtrue(pa, ia, xa, seed=0)
        axa = torch.from_numpy(xa).cuda()
        bda = torch.from_numpy(ba).cuda()
        aniso = {}
        for prec in ("fp32", "fp64"):
            F = Factorization(sym_for[prec], prec)
            if not args.no_autotune:
                F.autotune(axa)
            r = time_path(F, axa, bda, args.steps, args.warmup, args.tol)
            xs = r.pop("x").cpu().numpy()
            aniso[prec] = {"factor_ms": r["factor_s"] * 1e3, "refine_ms": r["refine_s"] * 1e3,
                           "time_to_solve_ms": (r["factor_s"] + r["refine_s"]) * 1e3, "ir_iters": r["iters"],
                           "eta": r["eta"], "converged": bool(r["eta"] <= args.tol),
                           "rel_err_xtrue": float(np.abs(xs - xta).max() / np.abs(xta).max()),
                           "fp64_fallbacks": F.info().get("fp64_fallbacks", 0)}
            log(f"anisotropic {prec}: factor {r['factor_s']*1e3:.2f} ms refine {r['refine_s']*1e3:.2f} ms "
                f"(it={r['iters']}, eta={r['eta']:.2e})")
            F.close()
            del F
            torch.cuda.empty_cache()
        aniso["tts_ratio_fp32ir_over_fp64"] = aniso["fp32"]["time_to_solve_ms"] / aniso["fp64"]["time_to_solve_ms"]
    from oracle import oracle as O
    eta32_cpu = O.backward_error(p, i, x, res["fp32"]["x"], b)
    eta64_cpu = O.backward_error(p, i, x, res["fp64"]["x"], b)
    x32_vs_64 = float(np.abs(res["fp32"]["x"] - res["fp64"]["x"]).max() / np.abs(res["fp64"]["x"]).max())

    r32, r64 = res["fp32"], res["fp64"]
    tts32 = r32["factor_s"] + r32["refine_s"]
    tts64 = r64["factor_s"] + r64["refine_s"]
    value = S.flops / r32["factor_s"] / 1e9

    # dominant kernel class of the timed FP32 path: live HIP-event timing
    prof = r32["profile"]
    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    if dom in ("gemm", "potrf_trsm", "inverse", "panel"):
        ach = d["flops"] / (d["ms"] * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": dom, "achieved": ach, "peak": PEAK_TFLOPS["fp32"],
                "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS["fp32"], "traffic": None,
                "launches": d["launches"], "avg_launch_ms": d["ms"] / max(d["launches"], 1)}
    else:
        ach = d["bytes"] / (d["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": None,
                "launches": d["launches"], "avg_launch_ms": d["ms"] / max(d["launches"], 1)}

    # HBM traffic per launch of the dominant kernel, measured by two
    # rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) scaled on a copy of
    # known size; the committed summary holds the figure
    # (the summary holds one figure per FP32 partition it was measured at)
    kname = {"gemm": "fp32_gemm"}.get(dom)
    tpath = os.path.join(ROOT, "profiles", "gemm_blk_extendadd_traffic_c3_fp32.md")
    part32 = tuning["fp32"]["budget"] if "fp32" in tuning else args.max_zeros
    if kname and args.workload == "c3" and os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith(f"traffic_bytes_per_launch {kname} rn{part32} ="):
                roof["traffic"] = float(line.split("=")[1])
                roof["traffic_unit"] = (f"bytes per launch (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE, copy-calibrated, "
                                        f"partition rn {part32})")
                roof["algorithmic_bytes_per_launch"] = d["bytes"] / max(d["launches"], 1)

    cpu = None
    if not args.no_cpu:
        cs = cpu_sample(S, p, i, x, args.cpu_budget_gflop * 1e9)
        cpu = {"value": cs["value"], "unit": "GFLOP/s", "cores": 1, "kind": "port",
               "sample": f"CPU oracle (FP64 restatement of the reference CPU path) on the {cs['subtrees']} "
                         f"largest complete subtrees of the same front tree that fit in "
                         f"{args.cpu_budget_gflop:g} GFLOP ({cs['fronts']} fronts, {cs['flops']:.3e} flops), "
                         f"{cs['seconds']:.1f} s, 1 thread"}
        log(f"cpu sample: {cpu['value']:.2f} GFLOP/s over {cs['seconds']:.1f}s")

    out = {
        "metric": "factor GFLOP/s + time-to-solve at 1e-12 backward err: FP32+IR vs FP64",
        "value": value,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tts32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": desc, "n": S.n, "nnz_lower": int(len(i)), "ordering": "amd",
                   "partition": (f"rn {tuning['fp32']['budget']} (tuned; fp64: rn {tuning['fp64']['budget']})"
                                 if tuning else f"{args.mode} {args.max_zeros}"),
                   "fronts": S.nnodes, "lnz": S.lnz,
                   "max_front": S.max_front, "flops_counts": S.flops, "flops_padded": S.flops_padded},
        "autotune_64_tile_launches": {"fp32": r32["autotune_64_tile_launches"],
                                      "fp64": r64["autotune_64_tile_launches"]},
        "fp32_ir": {"factor_ms": r32["factor_s"] * 1e3, "refine_ms": r32["refine_s"] * 1e3,
                    "time_to_solve_ms": tts32 * 1e3, "ir_iters": r32["iters"], "eta": r32["eta"],
                    "factor_ms_host_values": r32["factor_s_host_values"] * 1e3,
                    "eta_cpu_check": eta32_cpu, "rel_err_xtrue": r32["rel_err_xtrue"],
                    "rel_diff_vs_fp64_solution": x32_vs_64},
        "fp64": {"factor_ms": r64["factor_s"] * 1e3, "refine_ms": r64["refine_s"] * 1e3,
                 "time_to_solve_ms": tts64 * 1e3, "ir_iters": r64["iters"], "eta": r64["eta"],
                 "eta_cpu_check": eta64_cpu, "rel_err_xtrue": r64["rel_err_xtrue"],
                 "gflops": S.flops / r64["factor_s"] / 1e9},
        "tts_ratio_fp32ir_over_fp64": tts32 / tts64,
        "bf16_variants": {
            prec: {"factor_ms": res[prec]["factor_s"] * 1e3, "refine_ms": res[prec]["refine_s"] * 1e3,
                   "time_to_solve_ms": (res[prec]["factor_s"] + res[prec]["refine_s"]) * 1e3,
                   "ir_iters": res[prec]["iters"], "eta": res[prec]["eta"],
                   "converged": bool(res[prec]["eta"] <= args.tol),
                   "eta_cpu_check": O.backward_error(p, i, x, res[prec]["x"], b),
                   "tts_ratio_over_fp64": (res[prec]["factor_s"] + res[prec]["refine_s"]) / tts64,
                   "gemm": "split bf16 (hi*hi+hi*lo+lo*hi)" if prec == "bf16x3" else "single bf16 product"}
            for prec in ("bf16x3", "bf16") if prec in res},
        "anisotropic_variant": aniso,
        "kernel_classes_fp32": {k: {"ms": v["ms"], "launches": v["launches"]} for k, v in prof.items()},
        "kernel_report": {"fp32": r32["report"], "fp64": r64["report"]},
        "levels": level_report(S),
        "partition_tuning": tuning or None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "analyze_s": t_an,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
