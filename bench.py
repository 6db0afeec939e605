#!/usr/bin/env python3
"""Benchmark of the MI355X block-sparse hot path.

Default workload (one GPU): the D=1024 block-sparse contraction + block SVD
step (quantit_amd/workloads.py, contract_svd): X(b,a',s,w') . Abar(a',s,b')
over (a', s), then the truncating block SVD (cutoff 1e-10) of the rank-3
result with split 1.  Inputs are resident in HBM before the timed region.

value = (host-counted contraction FLOPs sum 2MNK + block-SVD model FLOPs
sum 14 m n^2 + 8 n^3) / step time, in TFLOP/s.  The per-kernel times of the
timed steps come from HIP events recorded on the context stream around every
launch; the roofline object describes the kernel with the largest share.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6    # MI355X dense FP64 MFMA
PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 MFMA
PEAK_HBM_GBS = 8000.0


def svd_model_flops(parts, split):
    """Sum over (row charge, column charge) sectors of 14 m n^2 + 8 n^3, m >= n,
    from the block structure (the sector grouping of compact_dense)."""
    secs, cvals, blocks = parts["secs"], parts["cvals"], parts["blocks"]
    rank = len(secs)
    sectors = {}
    for idx in blocks:
        rc = sum(cvals[d][idx[d]][0] for d in range(split))
        cc = sum(cvals[d][idx[d]][0] for d in range(split, rank))
        rows = int(np.prod([secs[d][idx[d]] for d in range(split)]))
        cols = int(np.prod([secs[d][idx[d]] for d in range(split, rank)]))
        s = sectors.setdefault((rc, cc), (dict(), dict()))
        s[0][idx[:split]] = rows
        s[1][idx[split:]] = cols
    total = 0.0
    for r, c in sectors.values():
        m, n = sum(r.values()), sum(c.values())
        m, n = max(m, n), min(m, n)
        total += 14.0 * m * n * n + 8.0 * n ** 3
    return total


def cpu_baseline(X, A, dx, da, split, budget_s=20.0):
    """The oracle (numpy restatement of the reference loop order) timed on the
    host on whole steps of the same workload, as many as fit the budget."""
    from oracle.btensor_oracle import BTensor, svd_truncated, tensordot
    try:
        from threadpoolctl import threadpool_info
        threads = max([p.get("num_threads", 1) for p in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    Xo = BTensor(X["groups"], X["secs"], X["cvals"], X["sel"], X["blocks"])
    Ao = BTensor(A["groups"], A["secs"], A["cvals"], A["sel"], A["blocks"])
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        Y = tensordot(Xo, Ao, dx, da)
        svd_truncated(Y, split, 1e-10)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 20:
            break
    return float(np.median(times)), len(times), threads


def copy_probe(mib):
    """One plain device-to-device copy of a known size: the scale for the
    FETCH_SIZE / WRITE_SIZE counters (expected mib*1024 KiB each)."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes": n * 8}))


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (lines 'traffic <kernel> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_contract_svd_D1024.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


# Whole-solve workloads: (L, bonds, max bond, sweeps, seed, description)
DMRG_WORKLOADS = {
    "chain_L32_D128": dict(L=32, geometry="chain", width=0, max_bond=128, sweeps=6, seed=0,
                           desc="Heisenberg chain J=1 OBC L=32, U(1) Sz=0, two-site DMRG, D<=128, cutoff 1e-10, "
                                "6 sweeps, random bMPS seed 0"),
    "cylinder4x16_D1024": dict(L=64, geometry="cylinder", width=4, max_bond=1024, sweeps=6, seed=1,
                               desc="Heisenberg 4x16 cylinder (periodic width 4, open length 16, snake order) J=1, "
                                    "U(1) Sz=0, two-site DMRG to D=1024, cutoff 1e-10, random bMPS seed 1"),
}


def _dmrg_inputs(qa, ctx, w):
    from quantit_amd import models
    L = w["L"]
    bonds = models.chain_bonds(L) if w["geometry"] == "chain" else models.cylinder_bonds(w["width"], L // w["width"])
    mpo_parts = models.bmpo_parts(models.spin_mpo(L, bonds))
    mps_parts = models.random_bmps_parts(L, 4, q=0, seed=w["seed"])
    return mpo_parts, mps_parts


def bench_dmrg(args):
    """One DMRG solve of a whole-solve workload: seconds per sweep (timed by
    the library around each sweep), energies, FLOPs from the plans."""
    import quantit_amd as qa
    w = DMRG_WORKLOADS[args.workload]
    sweeps = args.sweeps or w["sweeps"]
    ctx = qa.Context(0, arena_bytes=64 << 30,
                     precision=qa.PRECISION_SPLIT_FP32 if args.precision == "split" else qa.PRECISION_FP64)
    ctx.set_per_block(args.per_block)
    mpo_parts, mps_parts = _dmrg_inputs(qa, ctx, w)
    opts = qa.DmrgOptions(cutoff=1e-10, convergence_criterion=0.0, maximum_bond=w["max_bond"], minimum_bond=4,
                          maximum_iterations=sweeps)

    def solve(profile):
        mpo = [qa.BTensor.from_parts(ctx, **p) for p in mpo_parts]
        mps = qa.mps_prepare([qa.BTensor.from_parts(ctx, **p) for p in mps_parts])
        ctx.sync()
        ctx.kernel_times(reset=True)
        ctx.set_profiling(profile)
        t0 = time.perf_counter()
        E, mps, log = qa.dmrg(mpo, mps, opts)
        ctx.sync()
        wall = time.perf_counter() - t0
        ctx.set_profiling(False)
        return E, log, wall, ctx.kernel_times(reset=True)

    # warmup: a short solve of the same workload (first launches, arena growth)
    if args.warmup:
        wopts = opts
        opts = qa.DmrgOptions(cutoff=1e-10, convergence_criterion=0.0, maximum_bond=min(w["max_bond"], 32),
                              minimum_bond=4, maximum_iterations=1)
        solve(False)
        opts = wopts
    E, log, wall, _ = solve(False)
    kt = {}
    if w["max_bond"] <= 256:   # per-kernel breakdown from a second, profiled solve
        _, _, _, kt = solve(True)
    ms = log["sweep_ms"]
    flops = log["contract_flops"] + log["svd_flops"]
    out = {
        "metric": "DMRG sweep s & contraction TFLOP/s @|dE/E|<=1e-10; served solves/s + p99 latency",
        "value": float(np.mean(ms)) / 1e3,
        "unit": "s/sweep",
        "n_gpus": 1,
        "steps": len(ms),
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(ms)),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (split-fp32 contraction)",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w['desc']}", "precision": args.precision,
                   "per_block_launch": args.per_block},
        "energy": E,
        "sweep_energies": log["energies"],
        "sweep_ms": ms,
        "max_bond": log["max_bond"],
        "solve_wall_s": wall,
        "solve_tflops": flops / wall / 1e12,
        "contract_gflop": log["contract_flops"] / 1e9,
        "svd_model_gflop": log["svd_flops"] / 1e9,
        "kernel_ms_per_solve": {k: v[0] for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:12]},
        "roofline": None,
        "cpu_baseline": None,
    }
    if kt:
        # host_* entries time host phases (sector matching, planning), not kernels
        dom, (d_ms, d_n) = max(((k, v) for k, v in kt.items() if not k.startswith("host_")), key=lambda kv: kv[1][0])
        if dom.startswith("gemm"):
            per = log["contract_flops"] / max(d_n, 1)
        else:
            per = log["svd_flops"] / max(d_n, 1)
        ach = per / ((d_ms / max(d_n, 1)) * 1e-3) / 1e12
        peak = PEAK_FP32_TFLOPS if dom == "gemm_split" else PEAK_FP64_TFLOPS
        out["roofline"] = {"kernel": dom, "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                           "frac": ach / peak, "traffic": None,
                           "note": "model FLOPs of the solve attributed to the kernel's launches"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = dmrg_cpu_baseline(mpo_parts, mps_parts, opts, args.cpu_budget, E, log)
    print(json.dumps(out))


def dmrg_cpu_baseline(mpo_parts, mps_parts, opts, budget, E_dev, log_dev):
    """The oracle's DMRG (numpy restatement of sources/dmrg.cpp) on the same
    MPO and starting state, sweep by sweep until the budget is spent."""
    from oracle import dmrg_oracle as dor
    from oracle.btensor_oracle import BTensor as OB
    try:
        from threadpoolctl import threadpool_info
        threads = max([p.get("num_threads", 1) for p in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    mpo = [OB(**p) for p in mpo_parts]
    mps = dor.prepare_state([OB(**p) for p in mps_parts])
    times, energies = [], []
    t_start = time.perf_counter()
    last = [t_start]

    class Budget(Exception):
        pass

    def on_sweep(it, E, m):
        now = time.perf_counter()
        times.append(now - last[0])
        energies.append(E)
        last[0] = now
        if now - t_start > budget:
            raise Budget()

    o = dor.DmrgOptions(cutoff=opts.cutoff, convergence_criterion=opts.convergence_criterion,
                        maximum_bond=opts.maximum_bond, minimum_bond=opts.minimum_bond,
                        maximum_iterations=opts.maximum_iterations)
    try:
        dor.dmrg(mpo, mps, o, log=on_sweep)
    except Budget:
        pass
    n = len(times)
    dev_e = log_dev["energies"][:n]
    dev = [abs(a - b) / abs(b) for a, b in zip(dev_e, energies)]
    return {"value": float(np.mean(times)), "unit": "s/sweep", "cores": threads, "kind": "port",
            "sample": f"first {n} sweeps of the same solve (numpy oracle, same MPO and start state)",
            "sweep_s": times, "energies": energies, "max_rel_dE_vs_device": max(dev) if dev else None}


def _loguniform_ints(rng, n, lo=4, hi=768):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n))).astype(int).clip(lo, hi)


def grouped_inputs(seed=2, n_gemm=2048, n_svd=512):
    """Grouped batch of independent blocks as block-diagonal btensors: item i
    gets its own U(1) charge i, so A(m_i x k_i) . B(k_i x n_i) are the only
    matching block pairs of one tensordot, and S(m_i x n_i) are the sectors
    of one block SVD.  Dimensions log-uniform integers in [4, 768]."""
    rng = np.random.default_rng(seed)
    m, n, k = (_loguniform_ints(rng, n_gemm) for _ in range(3))
    sm, sn = (_loguniform_ints(rng, n_svd) for _ in range(2))

    def diag(rows, cols, fill):
        return dict(groups=(0,), secs=[list(map(int, rows)), list(map(int, cols))],
                    cvals=[[(i,) for i in range(len(rows))], [(-i,) for i in range(len(cols))]], sel=(0,),
                    blocks={(i, i): fill(int(r), int(c)) for i, (r, c) in enumerate(zip(rows, cols))})

    A = diag(m, k, lambda r, c: rng.standard_normal((r, c)))
    B = dict(groups=(0,), secs=[list(map(int, k)), list(map(int, n))],
             cvals=[[(i,) for i in range(n_gemm)], [(-i,) for i in range(n_gemm)]], sel=(0,),
             blocks={(i, i): rng.standard_normal((int(k[i]), int(n[i]))) for i in range(n_gemm)})
    S = diag(sm, sn, lambda r, c: rng.standard_normal((r, c)))
    return A, B, S, (m, n, k), (sm, sn)


def bench_grouped(args):
    """2048 independent GEMMs (one grouped launch) and 512 independent SVDs
    (one batched block SVD), against one launch per block pair."""
    import quantit_amd as qa
    A, B, S, (m, n, k), (sm, sn) = grouped_inputs()
    ctx = qa.Context(0, arena_bytes=16 << 30,
                     precision=qa.PRECISION_SPLIT_FP32 if args.precision == "split" else qa.PRECISION_FP64)
    a = qa.BTensor.from_parts(ctx, **A)
    b = qa.BTensor.from_parts(ctx, **B)
    s = qa.BTensor.from_parts(ctx, **S)
    f_gemm = float(np.sum(2.0 * m * n * k))
    q_gemm = 8.0 * (m * k + k * n + m * n)
    peak = (PEAK_FP32_TFLOPS if args.precision == "split" else PEAK_FP64_TFLOPS) * 1e12
    t_roof = float(np.sum(np.maximum(2.0 * m * n * k / peak, q_gemm / (PEAK_HBM_GBS * 1e9))))
    big, small = np.maximum(sm, sn).astype(float), np.minimum(sm, sn).astype(float)
    f_svd = float(np.sum(14.0 * big * small ** 2 + 8.0 * small ** 3))

    def timed(fn, steps):
        for _ in range(max(args.warmup, 1)):
            fn()
        ctx.sync()
        ctx.kernel_times(reset=True)
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.sync()
        dt = (time.perf_counter() - t0) / steps
        ctx.set_profiling(False)
        return dt, ctx.kernel_times(reset=True)

    steps = max(1, min(args.steps, 10))
    t_g, kt_g = timed(lambda: a.tensordot(b, [1], [0]), steps)
    ctx.set_per_block(True)
    t_pb, _ = timed(lambda: a.tensordot(b, [1], [0]), 1)
    ctx.set_per_block(False)
    t_s, kt_s = timed(lambda: qa.svd(s, 1), max(1, steps // 2))
    gname = "gemm_split" if args.precision == "split" else "gemm_f64"
    g_ms, g_n = kt_g.get(gname, (0.0, 1))
    g_avg = g_ms / max(g_n, 1) * 1e-3
    out = {
        "metric": "DMRG sweep s & contraction TFLOP/s @|dE/E|<=1e-10; served solves/s + p99 latency",
        "value": (f_gemm + f_svd) / (t_g + t_s) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": (t_g + t_s) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (split-fp32 contraction)",
        "data": "synthetic",
        "config": {"workload": "grouped_batch: 2048 independent GEMMs (m, n, k log-uniform in [4, 768], seed 2) as "
                               "one block-diagonal tensordot + 512 independent SVDs (m, n log-uniform in [4, 768]) "
                               "as one block SVD", "precision": args.precision},
        "gemm_batch": {"ms": t_g * 1e3, "tflops": f_gemm / t_g / 1e12, "gflop": f_gemm / 1e9,
                       "per_block_launch_ms": t_pb * 1e3, "speedup_vs_per_block": t_pb / t_g,
                       "roofline_additive_ms": t_roof * 1e3, "roofline_frac": t_roof / g_avg if g_avg else None},
        "svd_batch": {"ms": t_s * 1e3, "model_gflop": f_svd / 1e9, "model_tflops": f_svd / t_s / 1e12,
                      "kernel_ms": {kk: round(v[0] / max(1, steps // 2), 3)
                                    for kk, v in sorted(kt_s.items(), key=lambda kv: -kv[1][0])[:8]}},
        "roofline": {"kernel": gname, "bound": "mfma", "achieved": f_gemm / g_avg / 1e12 if g_avg else None,
                     "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": (f_gemm / g_avg) / peak if g_avg else None, "traffic": None,
                     "note": "grouped GEMM launch; additive per-item roofline in gemm_batch.roofline_frac"},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = grouped_cpu_baseline(A, B, S, args.cpu_budget)
    print(json.dumps(out))


def grouped_cpu_baseline(A, B, S, budget):
    """The oracle's tensordot and block SVD on leading slices of the batch,
    as many items as fit the budget; rate in TFLOP/s of the same model."""
    from oracle.btensor_oracle import BTensor as OB, svd_split, tensordot
    try:
        from threadpoolctl import threadpool_info
        threads = max([p.get("num_threads", 1) for p in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1

    def head(P, n):
        return dict(groups=P["groups"], secs=[P["secs"][0][:n], P["secs"][1][:n]],
                    cvals=[P["cvals"][0][:n], P["cvals"][1][:n]], sel=P["sel"],
                    blocks={kk: v for kk, v in P["blocks"].items() if kk[0] < n})

    flops, secs, n_g, n_s = 0.0, 0.0, 0, 0
    for n in (64, 256, 1024, 2048):
        a, b = OB(**head(A, n)), OB(**head(B, n))
        t0 = time.perf_counter()
        tensordot(a, b, [1], [0])
        dt = time.perf_counter() - t0
        m = np.array(A["secs"][0][:n]); k = np.array(A["secs"][1][:n]); nn = np.array(B["secs"][1][:n])
        flops, secs, n_g = float(np.sum(2.0 * m * nn * k)), dt, n
        if dt > budget / 4:
            break
    g_rate = flops / secs
    for n in (16, 64, 256, 512):
        s = OB(**head(S, n))
        t0 = time.perf_counter()
        svd_split(s, 1)
        dt = time.perf_counter() - t0
        r, c = np.array(S["secs"][0][:n], float), np.array(S["secs"][1][:n], float)
        big, small = np.maximum(r, c), np.minimum(r, c)
        sf, ss, n_s = float(np.sum(14 * big * small ** 2 + 8 * small ** 3)), dt, n
        if dt > budget / 4:
            break
    return {"value": (flops + sf) / (secs + ss) / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"oracle tensordot over the first {n_g} GEMM items and block SVD over the first {n_s} SVD items",
            "gemm_tflops": g_rate / 1e12, "svd_model_tflops": sf / ss / 1e12}


def bench_serving(args):
    """Seeded Poisson trace of XXZ ground-state requests served by solve
    slots with admission by HBM footprint: p50/p99 latency and solves/s."""
    from quantit_amd import serving
    if args.full_trace:
        L_range, D_choices = (16, 128), (64, 128, 256, 512)
    else:  # bounded default: a run of a few minutes
        L_range, D_choices = (16, 64), (64, 128, 256)
    trace = serving.make_trace(args.requests, rate=args.rate, seed=5, L_range=L_range, D_choices=D_choices)
    srv = serving.Server(n_slots=args.slots, hbm_budget=int(args.hbm_gib * (1 << 30)),
                         max_sweeps=args.sweeps or 20)

    def progress(r, x):
        msg = (f"request {r.rid}: L={r.L} D={r.D} delta={r.delta:.3f} "
               + (f"E={x.energy:.12f} sweeps={x.sweeps} latency={x.latency:.2f}s" if isinstance(x, serving.Result)
                  else f"error {x}"))
        print(msg, file=sys.stderr, flush=True)

    res, wall = srv.run(trace, progress=progress)
    summ = serving.summarize(res, wall)
    ok = [r for r in res if isinstance(r, serving.Result)]
    out = {
        "metric": "DMRG sweep s & contraction TFLOP/s @|dE/E|<=1e-10; served solves/s + p99 latency",
        "value": summ["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(trace),
        "warmup": 0,
        "ms_per_step": 1e3 * wall / max(1, len(ok)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"serving: {len(trace)} XXZ chain requests, Poisson rate {args.rate}/s (seed 5), "
                               f"L ~ U{{{L_range[0]}..{L_range[1]}}}, D ~ U{set(D_choices)}, delta ~ U[0.5,1.5], "
                               f"Sz=0, |dE/E| <= 1e-10 or {args.sweeps or 20} sweeps",
                   "slots": args.slots, "hbm_budget_gib": args.hbm_gib},
        "p50_latency_s": summ["p50_latency_s"],
        "p99_latency_s": summ["p99_latency_s"],
        "failed": summ["failed"],
        "mean_sweeps": summ["mean_sweeps"],
        "requests": [{"L": r.L, "D": r.D, "delta": round(r.delta, 3), "latency_s": round(x.latency, 3),
                      "solve_s": round(x.finish - x.start, 3), "sweeps": x.sweeps, "energy": x.energy}
                     if isinstance(x, serving.Result) else {"L": r.L, "D": r.D, "error": str(x)}
                     for r, x in zip(trace, res)],
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "split"], default="fp64")
    ap.add_argument("--per-block", action="store_true", help="one GEMM launch per block pair (baseline form)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-split", action="store_true", help="skip the split-FP32 side measurement")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--requests", type=int, default=8, help="serving: trace length")
    ap.add_argument("--rate", type=float, default=4.0, help="serving: Poisson arrival rate (1/s)")
    ap.add_argument("--slots", type=int, default=4, help="serving: concurrent solve slots")
    ap.add_argument("--hbm-gib", type=float, default=160.0, help="serving: HBM budget for admission")
    ap.add_argument("--full-trace", action="store_true",
                    help="serving: L up to 128 and D up to 512 (long run) instead of L <= 64, D <= 256")
    ap.add_argument("--workload", choices=["contract_svd_D1024", "grouped_batch", "serving"] + sorted(DMRG_WORKLOADS),
                    default="contract_svd_D1024")
    ap.add_argument("--sweeps", type=int, default=0, help="DMRG workloads: sweeps (0 = the workload's own)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload in DMRG_WORKLOADS:
        bench_dmrg(args)
        return
    if args.workload == "grouped_batch":
        bench_grouped(args)
        return
    if args.workload == "serving":
        bench_serving(args)
        return

    import quantit_amd as qa
    from quantit_amd import workloads

    X, A, dx, da, split = workloads.contract_svd(1024)
    ctx = qa.Context(0, arena_bytes=16 << 30,
                     precision=qa.PRECISION_SPLIT_FP32 if args.precision == "split" else qa.PRECISION_FP64)
    ctx.set_per_block(args.per_block)
    x = qa.BTensor.from_parts(ctx, **X)
    a = qa.BTensor.from_parts(ctx, **A)

    def step():
        y = x.tensordot(a, dx, da)
        return qa.svd(y, split, tol=1e-10)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    contract_flops = qa.last_tensordot_flops()
    y = x.tensordot(a, dx, da)
    svd_flops = svd_model_flops(dict(zip(["groups", "secs", "cvals", "sel", "blocks"], y.to_parts())), split)
    ctx.kernel_times(reset=True)
    ctx.set_profiling(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    ctx.set_profiling(False)
    kt = ctx.kernel_times(reset=True)
    ms = (t1 - t0) * 1e3 / args.steps
    flops = contract_flops + svd_flops
    value = flops / (ms * 1e-3) / 1e12

    gemm_name = "gemm_split" if args.precision == "split" else ("gemm_per_block" if args.per_block else "gemm_f64")
    kernels = {k: v for k, v in kt.items() if not k.startswith("host_")}  # host_*: host phases
    dominant = max(kernels.items(), key=lambda kv: kv[1][0])[0] if kernels else gemm_name
    g_ms, g_n = kt.get(gemm_name, (0.0, 0))
    per_launch_gemm = (contract_flops * args.steps / max(g_n, 1))
    gemm_avg_s = (g_ms / max(g_n, 1)) * 1e-3
    peak = PEAK_FP32_TFLOPS if args.precision == "split" else PEAK_FP64_TFLOPS
    if dominant.startswith("gemm"):
        achieved = per_launch_gemm / max(gemm_avg_s, 1e-12) / 1e12
        roof = {"kernel": gemm_name, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": read_traffic(gemm_name)}
    else:
        # block SVD: the timers svd_eig, svd_eig_it1, ... all time k_gram_eig (the
        # solve and the later Gram iterations); the model FLOPs of a step are
        # spread over all its launches, whose mean duration is the kernel's
        # rocprof average
        group = [k for k in kt if k.startswith("svd_eig")] if dominant.startswith("svd_eig") else [dominant]
        d_ms = sum(kt[k][0] for k in group)
        d_n = sum(kt[k][1] for k in group)
        per_launch = svd_flops * args.steps / max(d_n, 1)
        achieved = per_launch / max((d_ms / max(d_n, 1)) * 1e-3, 1e-12) / 1e12
        roof = {"kernel": dominant, "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": read_traffic(dominant),
                "avg_launch_us": d_ms / max(d_n, 1) * 1e3, "launches_per_step": d_n / args.steps,
                "note": "block-SVD model FLOPs (14mn^2+8n^3) of a step over its k_gram_eig launches "
                        "(" + ", ".join(sorted(group)) + "), per launch over the mean launch time"}

    out = {
        "metric": "DMRG sweep s & contraction TFLOP/s @|dE/E|<=1e-10; served solves/s + p99 latency",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (split-fp32 contraction)",
        "data": "synthetic",
        "config": {"workload": "contract_svd_D1024: X(1024,1024,2,26).Abar(1024,2,1024) over 2 legs + truncating "
                               "block SVD (split 1, cutoff 1e-10), U(1)",
                   "precision": args.precision, "per_block_launch": args.per_block},
        "contraction_gflop": contract_flops / 1e9,
        "svd_model_gflop": svd_flops / 1e9,
        "contraction_tflops": (contract_flops / ((g_ms / args.steps) * 1e-3) / 1e12) if g_ms else None,
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])},
        "kernel_launches_per_step": {k: v[1] / args.steps for k, v in kt.items()},
        "roofline": roof,
    }
    if not args.no_split and args.precision == "fp64" and not args.per_block:
        # side measurement: the same step with hi/lo split-FP32 MFMA contractions
        y64 = x.tensordot(a, dx, da).to_parts()[4]
        ctx.set_precision(qa.PRECISION_SPLIT_FP32)
        ysp = x.tensordot(a, dx, da).to_parts()[4]
        err = max(float(np.abs(ysp[k] - y64[k]).max()) for k in y64) / max(float(np.abs(v).max()) for v in y64.values())
        for _ in range(2):
            step()
        ctx.sync()
        ctx.kernel_times(reset=True)
        ctx.set_profiling(True)
        n_sp = max(args.steps // 2, 1)
        t0 = time.perf_counter()
        for _ in range(n_sp):
            step()
        ctx.sync()
        t_sp = (time.perf_counter() - t0) * 1e3 / n_sp
        ctx.set_profiling(False)
        kts = ctx.kernel_times(reset=True)
        ctx.set_precision(qa.PRECISION_FP64)
        gs_ms = kts.get("gemm_split", (0.0, 0))[0] / n_sp
        out["split_fp32"] = {"ms_per_step": t_sp, "contraction_ms": gs_ms,
                             "contraction_tflops": contract_flops / (gs_ms * 1e-3) / 1e12 if gs_ms else None,
                             "contraction_max_abs_err_rel_to_max": err}
    if not args.no_cpu_baseline:
        t_cpu, n_cpu, threads = cpu_baseline(X, A, dx, da, split, args.cpu_budget)
        out["cpu_baseline"] = {"value": flops / t_cpu / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
                               "sample": f"{n_cpu} whole steps of the same workload (numpy oracle), median "
                                         f"{t_cpu * 1e3:.1f} ms/step"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
