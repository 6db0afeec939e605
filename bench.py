#!/usr/bin/env python3
"""Benchmark of the MI355X block-sparse DMRG hot path.

Default (one GPU): the Heisenberg 4x16 cylinder (periodic width 4, open
length 16, snake-ordered MPS, U(1) Sz = 0, J = 1), two-site DMRG to D = 1024,
cutoff 1e-10, random Sz = 0 start from seed 1, 6 sweeps.  ``value`` is the
wall time of the last sweep (all bonds at D ~ 1024) of the FP64 batched
form; inputs are resident in HBM before the solve starts and the sweeps are
timed by the library around each sweep.  The same JSON line carries:

* the per-block FP64 form (one GEMM launch per block pair, as the
  reference's loop) and the batched/per-block ratio;
* the split-FP32 contraction form and its energy error against FP64;
* ``roofline`` for the kernel with the largest share of a profiled solve;
* ``cpu_baseline``: the reference's per-block libtorch calls
  (oracle/torch_cpu.py) timed on the host cores on centre two-site updates
  of the same state, extrapolated to the sweep by the plan's FLOPs;
* sub-objects for the L=32 D=128 chain, the D=1024 contraction + block SVD
  step and the grouped batch (2048 GEMMs + 512 SVDs).

Other workloads run alone with ``--workload``.  Prints one JSON line.
``--gpus`` above 1 is refused: this build runs on one GPU.
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

METRIC = "DMRG sweep s & contraction TFLOP/s @|dE/E|<=1e-10; served solves/s + p99 latency"
PEAK_FP64_TFLOPS = 78.6    # MI355X dense FP64 MFMA
PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 MFMA
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def threads_used():
    try:
        import torch
        return int(torch.get_num_threads())
    except Exception:
        return os.cpu_count() or 1


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


def read_traffic(summary, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 counter
    summary under profiles/ (lines 'traffic <kernel> <bytes>'), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", summary)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


# ---- whole-solve workloads --------------------------------------------------------
DMRG_WORKLOADS = {
    "chain_L32_D128": dict(L=32, geometry="chain", width=0, max_bond=128, sweeps=6, seed=0,
                           desc="Heisenberg chain J=1 OBC L=32, U(1) Sz=0, two-site DMRG, D<=128, cutoff 1e-10, "
                                "6 sweeps, random bMPS seed 0"),
    "cylinder4x16_D1024": dict(L=64, geometry="cylinder", width=4, max_bond=1024, sweeps=6, seed=1,
                               desc="Heisenberg 4x16 cylinder (periodic width 4, open length 16, snake order) J=1, "
                                    "U(1) Sz=0, two-site DMRG to D=1024, cutoff 1e-10, random bMPS seed 1, 6 sweeps"),
}


def dmrg_inputs(w):
    from quantit_amd import models
    L = w["L"]
    bonds = models.chain_bonds(L) if w["geometry"] == "chain" else models.cylinder_bonds(w["width"], L // w["width"])
    return models.bmpo_parts(models.spin_mpo(L, bonds)), models.random_bmps_parts(L, 4, q=0, seed=w["seed"])


def solve(qa, ctx, mpo_parts, mps_parts, opts, profile=False):
    """One DMRG solve from the prepared random state; the tensors are uploaded
    and prepared before the timed call."""
    mpo = [qa.BTensor.from_parts(ctx, **p) for p in mpo_parts]
    mps = qa.mps_prepare([qa.BTensor.from_parts(ctx, **p) for p in mps_parts])
    ctx.sync()
    ctx.kernel_times(reset=True)
    f0 = ctx.svd_flops()
    ctx.set_profiling(profile)
    t0 = time.perf_counter()
    E, mps, lg = qa.dmrg(mpo, mps, opts)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    f1 = ctx.svd_flops()
    kt = ctx.kernel_times(reset=True)
    return {"E": E, "log": lg, "wall": wall, "kt": kt, "mps": mps, "mpo": mpo,
            "svd_flops": f1[0] - f0[0], "svd_flops_wide": f1[1] - f0[1]}


def dmrg_opts(qa, w, sweeps):
    return qa.DmrgOptions(cutoff=1e-10, convergence_criterion=0.0, maximum_bond=w["max_bond"], minimum_bond=4,
                          maximum_iterations=sweeps)


# Timers that time one kernel each (host_* are host phases; the rest group
# several kernels behind one name or are host waits).
SVD_WIDE_TIMERS = ("svd_round",)
SVD_ALL_TIMERS = ("svd_eig", "svd_eig_it1", "svd_eig_it2", "svd_eig_it3+")


def kernel_roofline(kt, contract_flops, svd_flops, svd_flops_wide, traffic_summary):
    """Roofline object of the timer with the largest share.  Algorithmic work
    per launch: contraction timers -> sum 2 M N K of the solve over their
    launches; k_svd_round -> the SVD model FLOPs of the sectors it rotates
    (14 m n^2 + 8 n^3 each) over its launches; k_gram_eig -> the model FLOPs
    of all sectors over its launches."""
    kernels = {k: v for k, v in kt.items() if not k.startswith("host_")}
    if not kernels:
        return None
    dom, (ms, n) = max(kernels.items(), key=lambda kv: kv[1][0])
    if dom.startswith("gemm"):
        group = [k for k in kernels if k.startswith("gemm") and k != "gemm_split_k_sum" and k != "gemm_rows"]
        work, kname, peak = contract_flops, "k_gemm_f64", PEAK_FP64_TFLOPS
    elif dom in SVD_WIDE_TIMERS:
        group, work, kname, peak = list(SVD_WIDE_TIMERS), svd_flops_wide, "k_svd_round", PEAK_FP64_TFLOPS
    elif dom in SVD_ALL_TIMERS:
        group, work, kname, peak = [k for k in SVD_ALL_TIMERS if k in kernels], svd_flops, "k_gram_eig", \
            PEAK_FP64_TFLOPS
    else:
        group, work, kname, peak = [dom], None, dom, PEAK_FP64_TFLOPS
    g_ms = sum(kernels[k][0] for k in group if k in kernels)
    g_n = sum(kernels[k][1] for k in group if k in kernels)
    avg_s = g_ms / max(g_n, 1) * 1e-3
    ach = (work / max(g_n, 1)) / avg_s / 1e12 if work else None
    return {"kernel": kname, "timers": group, "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
            "frac": ach / peak if ach is not None else None,
            "traffic": read_traffic(traffic_summary, kname) if traffic_summary else None,
            "avg_launch_us": avg_s * 1e6, "launches": g_n, "kernel_ms": g_ms,
            "note": "algorithmic FLOPs of the solve attributed to the kernel's launches, over the mean launch time "
                    "(HIP events on the context stream)"}


def cylinder_cpu_baseline(qa, res, mpo_parts, n_updates, budget):
    """The reference's per-block libtorch loop (oracle/torch_cpu.py) on the
    host: the device's final D=1024 state with its centre moved to the middle
    is downloaded, the environments around the centre are built, and
    `n_updates` forward two-site updates (one-step Lanczos, truncated block
    SVD, environment update) are timed.  The sweep estimate scales the
    measured FLOP rate to the last sweep's FLOPs from the device plans."""
    import torch
    from oracle import torch_cpu as tc
    from oracle.btensor_oracle import BTensor as OB
    L = len(mpo_parts)
    oc = L // 2 - n_updates // 2
    mps = qa.mps_move_oc(res["mps"], 1, oc)
    host = [tc.from_oracle(OB(*t.to_parts())) for t in mps]
    mpo = [tc.from_oracle(OB(**p)) for p in mpo_parts]
    t0 = time.perf_counter()
    env = tc.environments(mpo, host, oc)
    t_env = time.perf_counter() - t0
    times, flops, energies = [], [], []
    for i in range(n_updates):
        f = [0.0]
        t0 = time.perf_counter()
        energies.append(tc.two_site_update(mpo, host, env, oc + i, 1, 1e-10, 4, 1024, flops=f))
        times.append(time.perf_counter() - t0)
        flops.append(f[0])
        if sum(times) + t_env > budget:
            break
    rate = sum(flops) / sum(times)
    return {"rate_gflops": rate / 1e9, "update_s": times, "update_gflop": [f / 1e9 for f in flops],
            "energies": energies, "env_build_s": t_env, "threads": int(torch.get_num_threads())}


def headline(args):
    import quantit_amd as qa
    w = DMRG_WORKLOADS["cylinder4x16_D1024"]
    sweeps = args.sweeps or w["sweeps"]
    ctx = qa.Context(0, arena_bytes=64 << 30)
    mpo_parts, mps_parts = dmrg_inputs(w)
    opts = dmrg_opts(qa, w, sweeps)
    log("cylinder: warmup (full solve: first launches, GEMM tile tuning of every contraction class)")
    solve(qa, ctx, mpo_parts, mps_parts, opts)
    log("cylinder: fp64 batched")
    fp = solve(qa, ctx, mpo_parts, mps_parts, opts)
    ms = fp["log"]["sweep_ms"]
    last_s = ms[-1] / 1e3
    out = {
        "metric": METRIC, "value": last_s, "unit": "s/sweep", "n_gpus": 1, "steps": len(ms), "warmup": 1,
        "ms_per_step": ms[-1], "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cylinder4x16_D1024: " + w["desc"] + "; value = last sweep (bonds at D~1024), "
                               "FP64 batched form"},
        "energy": fp["E"], "sweep_energies": fp["log"]["energies"], "sweep_ms": ms,
        "max_bond": fp["log"]["max_bond"], "solve_wall_s": fp["wall"],
        "contract_gflop": fp["log"]["contract_flops"] / 1e9, "svd_model_gflop": fp["log"]["svd_flops"] / 1e9,
        "gemm_tuning_classes": qa.gemm_tuning_size(),
    }
    E64 = fp["log"]["energies"]
    if not args.quick:
        log("cylinder: per-block fp64")
        ctx.set_per_block(True)
        pb = solve(qa, ctx, mpo_parts, mps_parts, opts)
        ctx.set_per_block(False)
        pbl = pb["log"]["sweep_ms"][-1] / 1e3
        out["per_block_fp64"] = {"last_sweep_s": pbl, "sweep_ms": pb["log"]["sweep_ms"],
                                 "max_rel_dE_vs_batched": max(abs(a - b) / abs(b) for a, b in
                                                              zip(pb["log"]["energies"], E64))}
        out["batched_over_per_block"] = pbl / last_s
        log("cylinder: split fp32")
        ctx.set_precision(qa.PRECISION_SPLIT_FP32)
        sp = solve(qa, ctx, mpo_parts, mps_parts, opts)
        ctx.set_precision(qa.PRECISION_FP64)
        out["split_fp32"] = {"last_sweep_s": sp["log"]["sweep_ms"][-1] / 1e3, "sweep_ms": sp["log"]["sweep_ms"],
                             "rel_dE_vs_fp64_per_sweep": [abs(a - b) / abs(b) for a, b in
                                                         zip(sp["log"]["energies"], E64)],
                             "per_block_over_split": pbl / (sp["log"]["sweep_ms"][-1] / 1e3)}
    log("cylinder: profiled fp64")
    pr = solve(qa, ctx, mpo_parts, mps_parts, opts, profile=True)
    kt = pr["kt"]
    out["kernel_ms_per_solve"] = {k: round(v[0], 3) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:16]}
    out["roofline"] = kernel_roofline(kt, pr["log"]["contract_flops"], pr["svd_flops"], pr["svd_flops_wide"],
                                      "pmc_cylinder4x16_D1024.txt")
    if args.lanczos:
        log("cylinder: k-step lanczos")
        ctx.set_lanczos(args.lanczos, 1e-12)
        lz = solve(qa, ctx, mpo_parts, mps_parts, opts)
        ctx.set_lanczos(0)
        out["lanczos"] = {"max_krylov": args.lanczos, "sweep_ms": lz["log"]["sweep_ms"],
                          "sweep_energies": lz["log"]["energies"], "vectors": ctx.lanczos_stats()}
    if args.no_cpu_baseline:
        out["cpu_baseline"] = None
    else:
        log("cylinder: cpu baseline")
        cb = cylinder_cpu_baseline(qa, fp, mpo_parts, 4, args.cpu_budget)
        last_flops = fp["log"]["contract_flops"] + fp["log"]["svd_flops"]
        # FLOPs of the last sweep: the device plans of the profiled solve's
        # last sweep are not separated, so the solve's FLOPs are split by the
        # sweep times (the last sweep dominates)
        frac = ms[-1] / sum(ms)
        sweep_flops = last_flops * frac
        out["cpu_baseline"] = {
            "value": sweep_flops / (cb["rate_gflops"] * 1e9), "unit": "s/sweep", "cores": cb["threads"],
            "kind": "port",
            "sample": f"{len(cb['update_s'])} centre two-site updates of the device's final D=1024 state (sites "
                      f"{len(mpo_parts) // 2 - 2}..), reference per-block libtorch CPU calls (mm/addmm_/svd), "
                      f"{cb['rate_gflops']:.1f} GFLOP/s, scaled to the last sweep's "
                      f"{sweep_flops / 1e9:.0f} GFLOP",
            "update_s": cb["update_s"], "update_gflop": cb["update_gflop"], "energies": cb["energies"],
            "env_build_s": cb["env_build_s"]}
    del fp, pr
    if not args.quick:
        out["chain_L32_D128"] = sub_chain(qa, args)
        out["contract_svd_D1024"] = sub_contract_svd(qa, args)
        out["grouped_batch"] = sub_grouped(qa, args)
        out["serving"] = sub_serving(args)
    print(json.dumps(out))


def sub_chain(qa, args):
    """L=32 chain, D=128, 6 sweeps: device s/sweep, energies, and the
    libtorch-CPU restatement of the same solve (first sweeps within budget)."""
    log("chain: fp64")
    w = DMRG_WORKLOADS["chain_L32_D128"]
    ctx = qa.Context(0, arena_bytes=8 << 30)
    mpo_parts, mps_parts = dmrg_inputs(w)
    opts = dmrg_opts(qa, w, w["sweeps"])
    solve(qa, ctx, mpo_parts, mps_parts, opts)
    r = solve(qa, ctx, mpo_parts, mps_parts, opts)
    ms = r["log"]["sweep_ms"]
    res = {"workload": w["desc"], "sweep_ms": ms, "mean_s_per_sweep": float(np.mean(ms)) / 1e3,
           "energies": r["log"]["energies"], "max_bond": r["log"]["max_bond"]}
    p = solve(qa, ctx, mpo_parts, mps_parts, opts, profile=True)
    res["roofline"] = kernel_roofline(p["kt"], p["log"]["contract_flops"], p["svd_flops"], p["svd_flops_wide"], None)
    res["kernel_ms_per_solve"] = {k: round(v[0], 3) for k, v in sorted(p["kt"].items(), key=lambda kv: -kv[1][0])[:8]}
    # the same solve with the k-step Lanczos eigensolver at tolerance 1e-12
    # (32 vectors, matvec replayed as a HIP graph), beside the reference's
    # one-step update above
    log("chain: k-step Lanczos, tol 1e-12")
    ctx.set_lanczos(32, 1e-12)
    solve(qa, ctx, mpo_parts, mps_parts, opts)
    st0 = ctx.lanczos_stats()
    lz = solve(qa, ctx, mpo_parts, mps_parts, opts)
    st1 = ctx.lanczos_stats()
    res["lanczos_tol_1e-12"] = {"max_krylov": 32, "sweep_ms": lz["log"]["sweep_ms"],
                                "mean_s_per_sweep": float(np.mean(lz["log"]["sweep_ms"])) / 1e3,
                                "energies": lz["log"]["energies"],
                                "krylov_vectors_per_update": (st1["vectors"] - st0["vectors"]) /
                                max(1, st1["updates"] - st0["updates"]),
                                "graph_replays": st1["graph_replays"] - st0["graph_replays"]}
    ctx.set_graphs(False)   # the same solve with every matvec launched eagerly
    eg = solve(qa, ctx, mpo_parts, mps_parts, opts)
    ctx.set_graphs(True)
    res["lanczos_tol_1e-12"]["eager_mean_s_per_sweep"] = float(np.mean(eg["log"]["sweep_ms"])) / 1e3
    res["lanczos_tol_1e-12"]["eager_energies_equal"] = eg["log"]["energies"] == lz["log"]["energies"]
    ctx.set_lanczos(0, 1e-12)
    if not args.no_cpu_baseline:
        log("chain: cpu baseline")
        from oracle import torch_cpu as tc
        from oracle import dmrg_oracle as dor
        from oracle.btensor_oracle import BTensor as OB
        mpo = [tc.from_oracle(OB(**q)) for q in mpo_parts]
        mps = [tc.from_oracle(t) for t in dor.prepare_state([OB(**q) for q in mps_parts])]
        env = tc.environments(mpo, mps, 0)
        L = len(mpo)
        oc, step, times, energies = 0, 1, [], []
        t_start = time.perf_counter()
        for sweep in range(w["sweeps"]):
            t0 = time.perf_counter()
            for _ in range(2 * (L - 2)):
                E = tc.two_site_update(mpo, mps, env, oc, step, 1e-10, 4, 128)
                oc += step
                if oc == 0 or oc == L - 2:
                    step = -step
            times.append(time.perf_counter() - t0)
            energies.append(E)
            if time.perf_counter() - t_start > args.cpu_budget:
                break
        res["cpu_baseline"] = {"value": float(np.mean(times)), "unit": "s/sweep", "cores": threads_used(),
                               "kind": "port", "sample": f"first {len(times)} sweeps of the same solve, reference "
                                                         f"per-block libtorch CPU calls", "sweep_s": times,
                               "max_rel_dE_vs_device": max(abs(a - b) / abs(b) for a, b in
                                                           zip(energies, r["log"]["energies"]))}
    ctx.close()
    return res


def sub_contract_svd(qa, args, steps=None):
    """X(1024,1024,2,26).Abar(1024,2,1024) over two legs + truncating block SVD."""
    from quantit_amd import workloads
    log("contract_svd: fp64")
    steps = steps or min(args.steps, 20)
    X, A, dx, da, split = workloads.contract_svd(1024)
    ctx = qa.Context(0, arena_bytes=16 << 30)
    x = qa.BTensor.from_parts(ctx, **X)
    a = qa.BTensor.from_parts(ctx, **A)

    def step():
        y = x.tensordot(a, dx, da)
        return qa.svd(y, split, tol=1e-10)

    for _ in range(3):
        step()
    ctx.sync()
    cflops = qa.last_tensordot_flops()
    y = x.tensordot(a, dx, da)
    sflops = svd_model_flops(dict(zip(["groups", "secs", "cvals", "sel", "blocks"], y.to_parts())), split)
    f0 = ctx.svd_flops()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    dt = (time.perf_counter() - t0) / steps
    ctx.kernel_times(reset=True)
    ctx.set_profiling(True)
    for _ in range(steps):
        step()
    ctx.sync()
    ctx.set_profiling(False)
    kt = ctx.kernel_times(reset=True)
    f1 = ctx.svd_flops()
    wide = (f1[1] - f0[1]) / 2.0   # both loops counted
    res = {"ms_per_step": dt * 1e3, "tflops": (cflops + sflops) / dt / 1e12, "contraction_gflop": cflops / 1e9,
           "svd_model_gflop": sflops / 1e9,
           "kernel_ms_per_step": {k: round(v[0] / steps, 4) for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])[:10]},
           "roofline": kernel_roofline(kt, cflops * steps, sflops * steps, wide, "pmc_contract_svd_D1024.txt")}
    g = sum(v[0] for k, v in kt.items() if k.startswith("gemm_f64"))
    if g:
        res["contraction_tflops"] = cflops * steps / (g * 1e-3) / 1e12
    if not args.no_cpu_baseline:
        log("contract_svd: cpu baseline")
        from oracle import torch_cpu as tc
        from oracle.btensor_oracle import BTensor as OB
        Xo, Ao = tc.from_oracle(OB(**X)), tc.from_oracle(OB(**A))
        times = []
        t_start = time.perf_counter()
        while len(times) < 5 and time.perf_counter() - t_start < args.cpu_budget:
            t0 = time.perf_counter()
            Y = tc.tensordot(Xo, Ao, dx, da)
            tc.svd_truncated(Y, split, 1e-10, 1, 2 ** 62)
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        res["cpu_baseline"] = {"value": (cflops + sflops) / t / 1e12, "unit": "TFLOP/s", "cores": threads_used(),
                               "kind": "port", "sample": f"{len(times)} whole steps, reference per-block libtorch "
                                                         f"CPU calls, median {t * 1e3:.1f} ms/step"}
    ctx.close()
    return res


def sub_serving(args, n=8, rate=4.0):
    """A bounded serving trace for the default line: n seeded Poisson XXZ
    requests (L in 16..64, D in {64, 128, 256}) at `rate` per second on
    --slots slots in the served mode (k-step Lanczos, --serve-krylov); the full trace and the rate
    sweep run with --workload serving --full-trace --rates ..."""
    from quantit_amd import serving
    log("serving: bounded trace")
    trace = serving.make_trace(n, rate=rate, seed=5, L_range=(16, 64), D_choices=(64, 128, 256))
    srv = serving.Server(n_slots=args.slots, hbm_budget=160 << 30, max_sweeps=40, max_krylov=args.serve_krylov)
    res, wall = srv.run(trace)
    summ = serving.summarize(res, wall)
    summ["hbm_reserved_peak_gib"] = (srv.budget - srv.min_free) / float(1 << 30)
    summ.update({"requests": n, "rate": rate, "slots": args.slots, "max_krylov": args.serve_krylov,
                 "trace": "L ~ U{16..64} even, D ~ U{64,128,256}, delta ~ U[0.5,1.5], Sz=0, |dE/E| <= 1e-10"})
    return summ


def sub_grouped(qa, args):
    """2048 independent GEMMs (one grouped launch) and 512 independent SVDs
    (one batched block SVD), against one launch per block pair."""
    from quantit_amd.workloads import grouped_batch
    log("grouped: fp64")
    A, B, S, (m, n, k), (sm, sn) = grouped_batch()
    ctx = qa.Context(0, arena_bytes=16 << 30)
    a = qa.BTensor.from_parts(ctx, **A)
    b = qa.BTensor.from_parts(ctx, **B)
    s = qa.BTensor.from_parts(ctx, **S)
    f_gemm = float(np.sum(2.0 * m * n * k))
    q_gemm = 8.0 * (m * k + k * n + m * n)
    peak = PEAK_FP64_TFLOPS * 1e12
    t_roof = float(np.sum(np.maximum(2.0 * m * n * k / peak, q_gemm / (PEAK_HBM_GBS * 1e9))))
    big, small = np.maximum(sm, sn).astype(float), np.minimum(sm, sn).astype(float)
    f_svd = float(np.sum(14.0 * big * small ** 2 + 8.0 * small ** 3))

    def timed(fn, steps):
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

    t_g, kt_g = timed(lambda: a.tensordot(b, [1], [0]), 5)
    ctx.set_per_block(True)
    t_pb, _ = timed(lambda: a.tensordot(b, [1], [0]), 1)
    ctx.set_per_block(False)
    t_s, kt_s = timed(lambda: qa.svd(s, 1), 2)
    g_ms, g_n = kt_g.get("gemm_f64", (0.0, 1))
    g_avg = g_ms / max(g_n, 1) * 1e-3
    res = {"gemm_batch": {"ms": t_g * 1e3, "tflops": f_gemm / t_g / 1e12, "gflop": f_gemm / 1e9,
                          "kernel_ms": g_avg * 1e3, "per_block_launch_ms": t_pb * 1e3,
                          "speedup_vs_per_block": t_pb / t_g, "roofline_additive_ms": t_roof * 1e3,
                          "roofline_frac": t_roof / g_avg if g_avg else None},
           "svd_batch": {"ms": t_s * 1e3, "model_gflop": f_svd / 1e9, "model_tflops": f_svd / t_s / 1e12,
                         "kernel_ms": {kk: round(v[0] / 2, 3) for kk, v in sorted(kt_s.items(),
                                                                                 key=lambda kv: -kv[1][0])[:8]}}}
    if not args.no_cpu_baseline:
        log("grouped: cpu baseline")
        import torch
        t0 = time.perf_counter()
        n_g = 0
        for i in range(len(m)):
            torch.mm(torch.from_numpy(A["blocks"][(i, i)]), torch.from_numpy(B["blocks"][(i, i)]))
            n_g = i + 1
            if time.perf_counter() - t0 > args.cpu_budget / 2:
                break
        tg = time.perf_counter() - t0
        fg = float(np.sum(2.0 * m[:n_g] * n[:n_g] * k[:n_g]))
        t0 = time.perf_counter()
        n_s = 0
        for i in range(len(sm)):
            torch.linalg.svd(torch.from_numpy(S["blocks"][(i, i)]), full_matrices=False)
            n_s = i + 1
            if time.perf_counter() - t0 > args.cpu_budget / 2:
                break
        ts = time.perf_counter() - t0
        fs = float(np.sum(14.0 * big[:n_s] * small[:n_s] ** 2 + 8.0 * small[:n_s] ** 3))
        res["cpu_baseline"] = {"value": (fg + fs) / (tg + ts) / 1e12, "unit": "TFLOP/s", "cores": threads_used(),
                               "kind": "port", "sample": f"libtorch CPU mm over the first {n_g} GEMM items and "
                                                         f"linalg.svd over the first {n_s} SVD items",
                               "gemm_tflops": fg / tg / 1e12, "svd_model_tflops": fs / ts / 1e12}
    ctx.close()
    return res


def bench_dmrg(args):
    """One whole-solve workload alone (chain or cylinder), optionally per-block
    or split-FP32: s/sweep, energies, per-kernel times of a profiled solve."""
    import quantit_amd as qa
    w = DMRG_WORKLOADS[args.workload]
    sweeps = args.sweeps or w["sweeps"]
    ctx = qa.Context(0, arena_bytes=64 << 30,
                     precision=qa.PRECISION_SPLIT_FP32 if args.precision == "split" else qa.PRECISION_FP64)
    ctx.set_per_block(args.per_block)
    if args.lanczos:
        ctx.set_lanczos(args.lanczos, 1e-12)
    ctx.set_graphs(not args.no_graphs)
    mpo_parts, mps_parts = dmrg_inputs(w)
    opts = dmrg_opts(qa, w, sweeps)
    solve(qa, ctx, mpo_parts, mps_parts, opts)   # warmup: first launches and GEMM tile tuning
    s0 = ctx.svd_stats()
    r = solve(qa, ctx, mpo_parts, mps_parts, opts)
    s1 = ctx.svd_stats()
    p = solve(qa, ctx, mpo_parts, mps_parts, opts, profile=True) if args.profile else None
    ms = r["log"]["sweep_ms"]
    out = {"metric": METRIC, "value": ms[-1] / 1e3, "unit": "s/sweep", "n_gpus": 1, "steps": len(ms),
           "warmup": 1, "ms_per_step": ms[-1], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if args.precision == "fp64" else "f64 (split-fp32 contraction)", "data": "synthetic",
           "config": {"workload": f"{args.workload}: {w['desc']}; value = last sweep", "precision": args.precision,
                      "per_block_launch": args.per_block, "lanczos": args.lanczos},
           "energy": r["E"], "sweep_energies": r["log"]["energies"], "sweep_ms": ms, "max_bond": r["log"]["max_bond"],
           "solve_wall_s": r["wall"], "contract_gflop": r["log"]["contract_flops"] / 1e9,
           "svd_model_gflop": r["log"]["svd_flops"] / 1e9, "roofline": None, "cpu_baseline": None}
    # block SVD health of the timed solve: calls, the largest Jacobi sweep
    # count of one call so far, sectors left rotating at the sweep cap
    out["svd_stats"] = {"calls": s1["calls"] - s0["calls"], "max_sweeps": s1["max_sweeps"],
                        "unconverged": s1["unconverged"] - s0["unconverged"]}
    if args.lanczos:
        out["lanczos_vectors"] = ctx.lanczos_stats()
    if p:
        out["kernel_ms_per_solve"] = {k: round(v[0], 3) for k, v in sorted(p["kt"].items(), key=lambda kv: -kv[1][0])}
        out["roofline"] = kernel_roofline(p["kt"], p["log"]["contract_flops"], p["svd_flops"], p["svd_flops_wide"],
                                          None)
    print(json.dumps(out))


def bench_serving(args):
    """Seeded Poisson traces of XXZ ground-state requests served by solve
    slots with admission by HBM footprint, one trace per arrival rate:
    p50/p99 latency, solves/s; unconverged solves count as failures."""
    from quantit_amd import serving
    if args.full_trace:
        L_range, D_choices = (16, 128), (64, 128, 256, 512)
    else:
        L_range, D_choices = (16, 64), (64, 128, 256)
    rates = [float(r) for r in args.rates.split(",")] if args.rates else [args.rate]
    per_rate = []
    for rate in rates:
        trace = serving.make_trace(args.requests, rate=rate, seed=5, L_range=L_range, D_choices=D_choices)
        srv = serving.Server(n_slots=args.slots, hbm_budget=int(args.hbm_gib * (1 << 30)),
                             max_sweeps=args.sweeps or 40, max_krylov=args.serve_krylov)

        def progress(r, x):
            log(f"rate {rate}: request {r.rid} L={r.L} D={r.D} "
                + (f"E={x.energy:.12f} sweeps={x.sweeps} converged={x.converged} latency={x.latency:.2f}s"
                   if isinstance(x, serving.Result) else f"error {x}"))

        res, wall = srv.run(trace, progress=progress)
        summ = serving.summarize(res, wall)
        summ["rate"] = rate
        # peak of the admission model's reservations (footprint_bytes of the
        # solves running at once): the HBM the served mix held at most
        summ["hbm_reserved_peak_gib"] = (srv.budget - srv.min_free) / float(1 << 30)
        per_rate.append(summ)
    head = per_rate[-1]
    out = {"metric": METRIC, "value": head["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
           "steps": args.requests, "warmup": 0, "ms_per_step": 1e3 * head["wall_s"] / max(1, head["solves"]),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"serving: {args.requests} XXZ chain requests per rate, Poisson arrivals (seed 5), "
                                  f"L ~ U{{{L_range[0]}..{L_range[1]}}}, D ~ U{set(D_choices)}, delta ~ U[0.5,1.5], "
                                  f"Sz=0, converged to |dE/E| <= 1e-10 within {args.sweeps or 40} sweeps",
                      "slots": args.slots, "hbm_budget_gib": args.hbm_gib, "lanczos": args.serve_krylov},
           "p50_latency_s": head["p50_latency_s"], "p99_latency_s": head["p99_latency_s"], "per_rate": per_rate,
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "split"], default="fp64")
    ap.add_argument("--per-block", action="store_true", help="one GEMM launch per block pair (baseline form)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--quick", action="store_true", help="headline only: FP64 batched, profiled solve, CPU baseline")
    ap.add_argument("--profile", action="store_true", help="single workloads: add a profiled solve")
    ap.add_argument("--lanczos", type=int, default=0, help="k-step Lanczos with this many Krylov vectors (0: one-step)")
    ap.add_argument("--no-graphs", action="store_true", help="k-step Lanczos: run every matvec eagerly")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--requests", type=int, default=8, help="serving: trace length per rate")
    ap.add_argument("--rate", type=float, default=4.0, help="serving: Poisson arrival rate (1/s)")
    ap.add_argument("--rates", type=str, default="", help="serving: comma-separated rates to sweep")
    ap.add_argument("--slots", type=int, default=8, help="serving: concurrent solve slots")
    ap.add_argument("--serve-krylov", type=lambda v: v if v == "auto" else int(v), default="auto",
                    help="serving: Krylov vectors of the k-step Lanczos update (0: the one-step update; "
                         "auto: 4 for D <= 256, 6 above)")
    ap.add_argument("--hbm-gib", type=float, default=160.0, help="serving: HBM budget for admission")
    ap.add_argument("--full-trace", action="store_true",
                    help="serving: L up to 128 and D up to 512 instead of L <= 64, D <= 256")
    ap.add_argument("--workload", choices=["headline", "contract_svd_D1024", "grouped_batch", "serving"]
                    + sorted(DMRG_WORKLOADS), default="headline")
    ap.add_argument("--sweeps", type=int, default=0, help="DMRG workloads: sweeps (0 = the workload's own)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import quantit_amd as qa
    if args.workload == "headline":
        headline(args)
    elif args.workload in DMRG_WORKLOADS:
        bench_dmrg(args)
    elif args.workload == "grouped_batch":
        r = sub_grouped(qa, args)
        print(json.dumps({"metric": METRIC, "value": r["gemm_batch"]["tflops"], "unit": "TFLOP/s", "n_gpus": 1,
                          "steps": 5, "warmup": 1, "ms_per_step": r["gemm_batch"]["ms"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": "grouped_batch: 2048 GEMMs + 512 SVDs, dims log-uniform in [4, 768]"},
                          **r, "roofline": None, "cpu_baseline": r.get("cpu_baseline")}))
    elif args.workload == "contract_svd_D1024":
        r = sub_contract_svd(qa, args)
        print(json.dumps({"metric": METRIC, "value": r["tflops"], "unit": "TFLOP/s", "n_gpus": 1,
                          "steps": min(args.steps, 20), "warmup": 3, "ms_per_step": r["ms_per_step"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic",
                          "config": {"workload": "contract_svd_D1024: X(1024,1024,2,26).Abar(1024,2,1024) over 2 legs "
                                                 "+ truncating block SVD (split 1, cutoff 1e-10), U(1)"},
                          **r, "cpu_baseline": r.get("cpu_baseline")}))
    else:
        bench_serving(args)


if __name__ == "__main__":
    main()
