"""Serving loop for independent DMRG ground-state solves on one GPU.

Requests (XXZ chains of different lengths, anisotropies and bond
dimensions) arrive over time.  Each one is admitted when a solve slot is free
and its HBM footprint fits the memory still unreserved.  The footprint model
covers the MPS, both environment chains and the block-SVD workspace at the
requested bond dimension.  Every slot is a context of its own: one HIP stream
and one arena.  A slot is driven by its own host thread; the library releases
the GIL inside each call, so the solves of different slots overlap on the
device.

The reference has no serving layer: a solve is ``dmrg(H, psi, options,
logger)`` (include/dmrg.h:39-40) run to the convergence criterion of
sources/dmrg.cpp:247-254, and every request here is exactly that call
(``quantit_amd.dmrg``).  Results do not depend on what else runs: each
solve's kernels read and write only its own arena, and no kernel uses
floating-point atomics, so a request gives the same bits alone and in a mix.
"""
from __future__ import annotations

import dataclasses
import queue
import threading
import time

import numpy as np

from . import Context, DmrgOptions, dmrg, mps_prepare
from . import models

MPO_BOND = 5          # XXZ chain MPO bond dimension
ARENA_SLACK = 1.6     # allocator fragmentation and transient copies


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float        # seconds after the start of the trace
    L: int
    D: int
    delta: float
    seed: int


@dataclasses.dataclass
class Result:
    rid: int
    energy: float
    sweeps: int
    start: float
    finish: float
    arrival: float
    slot: int
    converged: bool = True   # |dE/E| between the last two sweeps within the criterion

    @property
    def latency(self):
        return self.finish - self.arrival


def footprint_bytes(L, D, w=MPO_BOND, d=2, krylov=0):
    """HBM bytes a solve holds at bond dimension D.

    Per site: the MPS tensor (D d D), the left and right environments
    (D w D each) and the two-site Hamiltonian (w d^2 w d^2, negligible).
    Per update: theta and the Lanczos vectors (4 of D d d D), the SVD work
    area (Y, ping-pong copy, W for the (D d) x (d D) matrix) and the
    contraction intermediates (L.theta: D w d d D, twice).  With the k-step
    Lanczos eigensolver (krylov >= 2) also the Krylov basis (krylov vectors
    of D d d D) and the captured matvec's own input, intermediates and
    output (D d d D + 2 of D w d d D)."""
    site = 8.0 * (D * d * D + 2 * D * w * D)
    update = 8.0 * (4 * D * d * d * D + 3 * (d * D) ** 2 + 2 * D * w * d * d * D)
    if krylov >= 2:
        update += 8.0 * ((krylov + 1) * D * d * d * D + 2 * D * w * d * d * D)
    return int(ARENA_SLACK * (L * site + update)) + (64 << 20)


def make_trace(n, rate, seed=5, L_range=(16, 128), D_choices=(64, 128, 256, 512), delta_range=(0.5, 1.5)):
    """Seeded Poisson arrivals of XXZ requests: L ~ U{L_range}, even;
    D ~ U(D_choices); delta ~ U(delta_range)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        L = int(rng.integers(L_range[0] // 2, L_range[1] // 2 + 1)) * 2
        out.append(Request(i, float(t[i]), L, int(rng.choice(D_choices)), float(rng.uniform(*delta_range)),
                           int(rng.integers(0, 2 ** 31))))
    return out


def solve(ctx, req, cutoff=1e-10, criterion=1e-10, max_sweeps=40):
    """One ground-state solve on the given context: XXZ chain with open ends,
    U(1) Sz = 0, random bMPS start, two-site DMRG to |dE/E| <= criterion
    between sweeps (the reference's test, sources/dmrg.cpp:247-254).
    Returns (energy, sweeps, converged)."""
    dense = models.spin_mpo(req.L, models.chain_bonds(req.L), J=1.0, delta=req.delta)
    mpo = models.to_bmpo(ctx, dense)
    mps = mps_prepare(models.random_bmps(ctx, req.L, 4, q=0, seed=req.seed))
    opts = DmrgOptions(cutoff=cutoff, convergence_criterion=criterion, maximum_bond=req.D, minimum_bond=4,
                       maximum_iterations=max_sweeps)
    E, mps, log = dmrg(mpo, mps, opts)
    ctx.sync()
    es = log["energies"]
    converged = len(es) >= 2 and abs((es[-1] - es[-2]) / es[-1]) <= criterion
    return E, len(es), converged


class Server:
    """n_slots solve slots sharing an HBM budget (bytes)."""

    def __init__(self, n_slots=8, hbm_budget=200 << 30, device=0, solve_fn=None, context_fn=None, max_krylov="auto",
                 **solve_kw):
        """max_krylov: Krylov vectors of the k-step Lanczos update (2..64),
        0 for the reference's one-step update (which needs many more sweeps),
        or "auto": 4 for requests with D <= 256 and 6 above, the sizes that
        served the measured traces fastest (a sweep at D = 512 is costly
        enough that fewer sweeps with more vectors per update win)."""
        self.n_slots = n_slots
        self.max_krylov = max_krylov
        self.budget = hbm_budget
        self.device = device
        self.solve_kw = solve_kw
        self.solve_fn = solve_fn or solve
        self.context_fn = context_fn or self._context
        self.free = hbm_budget
        self.min_free = hbm_budget
        self.cv = threading.Condition()

    def _context(self, nbytes):
        ctx = Context(self.device, arena_bytes=nbytes)
        # several solves at once: one stream each (a second stream per solve
        # would share hardware queues with the other solves' streams)
        ctx.set_overlap(self.n_slots == 1)
        return ctx

    def krylov_for(self, r):
        """Krylov vectors used for request r (0: one-step update)."""
        if self.max_krylov == "auto":
            return 4 if r.D <= 256 else 6
        return int(self.max_krylov)

    def run(self, trace, time_scale=1.0, progress=None):
        """Replays the trace (arrival times scaled by time_scale) and returns
        the results in request order.  Admission: first come first served
        among the arrived requests whose footprint fits the unreserved
        budget; a request larger than the whole budget is rejected with
        MemoryError in its result slot."""
        pending = sorted(trace, key=lambda r: r.arrival)
        ready = queue.Queue()
        results = {}
        t0 = time.perf_counter()
        state = {"next": 0}

        def admit_loop():
            waiting = []
            while state["next"] < len(pending) or waiting:
                now = time.perf_counter() - t0
                while state["next"] < len(pending) and pending[state["next"]].arrival * time_scale <= now:
                    waiting.append(pending[state["next"]])
                    state["next"] += 1
                with self.cv:
                    for r in list(waiting):
                        need = footprint_bytes(r.L, r.D, krylov=self.krylov_for(r))
                        if need > self.budget:
                            waiting.remove(r)
                            results[r.rid] = MemoryError(f"request {r.rid} needs {need} bytes")
                        elif need <= self.free:
                            self.free -= need
                            self.min_free = min(self.min_free, self.free)
                            waiting.remove(r)
                            ready.put((r, need))
                    if waiting or state["next"] < len(pending):
                        nxt = pending[state["next"]].arrival * time_scale - now if state["next"] < len(pending) else 0.05
                        self.cv.wait(timeout=max(0.001, min(0.05, nxt)))
            for _ in range(self.n_slots):
                ready.put(None)

        def slot_loop(slot):
            while True:
                item = ready.get()
                if item is None:
                    return
                r, need = item
                start = time.perf_counter() - t0
                ctx = None
                try:
                    ctx = self.context_fn(need)
                    kr = self.krylov_for(r)
                    if kr and hasattr(ctx, "set_lanczos"):
                        ctx.set_lanczos(kr, 1e-12)
                    got = self.solve_fn(ctx, r, **self.solve_kw)
                    E, sweeps = got[0], got[1]
                    conv = bool(got[2]) if len(got) > 2 else True
                    results[r.rid] = Result(r.rid, E, sweeps, start, time.perf_counter() - t0,
                                            r.arrival * time_scale, slot, conv)
                except Exception as e:  # reported per request, the server goes on
                    # without its traceback, which would keep the solve's
                    # frames (and their device tensors) alive
                    results[r.rid] = e.with_traceback(None)
                finally:
                    if ctx is not None and hasattr(ctx, "close"):
                        ctx.close()
                    with self.cv:
                        self.free += need
                        self.cv.notify_all()
                if progress is not None:
                    progress(r, results[r.rid])

        threads = [threading.Thread(target=admit_loop)]
        threads += [threading.Thread(target=slot_loop, args=(i,)) for i in range(self.n_slots)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        return [results[r.rid] for r in trace], time.perf_counter() - t0


def summarize(results, wall):
    """Counts, throughput and latency percentiles of a served trace.  A solve
    that ended at its sweep cap without meeting the convergence criterion is
    a failure (``unconverged``), not a served result."""
    done = [r for r in results if isinstance(r, Result)]
    ok = [r for r in done if r.converged]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    return {"solves": len(ok), "failed": len(results) - len(done), "unconverged": len(done) - len(ok),
            "wall_s": wall, "solves_per_s": len(ok) / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_sweeps": float(np.mean([r.sweeps for r in done])) if done else 0.0}
