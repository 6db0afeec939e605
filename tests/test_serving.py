"""Admission and scheduling of the serving loop (quantit_amd/serving.py),
on the CPU with a stand-in solve: requests wait for memory, never exceed the
budget together, run at most n_slots at a time, oversize requests are
rejected, and the trace is seeded."""
import threading
import time

import pytest

from quantit_amd import serving


def test_trace_is_seeded_and_in_range():
    a = serving.make_trace(20, rate=8.0, seed=3)
    b = serving.make_trace(20, rate=8.0, seed=3)
    assert [(r.arrival, r.L, r.D, r.delta, r.seed) for r in a] == [(r.arrival, r.L, r.D, r.delta, r.seed) for r in b]
    assert all(16 <= r.L <= 128 and r.L % 2 == 0 and r.D in (64, 128, 256, 512) and 0.5 <= r.delta <= 1.5 for r in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))


def test_footprint_grows_with_size():
    assert serving.footprint_bytes(32, 128) < serving.footprint_bytes(64, 128) < serving.footprint_bytes(64, 256)


def test_krylov_size_per_request():
    """auto: 4 vectors up to D = 256, 6 above; an explicit size applies to
    every request; the footprint counts the request's own Krylov basis."""
    small, large = serving.Request(0, 0.0, 32, 256, 1.0, 0), serving.Request(1, 0.0, 32, 512, 1.0, 1)
    auto = serving.Server(n_slots=1, hbm_budget=1 << 40)
    assert auto.krylov_for(small) == 4 and auto.krylov_for(large) == 6
    fixed = serving.Server(n_slots=1, hbm_budget=1 << 40, max_krylov=8)
    assert fixed.krylov_for(small) == fixed.krylov_for(large) == 8
    assert serving.footprint_bytes(32, 512, krylov=6) > serving.footprint_bytes(32, 512, krylov=4)


def test_admission_respects_budget_and_slots():
    trace = [serving.Request(i, 0.001 * i, 32, 128, 1.0, i) for i in range(12)]
    need = serving.footprint_bytes(32, 128, krylov=4)   # the server's default k-step mode
    budget = int(2.5 * need)          # at most two fit at once
    live, peak, lock = [0], [0], threading.Lock()

    def fake_solve(ctx, req):
        with lock:
            live[0] += 1
            peak[0] = max(peak[0], live[0])
        time.sleep(0.02)
        with lock:
            live[0] -= 1
        return -float(req.rid), 3

    srv = serving.Server(n_slots=4, hbm_budget=budget, solve_fn=fake_solve, context_fn=lambda n: object())
    res, wall = srv.run(trace)
    assert [r.energy for r in res] == [-float(i) for i in range(12)]
    assert peak[0] <= 2 and srv.min_free >= 0
    assert srv.free == budget
    s = serving.summarize(res, wall)
    assert s["solves"] == 12 and s["failed"] == 0 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_oversize_request_is_rejected_not_blocking():
    big = serving.Request(0, 0.0, 128, 512, 1.0, 1)
    small = serving.Request(1, 0.0, 16, 64, 1.0, 2)
    srv = serving.Server(n_slots=2, hbm_budget=serving.footprint_bytes(16, 64, krylov=4) * 2,
                         solve_fn=lambda ctx, req: (1.0, 1), context_fn=lambda n: object())
    res, _ = srv.run([big, small])
    assert isinstance(res[0], MemoryError) and isinstance(res[1], serving.Result)


def test_solver_failure_is_reported_per_request():
    def failing(ctx, req):
        if req.rid == 1:
            raise ArithmeticError("NaN in the Lanczos step")
        return 0.5, 2

    trace = [serving.Request(i, 0.0, 16, 64, 1.0, i) for i in range(3)]
    srv = serving.Server(n_slots=2, hbm_budget=1 << 40, solve_fn=failing, context_fn=lambda n: object())
    res, _ = srv.run(trace)
    assert isinstance(res[1], ArithmeticError) and res[0].energy == 0.5 and res[2].energy == 0.5


def test_unconverged_solves_are_failures():
    """A solve that stops at the sweep cap short of the criterion is reported
    as unconverged and left out of solves/s and the latency percentiles."""
    def capped(ctx, req):
        return -1.0, 40, req.rid % 2 == 0

    trace = [serving.Request(i, 0.0, 16, 64, 1.0, i) for i in range(4)]
    srv = serving.Server(n_slots=2, hbm_budget=1 << 40, solve_fn=capped, context_fn=lambda n: object())
    res, wall = srv.run(trace)
    s = serving.summarize(res, wall)
    assert s["solves"] == 2 and s["unconverged"] == 2 and s["failed"] == 0
    assert [r.converged for r in res] == [True, False, True, False]
