"""Serving loop on device: the same request gives the same bits alone and in
a mix of concurrent solves (one context and stream per slot), and the
energies agree with exact diagonalisation for the small chains."""
import numpy as np
import pytest

from oracle import dmrg_oracle as dor

pytestmark = pytest.mark.gpu


def _exact(req):
    from quantit_amd import models
    H = dor.mpo_to_dense_matrix(models.spin_mpo(req.L, models.chain_bonds(req.L), delta=req.delta))
    ups = np.array([req.L - bin(n).count("1") for n in range(2 ** req.L)])
    sel = np.nonzero(2 * ups == req.L)[0]
    return float(np.linalg.eigvalsh(H[np.ix_(sel, sel)])[0])


def test_results_identical_alone_and_mixed():
    from quantit_amd import serving
    # D = 2^(L/2): the chains are represented exactly, so the energies are the exact ones
    trace = [serving.Request(0, 0.0, 12, 64, 0.7, 11), serving.Request(1, 0.0, 10, 32, 1.3, 12),
             serving.Request(2, 0.01, 12, 64, 1.0, 13), serving.Request(3, 0.02, 8, 16, 0.5, 14)]
    alone = []
    for r in trace:
        res, _ = serving.Server(n_slots=1, hbm_budget=8 << 30).run([r])
        alone.append(res[0])
    mixed, wall = serving.Server(n_slots=4, hbm_budget=8 << 30).run(trace)
    for a, m, r in zip(alone, mixed, trace):
        assert isinstance(a, serving.Result) and isinstance(m, serving.Result)
        assert a.energy == m.energy and a.sweeps == m.sweeps   # bit-identical
        e0 = _exact(r)
        assert abs(a.energy - e0) <= 1e-9 * abs(e0), (r, a.energy, e0)
    assert len({m.slot for m in mixed}) > 1   # the mix really ran on several slots


def _xx_exact(L):
    """XX chain (delta = 0) with open ends: free fermions with hopping 1/2,
    single-particle energies cos(pi k / (L + 1)), half filling."""
    eps = np.cos(np.pi * np.arange(1, L + 1) / (L + 1))
    return float(np.sort(eps)[:L // 2].sum())


def test_served_chains_in_the_trace_range_converge_and_match():
    """Requests with L in the trace's range (16..128) in k-step Lanczos mode:
    alone and mixed give the same bits, every solve meets |dE/E| <= 1e-10
    between sweeps, and the XX chains (delta = 0) reach the free-fermion
    ground-state energy."""
    from quantit_amd import serving
    trace = [serving.Request(0, 0.0, 16, 64, 0.0, 21), serving.Request(1, 0.0, 32, 128, 0.0, 22),
             serving.Request(2, 0.0, 24, 64, 1.2, 23), serving.Request(3, 0.0, 48, 64, 0.8, 24)]
    alone = []
    for r in trace:
        res, _ = serving.Server(n_slots=1, hbm_budget=16 << 30, max_krylov=32).run([r])
        alone.append(res[0])
    mixed, _ = serving.Server(n_slots=4, hbm_budget=16 << 30, max_krylov=32).run(trace)
    for a, m, r in zip(alone, mixed, trace):
        assert isinstance(a, serving.Result) and isinstance(m, serving.Result), (a, m)
        assert a.converged and m.converged
        assert a.energy == m.energy and a.sweeps == m.sweeps
    for a, r in zip(alone, trace):
        if r.delta == 0.0:
            e0 = _xx_exact(r.L)
            assert abs(a.energy - e0) <= 1e-9 * abs(e0), (r.L, a.energy, e0)


@pytest.mark.parametrize("L,D,krylov", [(16, 64, 0), (32, 128, 0), (48, 256, 0), (32, 256, 32)])
def test_footprint_covers_the_arena_peak(L, D, krylov):
    """The admission model's footprint_bytes bounds the arena's high-water
    mark of a solve of that size (and is not loose by more than 10x), in the
    one-step and the k-step Lanczos modes."""
    import quantit_amd as qa
    from quantit_amd import serving
    need = serving.footprint_bytes(L, D, krylov=krylov)
    ctx = qa.Context(0, arena_bytes=need)
    ctx.set_overlap(False)
    if krylov:
        ctx.set_lanczos(krylov, 1e-12)
    serving.solve(ctx, serving.Request(0, 0.0, L, D, 1.0, 5), max_sweeps=6)
    mem = ctx.memory()
    print(L, D, krylov, "peak", mem["peak"], "footprint", need, flush=True)
    assert mem["peak"] <= need
    assert mem["peak"] >= need / 10
    ctx.close()
