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
