"""Pins the libtorch-CPU restatement used as the benchmark's CPU baseline
(oracle/torch_cpu.py) against the numpy oracle: contractions, truncated block
SVD values and a whole sweep of two-site updates.  CPU only."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import btensor_oracle as bo
from oracle import dmrg_oracle as dor
from oracle import torch_cpu as tc
from tests.helpers import inv_leg, rand_btensor, rand_leg


def test_tensordot_matches_numpy_oracle():
    rng = np.random.default_rng(0)
    legs = [rand_leg(rng, 3, smax=5) for _ in range(3)]
    a = rand_btensor(rng, (0,), legs, (0,))
    b = rand_btensor(rng, (0,), [inv_leg((0,), legs[2]), inv_leg((0,), legs[1]), rand_leg(rng, 3)], (1,))
    ref = bo.tensordot(a, b, [1, 2], [1, 0])
    got = tc.to_oracle(tc.tensordot(tc.from_oracle(a), tc.from_oracle(b), [1, 2], [1, 0]))
    assert sorted(got.blocks) == sorted(ref.blocks)
    for k in ref.blocks:
        assert np.abs(got.blocks[k] - ref.blocks[k]).max() <= 1e-13


def test_truncated_svd_matches_numpy_oracle():
    rng = np.random.default_rng(1)
    rows = [(6, (0,)), (5, (1,)), (4, (-1,))]
    t = rand_btensor(rng, (0,), [rows, [(1, (1,)), (1, (-1,))], inv_leg((0,), rows)], (1,))
    rU, rd, rV = bo.svd_truncated(t, 2, 1e-2, 1, 7)
    U, d, V = (tc.to_oracle(x) for x in tc.svd_truncated(tc.from_oracle(t), 2, 1e-2, 1, 7))
    assert d.secs == rd.secs and sorted(d.blocks) == sorted(rd.blocks)
    for k in rd.blocks:
        assert np.abs(d.blocks[k] - rd.blocks[k]).max() <= 1e-13
    assert U.secs == rU.secs and V.secs == rV.secs


def test_sweep_matches_numpy_oracle():
    rng = np.random.default_rng(2)
    L = 10
    mpo = dor.guess_bmpo(dor.drop_zero_channels(dor.xxz_field_dense(0.8, rng.uniform(-1, 1, L))),
                         dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(4, mpo, (0,), 3))
    trace = []
    dor.dmrg(mpo, [t.copy() for t in mps], dor.DmrgOptions(cutoff=1e-8, convergence_criterion=0.0,
                                                            maximum_bond=16, maximum_iterations=1), trace=trace)
    tm = [tc.from_oracle(t) for t in mpo]
    ts = [tc.from_oracle(t) for t in mps]
    env = tc.environments(tm, ts, 0)
    oc, step = 0, 1
    for rec in trace:
        E = tc.two_site_update(tm, ts, env, oc, step, 1e-8, 4, 16)
        assert abs(E - rec["E"]) <= 1e-10 * abs(rec["E"])
        oc += step
        if oc == 0 or oc == L - 2:
            step = -step
