"""The library's own solve inputs on device: MPOs from quantit_amd.models,
random bMPS prepared by qtt_mps_prepare (normalise + move_oc), then DMRG.

* mps_prepare against the oracle's prepare_state on the same random state:
  sections and charges bit-exact, dense state equal up to a global sign,
  norm 1 and right-canonical sites.
* move_oc round trip leaves the dense state unchanged.
* ground-state energies against exact diagonalisation (chain, XXZ with
  random fields, 4x3 cylinder) within 1e-9 relative.
"""
import numpy as np
import pytest

from oracle import dmrg_oracle as dor
from oracle.btensor_oracle import BTensor as OB
from tests.helpers import assert_same_structure, from_dev

pytestmark = pytest.mark.gpu


def _dense_state(mps):
    T = mps[0].to_dense()[0]
    for A in mps[1:]:
        T = np.tensordot(T, A.to_dense(), ([T.ndim - 1], [0]))
    return T.reshape(-1)


def _oracle(parts):
    return [OB(p["groups"], p["secs"], p["cvals"], p["sel"], p["blocks"]) for p in parts]


def test_mps_prepare_matches_oracle(ctx):
    import quantit_amd as qa
    from quantit_amd import models
    parts = models.random_bmps_parts(10, 4, q=0, seed=11)
    dev = qa.mps_prepare([qa.BTensor.from_parts(ctx, **p) for p in parts])
    got = [from_dev(t) for t in dev]
    ref = dor.prepare_state(_oracle(parts))
    for g, r in zip(got, ref):
        assert_same_structure(g, r)
    a, b = _dense_state(got), _dense_state(ref)
    assert abs(abs(float(a @ b)) - 1.0) < 1e-12
    assert abs(qa.contract_norm(dev, dev) - 1.0) < 1e-12
    for A in got[1:]:   # right-canonical: sum_{s,r} A[l,s,r] A[l',s,r] = delta
        D = A.to_dense().reshape(A.to_dense().shape[0], -1)
        assert np.abs(D @ D.T - np.eye(D.shape[0])).max() < 1e-12


def test_mps_prepare_with_unit_sectors(ctx):
    """Regression: a one-element sector at the chain edge makes U and d of the
    move_oc SVD share a single section of size 1 (elementwise mul)."""
    import quantit_amd as qa
    from quantit_amd import models
    mps = qa.mps_prepare(models.random_bmps(ctx, 32, 4, q=0, seed=0))
    assert abs(qa.contract_norm(mps, mps) - 1.0) < 1e-12
    for A in [from_dev(t) for t in mps[-3:]]:
        D = A.to_dense().reshape(A.to_dense().shape[0], -1)
        assert np.abs(D @ D.T - np.eye(D.shape[0])).max() < 1e-12


def test_move_oc_round_trip(ctx):
    import quantit_amd as qa
    from quantit_amd import models
    mps = qa.mps_prepare(models.random_bmps(ctx, 9, 4, q=1, seed=2))
    before = _dense_state([from_dev(t) for t in mps])
    mps = qa.mps_move_oc(mps, 0, 6)
    mps = qa.mps_move_oc(mps, 6, 0)
    after = _dense_state([from_dev(t) for t in mps])
    assert abs(abs(float(before @ after)) - 1.0) < 1e-12
    with pytest.raises(ValueError):
        qa.mps_move_oc(mps, 0, 9)


def _ground(L, bonds, J=1.0, delta=1.0, fields=None):
    """Lowest energy in the Sz = 0 sector (basis element 0 of a site = up)."""
    from quantit_amd import models
    H = dor.mpo_to_dense_matrix(models.spin_mpo(L, bonds, J, delta, fields))
    ups = np.array([L - bin(n).count("1") for n in range(2 ** L)])
    sel = np.nonzero(2 * ups - L == 0)[0]
    return float(np.linalg.eigvalsh(H[np.ix_(sel, sel)])[0])


@pytest.mark.parametrize("case", ["chain12", "xxz_fields12", "cylinder4x3"])
def test_dmrg_on_library_inputs_matches_exact(ctx, case):
    import quantit_amd as qa
    from quantit_amd import models
    rng = np.random.default_rng(7)
    if case == "chain12":
        L, bonds, kw = 12, models.chain_bonds(12), {}
    elif case == "xxz_fields12":
        L, bonds, kw = 12, models.chain_bonds(12), {"delta": 0.7, "fields": rng.uniform(-0.5, 0.5, 12)}
    else:
        L, bonds, kw = 12, models.cylinder_bonds(4, 3), {}
    mpo = models.to_bmpo(ctx, models.spin_mpo(L, bonds, **kw))
    mps = qa.mps_prepare(models.random_bmps(ctx, L, 4, q=0, seed=1))
    opts = qa.DmrgOptions(cutoff=1e-12, convergence_criterion=1e-12, maximum_bond=128, maximum_iterations=12)
    E, mps, log = qa.dmrg(mpo, mps, opts)
    e0 = _ground(L, bonds, **kw)
    assert abs(E - e0) <= 1e-9 * abs(e0), (E, e0)
    assert abs(qa.contract_energy(mps, mps, mpo) - E) <= 1e-9 * abs(e0)
