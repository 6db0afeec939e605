"""Two-site DMRG on device against the oracle's restatement of
sources/dmrg.cpp (one-step Lanczos, block SVD truncation, environment
updates) on the same MPO and the same initial MPS.

* random-field XXZ chain (no exact Schmidt multiplets, tie-free cuts):
  per-sweep energies within 1e-10 relative, every MPS tensor's sections and
  charges bit-exact, final states equal up to a global sign (overlap).
* Heisenberg chains: energy against exact diagonalisation (the SU(2)
  multiplets make the kept counts at a cut depend on the last bit of the
  SVD, for the reference as well; energies are compared, not structure).
"""
import numpy as np
import pytest

from oracle import dmrg_oracle as dor
from tests.helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu


def _mps_dense(mps):
    """Dense state vector of an open MPS (edges of size 1)."""
    T = mps[0].to_dense()[0]  # (s, r)
    for A in mps[1:]:
        T = np.tensordot(T, A.to_dense(), ([T.ndim - 1], [0]))
    return T.reshape(-1)


def _setup(dense_mpo, seed, bond=4):
    mpo = dor.guess_bmpo(dor.drop_zero_channels(dense_mpo), dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(bond, mpo, (0,), seed))
    return mpo, mps


def test_dmrg_random_field_xxz_matches_oracle(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(3)
    L = 12
    dense = dor.xxz_field_dense(0.8, rng.uniform(-1.0, 1.0, L))
    mpo, mps = _setup(dense, seed=5)
    opts = dor.DmrgOptions(cutoff=1e-8, convergence_criterion=0.0, maximum_bond=24, minimum_bond=4,
                           maximum_iterations=3)
    E_ref, mps_ref, Es_ref = dor.dmrg(mpo, [t.copy() for t in mps], opts)
    dmpo = [to_dev(ctx, t) for t in mpo]
    dmps = [to_dev(ctx, t) for t in mps]
    E, new, log = qa.dmrg(dmpo, dmps, qa.DmrgOptions(cutoff=1e-8, convergence_criterion=0.0, maximum_bond=24,
                                                     minimum_bond=4, maximum_iterations=3))
    assert len(log["energies"]) == len(Es_ref)
    for a, b in zip(log["energies"], Es_ref):
        assert abs(a - b) <= 1e-10 * abs(b)
    out = [from_dev(t) for t in new]
    for a, b in zip(out, mps_ref):
        assert a.secs == b.secs and a.cvals == b.cvals and a.sel == b.sel
        assert sorted(a.blocks) == sorted(b.blocks)
    ov = abs(float(_mps_dense(out) @ _mps_dense(mps_ref)))
    assert abs(ov - 1.0) < 1e-9
    # exact ground state of the same Hamiltonian (variational upper bound)
    H = dor.mpo_to_dense_matrix(dense)
    sz = np.array([bin(i).count("1") for i in range(2 ** L)])  # down spins (element 1 = down)
    sec = np.where(sz == L // 2)[0]
    w = np.linalg.eigvalsh(H[np.ix_(sec, sec)])
    assert E >= w[0] - 1e-9


@pytest.mark.parametrize("L", [8, 10])
def test_dmrg_heisenberg_exact(ctx, L):
    import quantit_amd as qa
    dense = dor.heisenberg_dense(-1.0, L)
    mpo, mps = _setup(dense, seed=0)
    dmpo = [to_dev(ctx, t) for t in mpo]
    dmps = [to_dev(ctx, t) for t in mps]
    E, new, log = qa.dmrg(dmpo, dmps, qa.DmrgOptions(cutoff=1e-12, convergence_criterion=0.0, maximum_bond=64,
                                                     minimum_bond=4, maximum_iterations=8))
    H = dor.mpo_to_dense_matrix(dense)
    sz = np.array([bin(i).count("1") for i in range(2 ** L)])
    sec = np.where(sz == L // 2)[0]
    e0 = np.linalg.eigvalsh(H[np.ix_(sec, sec)])[0]
    assert abs(E - e0) <= 1e-10 * abs(e0)
    # the library's own <psi|H|psi> of the final state agrees
    Ec = qa.contract_energy(new, new, dmpo)
    nrm = qa.contract_norm(new, new)
    assert abs(Ec / nrm - e0) <= 1e-9 * abs(e0)
