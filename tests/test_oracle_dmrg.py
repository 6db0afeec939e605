"""Pins the DMRG side of the oracle: the 2x2 eigensolve known answer
(include/dmrg.h:115-139), the exact 3-site Heisenberg matrix
(include/models.h:62-73), the MPO builders against the reference's 4x8
cylinder MPO fixtures (tests/2dHeisenberg/mpotensor_*.txt, copied under
tests/golden/), energies against exact diagonalisation and the XX chain's
free-fermion closed form.  CPU only."""
import os

import numpy as np
import pytest

from oracle import dmrg_oracle as dor
from oracle.btensor_oracle import tensordot

GOLD = os.path.join(os.path.dirname(__file__), "golden", "2dHeisenberg")


def test_eig2x2_known_answer():
    rng = np.random.default_rng(0)
    for _ in range(20):
        angle = rng.random() * 2 * np.pi
        E0, E1 = -rng.random(), rng.random()
        p00, p01 = np.cos(angle), np.sin(angle)
        p10, p11 = np.sin(angle), -np.cos(angle)
        a0 = p00 ** 2 * E0 + p10 ** 2 * E1
        a1 = p01 ** 2 * E0 + p11 ** 2 * E1
        b = p01 * p00 * E0 + p11 * p10 * E1
        E, o, n = dor.eig2x2(a0, a1, b)
        assert abs(E - E0) < 1e-12
        assert (abs(o - p00) < 1e-8 and abs(n - p01) < 1e-8) or (abs(o + p00) < 1e-8 and abs(n + p01) < 1e-8)


def test_eig2x2_degenerate_branch():
    # b = 0 and a0 < a1: the current state is already the eigenvector: o = 1, n = 0
    E, o, n = dor.eig2x2(-1.0, 2.0, 0.0)
    assert E == -1.0 and abs(o - 1.0) < 1e-15 and n == 0.0
    # b = 0 and a1 < a0: zero overlap with the old state: o = 0, n = 1
    E, o, n = dor.eig2x2(2.0, -1.0, 0.0)
    assert E == -1.0 and o == 0.0 and n == 1.0


def test_heisenberg_three_site_matrix_known_answer():
    H = dor.mpo_to_dense_matrix(dor.heisenberg_impl(1.0, 3))
    ref = np.zeros((8, 8))
    ref[0, 0] = ref[2, 1] = ref[1, 2] = 2
    ref[4, 2] = ref[2, 4] = ref[5, 3] = ref[3, 5] = 2
    ref[7, 7] = ref[6, 5] = ref[5, 6] = 2
    ref[2, 2] = ref[5, 5] = -2
    assert np.array_equal(H, ref)


def _fixture_mpo():
    return [dor.parse_mpo_fixture(open(os.path.join(GOLD, f"mpotensor_{i}.txt")).read()) for i in range(1, 33)]


def _elem(mpo, sig, sigp):
    v = np.ones(1)
    for W, s, sp in zip(mpo, sig, sigp):
        v = v @ W[:, s, :, sp]
    return float(v[0])


# bond list of the fixture MPO: four 8-site rings (sites 8r..8r+7) coupled
# site by site, the last ring lacking its (24, 31) closing bond
FIXTURE_BONDS = sorted(
    [(8 * r + i, 8 * r + (i + 1) % 8) if i < 7 else (8 * r, 8 * r + 7) for r in range(4) for i in range(8)
     if not (r == 3 and i == 7)] + [(i, i + 8) for i in range(24)])


def test_fixture_geometry_and_builder_agree():
    """The FSA builder fed the fixture's bond list reproduces the fixture MPO's
    matrix elements (diagonal and exchange) on random configurations."""
    fx = _fixture_mpo()
    # element 0 of the fixture's physical index carries Sz = +1/2 (as pauli() sz[0,0] = +1)
    mine = dor.longrange_heisenberg_dense(32, FIXTURE_BONDS)
    rng = np.random.default_rng(1)
    for _ in range(200):
        s = rng.integers(0, 2, 32)
        assert abs(_elem(fx, s, s) - _elem(mine, s, s)) < 1e-12
        i, j = FIXTURE_BONDS[rng.integers(len(FIXTURE_BONDS))]
        if s[i] != s[j]:
            sp = s.copy()
            sp[i], sp[j] = s[j], s[i]
            assert abs(_elem(fx, s, sp) - 0.5) < 1e-12
            assert abs(_elem(mine, s, sp) - 0.5) < 1e-12
    # diagonal part identifies the bond set exactly (least squares over configurations)
    pairs = [(i, j) for i in range(32) for j in range(i + 1, 32)]
    rows, rhs = [], []
    for _ in range(1200):
        s = rng.integers(0, 2, 32)
        z = 1 - 2 * s
        rows.append([z[i] * z[j] for i, j in pairs])
        rhs.append(_elem(fx, s, s))
    sol = np.linalg.lstsq(np.array(rows, float), np.array(rhs), rcond=None)[0]
    assert sorted(p for k, p in enumerate(pairs) if abs(sol[k]) > 1e-8) == FIXTURE_BONDS
    assert np.allclose([sol[pairs.index(b)] for b in FIXTURE_BONDS], 0.25)


def test_fixture_charges_match_numsQNs():
    """Charge inference on the fixture MPO reproduces the per-index bond
    charges listed in numsQNs_16.txt (2Sz units)."""
    fx = _fixture_mpo()
    bmpo = dor.guess_bmpo(fx, [(1,), (-1,)])
    inferred = [c[0] for s, c in zip(bmpo[15].secs[2], bmpo[15].cvals[2]) for _ in range(s)]
    listed = [int(x) for x in open(os.path.join(GOLD, "numsQNs_16.txt")).read().splitlines()[3].split(",")]
    assert sorted(inferred) == sorted(listed) or sorted(inferred) == sorted(-x for x in listed)


def test_bmpo_roundtrip_and_energy_consistency():
    L = 8
    dense = dor.heisenberg_dense(-1.0, L)
    bmpo = dor.guess_bmpo(dor.drop_zero_channels(dense), dor.u1_spin_half_phys())
    H = dor.mpo_to_dense_matrix(dense)
    Hb = dor.mpo_to_dense_matrix([W.to_dense() for W in bmpo])
    assert np.allclose(H, Hb)
    mps = dor.prepare_state(dor.random_bmps(3, bmpo, (0,), 7))
    psi = mps[0].to_dense()[0]
    for A in mps[1:]:
        psi = np.tensordot(psi, A.to_dense(), ([psi.ndim - 1], [0]))
    psi = psi.reshape(-1)
    E_dense = float(psi @ H @ psi) / float(psi @ psi)
    assert abs(dor.contract_norm(mps, mps) - 1.0) < 1e-12
    assert abs(dor.contract_energy(mps, mps, bmpo) - E_dense) < 1e-12
    # <theta|H_eff|theta> at the centre equals the global energy (canonical form)
    env = dor.generate_env(bmpo, mps, 0)
    W12 = dor.two_site_hamil(bmpo)
    theta = tensordot(mps[0], mps[1], [2], [0])
    Ht = dor.matvec(theta, W12[0], env[-1], env[2])
    e = float(next(iter(tensordot(Ht, theta.conj(), [0, 1, 2, 3], [0, 1, 2, 3]).blocks.values())))
    assert abs(e - E_dense) < 1e-12


def _ed_sz0(dense, L):
    H = dor.mpo_to_dense_matrix(dense)
    sz = np.array([bin(i).count("1") for i in range(2 ** L)])
    sec = np.where(sz == L // 2)[0]
    return np.linalg.eigvalsh(H[np.ix_(sec, sec)])[0]


@pytest.mark.parametrize("name", ["heisenberg10", "cylinder4x3", "xx12"])
def test_dmrg_matches_exact(name):
    if name == "heisenberg10":
        L, dense = 10, dor.heisenberg_dense(-1.0, 10)
    elif name == "cylinder4x3":
        L, dense = 12, dor.longrange_heisenberg_dense(12, dor.cylinder_bonds(4, 3))
    else:
        L, dense = 12, dor.xxz_dense(0.0, 12)
    bmpo = dor.guess_bmpo(dor.drop_zero_channels(dense), dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(4, bmpo, (0,), 0))
    E, _, Es = dor.dmrg(bmpo, mps, dor.DmrgOptions(cutoff=1e-12, convergence_criterion=1e-13, maximum_bond=64,
                                                    maximum_iterations=12))
    e0 = _ed_sz0(dense, L)
    assert abs(E - e0) < 1e-9 * abs(e0)
    if name == "xx12":
        # free fermions with hopping 1/2 (S+S- + h.c.)/2: eps_k = cos(pi k/(L+1)), half filling
        eps = np.cos(np.pi * np.arange(1, L + 1) / (L + 1))
        assert abs(e0 - np.sort(eps)[:L // 2].sum()) < 1e-12


@pytest.mark.parametrize("oc", [9, 4])
def test_dmrg_from_other_centre_matches_exact(oc):
    """dmrg_impl from a centre other than 0 (sources/dmrg.cpp:224-236,
    :257-266): the sweeps converge to the exact ground state, and the centre
    ends where it started (one to the left for the last site)."""
    L, dense = 10, dor.heisenberg_dense(-1.0, 10)
    bmpo = dor.guess_bmpo(dor.drop_zero_channels(dense), dor.u1_spin_half_phys())
    mps, _ = dor.move_oc(dor.prepare_state(dor.random_bmps(4, bmpo, (0,), 1)), 0, oc)
    E, out, Es = dor.dmrg(bmpo, mps, dor.DmrgOptions(cutoff=1e-12, convergence_criterion=1e-13, maximum_bond=64,
                                                      maximum_iterations=12), oc=oc)
    e0 = _ed_sz0(dense, L)
    assert abs(E - e0) < 1e-9 * abs(e0)
    # the sweeps end on the start position; after a final backward update the
    # singular values sit one site to the right of it (the reference's own
    # comment at sources/dmrg.cpp:259-260)
    end = oc - 1 if oc == L - 1 else oc
    centre = end if oc in (0, L - 1) else end + 1
    for i in range(centre + 1, L):   # right-canonical beyond the centre
        D = out[i].to_dense().reshape(out[i].to_dense().shape[0], -1)
        assert np.abs(D @ D.T - np.eye(D.shape[0])).max() < 1e-10
    for i in range(centre):          # left-canonical before it
        D = out[i].to_dense()
        M = D.reshape(-1, D.shape[2])
        assert np.abs(M.T @ M - np.eye(M.shape[1])).max() < 1e-10


def _random_update(L=8, seed=3, bond=6):
    """theta, W12, environments at the left edge of a random-field XXZ chain
    (a non-degenerate two-site problem)."""
    rng = np.random.default_rng(seed)
    mpo = dor.guess_bmpo(dor.drop_zero_channels(dor.xxz_field_dense(0.7, rng.uniform(-1, 1, L))),
                         dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(bond, mpo, (0,), seed))
    env = dor.generate_env(mpo, mps, 0)
    W12 = dor.two_site_hamil(mpo)
    theta = tensordot(mps[0], mps[1], [2], [0])
    return theta, W12[0], env[-1], env[2]


def test_lanczos_two_vectors_is_the_one_step_update():
    """k-step Lanczos with two Krylov vectors: Ritz value = eig2x2's E of the
    one-step form (sources/dmrg.cpp:543-608), Ritz vector = o theta + n psi1."""
    from oracle.btensor_oracle import add as badd, scal as bscal
    theta, W, Le, Re = _random_update()
    E1, new1, _ = dor.two_sites_update(theta, W, Le, Re)
    E2, new2, nv = dor.lanczos_ground(theta, W, Le, Re, 2, 0.0)
    assert nv == 2
    assert abs(E1 - E2) <= 1e-12 * abs(E1)
    assert sorted(new1.blocks) == sorted(new2.blocks)
    diff = badd(new1, new2, -1.0)
    assert np.sqrt(sum(float(np.sum(b * b)) for b in diff.blocks.values())) <= 1e-10


def test_lanczos_reaches_dense_two_site_ground_state():
    """The dense effective Hamiltonian of the two-site problem, built by
    applying the matvec to every unit vector of theta's blocks: its lowest
    eigenvalue equals the converged Ritz value, and the Ritz vector is its
    eigenvector (residual below tolerance)."""
    theta, W, Le, Re = _random_update()
    E, vec, nv = dor.lanczos_ground(theta, W, Le, Re, 40, 1e-12)
    keys = sorted(theta.blocks)
    sizes = [theta.blocks[k].size for k in keys]
    n = sum(sizes)
    H = np.zeros((n, n))
    col = 0
    for k, sz in zip(keys, sizes):
        for e in range(sz):
            u = theta.copy()
            u.blocks = {kk: np.zeros_like(v) for kk, v in theta.blocks.items()}
            u.blocks[k].reshape(-1)[e] = 1.0
            hu = dor.matvec(u, W, Le, Re)
            H[:, col] = np.concatenate([hu.blocks[kk].reshape(-1) if kk in hu.blocks else np.zeros(s)
                                        for kk, s in zip(keys, sizes)])
            col += 1
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
    e0 = np.linalg.eigvalsh(H)[0]
    assert abs(E - e0) <= 1e-12 * abs(e0)
    assert set(vec.blocks) == set(keys) or set(vec.blocks) >= set(keys)
    x = np.concatenate([vec.blocks[k].reshape(-1) for k in keys])
    assert np.linalg.norm(H @ x - e0 * x) <= 1e-10 * abs(e0)


@pytest.mark.parametrize("name", ["heisenberg10", "xx12"])
def test_dmrg_with_lanczos_matches_exact(name):
    if name == "heisenberg10":
        L, dense = 10, dor.heisenberg_dense(-1.0, 10)
    else:
        L, dense = 12, dor.xxz_dense(0.0, 12)
    bmpo = dor.guess_bmpo(dor.drop_zero_channels(dense), dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(4, bmpo, (0,), 0))
    E, _, Es = dor.dmrg(bmpo, mps, dor.DmrgOptions(cutoff=1e-12, convergence_criterion=1e-12, maximum_bond=64,
                                                    maximum_iterations=8, max_krylov=30, lanczos_tol=1e-12))
    e0 = _ed_sz0(dense, L)
    assert abs(E - e0) < 1e-10 * abs(e0)
    assert len(Es) <= 5   # converged Lanczos: the energy settles within a few sweeps
