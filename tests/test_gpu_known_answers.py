"""Known answers of the reference's own tests, and every C-ABI entry point
called directly, on the device.

* C<5> contraction (include/blockTensor/btensor.h:1532-1608): the output
  block set is exactly {(0,1,0), (1,0,0), (1,1,1)}, each block the dense
  tensordot of its one block pair, and the full trace is the sum of the
  block traces.
* 2x2 eigenproblem (include/dmrg.h:115-139): a matrix built from a random
  eigenbasis gives back its lowest eigenvalue and eigenvector up to sign;
  NaN inputs follow the reference's degenerate branch.
* Ising two-site update (include/dmrg.h:141-205): o^2 + n^2 = 1, the updated
  state has unit norm, <H> from contract(psi, psi, H), from the two-site
  matvec and a0 agree, and the Lanczos vector is orthogonal to theta.
* permute, reshape, conj, mul, add, full_dot, truncate, matvec and
  one_step_lanczos through the C ABI against the oracle.
* error paths of the boundary (bad block indices, inconsistent truncate
  inputs, tensors of two contexts, a tensor outliving its context).
"""
import math

import numpy as np
import pytest

from oracle import dmrg_oracle as dor
from oracle import btensor_oracle as bo
from oracle.btensor_oracle import BTensor as OB
from tests.helpers import (assert_same_structure, assert_values_close, from_dev, inv_leg, rand_btensor, rand_leg,
                           to_dev)

pytestmark = pytest.mark.gpu

C5 = (5,)


def _c5_pair(rng):
    A = OB.from_sections(C5, [[(2, (0,)), (3, (1,))], [(2, (0,)), (3, (4,))]], (0,))
    A.blocks[(0, 0)] = rng.uniform(size=(2, 2))
    A.blocks[(1, 1)] = rng.uniform(size=(3, 3))
    B = OB.from_sections(C5, [[(3, (4,)), (2, (0,))], [(2, (0,)), (3, (1,))], [(1, (1,)), (3, (0,))]], (1,))
    B.blocks[(1, 0, 0)] = rng.uniform(size=(2, 2, 1))
    B.blocks[(0, 1, 0)] = rng.uniform(size=(3, 3, 1))
    B.blocks[(1, 1, 1)] = rng.uniform(size=(2, 3, 3))
    return A, B


def test_c5_contraction_block_set(ctx):
    A, B = _c5_pair(np.random.default_rng(0))
    C = from_dev(to_dev(ctx, A).tensordot(to_dev(ctx, B), [1], [1]))
    assert sorted(C.blocks) == [(0, 1, 0), (1, 0, 0), (1, 1, 1)]
    exp = {(0, 1, 0): np.tensordot(A.blocks[(0, 0)], B.blocks[(1, 0, 0)], ([1], [1])),
           (1, 0, 0): np.tensordot(A.blocks[(1, 1)], B.blocks[(0, 1, 0)], ([1], [1])),
           (1, 1, 1): np.tensordot(A.blocks[(1, 1)], B.blocks[(1, 1, 1)], ([1], [1]))}
    for k, v in exp.items():
        assert np.abs(C.blocks[k] - v).max() <= 1e-14
    assert C.sel == (1,)
    tr = from_dev(to_dev(ctx, A).tensordot(to_dev(ctx, A.inverse_cvals()), [0, 1], [0, 1]))
    want = sum(float(np.sum(b * b)) for b in A.blocks.values())
    assert list(tr.blocks) == [()] and abs(float(tr.blocks[()]) - want) <= 1e-14 * want


def test_eig2x2_known_answer(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(4)
    for _ in range(20):
        angle = rng.uniform() * 2 * math.pi
        E0, E1 = -rng.uniform(), rng.uniform()
        p00, p01, p10, p11 = math.cos(angle), math.sin(angle), math.sin(angle), -math.cos(angle)
        a0 = p00 ** 2 * E0 + p10 ** 2 * E1
        a1 = p01 ** 2 * E0 + p11 ** 2 * E1
        b = p01 * p00 * E0 + p11 * p10 * E1
        E, o, n = qa.eig2x2(ctx, a0, a1, b)
        assert abs(E - E0) <= 1e-12
        assert min(abs(o - p00) + abs(n - p01), abs(o + p00) + abs(n + p01)) <= 1e-9
        assert (E, o, n) == pytest.approx(dor.eig2x2(a0, a1, b), abs=1e-14)


def test_eig2x2_degenerate_and_nan(ctx):
    """Degenerate branch (sources/dmrg.cpp:557-565) and NaN inputs.  A NaN
    anywhere makes E0 and o NaN, and the degenerate branch then sets o = 0,
    n = 1 before the NaN checks (:566-569), so the reference returns (NaN, 0,
    1) without throwing; the device does the same."""
    import quantit_amd as qa
    assert qa.eig2x2(ctx, -1.0, 0.0, 0.0) == pytest.approx(dor.eig2x2(-1.0, 0.0, 0.0), abs=0)
    assert qa.eig2x2(ctx, 0.5, 0.5, 0.0) == pytest.approx(dor.eig2x2(0.5, 0.5, 0.0), abs=0)
    for a0, a1, b in ((float("nan"), 0.0, 0.1), (0.0, float("nan"), 0.1), (0.2, 0.1, float("nan")),
                      (float("inf"), 0.0, 0.1)):
        E, o, n = qa.eig2x2(ctx, a0, a1, b)
        rE, ro, rn = dor.eig2x2(a0, a1, b)
        assert (o, n) == (ro, rn) == (0.0, 1.0)
        assert math.isnan(E) == math.isnan(rE)


def _ising_two_site():
    """The reference's Ising MPO (Sz Sz with +-1 entries, include/dmrg.h:145-166)
    on two sites: left edge row 2, right edge column 0."""
    w = np.zeros((3, 2, 3, 2))
    w[1, 0, 0, 0] = w[2, 0, 1, 0] = 1
    w[1, 1, 0, 1] = w[2, 1, 1, 1] = -1
    w[0, 0, 0, 0] = w[0, 1, 0, 1] = 1
    w[2, 0, 2, 0] = w[2, 1, 2, 1] = 1
    return [w[2:3], w[:, :, 0:1, :]]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ising_two_site_lanczos_invariants(ctx, seed):
    import quantit_amd as qa
    mpo = dor.guess_bmpo(dor.drop_zero_channels(_ising_two_site()), dor.u1_spin_half_phys())
    mps = dor.random_bmps(4, mpo, (0,), seed)
    nrm = dor.contract_norm(mps, mps)
    mps[0] = bo.scal(mps[0], 1.0 / math.sqrt(nrm))
    theta = bo.tensordot(mps[0], mps[1], [2], [0])
    W12 = dor.two_site_hamil(mpo)[0]
    env = dor.generate_env(mpo, mps, 0)
    dth, dW, dL, dR = (to_dev(ctx, t) for t in (theta, W12, env[-1], env[2]))
    Hpsi, a0, a1, b = qa.one_step_lanczos(dth, dW, dL, dR)
    E, o, n = qa.eig2x2(ctx, a0, a1, b)
    assert abs(o * o + n * n - 1.0) <= 1e-14
    upd = dth.mul(qa.BTensor.from_parts(ctx, (0,), [], [], (0,), {(): np.array(o)})).add(
        Hpsi.mul(qa.BTensor.from_parts(ctx, (0,), [], [], (0,), {(): np.array(n)})))
    assert abs(upd.full_dot(upd) - 1.0) <= 1e-14
    assert abs(dth.full_dot(dth) - 1.0) <= 1e-14
    assert abs(Hpsi.full_dot(dth)) <= 1e-14
    assert abs(Hpsi.full_dot(Hpsi)) <= 1e-14   # theta is an eigenstate: b = 0, psi1 = 0
    H_av = dor.contract_energy(mps, mps, mpo)
    H_upd = dth.full_dot(qa.hamil2site_times_state(dth, dW, dL, dR))
    assert abs(H_av - H_upd) <= 1e-14 and abs(H_av - a0) <= 1e-14
    a1_test = Hpsi.full_dot(qa.hamil2site_times_state(Hpsi, dW, dL, dR))
    assert abs(a1 - a1_test) <= 1e-14
    _, ra0, ra1, rb = dor.one_step_lanczos(theta, W12, env[-1], env[2])
    assert (a0, a1, b) == pytest.approx((ra0, ra1, rb), abs=1e-14)


def test_one_step_lanczos_matches_oracle(ctx):
    """A non-degenerate update (random-field XXZ, random state): a0, a1, b
    and the Lanczos vector against the oracle; the Lanczos vector is unit
    norm and orthogonal to theta."""
    import quantit_amd as qa
    rng = np.random.default_rng(9)
    L = 8
    mpo = dor.guess_bmpo(dor.drop_zero_channels(dor.xxz_field_dense(0.7, rng.uniform(-1, 1, L))),
                         dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(6, mpo, (0,), 3))
    env = dor.generate_env(mpo, mps, 0)
    W12 = dor.two_site_hamil(mpo)
    theta = bo.tensordot(mps[0], mps[1], [2], [0])
    rpsi, ra0, ra1, rb = dor.one_step_lanczos(theta, W12[0], env[-1], env[2])
    dev = [to_dev(ctx, t) for t in (theta, W12[0], env[-1], env[2])]
    Hpsi, a0, a1, b = qa.one_step_lanczos(*dev)
    assert abs(a0 - ra0) <= 1e-12 * abs(ra0) and abs(a1 - ra1) <= 1e-12 * abs(ra1)
    assert abs(b - rb) <= 1e-12 * abs(rb) and rb > 1e-3
    got = from_dev(Hpsi)
    assert_same_structure(got, rpsi)
    assert_values_close(got, rpsi, 1e-12)
    assert abs(Hpsi.full_dot(Hpsi) - 1.0) <= 1e-12
    assert abs(Hpsi.full_dot(dev[0])) <= 1e-12
    mv = from_dev(qa.hamil2site_times_state(*dev))
    ref = dor.matvec(theta, W12[0], env[-1], env[2])
    assert_same_structure(mv, ref)
    assert_values_close(mv, ref, 1e-12)


# ---- structural and elementwise entry points ---------------------------------
def _rank4(rng, groups=(0,), nsec=3, smax=4):
    legs = [rand_leg(rng, nsec, smax=smax, groups=groups) for _ in range(4)]
    return rand_btensor(rng, groups, legs, tuple(0 for _ in groups))


@pytest.mark.parametrize("perm", [(0, 1, 2, 3), (3, 2, 1, 0), (1, 3, 0, 2), (2, 0, 3, 1)])
def test_permute_entry(ctx, perm):
    t = _rank4(np.random.default_rng(1))
    got = from_dev(to_dev(ctx, t).permute(perm))
    ref = t.permute(list(perm))
    assert_same_structure(got, ref)
    assert_values_close(got, ref, 0.0)
    assert np.array_equal(got.to_dense(), np.transpose(t.to_dense(), perm))


@pytest.mark.parametrize("groups_", [[2], [1, 3], [1], [3]])
def test_reshape_entry(ctx, groups_):
    t = _rank4(np.random.default_rng(2))
    got = from_dev(to_dev(ctx, t).reshape(groups_))
    ref = t.reshape(groups_)
    assert_same_structure(got, ref)
    assert_values_close(got, ref, 0.0)


def test_conj_entry(ctx):
    t = _rank4(np.random.default_rng(3), groups=(0, 3))
    got = from_dev(to_dev(ctx, t).conj())
    ref = t.conj()
    assert_same_structure(got, ref)
    assert_values_close(got, ref, 0.0)


def test_mul_add_full_dot_entries(ctx):
    rng = np.random.default_rng(5)
    x = _rank4(rng)
    y = OB(x.groups, x.secs, x.cvals, x.sel, {})
    keys = sorted(x.blocks)
    for k in keys[::2]:   # partly overlapping block sets
        y.blocks[k] = rng.standard_normal(x.blocks[k].shape)
    extra = [k for k in x.all_allowed_indices() if k not in x.blocks][:2]
    for k in extra:
        y.blocks[k] = rng.standard_normal(y.block_shape(k))
    dx, dy = to_dev(ctx, x), to_dev(ctx, y)
    got = from_dev(dx.add(dy, -0.75))
    ref = bo.add(x, y, -0.75)
    assert_same_structure(got, ref)
    assert_values_close(got, ref, 1e-15)
    # broadcast mul by a rank-1 tensor along the last leg (as d on V)
    last = OB(x.groups, [x.secs[3]], [[bo.q_neutral(x.groups)] * len(x.secs[3])], bo.q_neutral(x.groups))
    for i, s in enumerate(x.secs[3]):
        last.blocks[(i,)] = rng.uniform(size=(s,))
    gm = from_dev(dx.mul(to_dev(ctx, last)))
    rm = bo.mul(x, last)
    assert_same_structure(gm, rm)
    assert_values_close(gm, rm, 1e-15)
    want, _ = bo.full_dot(x, y)
    assert abs(dx.full_dot(dy) - want) <= 1e-13 * max(1.0, abs(want))


def test_truncate_entry(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(6)
    rows = [(5, (0,)), (4, (1,)), (3, (-1,))]
    t = rand_btensor(rng, (0,), [rows, inv_leg((0,), rows)], (0,))
    U, d, V = qa.svd(to_dev(ctx, t), 1)
    rU, rd, rV = bo.svd_split(t, 1)
    for mx, tol in ((5, 1e-14), (3, 1e-14), (100, 0.5)):
        tU, td, tV = (from_dev(x) for x in qa.truncate(U, d, V, mx, 1, tol))
        ed, (eU, eV), _ = bo.truncate_impl(from_dev(d), [from_dev(U), from_dev(V)], mx, 1, tol)
        for a, b in ((tU, eU), (td, ed), (tV, eV)):
            assert_same_structure(a, b)
            assert_values_close(a, b, 0.0)


# ---- error paths of the boundary ---------------------------------------------
def test_create_rejects_bad_block_index(ctx):
    import quantit_amd as qa
    with pytest.raises(ValueError):
        qa.BTensor.from_parts(ctx, (0,), [[2, 3], [2, 3]], [[(0,), (1,)], [(0,), (-1,)]], (0,),
                              {(0, 2): np.zeros((2, 3))})
    with pytest.raises(ValueError):
        qa.BTensor.from_parts(ctx, (0,), [[2, 3], [2, 3]], [[(0,), (1,)], [(0,), (-1,)]], (0,),
                              {(-1, 0): np.zeros((2, 2))})


def test_truncate_rejects_inconsistent_inputs(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(7)
    rows = [(4, (0,)), (3, (1,))]
    t = to_dev(ctx, rand_btensor(rng, (0,), [rows, inv_leg((0,), rows)], (0,)))
    U, d, V = qa.svd(t, 1)
    with pytest.raises(ValueError):
        qa.truncate(U, U, V, 4, 1, 1e-10)     # d of rank 2
    other = to_dev(ctx, rand_btensor(rng, (0,), [[(4, (0,)), (3, (1,)), (2, (2,))], [(4, (0,))]], (0,)))
    with pytest.raises(ValueError):
        qa.truncate(other, d, V, 4, 1, 1e-10)  # U's last leg has other sections


def test_tensors_of_two_contexts_rejected(ctx):
    import quantit_amd as qa
    c2 = qa.Context(0, arena_bytes=1 << 28)
    rng = np.random.default_rng(8)
    x = _rank4(rng)
    a, b = to_dev(ctx, x), to_dev(c2, x)
    for op in (lambda: a.tensordot(b, [0], [0]), lambda: a.add(b), lambda: a.mul(b), lambda: a.full_dot(b)):
        with pytest.raises(ValueError):
            op()
    del a, b
    c2.close()


def test_tensor_outliving_its_context():
    """A handle kept past qtt_context_destroy: calls on it fail cleanly, and
    releasing it afterwards frees the context (no use after free)."""
    import quantit_amd as qa
    c = qa.Context(0, arena_bytes=1 << 28)
    t = to_dev(c, _rank4(np.random.default_rng(9)))
    c.close()
    with pytest.raises(ValueError):
        t.conj()
    with pytest.raises(ValueError):
        t.to_parts()
    del t
    c2 = qa.Context(0, arena_bytes=1 << 28)
    u = to_dev(c2, _rank4(np.random.default_rng(10)))
    assert abs(u.full_dot(u)) > 0
    del u
    c2.close()


def test_truncate_erasures_follow_the_unsorted_list(ctx):
    """The hand-traced case of the oracle test: erasing sectors 3 then 2 of
    U = [(0,3), (1,2), (3,3)] leaves [(0,3)] (the reference's
    remove_unit_blocks walks the list its previous call left unsorted)."""
    import quantit_amd as qa
    z = [(0,)] * 4
    ublocks = {k: np.ones((1, 1)) for k in ((0, 3), (1, 2), (3, 3))}
    U = qa.BTensor.from_parts(ctx, (0,), [[1] * 4, [1] * 4], [z, z], (0,), ublocks)
    V = qa.BTensor.from_parts(ctx, (0,), [[1] * 4, [1] * 4], [z, z], (0,), ublocks)
    d = qa.BTensor.from_parts(ctx, (0,), [[1] * 4], [z], (0,),
                              {(i,): np.array([v]) for i, v in enumerate((1.0, 0.9, 1e-9, 1e-9))})
    tU, td, tV = qa.truncate(U, d, V, 10 ** 9, 1, 1e-8)
    assert tU.block_list() == [(0, 3)] and tV.block_list() == [(0, 3)]
    assert td.block_list() == [(0,), (1,)]

