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


@pytest.mark.parametrize("oc", [9, 4])
def test_dmrg_from_other_centre_matches_oracle(ctx, oc):
    """qtt_dmrg_from_center against the oracle's dmrg_impl from the same
    centre (sources/dmrg.cpp:224-236, :257-266): per-sweep energies, the
    final state's structure bit-exact and its dense vector equal up to sign."""
    import quantit_amd as qa
    rng = np.random.default_rng(11)
    L = 10
    dense = dor.xxz_field_dense(0.9, rng.uniform(-1.0, 1.0, L))
    mpo, mps = _setup(dense, seed=6)
    mps, _ = dor.move_oc(mps, 0, oc)
    opts = dor.DmrgOptions(cutoff=1e-8, convergence_criterion=0.0, maximum_bond=20, minimum_bond=4,
                           maximum_iterations=3)
    E_ref, mps_ref, Es_ref = dor.dmrg(mpo, [t.copy() for t in mps], opts, oc=oc)
    E, new, log = qa.dmrg([to_dev(ctx, t) for t in mpo], [to_dev(ctx, t) for t in mps],
                          qa.DmrgOptions(cutoff=1e-8, convergence_criterion=0.0, maximum_bond=20, minimum_bond=4,
                                         maximum_iterations=3), oc=oc)
    for a, b in zip(log["energies"], Es_ref):
        assert abs(a - b) <= 1e-10 * abs(b)
    out = [from_dev(t) for t in new]
    for a, b in zip(out, mps_ref):
        assert a.secs == b.secs and a.cvals == b.cvals and sorted(a.blocks) == sorted(b.blocks)
    ov = abs(float(_mps_dense(out) @ _mps_dense(mps_ref)))
    assert abs(ov - 1.0) < 1e-9


def _lanczos_ctx(ctx, k, tol):
    class Mode:
        def __enter__(self):
            ctx.set_lanczos(k, tol)

        def __exit__(self, *exc):
            ctx.set_lanczos(0)
    return Mode()


def test_lanczos_two_vectors_equals_one_step(ctx):
    """k-step Lanczos with two Krylov vectors is the reference's one-step
    update: the same per-sweep energies and the same MPS structure as the
    default mode."""
    import quantit_amd as qa
    rng = np.random.default_rng(12)
    L = 10
    mpo, mps = _setup(dor.xxz_field_dense(0.8, rng.uniform(-1.0, 1.0, L)), seed=2)
    opts = qa.DmrgOptions(cutoff=1e-8, convergence_criterion=0.0, maximum_bond=20, minimum_bond=4,
                          maximum_iterations=3)
    E1, m1, l1 = qa.dmrg([to_dev(ctx, t) for t in mpo], [to_dev(ctx, t) for t in mps], opts)
    with _lanczos_ctx(ctx, 2, 0.0):
        E2, m2, l2 = qa.dmrg([to_dev(ctx, t) for t in mpo], [to_dev(ctx, t) for t in mps], opts)
    for a, b in zip(l1["energies"], l2["energies"]):
        assert abs(a - b) <= 1e-12 * abs(a)
    for a, b in zip(m1, m2):
        ga, gb = from_dev(a), from_dev(b)
        assert ga.secs == gb.secs and sorted(ga.blocks) == sorted(gb.blocks)


def test_dmrg_lanczos_matches_oracle_and_exact(ctx):
    """k-step mode (up to 24 vectors, residual 1e-12) against the oracle's
    restatement on the same inputs, sweep by sweep, and against exact
    diagonalisation; each update builds more than two vectors."""
    import quantit_amd as qa
    rng = np.random.default_rng(13)
    L = 10
    dense = dor.xxz_field_dense(0.6, rng.uniform(-1.0, 1.0, L))
    mpo, mps = _setup(dense, seed=4)
    o = dict(cutoff=1e-10, convergence_criterion=0.0, maximum_bond=32, minimum_bond=4, maximum_iterations=3)
    E_ref, _, Es_ref = dor.dmrg(mpo, [t.copy() for t in mps],
                                dor.DmrgOptions(**o, max_krylov=24, lanczos_tol=1e-12))
    before = ctx.lanczos_stats()
    with _lanczos_ctx(ctx, 24, 1e-12):
        E, new, log = qa.dmrg([to_dev(ctx, t) for t in mpo], [to_dev(ctx, t) for t in mps], qa.DmrgOptions(**o))
    after = ctx.lanczos_stats()
    print("device", log["energies"], "oracle", Es_ref, "vectors/update",
          (after["vectors"] - before["vectors"]) / max(1, after["updates"] - before["updates"]), flush=True)
    for a, b in zip(log["energies"], Es_ref):
        assert abs(a - b) <= 1e-10 * abs(b)
    assert after["vectors"] - before["vectors"] > 2 * (after["updates"] - before["updates"])
    H = dor.mpo_to_dense_matrix(dense)
    sz = np.array([bin(i).count("1") for i in range(2 ** L)])
    sec = np.where(sz == L // 2)[0]
    e0 = np.linalg.eigvalsh(H[np.ix_(sec, sec)])[0]
    assert abs(E - e0) <= 1e-10 * abs(e0)


def test_lanczos_nan_in_state_raises(ctx):
    """A NaN in the two-site state reaches the k-step solve's tridiagonal
    step, whose flag comes back with the step's scalars: the update fails
    with status 4 (std::logic_error), as the one-step form's NaN check
    (sources/dmrg.cpp:566-569) makes it fail."""
    import quantit_amd as qa
    mpo, mps = _setup(dor.xxz_dense(1.0, 6), seed=2)
    key = sorted(mps[1].blocks)[0]
    blk = mps[1].blocks[key].copy()
    blk.flat[0] = np.nan
    mps[1].blocks[key] = blk
    o = dict(cutoff=1e-10, convergence_criterion=0.0, maximum_bond=16, minimum_bond=4, maximum_iterations=1)
    with _lanczos_ctx(ctx, 8, 1e-12):
        with pytest.raises(RuntimeError, match="nan"):
            qa.dmrg([to_dev(ctx, t) for t in mpo], [to_dev(ctx, t) for t in mps], qa.DmrgOptions(**o))


def test_lanczos_chain_L32_D128_converges(ctx):
    """The L=32, D=128 chain in k-step mode reaches |dE/E| <= 1e-10 between
    sweeps within 6 sweeps, below the one-step trajectory's sixth-sweep
    energy."""
    import quantit_amd as qa
    from quantit_amd import models
    L = 32
    mpo_parts = models.bmpo_parts(models.spin_mpo(L, models.chain_bonds(L)))
    mps_parts = models.random_bmps_parts(L, 4, q=0, seed=0)
    opts = qa.DmrgOptions(cutoff=1e-10, convergence_criterion=1e-10, maximum_bond=128, minimum_bond=4,
                          maximum_iterations=6)

    def run():
        mpo = [qa.BTensor.from_parts(ctx, **p) for p in mpo_parts]
        mps = qa.mps_prepare([qa.BTensor.from_parts(ctx, **p) for p in mps_parts])
        return qa.dmrg(mpo, mps, opts)

    E1, _, l1 = run()
    with _lanczos_ctx(ctx, 32, 1e-12):
        E2, _, l2 = run()
    es = l2["energies"]
    print("one-step", l1["energies"], "k-step", es, flush=True)
    assert len(es) < 6 or abs((es[-1] - es[-2]) / es[-1]) <= 1e-10
    assert abs((es[-1] - es[-2]) / es[-1]) <= 1e-10
    assert E2 <= E1 + 1e-12 * abs(E1)


def test_lanczos_graph_replays_equal_eager(ctx):
    """k-step mode with each update's matvec captured as a HIP graph and
    replayed for the later Krylov vectors (device work and host reads between
    the replays) against the same solve with every matvec run eagerly: the
    per-sweep energies and the final MPS are the same bits, and replays
    happened."""
    import quantit_amd as qa
    rng = np.random.default_rng(14)
    L = 10
    mpo, mps = _setup(dor.xxz_field_dense(0.6, rng.uniform(-1.0, 1.0, L)), seed=7)
    opts = qa.DmrgOptions(cutoff=1e-10, convergence_criterion=0.0, maximum_bond=32, minimum_bond=4,
                          maximum_iterations=3)
    runs = []
    try:
        for graphs in (False, True):
            ctx.set_graphs(graphs)
            before = ctx.lanczos_stats()["graph_replays"]
            with _lanczos_ctx(ctx, 24, 1e-12):
                E, new, lg = qa.dmrg([to_dev(ctx, t) for t in mpo], [to_dev(ctx, t) for t in mps], opts)
            runs.append((lg["energies"], [from_dev(t) for t in new], ctx.lanczos_stats()["graph_replays"] - before))
    finally:
        ctx.set_graphs(True)
    (e0, m0, r0), (e1, m1, r1) = runs
    assert r0 == 0 and r1 > 0
    assert e0 == e1
    for a, b in zip(m0, m1):
        assert sorted(a.blocks) == sorted(b.blocks)
        for k in a.blocks:
            assert np.array_equal(a.blocks[k], b.blocks[k])


@pytest.mark.parametrize("q", [0, 2])
def test_dmrg_random_start_matches_oracle(ctx, q):
    """dmrg(bMPO&, qnum, options) (sources/dmrg.cpp:109-118): the random
    start lies in the requested sector (every nonzero amplitude of the dense
    state has total 2Sz = q), and the sweeps from it agree with the oracle's
    dmrg on the same start (per-sweep energies within 1e-10 relative, MPS
    structure bit-exact)."""
    import quantit_amd as qa
    from oracle.btensor_oracle import BTensor as OT
    from quantit_amd import models
    rng = np.random.default_rng(11)
    L = 10
    mpo = dor.guess_bmpo(dor.drop_zero_channels(dor.xxz_field_dense(0.9, rng.uniform(-1.0, 1.0, L))),
                         dor.u1_spin_half_phys())
    phys = [[(n, (-c[0],)) for n, c in zip(W.secs[3], W.cvals[3])] for W in mpo]
    parts = models.random_bmps_general(4, phys, (q,), 9, (0,))
    start = [OT(p["groups"], p["secs"], p["cvals"], p["sel"], p["blocks"]) for p in parts]
    amp = _mps_dense(start)
    tot = np.array([L - 2 * bin(i).count("1") for i in range(2 ** L)])
    assert np.abs(amp).max() > 0 and np.all(amp[tot != q] == 0.0)
    kw = dict(cutoff=1e-8, convergence_criterion=0.0, maximum_bond=16, minimum_bond=4, maximum_iterations=3)
    E_ref, mps_ref, Es_ref = dor.dmrg(mpo, [t.copy() for t in start], dor.DmrgOptions(**kw))
    dmpo = [to_dev(ctx, t) for t in mpo]
    E, new, log = qa.dmrg(dmpo, q, qa.DmrgOptions(**kw), seed=9)
    assert len(log["energies"]) == len(Es_ref)
    for a, b in zip(log["energies"], Es_ref):
        assert abs(a - b) <= 1e-10 * abs(b)
    for a, b in zip((from_dev(t) for t in new), mps_ref):
        assert a.secs == b.secs and a.cvals == b.cvals and a.sel == b.sel
        assert sorted(a.blocks) == sorted(b.blocks)


def test_dmrg_logger_hooks(ctx):
    """dmrg with a dmrg_logger (include/dmrg_logger.h:27-108, called as
    sources/dmrg.cpp:239-268 does): init once with the options, it_log_all
    after every sweep with the sweep's energy and the live state (usable
    through the ABI during the call: <H> of that state equals the sweep's
    energy up to the truncation), end_log_all once at the end; the energies
    are the same bits as without a logger, and an exception raised by a hook
    comes back after dmrg returns."""
    import quantit_amd as qa
    rng = np.random.default_rng(6)
    L = 10
    dense = dor.xxz_field_dense(0.8, rng.uniform(-1.0, 1.0, L))
    mpo, mps = _setup(dense, seed=8)
    kw = dict(cutoff=1e-10, convergence_criterion=1e-9, maximum_bond=32, minimum_bond=4, maximum_iterations=8)
    dmpo = [to_dev(ctx, t) for t in mpo]
    E_plain, _, log_plain = qa.dmrg(dmpo, [to_dev(ctx, t) for t in mps], qa.DmrgOptions(**kw))

    class Rec(qa.DmrgLogger):
        def __init__(self):
            self.calls = []

        def init(self, options):
            self.calls.append(("init", options.maximum_iterations))

        def it_log_all(self, it, E, state):
            nrm = qa.contract_norm(state, state)
            h = qa.contract_energy(state, state, dmpo) / nrm
            self.calls.append(("it", it, E, h, max(qa.bond_dims(state))))

        def end_log_all(self, it, E, state):
            self.calls.append(("end", it, E, len(state)))

    rec = Rec()
    E, new, log = qa.dmrg(dmpo, [to_dev(ctx, t) for t in mps], qa.DmrgOptions(**kw), logger=rec)
    its = [c for c in rec.calls if c[0] == "it"]
    assert rec.calls[0] == ("init", 8) and rec.calls[-1][0] == "end"
    assert [c[1] for c in its] == list(range(len(its)))
    assert [c[2] for c in its] == log_plain["energies"] and E == E_plain
    for c in its:
        assert abs(c[3] - c[2]) <= 1e-7 * abs(c[2]), c
    assert rec.calls[-1] == ("end", len(its) - 1, E, L)
    assert len(new) == L and abs(qa.contract_norm(new, new) - 1.0) < 1e-10

    class Boom(qa.DmrgLogger):
        def it_log_all(self, it, E, state):
            raise KeyError("from the hook")

    with pytest.raises(KeyError):
        qa.dmrg(dmpo, [to_dev(ctx, t) for t in mps], qa.DmrgOptions(**kw), logger=Boom())
