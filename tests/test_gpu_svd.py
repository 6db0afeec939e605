"""Parity of the device block SVD / eigh / truncation with the oracle's
restatement of btensor_linalg.cpp.  Structure (sectors, charges, kept sizes,
block sets) bit-exact; singular values within 1e-12 of the largest singular
value; U and V compared through U.diag(d).V^T and projectors (gauge-free);
orthonormality within 1e-12."""
import numpy as np
import pytest

from oracle.btensor_oracle import (BTensor, eigh_split, svd_split, svd_truncated, tensordot, mul, rand_like)
from tests.helpers import assert_same_structure, from_dev, inv_leg, rand_btensor, rand_leg, to_dev

pytestmark = pytest.mark.gpu

G1 = (0,)


def _rank2(rng, nrow, ncol, smax, sel=(0,)):
    rows = rand_leg(rng, nrow, lo=-4, hi=4, smax=smax)
    cols = rand_leg(rng, ncol, lo=-4, hi=4, smax=smax)
    return rand_btensor(rng, G1, [rows, cols], sel)


def _check_svd(A, U, d, V, split):
    """U.diag(d).V^T == A and column orthonormality, in dense form."""
    Ad = A.to_dense()
    m = int(np.prod(Ad.shape[:split]))
    Ud = U.to_dense().reshape(m, -1)
    Vd = V.to_dense().reshape(-1, Ud.shape[1])
    dd = np.concatenate([d.blocks[(i,)] if (i,) in d.blocks else np.zeros(d.secs[0][i])
                         for i in range(len(d.secs[0]))])
    rec = (Ud * dd) @ Vd.T
    scale = max(np.linalg.norm(Ad), 1e-300)
    assert np.linalg.norm(rec - Ad.reshape(m, -1)) / scale < 1e-12
    keep = dd > dd.max() * 1e-10 if dd.size else dd
    Uk = Ud[:, keep]
    Vk = Vd[:, keep]
    assert np.abs(Uk.T @ Uk - np.eye(Uk.shape[1])).max() < 1e-12
    assert np.abs(Vk.T @ Vk - np.eye(Vk.shape[1])).max() < 1e-12
    return dd


@pytest.mark.parametrize("seed,nrow,ncol,smax", [(0, 4, 4, 6), (1, 6, 3, 12), (2, 3, 7, 20), (3, 5, 5, 90)])
def test_svd_rank2_matches_oracle(ctx, seed, nrow, ncol, smax):
    import quantit_amd as qa
    rng = np.random.default_rng(seed)
    A = _rank2(rng, nrow, ncol, smax)
    rU, rd, rV = svd_split(A, 1)
    U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1))
    assert_same_structure(d, rd)
    assert_same_structure(U, rU)
    assert_same_structure(V, rV)
    smax_all = max(float(b.max()) for b in rd.blocks.values())
    for k in rd.blocks:
        assert np.abs(d.blocks[k] - rd.blocks[k]).max() <= 1e-12 * smax_all
    dd = _check_svd(A, U, d, V, 1)
    # singular values equal those of the dense matrix (library routine)
    s_dense = np.linalg.svd(A.to_dense(), compute_uv=False)
    top = np.sort(dd)[::-1][:len(s_dense)]
    assert np.abs(top - s_dense[:len(top)]).max() <= 1e-12 * s_dense[0]


def test_svd_rank4_split2(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(5)
    l = rand_leg(rng, 4, smax=8)
    s = [(1, (1,)), (1, (-1,))]
    r = rand_leg(rng, 4, smax=8)
    theta = rand_btensor(rng, G1, [l, s, s, r], (0,))
    rU, rd, rV = svd_split(theta, 2)
    U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, theta), 2))
    assert_same_structure(U, rU)
    assert_same_structure(V, rV)
    assert_same_structure(d, rd)
    _check_svd(theta, U, d, V, 2)
    # projector onto the kept left space equals the oracle's
    Pu = U.to_dense().reshape(-1, sum(U.secs[2]))
    Pr = rU.to_dense().reshape(-1, sum(rU.secs[2]))
    assert np.abs(Pu @ Pu.T - Pr @ Pr.T).max() < 1e-11


def _spectral(rng, legs, sel, decay):
    """Rank-2 tensor whose sector spectra are well separated geometric series
    (no ties near any cut), built as U diag(s) V^T per sector."""
    A = rand_btensor(rng, G1, legs, sel)
    U, d, V = svd_split(A, 1)
    sc = 1.0
    for k in sorted(d.blocks):
        n = d.blocks[k].shape[0]
        d.blocks[k] = sc * decay ** np.arange(n) * (1 + 0.01 * rng.random())
        sc *= 0.37
    return tensordot(mul(U, d), V.conj(), [1], [1])


@pytest.mark.parametrize("tol,mn,mx", [(1e-3, 1, 10 ** 9), (1e-6, 1, 7), (1e-2, 5, 10 ** 9), (1e-8, 2, 3)])
def test_truncated_svd_structure_exact(ctx, tol, mn, mx):
    import quantit_amd as qa
    rng = np.random.default_rng(21)
    rows = rand_leg(rng, 5, lo=-4, hi=4, smax=6)
    A = _spectral(rng, [rows, inv_leg(G1, rows)], (0,), 0.3)
    rU, rd, rV = svd_truncated(A, 1, tol, mn, mx)
    U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1, tol=tol, min_size=mn, max_size=mx))
    assert_same_structure(d, rd)
    assert_same_structure(U, rU)
    assert_same_structure(V, rV)
    for k in rd.blocks:
        assert np.abs(d.blocks[k] - rd.blocks[k]).max() <= 1e-12


def test_svd_wide_sectors_gram_path(ctx):
    """Sectors much longer than wide go through the Gram preconditioning;
    both orientations (m << n and m >> n) and a graded spectrum."""
    import quantit_amd as qa
    rng = np.random.default_rng(51)
    rows = [(20, (1,)), (30, (-1,)), (5, (3,))]
    cols = [(100, (-1,)), (90, (-1,)), (120, (1,)), (70, (1,)), (40, (-3,))]
    for legs in ([rows, cols], [cols, rows]):
        A = rand_btensor(rng, G1, legs, (0,))
        # graded singular values down to 1e-12 relative
        for k in A.blocks:
            b = A.blocks[k]
            A.blocks[k] = b * np.logspace(0, -12, b.shape[1])[None, :] if b.shape[1] < b.shape[0] else \
                b * np.logspace(0, -12, b.shape[0])[:, None]
        rU, rd, rV = svd_split(A, 1)
        U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1))
        assert_same_structure(d, rd)
        assert_same_structure(U, rU)
        smax_all = max(float(b.max()) for b in rd.blocks.values())
        for k in rd.blocks:
            assert np.abs(d.blocks[k] - rd.blocks[k]).max() <= 1e-12 * smax_all
        _check_svd(A, U, d, V, 1)


def test_svd_contract_svd_workload_small(ctx):
    """The benchmark's contraction + SVD at D=64 against the oracle."""
    import quantit_amd as qa
    from oracle.btensor_oracle import BTensor as OB
    from quantit_amd import workloads
    X, A, dx, da, split = workloads.contract_svd(64)
    Xo, Ao = OB(X["groups"], X["secs"], X["cvals"], X["sel"], X["blocks"]), OB(A["groups"], A["secs"], A["cvals"],
                                                                                A["sel"], A["blocks"])
    Y = tensordot(Xo, Ao, dx, da)
    rU, rd, rV = svd_truncated(Y, split, 1e-10)
    y = qa.BTensor.from_parts(ctx, **X).tensordot(qa.BTensor.from_parts(ctx, **A), dx, da)
    U, d, V = (from_dev(x) for x in qa.svd(y, split, tol=1e-10))
    assert_same_structure(d, rd)
    assert_same_structure(U, rU)
    assert_same_structure(V, rV)
    smax_all = max(float(b.max()) for b in rd.blocks.values())
    for k in rd.blocks:
        assert np.abs(d.blocks[k] - rd.blocks[k]).max() <= 1e-12 * smax_all
    _check_svd(Y, U, d, V, split)


def test_svd_rank_deficient_orthonormal(ctx):
    """Sectors with exactly zero singular values: U and V stay orthonormal."""
    import quantit_amd as qa
    rng = np.random.default_rng(31)
    rows = rand_leg(rng, 3, smax=8)
    A = rand_btensor(rng, G1, [rows, inv_leg(G1, rows)], (0,))
    for k in A.blocks:
        b = A.blocks[k]
        if b.shape[0] > 1:
            b[-1] = b[0]  # duplicated row: rank deficient
    U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1))
    for t in (U, V):
        M = t.to_dense()
        assert np.abs(M.T @ M - np.eye(M.shape[1])).max() < 1e-12


@pytest.mark.parametrize("keep_all", [False, True])
def test_truncated_svd_completes_kept_null_columns(ctx, keep_all):
    """The truncating svd completes null columns only after the cut and only
    those it keeps: its U, d, V equal svd followed by truncate bit for bit,
    both when the null columns are dropped (min_size 1) and when min_size
    keeps every column; kept U and V are orthonormal and U d V^T = A."""
    import quantit_amd as qa
    rng = np.random.default_rng(37)
    rows = rand_leg(rng, 4, smax=24)
    A = rand_btensor(rng, G1, [rows, inv_leg(G1, rows)], (0,))
    for k in A.blocks:
        b = A.blocks[k]
        if b.shape[0] > 2:
            b[-1] = b[0]          # two null directions per sector
            b[-2] = 2.0 * b[1]
    total = sum(min(b.shape) for b in A.blocks.values())
    mins = total if keep_all else 1
    t = to_dev(ctx, A)
    fused = [from_dev(x) for x in qa.svd(t, 1, 1e-10, mins, 2 ** 40, 2.0)]
    U0, d0, V0 = qa.svd(t, 1)
    split = [from_dev(x) for x in qa.truncate(U0, d0, V0, 2 ** 40, mins, 1e-10, 2.0)]
    for a, b in zip(fused, split):
        assert_same_structure(a, b)
        for k in a.blocks:
            assert np.array_equal(a.blocks[k], b.blocks[k])
    U, d, V = fused
    kept = sum(b.size for b in d.blocks.values())
    assert kept == total if keep_all else kept < total
    _check_svd(A, U, d, V, 1)
    for x in (U, V):
        M = x.to_dense()
        assert np.abs(M.T @ M - np.eye(M.shape[1])).max() < 1e-12


def test_eigh_matches_dense(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(41)
    rows = [(5, (0,)), (3, (2,)), (7, (-2,)), (40, (4,))]
    A = rand_btensor(rng, G1, [rows, inv_leg(G1, rows)], (0,))
    for k in A.blocks:
        A.blocks[k] = A.blocks[k] + A.blocks[k].T
    rd, rU = eigh_split(A, 1)
    d, U = (from_dev(x) for x in qa.eigh(to_dev(ctx, A), 1))
    assert_same_structure(d, rd)
    assert_same_structure(U, rU)
    nrm = np.linalg.norm(A.to_dense(), 2)
    for k in rd.blocks:
        assert np.abs(d.blocks[k] - rd.blocks[k]).max() <= 1e-12 * nrm
    Ud = U.to_dense()
    dd = np.concatenate([d.blocks[(i,)] for i in range(len(d.secs[0]))])
    assert np.abs(A.to_dense() @ Ud - Ud * dd).max() <= 1e-11 * nrm


@pytest.mark.parametrize("m,n,decades", [(300, 200, 12), (200, 300, 8), (260, 150, 3)])
def test_svd_large_graded_sector(ctx, m, n, decades):
    """One sector wider than the one-work-group eigen-solver takes (k > 140),
    with a graded spectrum like a DMRG two-site tensor: singular values within
    1e-12 of the largest against the dense SVD (library routine), U and V
    orthonormal, reconstruction within 1e-12."""
    import quantit_amd as qa
    rng = np.random.default_rng(m + n)
    k = min(m, n)
    U0, _ = np.linalg.qr(rng.standard_normal((m, k)))
    V0, _ = np.linalg.qr(rng.standard_normal((n, k)))
    s = 10.0 ** (-decades * np.arange(k) / k)
    A = BTensor(G1, [[m], [n]], [[(0,)], [(0,)]], (0,))
    A.blocks[(0, 0)] = (U0 * s) @ V0.T
    U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1))
    dd = _check_svd(A, U, d, V, 1)
    ref = np.linalg.svd(A.blocks[(0, 0)], compute_uv=False)
    assert np.abs(np.sort(dd)[::-1] - ref).max() <= 1e-12 * ref[0]
    Ud = U.to_dense()
    Vd = V.to_dense()
    assert np.abs(Ud.T @ Ud - np.eye(k)).max() < 1e-10
    assert np.abs(Vd.T @ Vd - np.eye(k)).max() < 1e-12


def test_sweep_tail_check_equals_confirming_sweep(ctx, monkeypatch):
    """The all-pairs Gram test that ends the Jacobi sweeps in their tail
    (QTT_SVD_SWEEP_CHECK=1, the default) and the confirming sweep it replaces
    (=0) give the same singular values within 1e-12 of the largest, the same
    structure, and both reconstruct A, on two wide graded sectors."""
    import quantit_amd as qa
    rng = np.random.default_rng(91)
    A = BTensor(G1, [[300, 260], [280, 240]], [[(0,), (2,)], [(0,), (-2,)]], (0,))
    for i, (m, n) in enumerate(((300, 280), (260, 240))):
        k = min(m, n)
        U0, _ = np.linalg.qr(rng.standard_normal((m, k)))
        V0, _ = np.linalg.qr(rng.standard_normal((n, k)))
        A.blocks[(i, i)] = (U0 * 10.0 ** (-9 * np.arange(k) / k)) @ V0.T
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("QTT_SVD_SWEEP_CHECK", flag)
        U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1))
        _check_svd(A, U, d, V, 1)
        out[flag] = (U, d, V)
    (U1, d1, V1), (U0_, d0, V0_) = out["1"], out["0"]
    assert_same_structure(d1, d0)
    assert_same_structure(U1, U0_)
    smax = max(float(b.max()) for b in d0.blocks.values())
    for key in d0.blocks:
        assert np.abs(d1.blocks[key] - d0.blocks[key]).max() <= 1e-12 * smax


def test_eigh_wide_sector(ctx):
    """eigh of a sector wider than one work-group's LDS (k = 180): values
    against the dense eigvalsh within 1e-12 of the norm, A.U = U.Lambda."""
    import quantit_amd as qa
    rng = np.random.default_rng(77)
    n = 180
    M = rng.standard_normal((n, n))
    A = BTensor(G1, [[n], [n]], [[(0,)], [(0,)]], (0,))
    A.blocks[(0, 0)] = M + M.T
    d, U = (from_dev(x) for x in qa.eigh(to_dev(ctx, A), 1))
    w = np.linalg.eigvalsh(A.blocks[(0, 0)])
    nrm = np.linalg.norm(A.blocks[(0, 0)], 2)
    assert np.abs(np.sort(d.blocks[(0,)]) - w).max() <= 1e-12 * nrm
    Ud = U.to_dense()
    assert np.abs(A.blocks[(0, 0)] @ Ud - Ud * d.blocks[(0,)]).max() <= 1e-11 * nrm
    assert np.abs(Ud.T @ Ud - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("m,n,decades,zeros", [(300, 48, None, 0), (64, 200, 6, 0), (500, 65, 10, 0),
                                               (97, 260, 3, 0), (400, 140, None, 0), (139, 150, 12, 0),
                                               (300, 120, 4, 20), (90, 90, None, 0)])
def test_svd_one_group_widths(ctx, m, n, decades, zeros):
    """Sectors the one-work-group Gram eigen-solver takes (48 <= k <= 140,
    every register width of its wave-local sweeps, odd k, k equal to the
    widest): random (flat) and graded spectra, and a rank-deficient sector.
    Singular values within 1e-12 of the largest against the dense SVD
    (library routine), reconstruction and orthonormality within 1e-12."""
    import quantit_amd as qa
    rng = np.random.default_rng(7 * m + n)
    k = min(m, n)
    if decades is None:
        B = rng.standard_normal((m, n))
    else:
        U0, _ = np.linalg.qr(rng.standard_normal((m, k)))
        V0, _ = np.linalg.qr(rng.standard_normal((n, k)))
        s = 10.0 ** (-decades * np.arange(k) / k)
        s[k - zeros:] = 0.0
        B = (U0 * s) @ V0.T
    A = BTensor(G1, [[m], [n]], [[(0,)], [(0,)]], (0,))
    A.blocks[(0, 0)] = B
    U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1))
    dd = _check_svd(A, U, d, V, 1)
    ref = np.linalg.svd(B, compute_uv=False)
    got = np.sort(dd)[::-1]
    assert got.size <= k and np.abs(got - ref[:got.size]).max() <= 1e-12 * ref[0]
    assert np.all(ref[got.size:] <= 1e-12 * ref[0])
    for t in (U, V):
        M = t.to_dense()
        assert np.abs(M.T @ M - np.eye(M.shape[1])).max() < (1e-10 if decades and decades > 8 else 1e-12)


def _sym_case(seed):
    rng = np.random.default_rng(seed)
    rows = [(5, (0,)), (3, (2,)), (7, (-2,)), (9, (4,)), (2, (-4,))]
    A = rand_btensor(rng, G1, [rows, inv_leg(G1, rows)], (0,))
    for k in A.blocks:
        A.blocks[k] = A.blocks[k] + A.blocks[k].T
    return A


@pytest.mark.parametrize("tol,mn,mx,pw", [(1e-1, 1, 10 ** 6, 1.0), (2.0, 1, 10 ** 6, 1.0), (0.5, 4, 12, 2.0),
                                          (1e-8, 1, 6, 1.0), (5.0, 20, 10 ** 6, 1.0)])
def test_eigh_truncated_matches_oracle(ctx, tol, mn, mx, pw):
    """Truncating eigh (btensor_linalg.cpp:816-823) against the oracle's
    truncate_impl of its eigh: structure bit-exact (erased sectors included),
    kept eigenvalues within 1e-12 of the norm, A.U = U.Lambda on what is
    kept; the 2-tuple truncate gives the same bits as the truncating eigh."""
    import quantit_amd as qa
    from oracle.btensor_oracle import truncate_impl
    A = _sym_case(43)
    re, rS = eigh_split(A, 1)
    re2, (rS2,), _ = truncate_impl(re, [rS], mx, mn, tol, pw)
    dA = to_dev(ctx, A)
    e2, S2 = qa.eigh(dA, 1, tol=tol, min_size=mn, max_size=mx, pow=pw)
    he, hS = from_dev(e2), from_dev(S2)
    assert_same_structure(he, re2)
    assert_same_structure(hS, rS2)
    nrm = np.linalg.norm(A.to_dense(), 2)
    for k in re2.blocks:
        assert np.abs(he.blocks[k] - re2.blocks[k]).max() <= 1e-12 * nrm
    e, S = qa.eigh(dA, 1)
    t2 = qa.truncate_eigh(e, S, mx, mn, tol, pw)
    assert from_dev(t2[0]).blocks.keys() == he.blocks.keys()
    for k in he.blocks:
        assert np.array_equal(from_dev(t2[0]).blocks[k], he.blocks[k])
    for k in hS.blocks:
        assert np.array_equal(from_dev(t2[1]).blocks[k], hS.blocks[k])
    # eigenpairs of what is kept
    Ad = A.to_dense()
    for (kd,), w in he.blocks.items():
        cols = [ui for ui in hS.blocks if ui[-1] == kd]
        for ui in cols:
            r0 = sum(A.secs[0][:ui[0]])
            blk = hS.blocks[ui]
            Arow = Ad[r0:r0 + blk.shape[0]][:, r0:r0 + blk.shape[0]]
            assert np.abs(Arow @ blk - blk * w).max() <= 1e-11 * nrm


def test_eigh_truncated_known_answer(ctx):
    """The hand-traced case of the oracle tests on the device: the sector
    {-1, 4} is erased whole, {3, 5} and {0.5} are kept whole."""
    import quantit_amd as qa
    from tests.test_oracle_btensor import eigh_truncation_case
    A = eigh_truncation_case()
    e2, S2 = qa.eigh(to_dev(ctx, A), 1, tol=1.2, min_size=1, max_size=10, pow=1.0)
    he, hS = from_dev(e2), from_dev(S2)
    kept = sorted(np.round(np.concatenate([he.blocks[k] for k in he.blocks]), 12).tolist())
    assert kept == [0.5, 3.0, 5.0]
    assert len(he.blocks) == 2 and sorted(he.secs[0]) == [1, 2, 2]


def test_svd_matrix_modes(ctx):
    """svd(tensor, some, compute_uv) of a rank-2 tensor: some = True is the
    split-1 svd bit for bit; on square sectors some = False gives the same;
    compute_uv = False gives the same d and zero U, V; rectangular sectors
    are rejected in those modes with ValueError."""
    import quantit_amd as qa
    A = _sym_case(44)     # square sectors
    dA = to_dev(ctx, A)
    ref = [from_dev(x) for x in qa.svd(dA, 1)]
    for some, cuv in ((True, True), (False, True), (True, False), (False, False)):
        got = [from_dev(x) for x in qa.svd_matrix(dA, some=some, compute_uv=cuv)]
        for g, r in zip(got, ref):
            assert_same_structure(g, r)
        for k in ref[1].blocks:
            assert np.array_equal(got[1].blocks[k], ref[1].blocks[k])
        for gi, ri in ((got[0], ref[0]), (got[2], ref[2])):
            for k in ri.blocks:
                if cuv:
                    assert np.array_equal(gi.blocks[k], ri.blocks[k])
                else:
                    assert not np.any(gi.blocks[k])
    rng = np.random.default_rng(45)
    rows = [(5, (0,)), (3, (1,))]
    cols = [(4, (0,)), (3, (-1,))]
    R = to_dev(ctx, rand_btensor(rng, G1, [rows, cols], (0,)))
    assert len(from_dev(qa.svd_matrix(R)[1]).blocks) == 2
    with pytest.raises(ValueError):
        qa.svd_matrix(R, some=False)
    with pytest.raises(ValueError):
        qa.svd_matrix(R, compute_uv=False)
