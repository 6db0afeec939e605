"""Pins the numpy oracle of the block-sparse tensor path against the
reference's own known answers and against independent mathematics (dense
numpy contractions, dense SVD/eigh, closed-form truncation rule).  CPU only."""
import itertools

import numpy as np
import pytest

from oracle.btensor_oracle import (BTensor, ContractionError, add, compact_dense, compute_last_index, eigh_split,
                                   mul, q_add, q_inv, q_make, q_neutral, rand_like, reorder_by_cvals, scal,
                                   shape_product, svd_split, svd_truncated, tensordot, truncate_impl)
from tests.helpers import inv_leg, rand_btensor, rand_leg


# ---- charges ---------------------------------------------------------------
def test_charge_group_axioms():
    for groups in [(0,), (0, 0, 2), (0, 0, 6), (2, 4)]:
        rng = np.random.default_rng(0)
        for _ in range(50):
            a = tuple(int(rng.integers(-100, 100)) if n == 0 else int(rng.integers(0, n)) for n in groups)
            b = tuple(int(rng.integers(-100, 100)) if n == 0 else int(rng.integers(0, n)) for n in groups)
            c = tuple(int(rng.integers(-100, 100)) if n == 0 else int(rng.integers(0, n)) for n in groups)
            n0 = q_neutral(groups)
            assert q_add(groups, a, q_inv(groups, a)) == n0
            assert q_add(groups, a, n0) == a
            assert q_add(groups, q_add(groups, a, b), c) == q_add(groups, a, q_add(groups, b, c))
            assert q_add(groups, a, b) == q_add(groups, b, a)


def test_charge_int16_wrap_and_cyclic_construction():
    # Z is int16 with wrap-around (include/Conserved/quantity.h:175-239)
    assert q_add((0,), (32767,), (1,)) == (-32768,)
    # C<N>(int16) construction (quantity.h:80): negative values fold into [0, N)
    assert q_make((5,), (-1,)) == (4,)
    assert q_make((5,), (-5,)) == (0,)
    assert q_make((5,), (7,)) == (2,)
    assert q_inv((4,), (0,)) == (0,) and q_inv((4,), (1,)) == (3,)


# ---- reference known answer: reorder_by_cvals / compact_dense -------------
def _ref_A():
    """The C<5> tensor of include/blockTensor/LinearAlgebra.h:255-267."""
    g = (5,)
    c = lambda v: q_make(g, (v,))  # noqa: E731
    A = BTensor.from_sections(g, [[(2, c(0)), (3, c(1))],
                                  [(1, c(1)), (2, c(0)), (3, c(-1)), (1, c(1))],
                                  [(3, c(0)), (2, c(-2)), (2, c(-1))]], c(0))
    rng = np.random.default_rng(0)
    for idx in [(0, 0, 2), (0, 1, 0), (0, 3, 2), (1, 0, 1), (1, 1, 2), (1, 2, 0), (1, 3, 1)]:
        assert A.allowed(idx)
        A.blocks[idx] = rng.random(A.block_shape(idx))
    return A


def test_reorder_by_cvals_known_answer():
    A = _ref_A()
    order = [k for k, _ in reorder_by_cvals(A)]
    assert order == [(0, 1, 0), (1, 1, 2), (1, 0, 1), (1, 3, 1), (0, 0, 2), (0, 3, 2), (1, 2, 0)]
    B = A.reshape([1])
    order = [k for k, _ in reorder_by_cvals(B)]
    assert order == [(0, 2), (0, 3), (0, 11), (1, 1), (1, 5), (1, 6), (1, 10)]
    # compact_dense_single on the first sector reproduces every block from its slices
    sectors = compact_dense(B)
    dense, other, rows, cols = sectors[0]
    for rw, a, b in rows:
        for cl, c0, c1 in cols:
            if (rw, cl) in B.blocks:
                assert np.array_equal(dense[a:b, c0:c1], B.blocks[(rw, cl)])


# ---- tensordot --------------------------------------------------------------
PATTERNS = [((3, 3), [2], [0]), ((4, 3), [1, 2], [0, 1]), ((5, 6), [0, 2, 3], [0, 4, 5]), ((4, 4), [3, 0], [1, 2]),
            ((2, 3), [], []), ((3, 3), [0, 1, 2], [0, 1, 2])]


@pytest.mark.parametrize("pat", range(len(PATTERNS)))
def test_tensordot_equals_dense(pat):
    (ra, rb), da, db = PATTERNS[pat]
    rng = np.random.default_rng(pat)
    legs_a = [rand_leg(rng, 3, smax=3) for _ in range(ra)]
    legs_b = [rand_leg(rng, 3, smax=3) for _ in range(rb)]
    for x, y in zip(da, db):
        legs_b[y] = inv_leg((0,), legs_a[x])
    A = rand_btensor(rng, (0,), legs_a, (0,))
    B = rand_btensor(rng, (0,), legs_b, (0,))
    C = tensordot(A, B, da, db)
    assert C.check() == ""
    dense = np.tensordot(A.to_dense(), B.to_dense(), (da, db))
    assert np.allclose(C.to_dense(), dense, atol=1e-12)
    # every non-zero dense region lies in a block the conservation law allows
    for idx in C.all_allowed_indices():
        pass
    assert all(C.allowed(k) for k in C.blocks)


def test_tensordot_trace_invariance_24_orders():
    """Full contraction of X with itself is the same scalar whatever the order
    in which the four legs are listed (include/blockTensor/btensor.h:1551-1580)."""
    rng = np.random.default_rng(3)
    legs = [rand_leg(rng, 3, smax=3) for _ in range(4)]
    X = rand_btensor(rng, (0,), legs, (0,))
    Xc = X.conj()
    ref = float(np.sum(X.to_dense() ** 2))
    for p in itertools.permutations(range(4)):
        t = tensordot(X, Xc, list(p), list(p))
        assert abs(float(next(iter(t.blocks.values()))) - ref) < 1e-12 * ref


def test_tensordot_incompatible():
    rng = np.random.default_rng(4)
    A = rand_btensor(rng, (0,), [rand_leg(rng, 3), rand_leg(rng, 3)], (0,))
    with pytest.raises(ContractionError):
        tensordot(A, A, [0], [0])


# ---- reshape / permute --------------------------------------------------------
def test_reshape_round_trip_exact():
    rng = np.random.default_rng(5)
    legs = [rand_leg(rng, 3, smax=3) for _ in range(4)]
    X = rand_btensor(rng, (0,), legs, (1,))
    R = X.reshape([2])
    assert R.check() == ""
    assert [len(s) for s in R.secs] == [9, 9] or R.rank == 2
    back = R.reshape_as(X)
    assert back.secs == X.secs and back.cvals == X.cvals and back.sel == X.sel
    assert sorted(back.blocks) == sorted(X.blocks)
    for k in X.blocks:
        assert np.array_equal(back.blocks[k], X.blocks[k])
    # fused legs enumerate section combinations (no charge fusion), so the
    # dense layouts differ by a permutation of elements only
    assert np.array_equal(np.sort(R.to_dense().ravel()), np.sort(X.to_dense().ravel()))


def test_permute_matches_dense():
    rng = np.random.default_rng(6)
    X = rand_btensor(rng, (0,), [rand_leg(rng, 3) for _ in range(4)], (0,))
    P = X.permute([2, 0, 3, 1])
    assert P.check() == ""
    assert np.array_equal(P.to_dense(), np.transpose(X.to_dense(), [2, 0, 3, 1]))


def test_broadcast_mul_and_add():
    rng = np.random.default_rng(7)
    k = rand_leg(rng, 3)
    U = rand_btensor(rng, (0,), [rand_leg(rng, 3), k], (0,))
    d = rand_btensor(rng, (0,), [[(s, (0,)) for s, _ in k]], (0,))
    M = mul(U, d)
    dd = np.concatenate([d.blocks[(i,)] for i in range(len(k))])
    assert np.allclose(M.to_dense(), U.to_dense() * dd)
    S = add(U, scal(U, 2.0), -0.5)
    assert np.allclose(S.to_dense(), 0.0)


# ---- SVD / eigh / truncation ----------------------------------------------------
def test_svd_reference_tensor_reconstructs():
    """U.diag(d).V^T reconstruction and orthonormality on the reference's own
    C<5> example (include/blockTensor/LinearAlgebra.h:341-400)."""
    A = _ref_A()
    U, d, V = svd_split(A, 1)
    dense = A.to_dense().reshape(A.sizes()[0], -1)
    Ud = U.to_dense()
    Vd = V.to_dense().reshape(-1, Ud.shape[1])
    dd = np.concatenate([d.blocks[(i,)] for i in range(len(d.secs[0]))])
    assert np.allclose((Ud * dd) @ Vd.T, dense, atol=1e-13)
    assert np.allclose(Ud.T @ Ud, np.eye(Ud.shape[1]), atol=1e-13)
    assert U.check() == "" and V.check() == "" and d.check() == ""


def test_svd_values_equal_dense_svd():
    rng = np.random.default_rng(8)
    rows = rand_leg(rng, 5, smax=7)
    A = rand_btensor(rng, (0,), [rows, inv_leg((0,), rows)], (0,))
    U, d, V = svd_split(A, 1)
    s = np.sort(np.concatenate(list(d.blocks.values())))[::-1]
    sd = np.linalg.svd(A.to_dense(), compute_uv=False)
    assert np.allclose(s, sd[:len(s)], atol=1e-13)


def test_compute_last_index_closed_form():
    v = np.array([1.0, 0.5, 0.1, 1e-3, 1e-6, 1e-9])
    # discard while the tail weight sum s^2 stays <= tol^2
    assert compute_last_index(v, 1e-5, 2, 1, 100) == 3   # keeps 1e-3, drops 1e-6 and 1e-9
    assert compute_last_index(v, 2e-3, 2, 1, 100) == 2   # 1e-3^2 + ... < 4e-6
    assert compute_last_index(v, 1e-12, 2, 1, 100) == 5  # nothing can go
    assert compute_last_index(v, 1e-12, 2, 1, 3) == 2    # max_size takes precedence
    assert compute_last_index(v, 10.0, 2, 4, 100) == 3   # min_size takes precedence


def test_truncation_keeps_ties_and_removes_empty_sectors():
    g = (0,)
    d = BTensor(g, [[3, 2, 2]], [[(0,)] * 3], (0,))
    d.blocks[(0,)] = np.array([1.0, 0.5, 1e-7])
    d.blocks[(1,)] = np.array([0.5, 1e-8])          # tie with 0.5 of sector 0
    d.blocks[(2,)] = np.array([1e-9, 1e-10])        # whole sector below the cut
    U = BTensor(g, [[1, 1, 1], [3, 2, 2]], [[(0,), (1,), (2,)], [(0,), (1,), (2,)]], (0,))
    for i in range(3):
        U.blocks[(i, i)] = np.ones((1, d.secs[0][i]))
    d2, (U2,), thr = truncate_impl(d, [U], max_size=2, min_size=1, tol=1e-12)
    # max 2 -> cut at the second largest (0.5); its tie in sector 1 is kept too
    assert d2.secs[0][:2] == [2, 1]
    assert sorted(d2.blocks) == [(0,), (1,)]
    assert U2.blocks[(0, 0)].shape == (1, 2)
    assert d2.secs[0][2] == 2  # erased sector keeps its stale size
    # remove_unit_blocks (sources/btensor_linalg.cpp:690-709) never removes a
    # matching block that is last in the list: (2, 2) survives, full size
    assert (2, 2) in U2.blocks and U2.blocks[(2, 2)].shape == (1, 2)
    # ... while a matching block followed by a non-matching one is removed
    U3 = BTensor(g, [[1, 1, 1], [3, 2, 2]], [[(2,), (1,), (0,)], [(0,), (1,), (2,)]], (2,))
    for i in range(3):
        U3.blocks[(i, 2 - i)] = np.ones((1, d.secs[0][2 - i]))
    _, (U4,), _ = truncate_impl(d, [U3], max_size=2, min_size=1, tol=1e-12)
    assert sorted(U4.blocks) == [(1, 1), (2, 0)]


def test_truncated_svd_error_bound():
    """Discarded weight below tol (include/blockTensor/LinearAlgebra.h:473-567)."""
    rng = np.random.default_rng(9)
    rows = rand_leg(rng, 4, smax=8)
    A = rand_btensor(rng, (0,), [rows, inv_leg((0,), rows)], (0,))
    tol = 0.5
    U, d, V = svd_truncated(A, 1, tol)
    Ud, Vd = U.to_dense(), V.to_dense()
    dd = np.concatenate([d.blocks[(i,)] if (i,) in d.blocks else np.zeros(d.secs[0][i])
                         for i in range(len(d.secs[0]))])
    err = np.linalg.norm((Ud * dd) @ Vd.T - A.to_dense())
    assert err <= tol * (1 + 1e-12)


def test_eigh_equals_dense():
    rng = np.random.default_rng(10)
    rows = [(4, (0,)), (3, (1,)), (5, (-1,))]
    A = rand_btensor(rng, (0,), [rows, inv_leg((0,), rows)], (0,))
    for k in A.blocks:
        A.blocks[k] = A.blocks[k] + A.blocks[k].T
    d, U = eigh_split(A, 1)
    w = np.sort(np.concatenate(list(d.blocks.values())))
    assert np.allclose(w, np.linalg.eigvalsh(A.to_dense()), atol=1e-12)
    Ud = U.to_dense()
    dd = np.concatenate([d.blocks[(i,)] for i in range(len(d.secs[0]))])
    assert np.allclose(A.to_dense() @ Ud, Ud * dd, atol=1e-12)


def test_c5_contraction_block_set_known_answer():
    """include/blockTensor/btensor.h:1585-1608: C = A.tensordot(B, {1}, {1})
    with C<5> charges has exactly the blocks (0,1,0), (1,0,0), (1,1,1), each
    the dense tensordot of one block pair; the full self-contraction of A is
    the sum of its blocks' squared norms."""
    from oracle.btensor_oracle import tensordot as td
    rng = np.random.default_rng(0)
    C5 = (5,)
    A = BTensor.from_sections(C5, [[(2, (0,)), (3, (1,))], [(2, (0,)), (3, (4,))]], (0,))
    A.blocks[(0, 0)] = rng.uniform(size=(2, 2))
    A.blocks[(1, 1)] = rng.uniform(size=(3, 3))
    B = BTensor.from_sections(C5, [[(3, (4,)), (2, (0,))], [(2, (0,)), (3, (1,))], [(1, (1,)), (3, (0,))]], (1,))
    B.blocks[(1, 0, 0)] = rng.uniform(size=(2, 2, 1))
    B.blocks[(0, 1, 0)] = rng.uniform(size=(3, 3, 1))
    B.blocks[(1, 1, 1)] = rng.uniform(size=(2, 3, 3))
    C = td(A, B, [1], [1])
    assert sorted(C.blocks) == [(0, 1, 0), (1, 0, 0), (1, 1, 1)]
    assert np.allclose(C.blocks[(0, 1, 0)], np.tensordot(A.blocks[(0, 0)], B.blocks[(1, 0, 0)], ([1], [1])))
    assert np.allclose(C.blocks[(1, 0, 0)], np.tensordot(A.blocks[(1, 1)], B.blocks[(0, 1, 0)], ([1], [1])))
    assert np.allclose(C.blocks[(1, 1, 1)], np.tensordot(A.blocks[(1, 1)], B.blocks[(1, 1, 1)], ([1], [1])))
    tr = td(A, A.inverse_cvals(), [0, 1], [0, 1])
    assert np.isclose(float(tr.blocks[()]), sum(float(np.sum(b * b)) for b in A.blocks.values()))


def _trunc_case():
    """U (4 x 4 sections, one group with every charge 0) holding the blocks
    (0,3), (1,2), (3,3); d with one value per sector, sectors 2 and 3 below
    the cut."""
    z = [(0,)] * 4
    U = BTensor((0,), [[1] * 4, [1] * 4], [z, z], (0,))
    for k in ((0, 3), (1, 2), (3, 3)):
        U.blocks[k] = np.ones((1, 1))
    d = BTensor((0,), [[1] * 4], [z], (0,))
    for i, v in enumerate((1.0, 0.9, 1e-9, 1e-9)):
        d.blocks[(i,)] = np.array([v])
    return U, d


def test_truncate_erasures_follow_the_unsorted_list_known_answer():
    """remove_unit_blocks (sources/btensor_linalg.cpp:690-709) traced by hand.
    Sectors are visited last to first.  Erasing sector 3 from
    [(0,3), (1,2), (3,3)]: the first block matches and is swapped with the
    second, the third matches at the end, so the list becomes
    [(1,2), (0,3)] (a match directly after a match survives, and the list is
    no longer sorted).  Erasing sector 2 from that list: (1,2) matches and is
    swapped past (0,3); the result is [(0,3)].  Walking the re-sorted list
    [(0,3), (1,2)] instead would keep both."""
    U, d = _trunc_case()
    d2, (U2,), _ = truncate_impl(d, [U], 10 ** 9, 1, 1e-8)
    assert sorted(d2.blocks) == [(0,), (1,)]
    assert list(U2.blocks) == [(0, 3)]



def eigh_truncation_case():
    """Symmetric U(1) matrix with three sectors of known spectra: charge 0
    {-1, 4}, charge 2 {3, 5}, charge -2 {0.5} (rotated by fixed orthogonal
    matrices so that the blocks are not diagonal).  The first column section
    has charge 0: the reference's eigenvector tensor takes the inverse of
    that charge into its selection rule (btensor_linalg.cpp:294-330)."""
    c, s = np.cos(0.3), np.sin(0.3)
    R = np.array([[c, -s], [s, c]])
    rows = [(2, (0,)), (1, (-2,)), (2, (2,))]
    A = BTensor.from_sections((0,), [rows, [(n, (-q[0],)) for n, q in rows]], (0,))
    A.blocks[(0, 0)] = R @ np.diag([-1.0, 4.0]) @ R.T
    A.blocks[(1, 1)] = np.array([[0.5]])
    A.blocks[(2, 2)] = R.T @ np.diag([3.0, 5.0]) @ R
    return A


def _by_charge(e, S):
    """eigenvalue blocks keyed by the row charge of their sector"""
    return {S.cvals[-1][k[0]][0]: e.blocks[k] for k in e.blocks}


def test_eigh_truncation_erases_whole_sectors_known_answer():
    """truncate(e, S, ...) of an eigh (btensor_linalg.cpp:786-823) traced by
    hand: descending values 5, 4, 3, 0.5, -1; with tol 1.2 and pow 1 the
    tail sums give |-1| = 1 <= 1.2 and 1 + 0.5 > 1.2, so the cut is the value
    0.5 (minus 2 eps).  Within a sector the eigenvalues ascend and a sector
    keeps the values before the first one not above the cut: the sector
    {-1, 4} starts with -1 and is erased whole (its 4 included), {3, 5} and
    {0.5} are kept whole."""
    A = eigh_truncation_case()
    e, S = eigh_split(A, 1)
    full = _by_charge(e, S)
    assert sorted(np.round(np.concatenate(list(full.values())), 12).tolist()) == [-1.0, 0.5, 3.0, 4.0, 5.0]
    assert compute_last_index(np.array([5.0, 4.0, 3.0, 0.5, -1.0]), 1.2, 1.0, 1, 10) == 3
    e2, (S2,), thr = truncate_impl(e, [S], 10, 1, 1.2, 1.0)
    assert thr == 0.5 - 2 * 0.5 * np.finfo(float).eps
    kept = _by_charge(e2, S2)
    assert sorted(kept) == sorted(q for q, w in full.items() if w.min() > thr)
    assert len(kept) == 2 and all(np.array_equal(kept[q], full[q]) for q in kept)
    erased = [k for k in e.blocks if k not in e2.blocks]
    assert len(erased) == 1 and e2.secs[0][erased[0][0]] == 2   # size left as it was
    assert np.allclose(sorted(full[S.cvals[-1][erased[0][0]][0]]), [-1.0, 4.0])
    assert all(k[-1] != erased[0][0] for k in S2.blocks)
