"""Empty and ragged inputs through the C ABI, against the oracle: a
contraction whose block sets never meet on the contracted legs, an operand
with no blocks, the SVD and truncating SVD of a tensor with no blocks,
full_dot of disjoint tensors, and blocks of one row or one column against
long contracted legs (btensor::tensordot, sources/btensor.cpp:1971-2119;
svd, sources/btensor_linalg.cpp:503-534 and :805-809)."""
import numpy as np
import pytest

from oracle import btensor_oracle as bo
from oracle.btensor_oracle import BTensor
from tests.helpers import (assert_same_structure, assert_values_close, from_dev, inv_leg, rand_btensor, rand_leg,
                           to_dev)

pytestmark = pytest.mark.gpu

G1 = (0,)


def _theta_pair(seed):
    rng = np.random.default_rng(seed)
    legs = [rand_leg(rng, 4, smax=5) for _ in range(3)]
    A = rand_btensor(rng, G1, legs, (0,))
    B = rand_btensor(rng, G1, [inv_leg(G1, legs[2]), rand_leg(rng, 3, smax=4), rand_leg(rng, 3, smax=4)], (0,))
    return A, B


def test_tensordot_block_sets_that_never_meet(ctx):
    """A keeps only odd sections of the contracted leg, B only even ones: the
    output has the contraction's sections and no block."""
    A, B = _theta_pair(7)
    A.blocks = {k: v for k, v in A.blocks.items() if k[2] % 2 == 1}
    B.blocks = {k: v for k, v in B.blocks.items() if k[0] % 2 == 0}
    assert A.blocks and B.blocks
    ref = bo.tensordot(A, B, [2], [0])
    assert not ref.blocks
    out = from_dev(to_dev(ctx, A).tensordot(to_dev(ctx, B), [2], [0]))
    assert_same_structure(out, ref)
    assert not out.blocks


@pytest.mark.parametrize("empty", ["left", "right", "both"])
def test_tensordot_with_an_operand_without_blocks(ctx, empty):
    A, B = _theta_pair(8)
    if empty in ("left", "both"):
        A = BTensor(A.groups, A.secs, A.cvals, A.sel, {})
    if empty in ("right", "both"):
        B = BTensor(B.groups, B.secs, B.cvals, B.sel, {})
    ref = bo.tensordot(A, B, [2], [0])
    out = from_dev(to_dev(ctx, A).tensordot(to_dev(ctx, B), [2], [0]))
    assert_same_structure(out, ref)
    assert not out.blocks


def test_full_dot_of_disjoint_block_sets_is_zero(ctx):
    rng = np.random.default_rng(9)
    legs = [rand_leg(rng, 4, smax=4) for _ in range(3)]
    x = rand_btensor(rng, G1, legs, (0,))
    keys = sorted(x.blocks)
    y = BTensor(x.groups, x.secs, x.cvals, x.sel, {k: rng.standard_normal(x.blocks[k].shape) for k in keys[1::2]})
    x.blocks = {k: x.blocks[k] for k in keys[0::2]}
    want, _ = bo.full_dot(x, y)
    assert want == 0.0
    assert to_dev(ctx, x).full_dot(to_dev(ctx, y)) == 0.0


@pytest.mark.parametrize("tol", [None, 1e-6])
def test_svd_of_a_tensor_without_blocks(ctx, tol):
    import quantit_amd as qa
    rng = np.random.default_rng(10)
    M = rand_btensor(rng, G1, [rand_leg(rng, 3, smax=4), rand_leg(rng, 3, smax=4)], (0,))
    M = BTensor(M.groups, M.secs, M.cvals, M.sel, {})
    if tol is None:
        rU, rd, rV = bo.svd_split(M, 1)
        got = qa.svd(to_dev(ctx, M), 1)
    else:
        rU, rd, rV = bo.svd_truncated(M, 1, tol)
        got = qa.svd(to_dev(ctx, M), 1, tol)
    U, d, V = (from_dev(x) for x in got)
    for g, r in ((U, rU), (d, rd), (V, rV)):
        assert_same_structure(g, r)
        assert not g.blocks


@pytest.mark.parametrize("long_dim", [1000, 4099])
def test_contraction_of_single_row_and_column_blocks(ctx, long_dim):
    """A (1, K) . (K, 1) block pair with K up to 4099 beside a (3, 7) . (7, 3)
    one: a one-element output from a long contracted leg, within 1e-12 of the
    oracle."""
    rng = np.random.default_rng(long_dim)
    row = [(1, (0,)), (1, (1,)), (3, (-1,))]
    mid = [(long_dim, (0,)), (7, (1,))]
    A = rand_btensor(rng, G1, [row, mid], (0,))
    B = rand_btensor(rng, G1, [inv_leg(G1, mid), inv_leg(G1, row)], (0,))
    ref = bo.tensordot(A, B, [1], [0])
    out = from_dev(to_dev(ctx, A).tensordot(to_dev(ctx, B), [1], [0]))
    assert_same_structure(out, ref)
    assert_values_close(out, ref, 1e-12)
    dense = A.to_dense() @ B.to_dense()
    assert np.allclose(out.to_dense(), dense, atol=1e-11 * max(1.0, np.abs(dense).max()))


def test_svd_of_single_row_and_single_column_sectors(ctx):
    """Sectors of shape (1, n) and (m, 1) beside a square one: one singular
    value each, U V^T reconstruction and σ equal to the oracle's."""
    import quantit_amd as qa
    rng = np.random.default_rng(11)
    rows = [(1, (0,)), (37, (1,)), (5, (2,))]
    cols = [(41, (0,)), (1, (-1,)), (5, (-2,))]
    A = rand_btensor(rng, G1, [rows, cols], (0,))
    rU, rd, rV = bo.svd_split(A, 1)
    U, d, V = (from_dev(x) for x in qa.svd(to_dev(ctx, A), 1))
    assert_same_structure(d, rd)
    assert_same_structure(U, rU)
    assert_same_structure(V, rV)
    smax = max(float(b.max()) for b in rd.blocks.values())
    for k in rd.blocks:
        assert np.abs(d.blocks[k] - rd.blocks[k]).max() <= 1e-12 * smax
    Ad = A.to_dense()
    dd = np.concatenate([d.blocks[(i,)] for i in range(len(d.secs[0]))])
    rec = (U.to_dense() * dd) @ V.to_dense().T
    assert np.linalg.norm(rec - Ad) <= 1e-12 * np.linalg.norm(Ad)
