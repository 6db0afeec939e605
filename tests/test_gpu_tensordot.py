"""Parity of the device block contraction with the oracle's restatement of
btensor::tensordot (sources/btensor.cpp:1971-2119): output block set, sector
charges and shapes bit-exact; values within 1e-12 of the FP64 oracle
(relative to the tensor norm; FP64 reduction order differs), 1e-5 for the
split-FP32 mode."""
import numpy as np
import pytest

from oracle.btensor_oracle import BTensor, tensordot as ref_tensordot
from tests.helpers import (assert_same_structure, assert_values_close, from_dev, inv_leg, rand_btensor, rand_leg,
                           to_dev)

pytestmark = pytest.mark.gpu

G1 = (0,)
G3 = (0, 0, 2)


def _pair(rng, groups, ra, rb, dims_a, dims_b, nsec=4, smax=5):
    lo, hi = (-3, 3) if groups == G1 else (-1, 1)
    legs_a = [rand_leg(rng, nsec, lo=lo, hi=hi, smax=smax, groups=groups) for _ in range(ra)]
    legs_b = [rand_leg(rng, nsec, lo=lo, hi=hi, smax=smax, groups=groups) for _ in range(rb)]
    for x, y in zip(dims_a, dims_b):
        legs_b[y] = inv_leg(groups, legs_a[x])
    sel_a = tuple(int(v) for v in rng.integers(-1, 2, len(groups))) if groups == G1 else tuple(0 for _ in groups)
    sel_b = tuple(0 for _ in groups)
    if len(dims_a) == ra == rb:  # full contraction: B must carry the opposite rule
        sel_b = tuple(-v for v in sel_a)
    A = rand_btensor(rng, groups, legs_a, sel_a)
    B = rand_btensor(rng, groups, legs_b, sel_b)
    return A, B


CASES = [
    # (ranks, dims_a, dims_b)
    ((3, 3), [2], [0]),             # theta = A[oc] . A[oc+1]
    ((4, 3), [1, 2], [0, 1]),       # env-update pattern X(b,a',s,w) . A(a',s,b')
    ((3, 4), [0], [0]),             # L . theta
    ((5, 6), [0, 2, 3], [0, 4, 5]), # (w,a',s1,s2,b) . W12
    ((5, 3), [1, 4], [0, 1]),       # . R
    ((4, 4), [3, 0], [1, 2]),       # permuted contracted order
    ((2, 2), [], []),               # tensor product
    ((3, 3), [0, 1, 2], [0, 1, 2]), # full contraction -> rank 0
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("groups", [G1, G3])
def test_tensordot_matches_oracle(ctx, case, groups):
    (ra, rb), da, db = CASES[case]
    rng = np.random.default_rng(100 + case)
    A, B = _pair(rng, groups, ra, rb, da, db, nsec=3 if ra + rb > 8 else 4)
    ref = ref_tensordot(A, B, da, db)
    out = from_dev(to_dev(ctx, A).tensordot(to_dev(ctx, B), da, db))
    assert_same_structure(out, ref)
    assert_values_close(out, ref, 1e-12)
    # dense cross-check of the oracle itself on the same inputs
    dense = np.tensordot(A.to_dense(), B.to_dense(), (da, db))
    assert np.allclose(out.to_dense(), dense, atol=1e-11 * max(1.0, np.abs(dense).max()))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_tensordot_narrow_rows_equal_mfma_tiles(ctx, case, monkeypatch):
    """A contraction whose output blocks are all narrow (few columns, short
    summed K) takes the row kernel; the same contraction on MFMA tiles agrees
    to 1e-13, and the row kernel is seen to run where every block is narrow."""
    (ra, rb), da, db = CASES[case]
    rng = np.random.default_rng(300 + case)
    A, B = _pair(rng, G1, ra, rb, da, db, nsec=3 if ra + rb > 8 else 4)
    a, b = to_dev(ctx, A), to_dev(ctx, B)
    ctx.kernel_times(reset=True)
    ctx.set_profiling(True)
    try:
        rows = from_dev(a.tensordot(b, da, db))
    finally:
        ctx.set_profiling(False)
    kt = ctx.kernel_times(reset=True)
    monkeypatch.setenv("QTT_GEMM_ROWS", "0")
    tiles = from_dev(a.tensordot(b, da, db))
    assert_same_structure(rows, tiles)
    assert_values_close(rows, tiles, 1e-13)
    ref = ref_tensordot(A, B, da, db)
    assert_values_close(rows, ref, 1e-12)
    # with one free leg of B (sections of at most 5 columns) every output
    # block of these seeded cases is narrow, so the row kernel must have run
    if rb - len(db) == 1 and rows.blocks:
        assert "gemm_rows" in kt


@pytest.mark.parametrize("shape", ["many_pairs", "one_long_pair"])
def test_tensordot_split_k(ctx, shape, monkeypatch):
    """Few output tiles and a long contracted extent: the K range of each
    block (pairs in order, a pair possibly cut) is split into partials that
    a fixed-order sum combines.  Values against the oracle and against the
    unsplit launch, the split seen to run, two runs bit-identical."""
    rng = np.random.default_rng(12)
    la = rand_leg(rng, 3, lo=-3, hi=3, smin=17, smax=40, groups=G1)
    if shape == "many_pairs":
        lx = rand_leg(rng, 7, lo=-3, hi=3, smax=40, groups=G1)
        ly = rand_leg(rng, 7, lo=-3, hi=3, smax=40, groups=G1)
        A = rand_btensor(rng, G1, [la, lx, ly], (0,))
        B = rand_btensor(rng, G1, [inv_leg(G1, lx), inv_leg(G1, ly), inv_leg(G1, la)], (0,))
        da, db = [1, 2], [0, 1]
    else:  # one contracted sector of 700: a single pair per output block
        lx = [(700, (0,))]
        A = rand_btensor(rng, G1, [la, lx, inv_leg(G1, la)], (0,))
        B = rand_btensor(rng, G1, [inv_leg(G1, lx), [(30, (0,))]], (0,))
        da, db = [1], [0]
    ref = ref_tensordot(A, B, da, db)
    a, b = to_dev(ctx, A), to_dev(ctx, B)
    ctx.kernel_times(reset=True)
    ctx.set_profiling(True)
    try:
        x = from_dev(a.tensordot(b, da, db))
        y = from_dev(a.tensordot(b, da, db))
    finally:
        ctx.set_profiling(False)
    kt = ctx.kernel_times(reset=True)
    assert "gemm_split_k_sum" in kt
    assert_same_structure(x, ref)
    assert_values_close(x, ref, 1e-12)
    for k in x.blocks:
        assert np.array_equal(x.blocks[k], y.blocks[k])
    monkeypatch.setenv("QTT_GEMM_SPLIT_K", "0")
    whole = from_dev(a.tensordot(b, da, db))
    assert_values_close(x, whole, 1e-13)


def test_tensordot_large_blocks(ctx):
    """Blocks larger than one 64x64 tile and K spanning several LDS stages."""
    rng = np.random.default_rng(7)
    A, B = _pair(rng, G1, 3, 3, [2], [0], nsec=3, smax=150)
    ref = ref_tensordot(A, B, [2], [0])
    out = from_dev(to_dev(ctx, A).tensordot(to_dev(ctx, B), [2], [0]))
    assert_same_structure(out, ref)
    assert_values_close(out, ref, 1e-12)


def test_tensordot_per_block_mode_equal(ctx):
    rng = np.random.default_rng(8)
    A, B = _pair(rng, G1, 4, 3, [1, 2], [0, 1], nsec=4, smax=40)
    a, b = to_dev(ctx, A), to_dev(ctx, B)
    grouped = from_dev(a.tensordot(b, [1, 2], [0, 1]))
    ctx.set_per_block(True)
    try:
        plain = from_dev(a.tensordot(b, [1, 2], [0, 1]))
    finally:
        ctx.set_per_block(False)
    assert_same_structure(plain, grouped)
    assert_values_close(plain, grouped, 1e-13)


def test_tensordot_split_fp32(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(9)
    A, B = _pair(rng, G1, 4, 3, [1, 2], [0, 1], nsec=4, smax=60)
    ref = ref_tensordot(A, B, [1, 2], [0, 1])
    a, b = to_dev(ctx, A), to_dev(ctx, B)
    ctx.set_precision(qa.PRECISION_SPLIT_FP32)
    try:
        out = from_dev(a.tensordot(b, [1, 2], [0, 1]))
    finally:
        ctx.set_precision(qa.PRECISION_FP64)
    assert_same_structure(out, ref)
    assert_values_close(out, ref, 1e-5)


def test_tensordot_deterministic(ctx):
    rng = np.random.default_rng(10)
    A, B = _pair(rng, G1, 4, 3, [1, 2], [0, 1], nsec=4, smax=50)
    a, b = to_dev(ctx, A), to_dev(ctx, B)
    x = from_dev(a.tensordot(b, [1, 2], [0, 1]))
    y = from_dev(a.tensordot(b, [1, 2], [0, 1]))
    for k in x.blocks:
        assert np.array_equal(x.blocks[k], y.blocks[k])


def test_tensordot_incompatible_raises(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(11)
    A, B = _pair(rng, G1, 3, 3, [2], [0])
    a, b = to_dev(ctx, A), to_dev(ctx, B)
    with pytest.raises(qa.CheckError):
        a.tensordot(b, [0], [0])  # charges of these legs do not cancel


def _variant_cases(rng):
    from quantit_amd import workloads
    X, A, dx, da, _ = workloads.contract_svd(256)
    out = [(BTensor(**X), BTensor(**A), dx, da)]
    for (ra, rb), da2, db2 in CASES:
        a, b = _pair(rng, G1, ra, rb, da2, db2, nsec=4, smax=40 if ra + rb <= 7 else 6)
        out.append((a, b, da2, db2))
    # one long-K pair and one wide output block
    big_a = rand_btensor(rng, (0,), [[(300, (0,))], [(700, (0,))]], (0,))
    big_b = rand_btensor(rng, (0,), [[(700, (0,))], [(260, (0,))]], (0,))
    out.append((big_a, big_b, [1], [0]))
    return out


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_gemm_tile_variants_give_the_same_bits(ctx, variant, monkeypatch):
    """Every FP64 tile variant (64x64 double-buffered, 128x64, 64x128, 128x128,
    64x64 with 32-deep k steps) against the 64x64 kernel: bit-identical blocks
    (same per-element accumulation order)."""
    rng = np.random.default_rng(40 + variant)
    for a, b, da, db in _variant_cases(rng):
        monkeypatch.setenv("QTT_GEMM_VARIANT", "0")
        ref = from_dev(to_dev(ctx, a).tensordot(to_dev(ctx, b), da, db))
        monkeypatch.setenv("QTT_GEMM_VARIANT", str(variant))
        got = from_dev(to_dev(ctx, a).tensordot(to_dev(ctx, b), da, db))
        assert sorted(got.blocks) == sorted(ref.blocks)
        for k in ref.blocks:
            assert np.array_equal(got.blocks[k], ref.blocks[k]), (variant, k)


def test_gemm_tuning_fills_the_table_and_keeps_the_bits(ctx, monkeypatch, tmp_path):
    import quantit_amd as qa
    from quantit_amd import workloads
    X, A, dx, da, _ = workloads.contract_svd(512)
    x, a = qa.BTensor.from_parts(ctx, **X), qa.BTensor.from_parts(ctx, **A)
    monkeypatch.setenv("QTT_GEMM_VARIANT", "0")
    ref = from_dev(x.tensordot(a, dx, da))
    monkeypatch.delenv("QTT_GEMM_VARIANT")
    monkeypatch.setenv("QTT_GEMM_TUNE", "1")
    n0 = qa.gemm_tuning_size()
    first = from_dev(x.tensordot(a, dx, da))    # tunes this shape class
    second = from_dev(x.tensordot(a, dx, da))   # uses the table
    assert qa.gemm_tuning_size() >= n0
    for got in (first, second):
        for k in ref.blocks:
            assert np.array_equal(got.blocks[k], ref.blocks[k])
    qa.gemm_tuning_save(tmp_path / "tuning.json")
    assert "gemm_variants" in (tmp_path / "tuning.json").read_text()
