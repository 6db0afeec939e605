"""qtt_tensordot_into, the allocation-free contraction, against qtt_tensordot
(the allocating entry point) through the C ABI: bit-identical blocks for the
MFMA-tile path with packed operands, the split-K path and the narrow-row
path, on consecutive calls whose second input is smaller (the static scratch
keeps the first call's contents); no arena growth during the calls; sizes
beyond the scratch and malformed outputs rejected before any launch."""
import numpy as np
import pytest

from tests.helpers import from_dev

pytestmark = pytest.mark.gpu


def _zeros_like(ctx, t):
    """A tensor of t's structure and block list, every block zero."""
    import quantit_amd as qa
    groups, secs, cvals, sel, blocks = t.to_parts()
    return qa.BTensor.from_parts(ctx, groups, secs, cvals, sel, {k: np.zeros_like(v) for k, v in blocks.items()})


def _check_into(ctx, a, b, da, db):
    ref = a.tensordot(b, da, db)
    out = _zeros_like(ctx, ref)
    ctx.sync()
    used = ctx.memory()["used"]
    a.tensordot_into(b, da, db, out)
    ctx.sync()
    assert ctx.memory()["used"] == used
    r, o = from_dev(ref), from_dev(out)
    assert sorted(r.blocks) == sorted(o.blocks)
    for k in r.blocks:
        assert np.array_equal(r.blocks[k], o.blocks[k]), k
    return len(r.blocks)


def _contract_svd_inputs(ctx, D):
    import quantit_amd as qa
    from quantit_amd import workloads
    X, A, dx, da, _ = workloads.contract_svd(D)
    return qa.BTensor.from_parts(ctx, **X), qa.BTensor.from_parts(ctx, **A), dx, da


def test_into_equals_allocating_contraction_twice(ctx):
    """Environment-style contraction (packed operand, MFMA tiles) at D=256,
    then at D=96: the second, smaller call reuses the scratch the first one
    filled."""
    x, a, dx, da = _contract_svd_inputs(ctx, 256)
    assert _check_into(ctx, x, a, dx, da) > 10
    x, a, dx, da = _contract_svd_inputs(ctx, 96)
    assert _check_into(ctx, x, a, dx, da) > 5


def _dense_pair(ctx, m, k, n, seed):
    import quantit_amd as qa
    rng = np.random.default_rng(seed)
    A = qa.BTensor.from_parts(ctx, (0,), [[m], [k]], [[(0,)], [(0,)]], (0,), {(0, 0): rng.standard_normal((m, k))})
    B = qa.BTensor.from_parts(ctx, (0,), [[k], [n]], [[(0,)], [(0,)]], (0,), {(0, 0): rng.standard_normal((k, n))})
    return A, B


def test_into_split_k_and_narrow_rows_twice(ctx):
    """One long-K block pair (split-K partials in the scratch), then a
    shorter one; a narrow output (row kernel), then a smaller one."""
    for m, k, n in ((128, 6000, 96), (64, 2000, 64), (700, 100, 8), (300, 60, 4)):
        A, B = _dense_pair(ctx, m, k, n, m + k + n)
        assert _check_into(ctx, A, B, [1], [0]) == 1


def test_into_rejects_oversize_and_bad_outputs(ctx):
    import quantit_amd as qa
    rng = np.random.default_rng(1)
    # a block whose permuted view needs a 190 MB packing copy (> 128 MiB)
    big = qa.BTensor.from_parts(ctx, (0,), [[1100], [1100], [20]], [[(0,)], [(0,)], [(0,)]], (0,),
                                {(0, 0, 0): rng.standard_normal((1100, 1100, 20))})
    v = qa.BTensor.from_parts(ctx, (0,), [[1100]], [[(0,)]], (0,), {(0,): rng.standard_normal(1100)})
    ref = big.tensordot(v, [1], [0])
    out = _zeros_like(ctx, ref)
    with pytest.raises(ValueError, match="scratch"):
        big.tensordot_into(v, [1], [0], out)
    del big, ref
    x, a, dx, da = _contract_svd_inputs(ctx, 64)
    y = x.tensordot(a, dx, da)
    wrong = _zeros_like(ctx, y.permute([1, 0, 2]))
    with pytest.raises(ValueError, match="structure"):
        x.tensordot_into(a, dx, da, wrong)
    A, B = _dense_pair(ctx, 32, 32, 32, 3)
    with pytest.raises(ValueError, match="storage"):
        A.tensordot_into(B, [1], [0], A)
