"""Shared test helpers: seeded block-sparse inputs, conversion between the
oracle's BTensor and the library's handles, and structure/value comparison."""
from __future__ import annotations

import numpy as np

from oracle.btensor_oracle import BTensor, q_add, q_inv, q_neutral, rand_like


def rand_leg(rng, nsec, lo=-3, hi=3, smin=1, smax=4, groups=(0,)):
    """A leg with distinct charges drawn in [lo, hi] (per Z component) and
    random section sizes."""
    seen = set()
    out = []
    tries = 0
    while len(out) < nsec and tries < 1000:
        tries += 1
        ch = tuple(int(rng.integers(lo, hi + 1)) if n == 0 else int(rng.integers(0, n)) for n in groups)
        if ch in seen:
            continue
        seen.add(ch)
        out.append((int(rng.integers(smin, smax + 1)), ch))
    return out


def inv_leg(groups, leg):
    return [(s, q_inv(groups, c)) for s, c in leg]


def rand_btensor(rng, groups, legs, sel, dist="normal"):
    shape = BTensor.from_sections(groups, legs, sel)
    return rand_like(shape, rng, dist)


def to_dev(ctx, t: BTensor):
    import quantit_amd as qa
    return qa.BTensor.from_parts(ctx, t.groups, t.secs, t.cvals, t.sel, t.blocks)


def from_dev(d) -> BTensor:
    groups, secs, cvals, sel, blocks = d.to_parts()
    return BTensor(groups, secs, cvals, sel, blocks)


def assert_same_structure(dev: BTensor, ref: BTensor, check_block_set=True):
    """Bit-exact structure: sections, charges, selection rule, block set."""
    assert dev.groups == ref.groups
    assert dev.secs == ref.secs, (dev.secs, ref.secs)
    assert dev.cvals == ref.cvals
    assert dev.sel == ref.sel
    if check_block_set:
        assert sorted(dev.blocks) == sorted(ref.blocks), (sorted(dev.blocks), sorted(ref.blocks))
        for k in ref.blocks:
            assert dev.blocks[k].shape == ref.blocks[k].shape


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.linalg.norm(b), 1e-300)
    return float(np.linalg.norm(a - b) / den)


def assert_values_close(dev: BTensor, ref: BTensor, rtol):
    """Per-block relative Frobenius error against the block norm of the whole
    tensor (a block that is tiny compared to the tensor is compared in
    absolute terms)."""
    scale = max(np.sqrt(sum(float(np.sum(b * b)) for b in ref.blocks.values())), 1e-300)
    for k, rb in ref.blocks.items():
        err = float(np.linalg.norm(dev.blocks[k] - rb)) / scale
        assert err <= rtol, (k, err)
