"""Synthetic inputs of the benchmark workloads (plain numpy, no device code).

contract_svd_D1024 -- the standalone block-sparse contraction + SVD workload:
  the environment-update contraction X(b, a', s, w') . Abar(a', s, b') over
  (a', s) (the pattern of sources/dmrg.cpp:455-457), then the block SVD of
  the rank-3 result with split 1.
  * D legs (b, a', b'): U(1) charges q = -8..8, section sizes = histogram of
    1024 draws of clip(round(N(0, 3^2)), -8, 8) from numpy's default_rng(0).
  * w leg: the 26 bond charges of the width-4 cylinder MPO at its centre
    (tests/2dHeisenberg/numsQNs_16.txt:4 of QuantiT), consecutive equal
    charges forming one section as the reference's charge inference does
    (tests/2Dheisenberg.cpp:249-270).
  * s leg: spin-1/2, charges (+1, -1) in 2Sz units.
  * X selection rule 0, Abar legs inverse to X's (a', s) and rule 0.
  * entries i.i.d. N(0, 1): X from default_rng(1), Abar from default_rng(2),
    filled block by block in lexicographic block order.
"""
from __future__ import annotations

import itertools

import numpy as np

U1 = (0,)
CYL_W_CHARGES = [0, 0, 2, -2, 0, 2, -2, 0, 2, -2, 0, 2, -2, 0, 2, -2, 0, 2, -2, 0, 2, -2, 0, 2, -2, 0]
PHYS = [(1, (1,)), (1, (-1,))]


def d_leg(D=1024, seed=0, qmax=8, sigma=3.0):
    rng = np.random.default_rng(seed)
    x = np.clip(np.round(rng.normal(0.0, sigma, D)), -qmax, qmax).astype(int)
    counts = np.bincount(x + qmax, minlength=2 * qmax + 1)
    return [(int(c), (q,)) for q, c in zip(range(-qmax, qmax + 1), counts) if c > 0]


def runs_leg(charges):
    out = []
    for c in charges:
        if out and out[-1][1] == (c,):
            out[-1] = (out[-1][0] + 1, (c,))
        else:
            out.append((1, (c,)))
    return out


def inverse(leg):
    return [(s, tuple(-x for x in c)) for s, c in leg]


def allowed_blocks(legs, sel):
    """Lexicographically ordered block indices whose charges sum to sel (U(1))."""
    out = []
    for idx in itertools.product(*[range(len(l)) for l in legs]):
        if sum(legs[d][i][1][0] for d, i in enumerate(idx)) == sel[0]:
            out.append(idx)
    return out


def random_parts(legs, sel, seed):
    rng = np.random.default_rng(seed)
    blocks = {}
    for idx in allowed_blocks(legs, sel):
        shape = tuple(legs[d][i][0] for d, i in enumerate(idx))
        blocks[idx] = rng.standard_normal(shape)
    return dict(groups=U1, secs=[[s for s, _ in l] for l in legs], cvals=[[c for _, c in l] for l in legs],
                sel=sel, blocks=blocks)


def contract_svd(D=1024):
    """Returns (X parts, Abar parts, dims_x, dims_a, svd split)."""
    dl = d_leg(D)
    wl = runs_leg(CYL_W_CHARGES)
    X = random_parts([dl, dl, PHYS, wl], (0,), seed=1)
    A = random_parts([inverse(dl), inverse(PHYS), dl], (0,), seed=2)
    return X, A, [1, 2], [0, 1], 1


def _loguniform_ints(rng, n, lo=4, hi=768):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n))).astype(int).clip(lo, hi)


def grouped_batch(seed=2, n_gemm=2048, n_svd=512):
    """Grouped batch of independent blocks as block-diagonal btensors: item i
    gets its own U(1) charge i, so A(m_i x k_i) . B(k_i x n_i) are the only
    matching block pairs of one tensordot, and S(m_i x n_i) are the sectors
    of one block SVD.  Dimensions log-uniform integers in [4, 768]."""
    rng = np.random.default_rng(seed)
    m, n, k = (_loguniform_ints(rng, n_gemm) for _ in range(3))
    sm, sn = (_loguniform_ints(rng, n_svd) for _ in range(2))

    def diag(rows, cols, fill):
        return dict(groups=(0,), secs=[list(map(int, rows)), list(map(int, cols))],
                    cvals=[[(i,) for i in range(len(rows))], [(-i,) for i in range(len(cols))]], sel=(0,),
                    blocks={(i, i): fill(int(r), int(c)) for i, (r, c) in enumerate(zip(rows, cols))})

    A = diag(m, k, lambda r, c: rng.standard_normal((r, c)))
    B = dict(groups=(0,), secs=[list(map(int, k)), list(map(int, n))],
             cvals=[[(i,) for i in range(n_gemm)], [(-i,) for i in range(n_gemm)]], sel=(0,),
             blocks={(i, i): rng.standard_normal((int(k[i]), int(n[i]))) for i in range(n_gemm)})
    S = diag(sm, sn, lambda r, c: rng.standard_normal((r, c)))
    return A, B, S, (m, n, k), (sm, sn)
