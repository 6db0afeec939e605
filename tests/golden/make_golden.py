"""Writes the golden fixtures under tests/golden/fixtures/ with the numpy
oracle, in the library's btensor file format (quantit_amd/io.py): inputs and
expected outputs of a contraction (the C<5> case of
include/blockTensor/btensor.h:1585-1608 and a U(1) rank-4 case), a
truncated block SVD, a two-site one-step Lanczos update and three DMRG
sweeps of an 8-site chain.  Scalars go to results.json beside the tensors.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from oracle import btensor_oracle as bo  # noqa: E402
from oracle import dmrg_oracle as dor  # noqa: E402
from quantit_amd.io import save_parts  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "fixtures")


def save(case, name, t):
    d = os.path.join(OUT, case)
    os.makedirs(d, exist_ok=True)
    save_parts(os.path.join(d, name + ".npz"), t.groups, t.secs, t.cvals, t.sel, t.blocks)


def c5_contraction():
    rng = np.random.default_rng(0)
    C5 = (5,)
    A = bo.BTensor.from_sections(C5, [[(2, (0,)), (3, (1,))], [(2, (0,)), (3, (4,))]], (0,))
    A.blocks[(0, 0)] = rng.uniform(size=(2, 2))
    A.blocks[(1, 1)] = rng.uniform(size=(3, 3))
    B = bo.BTensor.from_sections(C5, [[(3, (4,)), (2, (0,))], [(2, (0,)), (3, (1,))], [(1, (1,)), (3, (0,))]], (1,))
    B.blocks[(1, 0, 0)] = rng.uniform(size=(2, 2, 1))
    B.blocks[(0, 1, 0)] = rng.uniform(size=(3, 3, 1))
    B.blocks[(1, 1, 1)] = rng.uniform(size=(2, 3, 3))
    save("td_c5", "a", A)
    save("td_c5", "b", B)
    save("td_c5", "out", bo.tensordot(A, B, [1], [1]))
    return {"dims_a": [1], "dims_b": [1]}


def _leg(rng, n, lo=-3, hi=3, smax=4):
    qs = rng.choice(np.arange(lo, hi + 1), size=n, replace=False)
    return [(int(rng.integers(1, smax + 1)), (int(q),)) for q in sorted(qs)]


def u1_contraction():
    rng = np.random.default_rng(1)
    l0, l1, l2, l3 = (_leg(rng, 4) for _ in range(4))
    inv = lambda leg: [(s, (-c[0],)) for s, c in leg]  # noqa: E731
    A = bo.rand_like(bo.BTensor.from_sections((0,), [l0, l1, l2, l3], (0,)), rng, "normal")
    B = bo.rand_like(bo.BTensor.from_sections((0,), [inv(l2), inv(l1), _leg(rng, 3)], (1,)), rng, "normal")
    save("td_u1", "a", A)
    save("td_u1", "b", B)
    save("td_u1", "out", bo.tensordot(A, B, [2, 1], [0, 1]))
    return {"dims_a": [2, 1], "dims_b": [0, 1]}


def truncated_svd():
    rng = np.random.default_rng(2)
    rows = [(6, (0,)), (5, (2,)), (4, (-2,))]
    phys = [(1, (1,)), (1, (-1,))]
    legs = [rows, phys, [(s, (-c[0] + 1,)) for s, c in rows] + [(3, (-3,))]]
    t = bo.BTensor.from_sections((0,), legs, (0,))
    for idx in t.all_allowed_indices():
        # graded singular values: blocks scaled by a decaying factor
        t.blocks[idx] = rng.standard_normal(t.block_shape(idx)) * 10.0 ** (-sum(idx))
    U, d, V = bo.svd_truncated(t, 2, 1e-3, 1, 9)
    for n, x in (("in", t), ("U", U), ("d", d), ("V", V)):
        save("svd_u1", n, x)
    return {"split": 2, "tol": 1e-3, "min_size": 1, "max_size": 9}


def two_site_update():
    rng = np.random.default_rng(3)
    L = 8
    mpo = dor.guess_bmpo(dor.drop_zero_channels(dor.xxz_field_dense(0.7, rng.uniform(-1, 1, L))),
                         dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(6, mpo, (0,), 3))
    env = dor.generate_env(mpo, mps, 0)
    W12 = dor.two_site_hamil(mpo)[0]
    theta = bo.tensordot(mps[0], mps[1], [2], [0])
    psi, a0, a1, b = dor.one_step_lanczos(theta, W12, env[-1], env[2])
    E, o, n = dor.eig2x2(a0, a1, b)
    for nm, x in (("theta", theta), ("W12", W12), ("L", env[-1]), ("R", env[2]), ("psi1", psi)):
        save("twosite", nm, x)
    return {"a0": a0, "a1": a1, "b": b, "E": E, "o": o, "n": n}


def dmrg_sweeps():
    rng = np.random.default_rng(4)
    L = 8
    mpo = dor.guess_bmpo(dor.drop_zero_channels(dor.xxz_field_dense(0.9, rng.uniform(-1, 1, L))),
                         dor.u1_spin_half_phys())
    mps = dor.prepare_state(dor.random_bmps(4, mpo, (0,), 5))
    for i, (w, a) in enumerate(zip(mpo, mps)):
        save("dmrg8", f"mpo{i}", w)
        save("dmrg8", f"mps{i}", a)
    opts = dict(cutoff=1e-8, convergence_criterion=0.0, maximum_bond=12, minimum_bond=4, maximum_iterations=3)
    E, out, Es = dor.dmrg(mpo, mps, dor.DmrgOptions(**opts))
    for i, a in enumerate(out):
        save("dmrg8", f"final{i}", a)
    return {"L": L, "options": opts, "energies": Es}


def main():
    results = {"td_c5": c5_contraction(), "td_u1": u1_contraction(), "svd_u1": truncated_svd(),
               "twosite": two_site_update(), "dmrg8": dmrg_sweeps()}
    with open(os.path.join(OUT, "results.json"), "w") as f:
        json.dump(results, f, indent=1, sort_keys=True)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
