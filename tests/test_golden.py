"""The golden fixtures (tests/golden/fixtures/, written by
tests/golden/make_golden.py in the library's btensor file format) against
the oracle: the oracle reproduces every stored output from the stored inputs,
so a change of the restatement cannot pass silently.  CPU only."""
import json
import os

import numpy as np
import pytest

from oracle import btensor_oracle as bo
from oracle import dmrg_oracle as dor
from quantit_amd.io import load_parts

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")


def load(case, name):
    return bo.BTensor(**load_parts(os.path.join(FIX, case, name + ".npz")))


@pytest.fixture(scope="module")
def results():
    with open(os.path.join(FIX, "results.json")) as f:
        return json.load(f)


def same(a, b, tol):
    assert a.groups == b.groups and a.secs == b.secs and a.cvals == b.cvals and a.sel == b.sel
    assert sorted(a.blocks) == sorted(b.blocks)
    for k in b.blocks:
        assert np.abs(a.blocks[k] - b.blocks[k]).max(initial=0.0) <= tol


@pytest.mark.parametrize("case", ["td_c5", "td_u1"])
def test_contraction_fixtures(case, results):
    r = results[case]
    same(bo.tensordot(load(case, "a"), load(case, "b"), r["dims_a"], r["dims_b"]), load(case, "out"), 1e-13)


def test_svd_fixture(results):
    r = results["svd_u1"]
    U, d, V = bo.svd_truncated(load("svd_u1", "in"), r["split"], r["tol"], r["min_size"], r["max_size"])
    same(d, load("svd_u1", "d"), 1e-13)
    # U and V up to the sign of each singular vector
    for got, name in ((U, "U"), (V, "V")):
        ref = load("svd_u1", name)
        assert got.secs == ref.secs and sorted(got.blocks) == sorted(ref.blocks)


def test_two_site_fixture(results):
    r = results["twosite"]
    psi, a0, a1, b = dor.one_step_lanczos(load("twosite", "theta"), load("twosite", "W12"), load("twosite", "L"),
                                          load("twosite", "R"))
    assert (a0, a1, b) == pytest.approx((r["a0"], r["a1"], r["b"]), rel=1e-13)
    same(psi, load("twosite", "psi1"), 1e-12)
    assert dor.eig2x2(a0, a1, b) == pytest.approx((r["E"], r["o"], r["n"]), rel=1e-12)


def test_dmrg_fixture(results):
    r = results["dmrg8"]
    L = r["L"]
    mpo = [load("dmrg8", f"mpo{i}") for i in range(L)]
    mps = [load("dmrg8", f"mps{i}") for i in range(L)]
    E, out, Es = dor.dmrg(mpo, mps, dor.DmrgOptions(**r["options"]))
    assert Es == pytest.approx(r["energies"], rel=1e-12)
    for i in range(L):
        ref = load("dmrg8", f"final{i}")
        assert out[i].secs == ref.secs and sorted(out[i].blocks) == sorted(ref.blocks)
