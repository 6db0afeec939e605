"""The device against the golden fixtures (tests/golden/fixtures/): every
stored output is reproduced from the stored inputs through the C ABI, with no
oracle involved at run time: structure bit-exact, values within 1e-12."""
import json
import os

import numpy as np
import pytest

from quantit_amd.io import load_parts

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")


def parts(case, name):
    return load_parts(os.path.join(FIX, case, name + ".npz"))


def dev(ctx, case, name):
    import quantit_amd as qa
    return qa.BTensor.from_parts(ctx, **parts(case, name))


def host(t):
    g, secs, cvals, sel, blocks = t.to_parts()
    return dict(groups=g, secs=secs, cvals=cvals, sel=sel, blocks=blocks)


def same(got, ref, tol):
    assert tuple(got["groups"]) == tuple(ref["groups"])
    assert got["secs"] == ref["secs"] and got["cvals"] == ref["cvals"] and tuple(got["sel"]) == tuple(ref["sel"])
    assert sorted(got["blocks"]) == sorted(ref["blocks"])
    scale = max(max((np.abs(b).max(initial=0.0) for b in ref["blocks"].values()), default=0.0), 1e-300)
    for k in ref["blocks"]:
        assert np.abs(got["blocks"][k] - ref["blocks"][k]).max(initial=0.0) <= tol * scale, k


@pytest.fixture(scope="module")
def results():
    with open(os.path.join(FIX, "results.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", ["td_c5", "td_u1"])
def test_contraction_fixtures(ctx, case, results):
    r = results[case]
    out = dev(ctx, case, "a").tensordot(dev(ctx, case, "b"), r["dims_a"], r["dims_b"])
    same(host(out), parts(case, "out"), 1e-12)


def test_svd_fixture(ctx, results):
    import quantit_amd as qa
    r = results["svd_u1"]
    U, d, V = qa.svd(dev(ctx, "svd_u1", "in"), r["split"], tol=r["tol"], min_size=r["min_size"],
                     max_size=r["max_size"])
    same(host(d), parts("svd_u1", "d"), 1e-12)
    for got, name in ((U, "U"), (V, "V")):
        g, ref = host(got), parts("svd_u1", name)
        assert g["secs"] == ref["secs"] and sorted(g["blocks"]) == sorted(ref["blocks"])


def test_two_site_fixture(ctx, results):
    import quantit_amd as qa
    r = results["twosite"]
    Hpsi, a0, a1, b = qa.one_step_lanczos(*(dev(ctx, "twosite", n) for n in ("theta", "W12", "L", "R")))
    assert (a0, a1, b) == pytest.approx((r["a0"], r["a1"], r["b"]), rel=1e-12)
    same(host(Hpsi), parts("twosite", "psi1"), 1e-11)
    assert qa.eig2x2(ctx, a0, a1, b) == pytest.approx((r["E"], r["o"], r["n"]), rel=1e-11)


def test_dmrg_fixture(ctx, results):
    import quantit_amd as qa
    r = results["dmrg8"]
    L = r["L"]
    o = r["options"]
    E, out, lg = qa.dmrg([dev(ctx, "dmrg8", f"mpo{i}") for i in range(L)],
                         [dev(ctx, "dmrg8", f"mps{i}") for i in range(L)],
                         qa.DmrgOptions(cutoff=o["cutoff"], convergence_criterion=o["convergence_criterion"],
                                        maximum_bond=o["maximum_bond"], minimum_bond=o["minimum_bond"],
                                        maximum_iterations=o["maximum_iterations"]))
    assert lg["energies"] == pytest.approx(r["energies"], rel=1e-10)
    for i in range(L):
        g, ref = host(out[i]), parts("dmrg8", f"final{i}")
        assert g["secs"] == ref["secs"] and sorted(g["blocks"]) == sorted(ref["blocks"])
