"""The plain C client of the ABI (tests/abi_client.c, built with the library)
run on the device: tensors from host arrays, a contraction, readback and an
error status, through nothing but include/quantit_amd.h."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

CLIENT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi_client")


def test_c_client_runs():
    assert os.path.exists(CLIENT), "tests/abi_client is built by quantit_amd.build"
    res = subprocess.run([CLIENT], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "abi client ok" in res.stdout
