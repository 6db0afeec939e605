"""The C ABI library loads without a GPU and exports every function that
include/quantit_amd.h declares; the Python mirror declares all of them.
No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "quantit_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qtt_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import quantit_amd as qa
    if not qa.LIB_PATH.exists():
        from quantit_amd.build import build
        build()
    return ctypes.CDLL(str(qa.LIB_PATH))


def test_header_declares_the_boundary():
    fns = header_functions()
    for name in ["qtt_tensordot", "qtt_svd", "qtt_svd_truncated", "qtt_eigh", "qtt_truncate", "qtt_dmrg",
                 "qtt_btensor_create", "qtt_context_create"]:
        assert name in fns


def test_library_exports_every_declared_symbol(lib):
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_python_mirror_declares_every_symbol():
    import quantit_amd as qa
    assert sorted(qa.SIGNATURES) == header_functions()


def test_error_channel_without_gpu(lib):
    """Argument errors come back as a status and a message, no exception crosses the ABI."""
    lib.qtt_last_error.restype = ctypes.c_char_p
    lib.qtt_btensor_info.restype = ctypes.c_int
    st = lib.qtt_btensor_info(None, None, None, None, None, None)
    assert st == 1
    assert b"null pointer" in lib.qtt_last_error()


def test_c_client_compiles_against_the_header(tmp_path):
    """The header is usable from plain C99: the client compiles and links
    against the library (running it needs a GPU)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    out = tmp_path / "client"
    res = subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                          os.path.join(ROOT, "tests", "abi_client.c"), "-o", str(out),
                          "-L", os.path.join(ROOT, "quantit_amd"), "-lquantit_amd", "-lm"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_gemm_tuning_table_json_round_trip(tmp_path):
    """The tile-variant table loads from and saves to JSON (no GPU needed)."""
    import quantit_amd as qa
    src = tmp_path / "in.json"
    src.write_text('{"gemm_variants": {\n  "f64 m7 n7 t9": 2,\n  "f64 m9 n9 t12": 4\n}}\n')
    before = qa.gemm_tuning_size()
    qa.gemm_tuning_load(src)
    assert qa.gemm_tuning_size() >= before + 0 and qa.gemm_tuning_size() >= 2
    out = tmp_path / "out.json"
    qa.gemm_tuning_save(out)
    text = out.read_text()
    assert '"f64 m7 n7 t9": 2' in text and '"f64 m9 n9 t12": 4' in text
    bad = tmp_path / "bad.json"
    bad.write_text('{"gemm_variants": {"x": 99}}')
    with pytest.raises(ValueError):
        qa.gemm_tuning_load(bad)
