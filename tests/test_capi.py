"""CPU checks of the C-ABI library: it loads, exports every function that
include/svd_gpu.h declares, and links nothing from the oracle.  No compute
call is made (this runs without a GPU)."""
import os
import subprocess

import pytest

import ddcsvd


def test_library_exports_declared_symbols():
    if not os.path.exists(ddcsvd.LIB_PATH):
        pytest.skip("libddcsvd.so not built (python ddc-svd_amd/build.py)")
    lib = ddcsvd.load()
    names = ddcsvd.declared_symbols()
    for required in ("svd_gpu", "svd_gpu_status", "svd_gpu_dev", "bidiag_gpu",
                     "bdsvd_values_gpu", "bdsvd_vectors_gpu", "svd_gpu_last_timings",
                     "svd_gpu_last_error"):
        assert required in names
    for name in names:
        assert hasattr(lib, name), f"{name} declared but not exported"


def test_library_does_not_contain_the_oracle():
    if not os.path.exists(ddcsvd.LIB_PATH):
        pytest.skip("libddcsvd.so not built")
    out = subprocess.run(["nm", "-D", ddcsvd.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in out and "ref_svd" not in out and "bidiag_seq" not in out


def test_missing_library_raises(monkeypatch):
    monkeypatch.setattr(ddcsvd, "_lib", None)
    monkeypatch.setattr(ddcsvd, "LIB_PATH", "/nonexistent/libddcsvd.so")
    with pytest.raises(FileNotFoundError):
        ddcsvd.load()
