"""The drop-in boundary from C: tests/c/whole_svd_dropin.c calls the
`void svd_gpu` symbol exactly as the reference's driver does
(test-whole-svd.c:75), compiled with gcc -std=gnu99 against
include/svd_gpu.h and linked with -lddcsvd (the link line INTEGRATION.md
gives).  The C program asserts residual, orthogonality and the reflector
layout of the overwritten A itself; this test compares what it writes back
with LAPACK, the oracle and the reference's golden fixtures.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "whole_svd_dropin.c")
LIBDIR = os.path.join(ROOT, "ddc-svd_amd", "lib")


def build_dropin(out_dir: str) -> str:
    exe = os.path.join(out_dir, "whole_svd_dropin")
    cmd = ["gcc", "-std=gnu99", "-O2", SRC, "-I", os.path.join(ROOT, "include"),
           "-L", LIBDIR, "-lddcsvd", f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_compiles_against_header(tmp_path):
    """CPU: the C caller compiles and links against the library's exports."""
    if not os.path.exists(os.path.join(LIBDIR, "libddcsvd.so")):
        pytest.skip("libddcsvd.so not built")
    exe = build_dropin(str(tmp_path))
    assert os.path.exists(exe)


def _run(exe, m, n, src, out):
    r = subprocess.run([exe, str(m), str(n), src, out], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    data = np.fromfile(out, dtype=np.float64)
    mn = min(m, n)
    return data[:mn], data[mn:].reshape(n, m).T


@pytest.mark.gpu
def test_dropin_literal_driver(tmp_path, po):
    """C1-literal: `test-whole-svd 512 256` is a 512 x 512 U[1,4) matrix
    from glibc rand() (the driver reads n from argv[1] too)."""
    exe = build_dropin(str(tmp_path))
    gpu_case("C drop-in svd_gpu 512x512 glibc rand")
    s, Am = _run(exe, 512, 512, "rand", str(tmp_path / "lit.bin"))
    A = gen.glibc_rand_matrix(512, 512)
    exact = np.sort(np.linalg.svd(A, compute_uv=False))
    assert np.max(np.abs(s - exact)) <= 1e-12 * exact[-1]
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_fixtures.npz"))
    # the reference's own values are accurate to about 1e-5 sigma_max
    assert np.max(np.abs(s - g["c1lit_sigma"])) <= 1e-4 * exact[-1]
    Ao, _, _ = po.oracle_bidiag(A)
    assert np.max(np.abs(Am - Ao)) <= 1e-11


@pytest.mark.gpu
def test_dropin_c1(tmp_path, po):
    """C1: 512 x 256 U(-1,1) seed 1, read from a file."""
    exe = build_dropin(str(tmp_path))
    A = gen.uniform_matrix(512, 256, 1)
    path = tmp_path / "a.bin"
    np.asfortranarray(A).T.tofile(path)  # column-major
    gpu_case("C drop-in svd_gpu 512x256 U(-1,1)")
    s, Am = _run(exe, 512, 256, f"file:{path}", str(tmp_path / "c1.bin"))
    exact = np.sort(np.linalg.svd(A, compute_uv=False))
    assert np.max(np.abs(s - exact)) <= 1e-12 * exact[-1]
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_fixtures.npz"))
    assert np.max(np.abs(s - g["c1_sigma"])) <= 1e-4 * exact[-1]
    Ao, _, _ = po.oracle_bidiag(A)
    assert np.max(np.abs(Am - Ao)) <= 1e-11


@pytest.mark.gpu
def test_dropin_wide_and_identity(tmp_path):
    """m < n and a repeated spectrum through the same C entry point."""
    exe = build_dropin(str(tmp_path))
    for name, A in (("wide 96x160", gen.normal_matrix(96, 160, 12)),
                    ("identity 200", np.eye(200))):
        path = tmp_path / "a.bin"
        np.asfortranarray(A).T.tofile(path)
        m, n = A.shape
        gpu_case(f"C drop-in svd_gpu {name}")
        s, _ = _run(exe, m, n, f"file:{path}", str(tmp_path / "o.bin"))
        exact = np.sort(np.linalg.svd(A, compute_uv=False))
        assert np.max(np.abs(s - exact)) <= 1e-12 * exact[-1]
