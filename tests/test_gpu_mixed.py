"""GPU tests of the mixed-precision SVD (svd_gpu_mixed): FP32-storage
bidiagonalization with an FP32 / split-BF16 / BF16 trailing update, FP64
bidiagonal solve, FP32 back-transformation, FP64 refinement steps.

Bounds (measured behaviour with margin; the FP64 reference is LAPACK on the
host for the Gaussian case and the exact singular values by construction for
the graded case):
  Gaussian 1024^2, FP32 update:  1 step  |dsigma|/sigma_max <= 1e-11,
                                         residual <= 1e-8
                                 2 steps |dsigma|/sigma_max <= 1e-13,
                                         residual <= 1e-12, orthogonality <= 1e-11
  A = Q diag(10^(-6 i/(n-1))) P^T, n = 1024, FP32 update, 1 step:
                                 |dsigma|/sigma_max <= 2e-9, residual <= 1e-7;
                                 sigma >= 1e-2: relative error <= 1e-8
                                 (the FP32 stage does not resolve the vectors of
                                 the smallest values: their relative gap 1.7e-3
                                 is below its backward error)
  split-BF16 update, 1 step:     |dsigma|/sigma_max <= 1e-6
  BF16 update, no refinement:    |dsigma|/sigma_max <= 1e-2
  Gaussian 1024^2, FP32 update, SVD_GPU_MIXED_SEPARATED, 2 steps:
                                 |dsigma|/sigma_max <= 1e-13, residual <= 1e-12,
                                 orthogonality <= 1e-11
  graded 1024^2, FP32 update, 2 steps, SVD_GPU_MIXED_RESOLVE:
                                 |dsigma|/sigma_max <= 1e-10, residual <= 1e-10
                                 (measured 5.4e-12 and 3.2e-12), orthogonality
                                 <= 1e-4 (measured 2.9e-6)
"""
import ctypes

import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _run(ddcsvd, A, update, steps):
    At = torch.from_numpy(np.asfortranarray(A)).cuda()
    s, U, V, diag = ddcsvd.svd_mixed_torch(At, update, steps)
    torch.cuda.synchronize()
    s, U, V = s.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy()
    n = A.shape[0]
    res = np.linalg.norm(A - (U * s) @ V.T) / np.linalg.norm(A)
    orth = max(np.linalg.norm(U.T @ U - np.eye(n)), np.linalg.norm(V.T @ V - np.eye(n)))
    return s, res, orth, diag


def test_mixed_gaussian_refinement_converges(ddcsvd):
    n = 1024
    gpu_case(f"mixed gaussian {n}")
    A = gen.normal_matrix(n, n, seed=5)
    sl = np.sort(np.linalg.svd(A, compute_uv=False))
    s1, r1, o1, d1 = _run(ddcsvd, A, "f32", 1)
    e1 = np.max(np.abs(s1 - sl)) / sl[-1]
    s2, r2, o2, _ = _run(ddcsvd, A, "f32", 2)
    e2 = np.max(np.abs(s2 - sl)) / sl[-1]
    print(f"[mixed] gaussian {n}: 1 step sigma {e1:.2e} residual {r1:.2e}; 2 steps sigma "
          f"{e2:.2e} residual {r2:.2e} orth {o2:.2e}; before refinement orth "
          f"{d1['orth_u_before']:.2e}", flush=True)
    assert e1 <= 1e-11 and r1 <= 1e-8
    assert e2 <= 1e-13 and r2 <= 1e-12 and o2 <= 1e-11


def test_mixed_graded_per_sigma_error(ddcsvd):
    n = 1024
    gpu_case(f"mixed graded {n}")
    sig = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    A = gen.graded_matrix(n, n, sig, seed=3)
    exact = np.sort(sig)
    s, res, orth, _ = _run(ddcsvd, A, "f32", 1)
    err = np.abs(s - exact)
    big = exact >= 1e-2
    rel_big = np.max(err[big] / exact[big])
    print(f"[mixed] graded {n}: max |dsigma|/sigma_max {err.max():.2e}, sigma >= 1e-2 max rel "
          f"{rel_big:.2e}, median rel {np.median(err / exact):.2e}, residual {res:.2e}, "
          f"orth {orth:.2e}", flush=True)
    assert err.max() <= 2e-9
    assert rel_big <= 1e-8
    assert res <= 1e-7


def test_mixed_graded_resolve_reaches_residual_bar(ddcsvd):
    # the graded spectrum's bottom values have absolute gaps below the FP32
    # stage's error; SVD_GPU_MIXED_RESOLVE gives them an FP64 Rayleigh-Ritz
    # step on their projected block after the refinement steps
    n = 1024
    gpu_case(f"mixed graded resolve {n}")
    sig = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    A = gen.graded_matrix(n, n, sig, seed=3)
    exact = np.sort(sig)
    At = torch.from_numpy(np.asfortranarray(A)).cuda()
    s, U, V, diag = ddcsvd.svd_mixed_torch(At, "f32", 2, resolve=True)
    torch.cuda.synchronize()
    s, U, V = s.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy()
    err = np.abs(s - exact)
    res = np.linalg.norm(A - (U * s) @ V.T) / np.linalg.norm(A)
    orth = max(np.linalg.norm(U.T @ U - np.eye(n)), np.linalg.norm(V.T @ V - np.eye(n)))
    print(f"[mixed] graded {n} resolve: {int(diag['resolved'])} values resolved in "
          f"{diag['resolve_ms']:.1f} ms; max |dsigma|/sigma_max {err.max():.2e}, residual "
          f"{res:.2e}, orth {orth:.2e}", flush=True)
    assert 0 < diag["resolved"] < n
    assert np.all(np.diff(s) >= 0)
    assert err.max() <= 1e-10
    assert res <= 1e-10
    assert orth <= 1e-4


def test_mixed_resolve_leaves_separated_spectra_alone(ddcsvd):
    # a Gaussian spectrum has no unresolved pairs after two steps of the
    # larger bound: the resolve stage finds nothing and changes nothing
    n = 512
    gpu_case(f"mixed resolve gaussian {n}")
    A = gen.normal_matrix(n, n, seed=8)
    At = torch.from_numpy(np.asfortranarray(A)).cuda()
    s0, U0, V0, _ = ddcsvd.svd_mixed_torch(At, "f32", 2, separated=True)
    s1, U1, V1, d1 = ddcsvd.svd_mixed_torch(At, "f32", 2, separated=True, resolve=True)
    torch.cuda.synchronize()
    print(f"[mixed] gaussian {n} resolve: {int(d1['resolved'])} values resolved", flush=True)
    assert d1["resolved"] == 0
    assert torch.equal(s0, s1) and torch.equal(U0, U1) and torch.equal(V0, V1)


@pytest.mark.parametrize("update,steps,bound", [("bf16x3", 1, 1e-6), ("bf16", 0, 1e-2)])
def test_mixed_bf16_updates(ddcsvd, update, steps, bound):
    n = 512
    gpu_case(f"mixed {update} {n}")
    A = gen.normal_matrix(n, n, seed=6)
    sl = np.sort(np.linalg.svd(A, compute_uv=False))
    s, res, orth, _ = _run(ddcsvd, A, update, steps)
    e = np.max(np.abs(s - sl)) / sl[-1]
    print(f"[mixed] {update} steps {steps}: sigma {e:.2e} residual {res:.2e}", flush=True)
    assert np.all(np.isfinite(s)) and e <= bound


def test_mixed_separated_flag_reaches_fp64_accuracy(ddcsvd):
    # Gaussian spectrum: its close values come in isolated pairs, and the
    # larger correction bound resolves them in two steps
    n = 1024
    gpu_case(f"mixed separated {n}")
    A = gen.normal_matrix(n, n, seed=7)
    sl = np.sort(np.linalg.svd(A, compute_uv=False))
    At = torch.from_numpy(np.asfortranarray(A)).cuda()
    s, U, V, _ = ddcsvd.svd_mixed_torch(At, "f32", 2, separated=True)
    torch.cuda.synchronize()
    s, U, V = s.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy()
    e = np.max(np.abs(s - sl)) / sl[-1]
    res = np.linalg.norm(A - (U * s) @ V.T) / np.linalg.norm(A)
    orth = max(np.linalg.norm(U.T @ U - np.eye(n)), np.linalg.norm(V.T @ V - np.eye(n)))
    print(f"[mixed] separated 2 steps: sigma {e:.2e} residual {res:.2e} orth {orth:.2e}",
          flush=True)
    assert e <= 1e-13 and res <= 1e-12 and orth <= 1e-11


@pytest.mark.parametrize("n", [1, 2, 3, 37, 777])
def test_mixed_odd_sizes(ddcsvd, n):
    # shapes off every tile boundary of the FP32 stages, two refinement steps
    gpu_case(f"mixed odd size {n}")
    A = gen.normal_matrix(n, n, seed=11 + n)
    sl = np.sort(np.linalg.svd(A, compute_uv=False))
    s, res, orth, _ = _run(ddcsvd, A, "f32", 2)
    e = np.max(np.abs(s - sl)) / sl[-1]
    print(f"[mixed] n={n}: sigma {e:.2e} residual {res:.2e} orth {orth:.2e}", flush=True)
    assert np.all(np.isfinite(s)) and e <= 1e-12 and res <= 1e-9 and orth <= 1e-9


def test_mixed_rejects_bad_arguments(ddcsvd):
    gpu_case("mixed argument checks")
    lib = ddcsvd.load()
    A = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    s = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    assert lib.svd_gpu_mixed(8, p(A), 8, p(s), p(A), 8, p(A), 8, 7, 1, None, None) == ddcsvd.E_ARG
    assert lib.svd_gpu_mixed(0, p(A), 8, p(s), p(A), 8, p(A), 8, 1, 1, None, None) == ddcsvd.E_ARG
    assert lib.svd_gpu_mixed(8, p(A), 4, p(s), p(A), 8, p(A), 8, 1, 1, None, None) == ddcsvd.E_ARG
    assert lib.svd_gpu_mixed(8, p(A), 8, p(s), p(A), 8, p(A), 8, 1, 9, None, None) == ddcsvd.E_ARG
    assert lib.svd_gpu_mixed(8, p(A), 8, p(s), p(A), 8, p(A), 8, 1 | 0x400, 1, None,
                             None) == ddcsvd.E_ARG
    # the resolve stage needs a refinement step to find the unresolved pairs
    assert lib.svd_gpu_mixed(8, p(A), 8, p(s), p(A), 8, p(A), 8, 1 | 0x200, 0, None,
                             None) == ddcsvd.E_ARG
