"""Python access to libddcsvd (the MI355X svd_gpu path) through its C-ABI.

The functions mirror the reference's C entry points (svd_gpu.h:5,
bidiag_par.h:30, Calculations-Parallel.h:41, parallel-twisted.h:18-19):
column-major float64 arrays, singular values ascending, column i of U and V
paired with sigma[i].  Every call goes to the HIP library; there is no CPU
fallback, and a missing library raises.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libddcsvd.so")
HEADER = os.path.join(ROOT, "include", "svd_gpu.h")

OK, E_ARG, E_NOMEM, E_HIP, E_NODEV, E_NUMERIC = 0, -1, -2, -3, -4, -5


class SvdGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libddcsvd status {code}: {msg}")
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def declared_symbols() -> list[str]:
    """Names of the functions include/svd_gpu.h declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)) - {"if", "defined"})


def load():
    """Load the built library (raises if it was not built)."""
    global _lib
    if _lib is None:
        try:
            # PyTorch-ROCm carries its own HIP runtime (same soname); loading it
            # first makes this library bind to that one instance instead of
            # bringing a second runtime into the process.
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python ddc-svd_amd/build.py`")
        lib = ctypes.CDLL(LIB_PATH)
        i, d, v = ctypes.c_int, _dp, ctypes.c_void_p
        lib.svd_gpu.argtypes = [i, i, d, d, d, d]
        lib.svd_gpu.restype = None
        lib.svd_gpu_status.argtypes = [i, i, d, d, d, d]
        lib.svd_gpu_status.restype = i
        lib.svd_gpu_dev.argtypes = [i, i, v, i, v, v, i, v, i, v]
        lib.svd_gpu_dev.restype = i
        lib.bidiag_gpu.argtypes = [i, i, d, d, d]
        lib.bidiag_gpu.restype = i
        lib.bidiag_form_u_gpu.argtypes = [i, i, d, d]
        lib.bidiag_form_u_gpu.restype = i
        lib.bidiag_form_v_gpu.argtypes = [i, i, d, d]
        lib.bidiag_form_v_gpu.restype = i
        lib.bdsvd_values_gpu.argtypes = [i, d, d, d, i]
        lib.bdsvd_values_gpu.restype = i
        lib.bdsvd_vectors_gpu.argtypes = [i, d, d, d, d, d]
        lib.bdsvd_vectors_gpu.restype = i
        lib.svd_gpu_backward.argtypes = [i, i, v, i, v, v, i, v, i, v, v, i, v, i,
                                         ctypes.c_double, v]
        lib.svd_gpu_backward.restype = i
        lib.svd_gpu_mixed.argtypes = [i, v, i, v, v, i, v, i, i, i, d, v]
        lib.svd_gpu_mixed.restype = i
        lib.svd_gpu_last_timings.argtypes = [d, i]
        lib.svd_gpu_last_timings.restype = i
        lib.svd_gpu_profile_kernels.argtypes = [i]
        lib.svd_gpu_profile_kernels.restype = i
        lib.svd_gpu_profile_left_gemv.argtypes = [i, i, v, i, i, v]
        lib.svd_gpu_profile_left_gemv.restype = i
        lib.svd_gpu_profile_right_gemv.argtypes = [i, i, v, i, i, v]
        lib.svd_gpu_profile_right_gemv.restype = i
        lib.svd_gpu_profile_fused_gemv.argtypes = [i, i, v, i, i, v]
        lib.svd_gpu_profile_fused_gemv.restype = i
        lib.svd_gpu_last_error.argtypes = []
        lib.svd_gpu_last_error.restype = ctypes.c_char_p
        c = ctypes.c_char
        lib.svd_gpu_dgemm.argtypes = [c, c, i, i, i, ctypes.c_double, v, i, v, i,
                                      ctypes.c_double, v, i, v]
        lib.svd_gpu_dgemm.restype = i
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _check(rc: int):
    if rc != OK:
        raise SvdGpuError(rc, load().svd_gpu_last_error().decode())


def last_error() -> str:
    return load().svd_gpu_last_error().decode()


TIMING_KEYS = ["copy_in", "bidiag", "values", "vectors", "backtransform", "copy_out",
               "left_gemv_ms", "right_gemv_ms", "left_gemv_launches", "right_gemv_launches",
               "fused_gemv_ms", "fused_gemv_launches"]


def last_timings() -> dict:
    """Stage times (ms) of the last svd call (and, with profile_kernels(True),
    the summed GEMV launch times and counts)."""
    out = np.zeros(len(TIMING_KEYS))
    load().svd_gpu_last_timings(_p(out), len(TIMING_KEYS))
    return dict(zip(TIMING_KEYS, out.tolist()))


def profile_kernels(on: bool) -> None:
    """Bracket every bidiagonalization GEMV launch with HIP events."""
    load().svd_gpu_profile_kernels(1 if on else 0)


def profile_left_gemv(A, steps: int, stream=None) -> None:
    """svd_gpu_profile_left_gemv on a column-major float64 CUDA matrix: the
    first `steps` left-GEMV launches of its bidiagonalization, alone."""
    m, n = A.shape
    lda = _colmajor_ld(A, "A")
    _check(load().svd_gpu_profile_left_gemv(m, n, ctypes.c_void_p(A.data_ptr()), lda, steps,
                                            ctypes.c_void_p(stream)))


def profile_right_gemv(A, steps: int, stream=None) -> None:
    """svd_gpu_profile_right_gemv on a column-major float64 CUDA matrix: the
    first `steps` right-GEMV launches of its bidiagonalization, alone."""
    m, n = A.shape
    lda = _colmajor_ld(A, "A")
    _check(load().svd_gpu_profile_right_gemv(m, n, ctypes.c_void_p(A.data_ptr()), lda, steps,
                                             ctypes.c_void_p(stream)))


def profile_fused_gemv(A, steps: int, stream=None) -> None:
    """svd_gpu_profile_fused_gemv on a column-major float64 CUDA matrix: the
    first `steps` one-pass GEMV launches of its bidiagonalization, alone."""
    m, n = A.shape
    lda = _colmajor_ld(A, "A")
    _check(load().svd_gpu_profile_fused_gemv(m, n, ctypes.c_void_p(A.data_ptr()), lda, steps,
                                             ctypes.c_void_p(stream)))


def svd(A: np.ndarray, raw: bool = False):
    """svd_gpu on host arrays.  Returns (sigma, U[:, :mn], V[:, :mn], A_after).

    raw=True calls the int-returning svd_gpu_status and returns the status
    instead of raising."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    mn = min(m, n)
    sigma = np.zeros(mn)
    U = np.zeros((m, m), order="F")
    V = np.zeros((n, n), order="F")
    rc = load().svd_gpu_status(m, n, _p(A), _p(sigma), _p(U), _p(V))
    if raw:
        return rc, sigma, U, V, A
    _check(rc)
    U = U if U.shape[1] == mn else U[:, :mn].copy(order="F")
    V = V if V.shape[1] == mn else V[:, :mn].copy(order="F")
    return sigma, U, V, A


def svd_dev(A, sigma, U, V, stream=None) -> None:
    """svd_gpu_dev on torch.float64 CUDA tensors (column-major views: pass
    A.t().contiguous().t()-style tensors or use `svd_torch`)."""
    m, n = A.shape
    lda = A.stride(1)
    rc = load().svd_gpu_dev(m, n, ctypes.c_void_p(A.data_ptr()), lda,
                            ctypes.c_void_p(sigma.data_ptr()), ctypes.c_void_p(U.data_ptr()),
                            U.stride(1), ctypes.c_void_p(V.data_ptr()), V.stride(1),
                            ctypes.c_void_p(stream))
    _check(rc)


def col_major(t):
    """A column-major (Fortran-ordered) copy of a 2-D torch tensor."""
    return t.t().contiguous().t()


def svd_torch(A):
    """Device SVD of a CUDA float64 tensor; returns (sigma, U, V) tensors."""
    import torch
    A = col_major(A)
    m, n = A.shape
    mn = min(m, n)
    sigma = torch.empty(mn, dtype=torch.float64, device=A.device)
    U = col_major(torch.empty((m, mn), dtype=torch.float64, device=A.device))
    V = col_major(torch.empty((n, mn), dtype=torch.float64, device=A.device))
    stream = torch.cuda.current_stream(A.device).cuda_stream
    svd_dev(A, sigma, U, V, stream)
    return sigma, U, V


def _colmajor_ld(t, name: str) -> int:
    if t.dtype != _torch().float64 or not t.is_cuda or t.dim() != 2:
        raise ValueError(f"{name} must be a column-major float64 CUDA matrix")
    if t.numel() == 0:  # empty views carry no meaningful strides
        return max(t.shape[0], 1)
    if t.stride(0) != 1:
        raise ValueError(f"{name} must be a column-major float64 CUDA matrix")
    return max(t.stride(1), t.shape[0], 1)


def _torch():
    import torch
    return torch


def svd_backward_dev(U, sigma, V, gU, gsigma, gV, gA, gap_guard: float = 1e-13,
                     stream=None) -> None:
    """svd_gpu_backward on column-major float64 CUDA tensors.  gU, gsigma and gV
    may be None (zero cotangent); gA (m x n) is written."""
    m, mn = U.shape
    n = V.shape[0]
    ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
    ldu, ldv, ldga = _colmajor_ld(U, "U"), _colmajor_ld(V, "V"), _colmajor_ld(gA, "gA")
    ldgu = _colmajor_ld(gU, "gU") if gU is not None else 0
    ldgv = _colmajor_ld(gV, "gV") if gV is not None else 0
    if gsigma is not None and (not gsigma.is_contiguous() or gsigma.numel() != mn):
        raise ValueError("gsigma must be a contiguous vector of min(m, n) values")
    if tuple(gA.shape) != (m, n) or V.shape[1] != mn or sigma.numel() != mn:
        raise ValueError("inconsistent shapes")
    rc = load().svd_gpu_backward(m, n, ptr(U), ldu, ptr(sigma), ptr(V), ldv, ptr(gU), ldgu,
                                 ptr(gsigma), ptr(gV), ldgv, ptr(gA), ldga, float(gap_guard),
                                 ctypes.c_void_p(stream))
    _check(rc)


def svd_backward_torch(U, sigma, V, gU, gsigma, gV, gap_guard: float = 1e-13):
    """dL/dA for A = U diag(sigma) V^T from CUDA float64 tensors; returns a
    column-major tensor."""
    torch = _torch()
    m, n = U.shape[0], V.shape[0]
    cm = lambda t: None if t is None else col_major(t.to(torch.float64))  # noqa: E731
    gA = col_major(torch.empty((m, n), dtype=torch.float64, device=U.device))
    gs = None if gsigma is None else gsigma.to(torch.float64).contiguous()
    stream = torch.cuda.current_stream(U.device).cuda_stream
    svd_backward_dev(cm(U), sigma.contiguous(), cm(V), cm(gU), gs, cm(gV), gA, gap_guard,
                     stream)
    return gA


MIXED_UPDATES = {"f32": 1, "bf16x3": 2, "bf16": 3}
MIXED_DIAG_KEYS = ["orth_u_before", "orth_v_before", "bidiag_ms", "values_vectors_ms",
                   "backtransform_ms", "refine_ms", "convert_ms", "last_sigma_change",
                   "resolved", "resolve_ms"]


MIXED_SEPARATED = 0x100
MIXED_RESOLVE = 0x200


def svd_mixed_torch(A, update: str = "f32", refine_steps: int = 1, separated: bool = False,
                    resolve: bool = False):
    """svd_gpu_mixed on a square CUDA float64 tensor: (sigma, U, V, diag).
    separated: SVD_GPU_MIXED_SEPARATED (larger first-order corrections, for
    spectra without long runs of close values); resolve: SVD_GPU_MIXED_RESOLVE
    (FP64 Rayleigh-Ritz on the values the last step left unresolved; diag
    then also holds "resolved" and "resolve_ms")."""
    import torch
    A = col_major(A)
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("svd_gpu_mixed needs a square matrix")
    sigma = torch.empty(n, dtype=torch.float64, device=A.device)
    U = col_major(torch.empty((n, n), dtype=torch.float64, device=A.device))
    V = col_major(torch.empty((n, n), dtype=torch.float64, device=A.device))
    diag = np.zeros(10)
    stream = torch.cuda.current_stream(A.device).cuda_stream
    _check(load().svd_gpu_mixed(n, ctypes.c_void_p(A.data_ptr()), A.stride(1),
                                ctypes.c_void_p(sigma.data_ptr()), ctypes.c_void_p(U.data_ptr()),
                                U.stride(1), ctypes.c_void_p(V.data_ptr()), V.stride(1),
                                MIXED_UPDATES[update] | (MIXED_SEPARATED if separated else 0)
                                | (MIXED_RESOLVE if resolve else 0),
                                refine_steps, _p(diag),
                                ctypes.c_void_p(stream)))
    keys = MIXED_DIAG_KEYS if resolve else MIXED_DIAG_KEYS[:8]
    return sigma, U, V, dict(zip(keys, diag.tolist()))


def bidiag(A: np.ndarray):
    """bidiag_gpu: returns (A_with_reflectors, alpha, beta)."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    # m >= n: n x n upper bidiagonal (beta: n-1 entries); m < n: the
    # m x (m+1) upper bidiagonal of the reference (beta: m entries)
    nbeta = n - 1 if m >= n else m
    alpha = np.zeros(min(m, n))
    beta = np.zeros(max(nbeta, 1))
    _check(load().bidiag_gpu(m, n, _p(A), _p(alpha), _p(beta)))
    return A, alpha, beta[:nbeta]


def form_u(A_mod: np.ndarray) -> np.ndarray:
    """bidiag_form_u_gpu: U (m x m) from the reflectors left by bidiag()."""
    A_mod = np.array(A_mod, dtype=np.float64, order="F", copy=True)
    m, n = A_mod.shape
    U = np.zeros((m, m), order="F")
    _check(load().bidiag_form_u_gpu(m, n, _p(A_mod), _p(U)))
    return U


def form_v(A_mod: np.ndarray) -> np.ndarray:
    """bidiag_form_v_gpu: V (n x n) from the reflectors left by bidiag()."""
    A_mod = np.array(A_mod, dtype=np.float64, order="F", copy=True)
    m, n = A_mod.shape
    V = np.zeros((n, n), order="F")
    _check(load().bidiag_form_v_gpu(m, n, _p(A_mod), _p(V)))
    return V


def bdsvd_values(alpha, beta, refine: bool = False) -> np.ndarray:
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    n = len(alpha)
    beta = np.ascontiguousarray(np.append(np.asarray(beta, dtype=np.float64)[: n - 1], 0.0))
    sigma = np.zeros(n)
    _check(load().bdsvd_values_gpu(n, _p(alpha), _p(beta), _p(sigma), 1 if refine else 0))
    return sigma


def bdsvd_vectors(alpha, beta):
    """Returns (sigma, X, Y): B X[:, i] = sigma_i Y[:, i]."""
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    n = len(alpha)
    beta = np.ascontiguousarray(np.append(np.asarray(beta, dtype=np.float64)[: n - 1], 0.0))
    sigma = np.zeros(n)
    X = np.zeros((n, n), order="F")
    Y = np.zeros((n, n), order="F")
    _check(load().bdsvd_vectors_gpu(n, _p(alpha), _p(beta), _p(sigma), _p(X), _p(Y)))
    return sigma, X, Y


def dgemm_torch(transa: str, transb: str, A, B, C, alpha: float = 1.0, beta: float = 0.0,
                stream=None) -> None:
    """C = alpha op(A) op(B) + beta C through svd_gpu_dgemm (the library's
    FP64 MFMA GEMM) on column-major float64 CUDA tensors; C is updated in
    place."""
    torch = _torch()
    ta, tb = transa.upper() == "T", transb.upper() == "T"
    M, N = C.shape
    K = A.shape[0] if ta else A.shape[1]
    lda, ldb, ldc = _colmajor_ld(A, "A"), _colmajor_ld(B, "B"), _colmajor_ld(C, "C")
    if stream is None:
        stream = torch.cuda.current_stream(C.device).cuda_stream
    _check(load().svd_gpu_dgemm(b"T" if ta else b"N", b"T" if tb else b"N", M, N, K, alpha,
                                ctypes.c_void_p(A.data_ptr()), lda,
                                ctypes.c_void_p(B.data_ptr()), ldb, beta,
                                ctypes.c_void_p(C.data_ptr()), ldc, ctypes.c_void_p(stream)))
