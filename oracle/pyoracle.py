"""ctypes access to the test oracle (liboracle.so), to the reference built from
its own sources (oracle/_ref/libddcref.so, present only where it was built),
and to the LAPACK routines that pin both.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  Nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libddcref.so")

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


_lib = None
_ref = None


def oracle_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.orc_svd.restype = ctypes.c_int
    return _lib


def ref_lib():
    """The reference compiled from its sources, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = ctypes.CDLL(REF_SO)
        _ref.ref_svd_seq.restype = ctypes.c_int
    return _ref


def _dims(m, n):
    mn = min(m, n)
    len_beta = n - 1 if m >= n else m
    return mn, len_beta


def _svd(fn, A, extra_args=()):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    mn, len_beta = _dims(m, n)
    sigma = np.zeros(mn)
    U = np.zeros((m, m), order="F")
    V = np.zeros((n, n), order="F")
    alpha = np.zeros(mn)
    beta = np.zeros(max(len_beta, 1))
    rc = fn(m, n, _ptr(A), _ptr(sigma), _ptr(U), _ptr(V), _ptr(alpha), _ptr(beta), *extra_args)
    if rc != 0:
        raise MemoryError("host SVD allocation failed")
    return dict(sigma=sigma, U=U[:, :mn].copy(order="F"), V=V[:, :mn].copy(order="F"),
                alpha=alpha, beta=beta[:len_beta], A_mod=A)


def oracle_svd(A, fix_multv: bool = True):
    """Oracle restatement of svd_gpu's host path (svd_gpu.c:53-131)."""
    lib = oracle_lib()
    return _svd(lib.orc_svd, A, (ctypes.c_int(1 if fix_multv else 0),))


def oracle_timings():
    out = np.zeros(4)
    oracle_lib().orc_last_timings(_ptr(out))
    return out


def ref_svd(A):
    lib = ref_lib()
    if lib is None:
        return None
    return _svd(lib.ref_svd_seq, A)


def oracle_bidiag(A):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    mn, len_beta = _dims(m, n)
    alpha = np.zeros(mn)
    beta = np.zeros(max(len_beta, 1))
    oracle_lib().orc_bidiag_seq(m, n, _ptr(A), _ptr(alpha), _ptr(beta))
    return A, alpha, beta[:len_beta]


def ref_bidiag(A):
    lib = ref_lib()
    if lib is None:
        return None
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    mn, len_beta = _dims(m, n)
    alpha = np.zeros(mn)
    beta = np.zeros(max(len_beta, 1))
    lib.ref_bidiag_seq(m, n, _ptr(A), _ptr(alpha), _ptr(beta))
    return A, alpha, beta[:len_beta]


def _form(lib, fname, A_mod, left: bool):
    A_mod = np.array(A_mod, dtype=np.float64, order="F", copy=True)
    m, n = A_mod.shape
    k = m if left else n
    Q = np.zeros((k, k), order="F")
    getattr(lib, fname)(ctypes.c_int(m), ctypes.c_int(n), _ptr(A_mod), _ptr(Q))
    return Q


def oracle_form_u(A_mod):
    return _form(oracle_lib(), "orc_form_u", A_mod, True)


def oracle_form_v(A_mod):
    return _form(oracle_lib(), "orc_form_v", A_mod, False)


def ref_form_u(A_mod):
    lib = ref_lib()
    return None if lib is None else _form(lib, "form_u", A_mod, True)


def ref_form_v(A_mod):
    lib = ref_lib()
    return None if lib is None else _form(lib, "form_v", A_mod, False)


def oracle_dc_values(b1, b2):
    b1 = np.ascontiguousarray(b1, dtype=np.float64)
    b2 = np.ascontiguousarray(b2, dtype=np.float64)
    s = np.zeros(len(b1))
    oracle_lib().orc_dc_values(len(b1), _ptr(b1), _ptr(b2), _ptr(s))
    return s


def ref_dc_values(b1, b2):
    lib = ref_lib()
    if lib is None:
        return None
    b1 = np.array(b1, dtype=np.float64)
    b2 = np.array(b2, dtype=np.float64)
    s = np.zeros(len(b1))
    lib.ref_dc_values(len(b1), _ptr(b1), _ptr(b2), _ptr(s))
    return s


# ---------------------------------------------------------------------------
# LAPACK through scipy's Cython capsules (dgebrd and dbdsdc are not wrapped in
# scipy.linalg.lapack).
# ---------------------------------------------------------------------------

def _lapack_fn(name, argtypes):
    from scipy.linalg import cython_lapack
    cap = cython_lapack.__pyx_capi__[name]
    pyget = ctypes.pythonapi.PyCapsule_GetPointer
    pyget.restype = ctypes.c_void_p
    pyget.argtypes = [ctypes.py_object, ctypes.c_char_p]
    pyname = ctypes.pythonapi.PyCapsule_GetName
    pyname.restype = ctypes.c_char_p
    pyname.argtypes = [ctypes.py_object]
    addr = pyget(cap, pyname(cap))
    return ctypes.CFUNCTYPE(None, *argtypes)(addr)


_ip = ctypes.POINTER(ctypes.c_int)
_cp = ctypes.c_char_p


def lapack_dgebrd(A):
    """LAPACK dgebrd: returns (a_out, d, e, tauq, taup)."""
    fn = _lapack_fn("dgebrd", [_ip, _ip, _dp, _ip, _dp, _dp, _dp, _dp, _dp, _ip, _ip])
    a = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = a.shape
    mn = min(m, n)
    d = np.zeros(mn)
    e = np.zeros(max(mn - 1, 1))
    tq = np.zeros(mn)
    tp = np.zeros(mn)
    lwork = ctypes.c_int(-1)
    wq = np.zeros(1)
    info = ctypes.c_int(0)
    mm, nn, lda = ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(m)
    args = (ctypes.byref(mm), ctypes.byref(nn), _ptr(a), ctypes.byref(lda), _ptr(d), _ptr(e),
            _ptr(tq), _ptr(tp))
    fn(*args, _ptr(wq), ctypes.byref(lwork), ctypes.byref(info))
    lwork = ctypes.c_int(int(wq[0]))
    work = np.zeros(lwork.value)
    fn(*args, _ptr(work), ctypes.byref(lwork), ctypes.byref(info))
    if info.value != 0:
        raise RuntimeError(f"dgebrd info={info.value}")
    return a, d, e[: mn - 1], tq, tp


def bidiag_values(alpha, beta, square=True):
    """Singular values (ascending) of the bidiagonal [alpha; beta] by LAPACK.
    square: n x n upper bidiagonal (len(beta) = n-1); else n x (n+1)."""
    n = len(alpha)
    if square:
        B = np.diag(alpha) + np.diag(beta[: n - 1], 1)
    else:
        B = np.zeros((n, n + 1))
        B[np.arange(n), np.arange(n)] = alpha
        B[np.arange(n), np.arange(1, n + 1)] = beta[:n]
    return np.sort(np.linalg.svd(B, compute_uv=False))


def lapack_dlasq1(d, e):
    """Singular values (ascending) of the upper bidiagonal (d, e) by LAPACK
    dlasq1 (dqds), accurate to high relative accuracy."""
    fn = _lapack_fn("dlasq1", [_ip, _dp, _dp, _dp, _ip])
    d = np.array(d, dtype=np.float64, copy=True)
    n = len(d)
    e = np.array(np.append(np.asarray(e, dtype=np.float64)[: n - 1], 0.0), copy=True)
    work = np.zeros(4 * n + 4)
    info = ctypes.c_int(0)
    nn = ctypes.c_int(n)
    fn(ctypes.byref(nn), _ptr(d), _ptr(e), _ptr(work), ctypes.byref(info))
    if info.value != 0:
        raise RuntimeError(f"dlasq1 info={info.value}")
    return np.sort(d)
