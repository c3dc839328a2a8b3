"""Repeated and clustered singular values through the C-ABI.

The reference computes every singular vector on its own
(parallel-twisted.c:554-637), so equal values share one vector and U, V lose
rank (A = I gives a rank-1 U).  The GPU path splits the bidiagonal at
negligible super-diagonal entries, hands inseparable values out to the
blocks, and replaces the vectors of close values in one block by a
Rayleigh-Ritz basis.  These tests use inputs whose spectra are known exactly
and assert, besides the values:

  orthogonality   ||U^T U - I||_F, ||V^T V - I||_F <= 2e-13 * mn
  residual        ||A - U S V^T||_F / ||A||_F <= 1e-13 (zero matrix: exact)
"""
import os

import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

pytestmark = pytest.mark.gpu


def check_exact(A, s, U, V, sig_exact, sig_tol=1e-12):
    m, n = A.shape
    mn = min(m, n)
    assert s.shape == (mn,)
    assert np.all(np.isfinite(U)) and np.all(np.isfinite(V))
    assert np.all(np.diff(s) >= 0), "singular values not ascending"
    smax = max(float(np.max(sig_exact)), np.finfo(float).tiny)
    err = np.max(np.abs(s - np.sort(sig_exact))) / smax
    ou = np.linalg.norm(U.T @ U - np.eye(mn))
    ov = np.linalg.norm(V.T @ V - np.eye(mn))
    na = np.linalg.norm(A)
    res = np.linalg.norm(A - (U * s) @ V.T) / na if na > 0 else np.linalg.norm((U * s) @ V.T)
    bound = 2e-13 * max(mn, 8)
    assert err <= sig_tol, f"sigma error {err:.3e}"
    assert ou <= bound, f"orth U {ou:.3e} > {bound:.1e}"
    assert ov <= bound, f"orth V {ov:.3e} > {bound:.1e}"
    assert res <= 1e-13, f"residual {res:.3e}"
    return dict(err=err, ou=ou, ov=ov, res=res)


@pytest.mark.parametrize("n", [1, 2, 64, 1000])
def test_identity(ddcsvd, n):
    A = np.eye(n)
    gpu_case(f"svd_gpu identity {n}x{n}")
    s, U, V, _ = ddcsvd.svd(A)
    check_exact(A, s, U, V, np.ones(n))


def test_identity_stacked_and_wide(ddcsvd):
    A = np.zeros((300, 200))
    A[:200] = np.eye(200)
    gpu_case("svd_gpu [I;0] 300x200")
    s, U, V, _ = ddcsvd.svd(A)
    check_exact(A, s, U, V, np.ones(200))
    gpu_case("svd_gpu [I 0] 200x300")
    s, U, V, _ = ddcsvd.svd(A.T.copy())
    check_exact(A.T, s, U, V, np.ones(200))


def test_scaled_permutation(ddcsvd):
    """Signed, scaled permutation with every value repeated: diag entries
    of B come out exactly, the super-diagonal is zero."""
    n = 96
    p = np.random.default_rng(5).permutation(n)
    vals = np.repeat([0.5, 1.0, 3.0, 7.0], n // 4) * np.where(np.arange(n) % 3, 1.0, -1.0)
    A = np.zeros((n, n))
    A[np.arange(n), p] = vals
    gpu_case("svd_gpu scaled permutation 96")
    s, U, V, _ = ddcsvd.svd(A)
    check_exact(A, s, U, V, np.abs(vals))


@pytest.mark.parametrize("n", [128, 512])
def test_haar_orthogonal(ddcsvd, n):
    A = gen.haar_orthogonal(n, 21 + n)
    gpu_case(f"svd_gpu Haar orthogonal {n}")
    s, U, V, _ = ddcsvd.svd(A)
    check_exact(A, s, U, V, np.ones(n))


@pytest.mark.parametrize("mult,n", [(3, 300), (5, 200), (2, 1024), (100, 200)])
def test_repeated_values(ddcsvd, mult, n):
    """Q diag(1,..,1, 2,..,2, ...) P^T: every value mult times.  The
    bidiagonal of such a matrix is not split; its values come in clusters of
    width ~ eps inside one unreduced block."""
    Q = gen.haar_orthogonal(n, 5 + n)
    P = gen.haar_orthogonal(n, 6 + n)
    sig = np.repeat(np.arange(1, n // mult + 1, dtype=float), mult)
    A = (Q * sig) @ P.T
    gpu_case(f"svd_gpu repeated x{mult} {n}")
    s, U, V, _ = ddcsvd.svd(A)
    check_exact(A, s, U, V, sig)


def test_long_repeats_and_tail(ddcsvd):
    """Thirty-fold values 1 and 2 beside a separated tail: the bidiagonal
    splits into blocks that each hold several of the repeated values."""
    sig = np.concatenate([np.ones(30), 2 * np.ones(30), np.linspace(3, 4, 40)])
    Q = gen.haar_orthogonal(100, 8)
    P = gen.haar_orthogonal(100, 9)
    A = (Q * sig) @ P.T
    gpu_case("svd_gpu 30-fold repeats 100")
    s, U, V, _ = ddcsvd.svd(A)
    check_exact(A, s, U, V, sig)


@pytest.mark.parametrize("shape,rank", [((40, 30), 1), ((200, 150), 50), ((150, 200), 20)])
def test_rank_deficient(ddcsvd, shape, rank):
    m, n = shape
    X = gen.normal_matrix(m, rank, 3 + m)
    Y = gen.normal_matrix(n, rank, 4 + n)
    A = X @ Y.T
    sig = np.linalg.svd(A, compute_uv=False)
    sig[rank:] = 0.0
    gpu_case(f"svd_gpu rank {rank} {m}x{n}")
    s, U, V, _ = ddcsvd.svd(A)
    check_exact(A, s, U, V, sig)


def test_zero_matrix(ddcsvd):
    for shape in ((20, 10), (10, 20), (7, 7)):
        A = np.zeros(shape)
        gpu_case(f"svd_gpu zero {shape}")
        s, U, V, _ = ddcsvd.svd(A)
        check_exact(A, s, U, V, np.zeros(min(shape)))


def test_bdsvd_vectors_diagonal_repeats(ddcsvd):
    """Diagonal bidiagonal with repeated |a| (and one exact zero): the
    vectors are signed unit vectors, one per diagonal entry."""
    a = np.array([1.0, -2.0, 1.0, 3.0, -1.0, 2.0, 0.5, 3.0, 0.0, 1.0])
    n = len(a)
    gpu_case("bdsvd_vectors diagonal with repeats")
    s, X, Y = ddcsvd.bdsvd_vectors(a, np.zeros(n - 1))
    B = np.diag(a)
    check_exact(B, s, Y, X, np.abs(a))
    # each vector is a unit vector
    assert np.allclose(np.sort(np.abs(X), axis=0)[-1], 1.0, rtol=0, atol=1e-15)


def test_bdsvd_vectors_tight_chain(ddcsvd):
    """Unreduced bidiagonal with all values within 2e-9 of each other
    (a = 1, b = 1e-9): one group of 40 in a single block."""
    n = 40
    a = np.ones(n)
    b = np.full(n - 1, 1e-9)
    B = np.diag(a) + np.diag(b, 1)
    gpu_case("bdsvd_vectors tight chain 40")
    s, X, Y = ddcsvd.bdsvd_vectors(a, b)
    check_exact(B, s, Y, X, np.linalg.svd(B, compute_uv=False))


def test_bdsvd_vectors_pairs_a_few_ulps_apart(ddcsvd):
    """Regression input: the bidiagonal of the x2-repeat 1024 matrix of
    test_repeated_values, as a bidiagonalization with another (equally
    valid) summation order left it (tests/golden/repeat2_1024_bidiag.npz).
    Its 512 pairs are unreduced; the pair at sigma = 472 lies 5.3 eps apart
    in sigma^2.  Shifts within ulps of such a pair made both inverse
    iterations collapse onto one vector (triplet residual 4e-9 against
    4e-11 elsewhere); the pair must come out as accurately as the rest."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "repeat2_1024_bidiag.npz"))
    a, b = g["alpha"], g["beta"]
    B = np.diag(a) + np.diag(b, 1)
    exact = np.linalg.svd(B, compute_uv=False)
    assert np.max(np.abs(np.sort(exact) - np.repeat(np.arange(1.0, 513.0), 2))) <= 1e-11
    gpu_case("bdsvd_vectors x2 repeats 1024 (fixture)")
    s, X, Y = ddcsvd.bdsvd_vectors(a, b)
    check_exact(B, s, Y, X, exact)
    tri = np.linalg.norm(B @ X - Y * s, axis=0)
    assert np.max(tri) <= 1e-12 * np.max(exact), f"worst triplet residual {np.max(tri):.3e}"


@pytest.mark.parametrize("n", [100, 200])
def test_bdsvd_vectors_inseparable_run_beyond_one_window(ddcsvd, n):
    """Unreduced bidiagonal (a = 1, b = 1e-14, above the split threshold)
    whose n values all lie within a few eps: one group of n members, more
    than one Rayleigh-Ritz window of 64, and no value separable from the
    others.  The windows' vectors must be orthogonal to each other too."""
    a = np.ones(n)
    b = np.full(n - 1, 1e-14)
    B = np.diag(a) + np.diag(b, 1)
    gpu_case(f"bdsvd_vectors inseparable run of {n}")
    s, X, Y = ddcsvd.bdsvd_vectors(a, b)
    check_exact(B, s, Y, X, np.linalg.svd(B, compute_uv=False))


@pytest.mark.parametrize("n", [3, 20, 100])
def test_bdsvd_vectors_zero_diagonal(ddcsvd, n):
    """B = the shift matrix (a = 0, b = 1), one unreduced block: values 0
    and 1 (n - 1 times).  The zero value's vectors are e_0 (right) and
    e_{n-1} (left).  At a shift of exactly 0 the transforms' quotients are
    0/0 and every twist index ties; the solve must still find them."""
    a = np.zeros(n)
    b = np.ones(n - 1)
    B = np.diag(a) + np.diag(b, 1)
    gpu_case(f"bdsvd_vectors zero diagonal {n}")
    s, X, Y = ddcsvd.bdsvd_vectors(a, b)
    check_exact(B, s, Y, X, np.r_[0.0, np.ones(n - 1)])
    assert abs(abs(X[0, 0]) - 1.0) <= 1e-14, "right null vector is not e_0"
    assert abs(abs(Y[n - 1, 0]) - 1.0) <= 1e-14, "left null vector is not e_(n-1)"


def test_bdsvd_vectors_graded_over_200_decades(ddcsvd):
    """a_i = 10^(-4i), b_i = 10^(-4i-1): values over 200 decades."""
    a = 10.0 ** (-4.0 * np.arange(50))
    b = 10.0 ** (-4.0 * np.arange(49) - 1.0)
    B = np.diag(a) + np.diag(b, 1)
    gpu_case("bdsvd_vectors graded over 200 decades")
    s, X, Y = ddcsvd.bdsvd_vectors(a, b)
    check_exact(B, s, Y, X, np.linalg.svd(B, compute_uv=False))
