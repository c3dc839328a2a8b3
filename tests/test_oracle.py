"""CPU tests that pin the oracle (the C restatement of the reference host
path) before it is trusted as the GPU path's checker.

Pins, per oracle function:
  orc_bidiag_seq     bit-equal to the reference compiled from its sources
                     (oracle/_ref) and to the fixtures; alpha/beta/reflectors
                     equal LAPACK dgebrd under the documented sign mapping
  orc_dc_values      bit-equal to the reference/fixtures; closed forms for
                     N = 1, 2; LAPACK where the reference is accurate
  orc_right_vectors  bit-equal to the reference/fixtures; LAPACK vectors up to
                     sign for a well-separated spectrum
  orc_mult_u/_v      bit-equal to the reference (m >= n); the residual is tiny
                     with the multV repair and O(1) without it for m != n
  orc_svd            all of the above end to end
"""
import os

import numpy as np
import pytest

from ddcsvd import gen

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.npz"))


def have_ref(po):
    return po.ref_lib() is not None


@pytest.mark.parametrize("shape", [(64, 64), (96, 64), (130, 100), (300, 300), (17, 17)])
def test_oracle_bit_equal_to_compiled_reference(po, shape):
    if not have_ref(po):
        pytest.skip("reference sources not built here (oracle/_ref)")
    m, n = shape
    A = gen.uniform_matrix(m, n, m + 3 * n)
    o = po.oracle_svd(A, fix_multv=False)
    r = po.ref_svd(A)
    for k in ("sigma", "alpha", "beta", "U", "V", "A_mod"):
        assert np.array_equal(o[k], r[k], equal_nan=True), k


def _bidiagonal(m, n, alpha, beta):
    B = np.zeros((m, n))
    for i in range(len(alpha)):
        B[i, i] = alpha[i]
    for i in range(len(beta)):
        B[i, i + 1] = beta[i]
    return B


@pytest.mark.parametrize("shape", [(40, 40), (60, 25), (17, 17), (3, 2), (1, 1)])
def test_oracle_form_u_v(po, shape):
    """form_u / form_v restated: orthogonal, U^T A V = B (closed form of the
    reduction), and bit-equal to the compiled reference where it is built."""
    m, n = shape
    A = gen.uniform_matrix(m, n, 4 + m)
    Am, al, be = po.oracle_bidiag(A)
    U, V = po.oracle_form_u(Am), po.oracle_form_v(Am)
    B = _bidiagonal(m, n, al, be)
    assert np.linalg.norm(U.T @ A @ V - B) <= 1e-14 * max(1.0, np.linalg.norm(A)) * max(m, 4)
    assert np.linalg.norm(U.T @ U - np.eye(m)) <= 1e-14 * max(m, 4)
    assert np.linalg.norm(V.T @ V - np.eye(n)) <= 1e-14 * max(n, 4)
    if have_ref(po):
        assert np.array_equal(U, po.ref_form_u(Am))
        assert np.array_equal(V, po.ref_form_v(Am))


def test_oracle_matches_golden_fixtures(po):
    A = gen.glibc_rand_matrix(512, 512, 1.0, 4.0, seed=1)
    Ao, ao, bo = po.oracle_bidiag(A)
    np.testing.assert_allclose(ao, GOLD["c1lit_alpha"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(bo, GOLD["c1lit_beta"], rtol=0, atol=1e-12)
    s = po.oracle_dc_values(ao, np.append(bo, 0.0))
    np.testing.assert_allclose(s, GOLD["c1lit_sigma"], rtol=0, atol=1e-10)
    for tag, (m, n, seed) in (("full64", (64, 64, 3)), ("full96", (96, 64, 4))):
        o = po.oracle_svd(gen.uniform_matrix(m, n, seed), fix_multv=False)
        for k, g in (("sigma", "sigma"), ("U", "U"), ("V", "V"), ("alpha", "alpha"),
                     ("beta", "beta"), ("A_mod", "Amod")):
            np.testing.assert_allclose(o[k], GOLD[f"{tag}_{g}"], rtol=0, atol=1e-12)
    for N in (1, 2, 3, 5, 17, 64):
        s = po.oracle_dc_values(GOLD[f"dc{N}_b1"], GOLD[f"dc{N}_b2"])
        np.testing.assert_allclose(s, GOLD[f"dc{N}_sigma"], rtol=0, atol=1e-12)


def test_oracle_bidiag_equals_lapack_dgebrd(po):
    """dgebrd pins the bidiagonalization: for m > n, alpha = d and beta = e;
    the left reflector j equals +-sqrt(tauq_j/2) [1; Q(j+1:, j)] and the right
    reflector +-sqrt(taup_j/2) [1; P(j, j+2:)]."""
    for (m, n) in ((120, 80), (80, 80)):
        A = gen.normal_matrix(m, n, 17)
        Ao, ao, bo = po.oracle_bidiag(A)
        a, d, e, tq, tp = po.lapack_dgebrd(A)
        scale = np.max(np.abs(A))
        dd = d.copy()
        if m == n:
            dd[-1] = -dd[-1]  # the reference reflects the final 1 x 1 column too
        assert np.max(np.abs(ao - dd)) <= 1e-11 * scale
        assert np.max(np.abs(bo - e)) <= 1e-11 * scale
        for j in range(n - 1):
            lv = np.concatenate([[1.0], a[j + 1:, j]]) * np.sqrt(tq[j] / 2)
            ov = Ao[j:, j]
            assert min(np.max(np.abs(ov - lv)), np.max(np.abs(ov + lv))) <= 1e-10
        for j in range(n - 2):
            rv = np.concatenate([[1.0], a[j, j + 2:]]) * np.sqrt(tp[j] / 2)
            ov = Ao[j, j + 1:]
            assert min(np.max(np.abs(ov - rv)), np.max(np.abs(ov + rv))) <= 1e-10


def test_oracle_dc_closed_forms_and_lapack(po):
    # N = 1: [a b] -> hypot(a, b)
    s = po.oracle_dc_values(np.array([3.0]), np.array([4.0]))
    assert s[0] == 5.0
    # N = 2 (2 x 3): eigenvalues of B B^T in closed form
    b1, b2 = np.array([1.5, -0.7]), np.array([0.3, 2.0])
    B = np.array([[b1[0], b2[0], 0.0], [0.0, b1[1], b2[1]]])
    G = B @ B.T
    tr, det = np.trace(G), np.linalg.det(G)
    lam = np.array([(tr - np.sqrt(tr * tr - 4 * det)) / 2, (tr + np.sqrt(tr * tr - 4 * det)) / 2])
    np.testing.assert_allclose(po.oracle_dc_values(b1, b2), np.sqrt(lam), rtol=1e-12)
    # small random problems: the reference solver is accurate there
    for N in (3, 5, 9):
        u = gen.uniform01(2 * N, 300 + N)
        b1, b2 = 2 * u[:N] - 1, 2 * u[N:] - 1
        B = np.zeros((N, N + 1))
        B[np.arange(N), np.arange(N)] = b1
        B[np.arange(N), np.arange(1, N + 1)] = b2
        exact = np.sort(np.linalg.svd(B, compute_uv=False))
        assert np.max(np.abs(po.oracle_dc_values(b1, b2) - exact)) <= 1e-9 * exact[-1]


def test_oracle_twisted_vectors_vs_lapack(po):
    import ctypes
    N = 8
    a = np.array([4.0, 3.1, 2.5, 2.0, 1.6, 1.2, 0.9, 0.5])
    b = np.array([0.3, -0.2, 0.25, 0.1, -0.15, 0.2, 0.05, 0.0])
    B = np.diag(a) + np.diag(b[:-1], 1)
    _, S, Vt = np.linalg.svd(B)
    o = np.argsort(S)
    s = np.sort(S)
    X = np.zeros(N * N)
    po.oracle_lib().orc_right_vectors(ctypes.c_int(N), ctypes.c_int(N), po._ptr(a), po._ptr(b),
                                      po._ptr(s), po._ptr(X))
    X = X.reshape(N, N)
    V = Vt.T[:, o]
    for i in range(N):
        assert abs(abs(X[i] @ V[:, i]) - 1) <= 1e-10


def test_oracle_multv_defect_and_repair(po):
    """m != n: the reference's multV indexing (bidiag_par.c:1025,1034,1036)
    gives a wrong V; the opt-in repair restores a tiny residual."""
    A = gen.uniform_matrix(96, 64, 4)
    bad = po.oracle_svd(A, fix_multv=False)
    good = po.oracle_svd(A, fix_multv=True)
    rb = np.linalg.norm(A - (bad["U"] * bad["sigma"]) @ bad["V"].T) / np.linalg.norm(A)
    rg = np.linalg.norm(A - (good["U"] * good["sigma"]) @ good["V"].T) / np.linalg.norm(A)
    assert rb > 0.1
    assert rg < 1e-10
    # m == n: both forms coincide
    A = gen.uniform_matrix(64, 64, 3)
    assert np.array_equal(po.oracle_svd(A, fix_multv=False)["V"], po.oracle_svd(A, fix_multv=True)["V"])


def test_checker_detects_planted_mistakes(po):
    """A transposed operand or a dropped term in the oracle's output must fail
    the comparisons the GPU tests use."""
    A = gen.uniform_matrix(40, 40, 8)
    o = po.oracle_svd(A)
    res = lambda U, s, V: np.linalg.norm(A - (U * s) @ V.T) / np.linalg.norm(A)
    assert res(o["U"], o["sigma"], o["V"]) < 1e-12
    assert res(o["V"], o["sigma"], o["U"]) > 1e-3           # U and V swapped
    assert res(o["U"][:, ::-1], o["sigma"], o["V"]) > 1e-3  # misordered pairs
    s2 = o["sigma"].copy()
    s2[0] = 0.0                                             # dropped term
    sv = np.sort(np.linalg.svd(A, compute_uv=False))
    assert np.max(np.abs(s2 - sv)) / sv[-1] > 1e-6


def test_generator_streams():
    # entry k of the stream is mix64(base + (k+1) * golden): check a few in
    # plain Python integer arithmetic
    M = 2 ** 64 - 1

    def mix(x):
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        return x ^ (x >> 31)

    seed = 5
    base = (seed * 0x632BE59BD9B4E019) & M
    u = gen.uniform01(10, seed)
    for k in (0, 3, 9):
        z = mix((base + (k + 1) * 0x9E3779B97F4A7C15) & M)
        assert u[k] == (z >> 11) * 2.0 ** -53
    A = gen.uniform_matrix(3, 2, seed)
    assert A[1, 0] == -1.0 + 2.0 * u[1] and A[0, 1] == -1.0 + 2.0 * u[3]
    Z = gen.normal_matrix(400, 400, 1)
    assert abs(Z.mean()) < 0.01 and abs(Z.std() - 1) < 0.01
    ms, ns = gen.batch_shapes(10000, 0)
    assert ms.min() >= 16 and ms.max() <= 256 and np.all(ns >= 8) and np.all(ns <= ms)
