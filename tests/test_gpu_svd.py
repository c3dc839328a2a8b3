"""GPU parity tests: the HIP svd_gpu path against the oracle restatement, the
reference's own fixtures, and LAPACK, through the C-ABI (libddcsvd.so).

Tolerances (FP64):
  singular values   |sigma - sigma_LAPACK| <= 1e-12 * sigma_max
  residual          ||A - U S V^T||_F / ||A||_F <= 1e-13 * max(1, sqrt(mn)) * 10
  orthogonality     ||U^T U - I||_F, ||V^T V - I||_F <= 2e-13 * mn
  vectors           up to sign, for singular values separated by a relative
                    gap g: ||u - u_LAPACK|| <= 1e-13 / g
  bidiagonalization alpha, beta and the stored reflectors within
                    1e-12 * max|A| of the serial reference (bidiag.c:33)
"""
import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

pytestmark = pytest.mark.gpu

GOLD = None


def golden():
    global GOLD
    if GOLD is None:
        import os
        GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fixtures.npz"))
    return GOLD


def lapack_svd(A):
    U, s, Vt = np.linalg.svd(A, full_matrices=False)
    o = np.argsort(s, kind="stable")
    return s[o], U[:, o], Vt.T[:, o]


def check_svd(A, s, U, V, sig_tol=1e-12, vec=True):
    m, n = A.shape
    mn = min(m, n)
    sl, Ul, Vl = lapack_svd(A)
    smax = max(sl[-1], np.finfo(float).tiny)
    assert s.shape == (mn,)
    assert np.all(np.diff(s) >= 0), "singular values not ascending"
    sig_err = np.max(np.abs(s - sl)) / smax
    res = np.linalg.norm(A - (U * s) @ V.T) / max(np.linalg.norm(A), np.finfo(float).tiny)
    ou = np.linalg.norm(U.T @ U - np.eye(mn))
    ov = np.linalg.norm(V.T @ V - np.eye(mn))
    assert sig_err <= sig_tol, f"sigma error {sig_err:.3e}"
    assert res <= 1e-12 * max(1.0, np.sqrt(mn)), f"residual {res:.3e}"
    assert ou <= 2e-13 * max(mn, 8), f"orth U {ou:.3e}"
    assert ov <= 2e-13 * max(mn, 8), f"orth V {ov:.3e}"
    if vec and mn > 1:
        gaps = np.full(mn, np.inf)
        d = np.diff(sl) / smax
        gaps[:-1] = np.minimum(gaps[:-1], d)
        gaps[1:] = np.minimum(gaps[1:], d)
        sep = gaps > 1e-6
        for M, L in ((U, Ul), (V, Vl)):
            dots = np.sign(np.sum(M * L, axis=0))
            dev = np.linalg.norm(M - L * dots, axis=0)
            bound = 1e-13 / gaps
            assert np.all(dev[sep] <= np.maximum(bound[sep], 1e-12)), \
                f"vector deviation {np.max(dev[sep] / bound[sep]):.3e} x bound"
    return dict(sig_err=sig_err, res=res, orth_u=ou, orth_v=ov)


# ---------------------------------------------------------------- bidiagonalization

# (2500, 200) and (4200, 130): several 2048-row chunks of the left GEMV and
# several 512-row blocks of the right GEMV
# (70, 66): the panel boundary k = 64 is column n-2, whose left GEMV partial
# sums come from the trailing update and go to the last left finalization.
# (8250, 120): steps k < 58 (more than 8192 rows) run the two-pass GEMVs
# (the panel start k = 32 with the trailing update's partial sums), k = 58
# onward the one-pass GEMV, its first step inside a panel.
BIDIAG_SHAPES = [(1, 1), (5, 1), (2, 2), (3, 2), (3, 3), (8, 3), (33, 33), (64, 64),
                 (96, 64), (100, 70), (70, 66), (130, 97), (300, 300), (2500, 200),
                 (4200, 130), (8250, 120)]


@pytest.mark.parametrize("shape", BIDIAG_SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_bidiag_matches_serial_reference(ddcsvd, po, shape):
    m, n = shape
    A = gen.uniform_matrix(m, n, 7 + m + n)
    Ao, ao, bo = po.oracle_bidiag(A)
    gpu_case(f"bidiag_gpu {m}x{n}")
    Ag, ag, bg = ddcsvd.bidiag(A)
    scale = np.max(np.abs(A))
    assert np.max(np.abs(ag - ao)) <= 1e-12 * scale * max(1, m)
    if n > 1:
        assert np.max(np.abs(bg - bo)) <= 1e-12 * scale * max(1, m)
    assert np.max(np.abs(Ag - Ao)) <= 1e-11


WIDE_SHAPES = [(1, 2), (1, 5), (2, 3), (3, 5), (5, 6), (33, 100), (64, 96), (200, 700),
               (130, 300), (65, 90)]


@pytest.mark.parametrize("shape", WIDE_SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_bidiag_wide_matches_serial_reference(ddcsvd, po, shape):
    """m < n: the m x (m+1) upper bidiagonal and the reflector layout of
    bidiag.c:133-160 (the oracle handles n >= m + 1 as the reference does)."""
    m, n = shape
    A = gen.uniform_matrix(m, n, 3 + m + n)
    Ao, ao, bo = po.oracle_bidiag(A)
    gpu_case(f"bidiag_gpu wide {m}x{n}")
    Ag, ag, bg = ddcsvd.bidiag(A)
    assert ag.shape == (m,) and bg.shape == (m,)
    scale = np.max(np.abs(A))
    assert np.max(np.abs(ag - ao)) <= 1e-12 * scale * max(1, n)
    assert np.max(np.abs(bg - bo)) <= 1e-12 * scale * max(1, n)
    assert np.max(np.abs(Ag - Ao)) <= 1e-11
    # the m x (m+1) bidiagonal has the singular values of A
    B = np.zeros((m, m + 1))
    B[np.arange(m), np.arange(m)] = ag
    B[np.arange(m), np.arange(1, m + 1)] = bg
    sb = np.linalg.svd(B, compute_uv=False)
    sa = np.linalg.svd(A, compute_uv=False)
    assert np.max(np.abs(sb - sa)) <= 1e-13 * sa[0] * max(1, n)


def test_svd_gpu_wide_leaves_reflectors_in_A(ddcsvd, po):
    """The drop-in call leaves a wide A holding its reflectors in the
    reference's m x (m+1) layout (bidiag_par.c:361-383, 405-418)."""
    A = gen.uniform_matrix(96, 160, 21)
    gpu_case("svd_gpu 96x160 A after the call")
    s, U, V, Aout = ddcsvd.svd(A)
    Ao, _, _ = po.oracle_bidiag(A)
    assert np.max(np.abs(Aout - Ao)) <= 1e-11
    check_svd(A, s, U, V)


def test_bidiag_matches_lapack_dgebrd(ddcsvd, po):
    m, n = 200, 150
    A = gen.normal_matrix(m, n, 5)
    gpu_case("bidiag_gpu vs dgebrd 200x150")
    Ag, ag, bg = ddcsvd.bidiag(A)
    _, d, e, _, _ = po.lapack_dgebrd(A)
    # same reflector sign convention: alpha = d for i < n-1 (m > n: also the last)
    assert np.max(np.abs(ag - d)) <= 1e-12 * np.max(np.abs(A)) * m
    assert np.max(np.abs(bg - e)) <= 1e-12 * np.max(np.abs(A)) * m


def test_bidiag_matches_golden_reflectors(ddcsvd):
    g = golden()
    for tag, (m, n, seed) in (("full64", (64, 64, 3)), ("full96", (96, 64, 4))):
        A = gen.uniform_matrix(m, n, seed)
        gpu_case(f"bidiag_gpu golden {tag}")
        Ag, ag, bg = ddcsvd.bidiag(A)
        assert np.max(np.abs(ag - g[f"{tag}_alpha"])) <= 1e-12 * m
        assert np.max(np.abs(bg - g[f"{tag}_beta"])) <= 1e-12 * m
        assert np.max(np.abs(Ag - g[f"{tag}_Amod"])) <= 1e-11


# ---------------------------------------------------------------- bidiagonal values

def _bd_sv(a, b):
    n = len(a)
    B = np.diag(a) + np.diag(b[: n - 1], 1)
    return np.sort(np.linalg.svd(B, compute_uv=False))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 17, 64, 257, 1000])
def test_bdsvd_values_random(ddcsvd, n):
    u = gen.uniform01(2 * n, 40 + n)
    a, b = 2 * u[:n] - 1, 2 * u[n:] - 1
    gpu_case(f"bdsvd_values n={n}")
    s = ddcsvd.bdsvd_values(a, b)
    sv = _bd_sv(a, b)
    assert np.max(np.abs(s - sv)) <= 1e-13 * max(sv[-1], 1e-300) * max(4, np.log2(n))


def test_bdsvd_values_closed_forms(ddcsvd):
    # 1 x 1: |a|; diagonal: sorted |a|; 2 x 2 [[a, b], [0, c]]
    gpu_case("bdsvd_values closed forms")
    assert np.allclose(ddcsvd.bdsvd_values([-3.0], []), [3.0], rtol=0, atol=0)
    for k in (1, 2, 5, 40):  # B = 0
        assert np.array_equal(ddcsvd.bdsvd_values(np.zeros(k), np.zeros(k - 1)), np.zeros(k))
    a = np.array([3.0, -1.0, 2.0, 0.5])
    s = ddcsvd.bdsvd_values(a, np.zeros(3))
    assert np.array_equal(s, np.sort(np.abs(a)))
    a_, b_, c_ = 2.0, 1.0, 1.5
    T = a_ * a_ + b_ * b_ + c_ * c_
    D = a_ * c_
    r = np.sqrt(T * T - 4 * D * D)
    expect = np.sqrt(np.array([(T - r) / 2, (T + r) / 2]))
    s = ddcsvd.bdsvd_values([a_, c_], [b_])
    assert np.max(np.abs(s - expect)) <= 4e-16 * expect[-1]


def test_bdsvd_values_golden_and_oracle(ddcsvd, po):
    """The reference's value solver (fixtures) is accurate only to about
    1e-5 sigma_max and fails outright on some inputs; the GPU solver agrees
    with LAPACK to 1e-13 and with the reference wherever the reference is
    accurate."""
    g = golden()
    for N in (1, 2, 3, 5, 17, 64):
        b1, b2 = g[f"dc{N}_b1"], g[f"dc{N}_b2"]
        B = np.zeros((N, N + 1))
        B[np.arange(N), np.arange(N)] = b1
        B[np.arange(N), np.arange(1, N + 1)] = b2
        exact = np.sort(np.linalg.svd(B, compute_uv=False))
        ref = g[f"dc{N}_sigma"]
        orc = po.oracle_dc_values(b1, b2)
        assert np.array_equal(orc, ref)
        # the GPU path takes n x n with beta[n-1] = 0; use an (N+1)-square
        # embedding of the N x (N+1) bidiagonal: append a zero row
        a = np.append(b1, 0.0)
        b = b2.copy()
        gpu_case(f"bdsvd_values golden N={N}")
        s = ddcsvd.bdsvd_values(a, b)[1:]
        assert np.max(np.abs(s - exact)) <= 1e-13 * exact[-1] * 8
        # where the reference is right the two agree; where its solver loses
        # a root (N = 64 here) the GPU is the one that matches LAPACK
        assert np.all(np.abs(s - exact) <= np.abs(ref - exact) + 1e-13 * exact[-1] * 8)
        good = np.abs(ref - exact) <= 1e-6 * exact[-1]
        assert np.max(np.abs(s - ref)[good]) <= 1e-6 * exact[-1]


def test_bdsvd_values_graded_relative(ddcsvd, po):
    """Graded spectrum 1 .. 1e-6: every value to high relative accuracy
    (reference: LAPACK dlasq1, dqds).  A second bidiagonal has a value near
    1e-222, whose square underflows: there the bound is absolute."""
    n = 400
    a = 10.0 ** (-6 * np.arange(n) / (n - 1))
    b = 0.5 * a[1:] * (2 * gen.uniform01(n - 1, 9) - 1)
    exact = po.lapack_dlasq1(a, b)
    gpu_case("bdsvd_values graded refine")
    s = ddcsvd.bdsvd_values(a, b, refine=True)
    assert np.max(np.abs(s - exact) / exact) <= 1e-13
    b = 1e-3 * (2 * gen.uniform01(n, 9) - 1)
    exact = po.lapack_dlasq1(a, b)
    gpu_case("bdsvd_values graded refine, underflowing value")
    s = ddcsvd.bdsvd_values(a, b, refine=True)
    ok = exact > 1e-150
    assert np.max(np.abs(s - exact)[ok] / exact[ok]) <= 1e-13
    assert np.max(np.abs(s - exact)) <= 4e-15 * exact[-1]


# ---------------------------------------------------------------- bidiagonal vectors

@pytest.mark.parametrize("n", [1, 2, 3, 8, 64, 300])
def test_bdsvd_vectors(ddcsvd, n):
    u = gen.uniform01(2 * n, 70 + n)
    a, b = 2 * u[:n] - 1, 2 * u[n:] - 1
    B = np.diag(a) + np.diag(b[: n - 1], 1)
    gpu_case(f"bdsvd_vectors n={n}")
    s, X, Y = ddcsvd.bdsvd_vectors(a, b)
    check_svd(B, s, Y, X)


def test_bdsvd_vectors_vs_golden_twisted(ddcsvd):
    """Fixture vectors of the reference's twisted solver (B^T B form).  The
    reference loses some of them entirely (its squared form); where its
    vector is right (matches LAPACK) the GPU vector matches it up to sign,
    and every GPU vector matches LAPACK."""
    g = golden()
    for N in (8, 64):
        a, b, X = g[f"tw{N}_a"], g[f"tw{N}_b"], g[f"tw{N}_X"]
        B = np.diag(a) + np.diag(b[: N - 1], 1)
        _, S, Vt = np.linalg.svd(B)
        Vl = Vt.T[:, np.argsort(S)]
        gpu_case(f"bdsvd_vectors golden N={N}")
        s, Xg, _ = ddcsvd.bdsvd_vectors(a, b)
        ref_ok = np.abs(np.sum(X.T * Vl, axis=0)) >= 1 - 1e-8
        assert np.all(np.abs(np.sum(Xg * Vl, axis=0)) >= 1 - 1e-10)
        dots = np.abs(np.sum(Xg.T * X, axis=1))
        assert np.all(dots[ref_ok] >= 1 - 1e-8)
        assert ref_ok.sum() >= N // 2


# ---------------------------------------------------------------- full svd_gpu

# (8193, 300), (300, 8193): step 0 has more than 8192 rows and runs the
# two-pass GEMVs, the later steps the one-pass GEMV, whose first step takes
# its panel dots from the two-pass step's finalization
SVD_SHAPES = [(1, 1), (5, 1), (1, 5), (2, 2), (3, 3), (7, 5), (5, 7), (33, 31), (64, 64),
              (96, 64), (64, 96), (130, 100), (300, 300), (512, 256), (768, 500), (256, 768),
              (4200, 300), (300, 4200), (8193, 300), (300, 8193)]


@pytest.mark.parametrize("shape", SVD_SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_svd_gpu_shapes(ddcsvd, shape):
    m, n = shape
    A = gen.uniform_matrix(m, n, 100 + m * 7 + n)
    gpu_case(f"svd_gpu {m}x{n}")
    s, U, V, _ = ddcsvd.svd(A)
    check_svd(A, s, U, V)


def test_svd_gpu_matches_oracle_and_reference_fixture(ddcsvd, po):
    """Config 1 (512 x 256 U(-1,1)): the oracle restatement of the reference
    (multV indexing repaired) and the reference's fixture agree with the GPU
    to the reference's accuracy; the GPU agrees with LAPACK to 1e-12."""
    g = golden()
    A = gen.uniform_matrix(512, 256, 1)
    o = po.oracle_svd(A, fix_multv=True)
    assert np.array_equal(o["sigma"], g["c1_sigma"])
    gpu_case("svd_gpu 512x256 config 1")
    s, U, V, Amod = ddcsvd.svd(A)
    check_svd(A, s, U, V)
    assert np.max(np.abs(s - g["c1_sigma"])) <= 1e-4 * s[-1]
    assert np.max(np.abs(Amod - o["A_mod"])) <= 1e-11
    # U columns of well-separated values agree with the reference's up to sign
    ue = g["c1_U_edge"]
    ug = np.hstack([U[:, :4], U[:, -4:]])
    assert np.all(np.abs(np.sum(ue * ug, axis=0)) >= 1 - 1e-4)


def test_svd_gpu_literal_reference_driver(ddcsvd):
    """test-whole-svd 512 256 as the reference runs it: 512 x 512 U[1,4) from
    glibc rand() seed 1 (test-whole-svd.c:38,69-73)."""
    g = golden()
    A = gen.glibc_rand_matrix(512, 512, 1.0, 4.0, seed=1)
    gpu_case("svd_gpu literal 512x512")
    s, U, V, _ = ddcsvd.svd(A)
    check_svd(A, s, U, V)
    assert np.max(np.abs(s - g["c1lit_sigma"])) <= 1e-4 * s[-1]


def test_svd_gpu_full_fixtures(ddcsvd):
    g = golden()
    for tag, (m, n, seed) in (("full64", (64, 64, 3)), ("full96", (96, 64, 4))):
        A = gen.uniform_matrix(m, n, seed)
        gpu_case(f"svd_gpu fixture {tag}")
        s, U, V, _ = ddcsvd.svd(A)
        check_svd(A, s, U, V)
        assert np.max(np.abs(s - g[f"{tag}_sigma"])) <= 1e-6 * s[-1]
        dots = np.abs(np.sum(U * g[f"{tag}_U"], axis=0))
        assert np.median(dots) >= 1 - 1e-6
        if m == n:  # the reference's V is wrong for m != n (multV indexing)
            dots = np.abs(np.sum(V * g[f"{tag}_V"], axis=0))
            assert np.median(dots) >= 1 - 1e-6


def test_svd_gpu_gaussian_2048x1024(ddcsvd):
    """The shape where the reference's value solver loses a root."""
    A = gen.normal_matrix(2048, 1024, 2)
    gpu_case("svd_gpu 2048x1024 N(0,1)")
    s, U, V, _ = ddcsvd.svd(A)
    check_svd(A, s, U, V)


def test_svd_gpu_graded_relative_accuracy(ddcsvd):
    n = 512
    sig = 10.0 ** (-6 * np.arange(n) / (n - 1))
    A = gen.graded_matrix(n, n, sig, 21)
    gpu_case("svd_gpu graded 512")
    s, U, V, _ = ddcsvd.svd(A)
    rel = np.abs(s - np.sort(sig)) / np.sort(sig)
    # A itself carries rounding of size eps*||A||, so small values are fixed
    # only to about eps*sigma_max absolute
    assert np.max(np.abs(s - np.sort(sig))) <= 1e-13
    assert np.median(rel) <= 1e-9
    res = np.linalg.norm(A - (U * s) @ V.T) / np.linalg.norm(A)
    assert res <= 1e-13


# ---------------------------------------------------------------- device API, edges

def test_svd_gpu_dev_torch(ddcsvd):
    import torch
    A = gen.normal_matrix(300, 200, 33)
    gpu_case("svd_gpu_dev torch 300x200")
    At = torch.from_numpy(np.ascontiguousarray(A)).to("cuda")
    s, U, V = ddcsvd.svd_torch(At)
    torch.cuda.synchronize()
    check_svd(A, s.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy())
    # the device input is not modified
    assert np.array_equal(At.cpu().numpy(), A)


def test_svd_gpu_zero_and_rank_deficient(ddcsvd):
    gpu_case("svd_gpu zero 20x10")
    s, U, V, _ = ddcsvd.svd(np.zeros((20, 10)))
    assert np.all(s == 0) and np.all(np.isfinite(U)) and np.all(np.isfinite(V))
    x = gen.normal_matrix(40, 1, 3)
    y = gen.normal_matrix(30, 1, 4)
    A = x @ y.T  # rank 1
    gpu_case("svd_gpu rank-1 40x30")
    s, U, V, _ = ddcsvd.svd(A)
    assert np.all(np.isfinite(U)) and np.all(np.isfinite(V))
    assert abs(s[-1] - np.linalg.norm(x) * np.linalg.norm(y)) <= 1e-13 * s[-1]
    assert np.max(s[:-1]) <= 1e-13 * s[-1]
    res = np.linalg.norm(A - (U * s) @ V.T) / np.linalg.norm(A)
    assert res <= 1e-13
    # the 29-fold zero cluster still gets 29 orthonormal vectors
    assert np.linalg.norm(U.T @ U - np.eye(30)) <= 2e-13 * 30
    assert np.linalg.norm(V.T @ V - np.eye(30)) <= 2e-13 * 30


@pytest.mark.parametrize("scale", [1e-300, 1e-160, 1e160, 1e300])
def test_svd_gpu_extreme_scales(ddcsvd, scale):
    """Entries whose squares under- or overflow: the library scales A by a
    power of two (exact) and sigma back.  Checked on the matrix brought back
    to unit scale (the checks' own squares would under- or overflow)."""
    A = scale * gen.normal_matrix(120, 100, 7)
    gpu_case(f"svd_gpu 120x100 scaled by {scale:g}")
    s, U, V, _ = ddcsvd.svd(A)
    assert np.all(np.isfinite(s)) and np.all(np.isfinite(U)) and np.all(np.isfinite(V))
    check_svd(A / scale, s / scale, U, V)


@pytest.mark.parametrize("scale", [1e-300, 1e-160, 1e160, 1e300])
def test_bdsvd_extreme_scales(ddcsvd, scale):
    """A bidiagonal at a scale whose squares under- or overflow: values and
    vectors as for the same bidiagonal at unit scale."""
    rng = np.random.default_rng(11)
    a0, b0 = rng.uniform(-1, 1, 64), rng.uniform(-1, 1, 63)
    a, b = scale * a0, scale * b0
    B0 = np.diag(a / scale) + np.diag(b / scale, 1)
    gpu_case(f"bdsvd 64 scaled by {scale:g}")
    sv = ddcsvd.bdsvd_values(a, b)
    s, X, Y = ddcsvd.bdsvd_vectors(a, b)
    ex = np.linalg.svd(B0, compute_uv=False)
    assert np.all(np.isfinite(sv)) and np.all(np.isfinite(s))
    assert np.max(np.abs(np.sort(sv) / scale - np.sort(ex))) <= 1e-12 * ex.max()
    check_svd(B0, s / scale, Y, X)


@pytest.mark.parametrize("p", [-1000, -520, 520, 1000])
def test_bidiag_extreme_scales(ddcsvd, p):
    """bidiag_gpu on 2^p G: the library scales by a power of two, so the
    reflectors left in A (dimensionless) are those of G and alpha, beta are
    G's times 2^p (no entry of G becomes subnormal at 2^-1000)."""
    G = gen.normal_matrix(100, 80, 3)
    A0, al0, be0 = ddcsvd.bidiag(G)
    gpu_case(f"bidiag_gpu 100x80 scaled by 2^{p}")
    A, al, be = ddcsvd.bidiag(np.ldexp(G, p))
    assert np.all(np.isfinite(A)) and np.all(np.isfinite(al)) and np.all(np.isfinite(be))
    np.testing.assert_allclose(A, A0, rtol=0, atol=4e-16 * np.abs(A0).max())
    np.testing.assert_allclose(np.ldexp(al, -p), al0, rtol=1e-15, atol=0)
    np.testing.assert_allclose(np.ldexp(be, -p), be0, rtol=1e-15, atol=0)


@pytest.mark.parametrize("shape", [(100, 80), (80, 100)])
@pytest.mark.parametrize("p", [-1000, 1000])
def test_svd_gpu_extreme_scales_leave_scale_free_reflectors(ddcsvd, shape, p):
    """svd_gpu on 2^p G, tall and wide: sigma is G's times 2^p and A is left
    holding G's reflectors (dimensionless), as at unit scale."""
    G = gen.normal_matrix(shape[0], shape[1], 3)
    s0, _, _, A0 = ddcsvd.svd(G)
    gpu_case(f"svd_gpu {shape[0]}x{shape[1]} scaled by 2^{p}")
    s, U, V, A = ddcsvd.svd(np.ldexp(G, p))
    assert np.all(np.isfinite(A)) and np.all(np.isfinite(U)) and np.all(np.isfinite(V))
    np.testing.assert_allclose(np.ldexp(s, -p), s0, rtol=1e-14, atol=0)
    np.testing.assert_allclose(A, A0, rtol=0, atol=4e-16 * np.abs(A0).max())


def test_svd_gpu_rejects_nonfinite(ddcsvd):
    A = np.ones((4, 3))
    A[1, 1] = np.nan
    gpu_case("svd_gpu NaN input")
    rc, *_ = ddcsvd.svd(A, raw=True)
    assert rc == ddcsvd.E_NUMERIC
    rc, *_ = ddcsvd.svd(np.zeros((0, 3)), raw=True)
    assert rc == ddcsvd.E_ARG


@pytest.mark.parametrize("shape", [(200, 120), (128, 128), (300, 1), (2, 2), (1, 1), (520, 300)])
def test_form_u_v_matches_oracle(ddcsvd, po, shape):
    """bidiag_form_u_gpu / bidiag_form_v_gpu against the oracle's form_u /
    form_v applied to the same reflectors, and U^T A V = B."""
    m, n = shape
    gpu_case(f"form_u/form_v {m}x{n}")
    A = gen.uniform_matrix(m, n, seed=70 + m)
    Am, al, be = ddcsvd.bidiag(A)
    U, V = ddcsvd.form_u(Am), ddcsvd.form_v(Am)
    Uo, Vo = po.oracle_form_u(Am), po.oracle_form_v(Am)
    assert np.max(np.abs(U - Uo)) <= 1e-13 * max(m, 8)
    assert np.max(np.abs(V - Vo)) <= 1e-13 * max(n, 8)
    B = np.zeros((m, n))
    B[np.arange(n), np.arange(n)] = al
    B[np.arange(n - 1), np.arange(1, n)] = be
    assert np.linalg.norm(U.T @ A @ V - B) <= 1e-13 * max(1.0, np.linalg.norm(A)) * max(m, 8) ** 0.5


def test_profile_left_gemv_hook(ddcsvd):
    """The counter-pass hook runs the left-GEMV launches of every column step
    on a workspace copy (the caller's matrix is unchanged) and rejects step
    counts outside 1..n-1 and wide shapes."""
    import torch
    A = gen.normal_matrix(700, 300, 35)
    At = ddcsvd.col_major(torch.from_numpy(A).to("cuda"))
    gpu_case("profile_left_gemv 700x300, 299 steps")
    ddcsvd.profile_left_gemv(At, 299)
    torch.cuda.synchronize()
    assert np.array_equal(At.cpu().numpy(), A)
    for steps in (0, 300):
        with pytest.raises(ddcsvd.SvdGpuError):
            ddcsvd.profile_left_gemv(At, steps)
    with pytest.raises(ddcsvd.SvdGpuError):
        ddcsvd.profile_left_gemv(ddcsvd.col_major(torch.zeros((3, 5), dtype=torch.float64,
                                                              device="cuda")), 1)


def test_profile_right_gemv_hook(ddcsvd):
    """The right-GEMV counter hook runs on a workspace copy (the caller's
    matrix is unchanged) and rejects step counts outside 1..n-2 and shapes
    with m < n or n < 3."""
    import torch
    A = gen.normal_matrix(700, 300, 36)
    At = ddcsvd.col_major(torch.from_numpy(A).to("cuda"))
    gpu_case("profile_right_gemv 700x300, 298 steps")
    ddcsvd.profile_right_gemv(At, 298)
    torch.cuda.synchronize()
    assert np.array_equal(At.cpu().numpy(), A)
    for steps in (0, 299):
        with pytest.raises(ddcsvd.SvdGpuError):
            ddcsvd.profile_right_gemv(At, steps)
    for shape in ((3, 5), (4, 2)):
        with pytest.raises(ddcsvd.SvdGpuError):
            ddcsvd.profile_right_gemv(ddcsvd.col_major(torch.zeros(shape, dtype=torch.float64,
                                                                   device="cuda")), 1)


def test_profile_fused_gemv_hook(ddcsvd):
    """The one-pass GEMV counter hook runs on a workspace copy (the caller's
    matrix is unchanged) and rejects step counts outside 1..n-2 and shapes
    with m < n, n < 3 or m > 8192."""
    import torch
    A = gen.normal_matrix(700, 300, 37)
    At = ddcsvd.col_major(torch.from_numpy(A).to("cuda"))
    gpu_case("profile_fused_gemv 700x300, 298 steps")
    ddcsvd.profile_fused_gemv(At, 298)
    torch.cuda.synchronize()
    assert np.array_equal(At.cpu().numpy(), A)
    for steps in (0, 299):
        with pytest.raises(ddcsvd.SvdGpuError):
            ddcsvd.profile_fused_gemv(At, steps)
    for shape in ((3, 5), (4, 2), (8193, 3)):
        with pytest.raises(ddcsvd.SvdGpuError):
            ddcsvd.profile_fused_gemv(ddcsvd.col_major(torch.zeros(shape, dtype=torch.float64,
                                                                   device="cuda")), 1)


@pytest.mark.parametrize("shape", [(4100, 1100), (1100, 4100)])
def test_host_buffers_staged_copies_match_device_path(ddcsvd, shape):
    """Host-buffer calls above 4 MiB move their data through the library's
    pinned staging buffers in 32 MiB slices (36 MB here: two slices, and a
    wide input whose A goes back through the transposed layout).  The
    results must equal the device-pointer path's bit for bit (same kernels),
    and the overwritten A the reflectors of bidiag_gpu, whose copies are not
    staged."""
    import torch
    m, n = shape
    A = gen.normal_matrix(m, n, 41)
    gpu_case(f"svd_gpu host buffers staged {m}x{n}")
    s, U, V, A_after = ddcsvd.svd(A)
    At = ddcsvd.col_major(torch.from_numpy(A).to("cuda"))
    sd, Ud, Vd = ddcsvd.svd_torch(At)
    torch.cuda.synchronize()
    assert np.array_equal(s, sd.cpu().numpy())
    assert np.array_equal(U, Ud.cpu().numpy())
    assert np.array_equal(V, Vd.cpu().numpy())
    Ab, _, _ = ddcsvd.bidiag(A)
    assert np.array_equal(A_after, Ab)


def test_svd_dev_null_stream_orders_after_torch_work(ddcsvd):
    """svd_gpu_dev with stream NULL (torch's default stream handle is 0) must
    run after the caller's pending work on that stream: A is produced by a
    chain of GEMMs still in flight when the call is made (the workspace is
    already allocated, so no allocation synchronizes the device)."""
    import torch
    m, n = 1024, 512
    A0 = gen.normal_matrix(m, n, 41)
    ddcsvd.svd_torch(torch.from_numpy(A0).cuda())  # warm the workspace at this size
    torch.cuda.synchronize()
    X = torch.from_numpy(gen.normal_matrix(4096, 4096, 42)).cuda() / 64.0
    gpu_case("svd_gpu_dev right after in-flight GEMMs on the null stream")
    Y = X
    for _ in range(24):
        Y = Y @ X                       # ~0.1 s of queued work
    A = ddcsvd.col_major(torch.from_numpy(A0).cuda() + 0.0 * Y[:m, :n])
    s, U, V = ddcsvd.svd_torch(A)
    torch.cuda.synchronize()
    check_svd(A0, s.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy(), vec=False)
