"""The library's FP64 MFMA GEMM (svd_gpu_dgemm, gemm.hip) against torch's
FP64 matmul: all four transpose combinations, shapes that are not multiples
of the 128 x 128 tile or of the K slice, leading dimensions larger than the
rows (odd ones included, which take the unaligned paths), alpha and beta.
Bound: |C - C_ref| <= 8 eps K max|A| max|B| |alpha| + eps |beta| max|C0|
entrywise (FP64 accumulation in a different order)."""
import numpy as np
import pytest

from conftest import gpu_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _colmajor(rows, cols, ld, gen):
    """rows x cols column-major view with leading dimension ld"""
    buf = torch.from_numpy(gen.standard_normal((cols, ld))).cuda()
    return buf.t()[:rows, :]


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("M,N,K,pad", [(128, 128, 64, 0), (300, 257, 64, 3), (1, 1, 1, 1),
                                       (129, 17, 5, 0), (513, 700, 130, 8), (64, 300, 0, 2)])
def test_dgemm_matches_torch(ddcsvd, ta, tb, M, N, K, pad):
    g = np.random.default_rng(M * 7 + N * 3 + K)
    A = _colmajor(K if ta == "T" else M, M if ta == "T" else K,
                  max(K if ta == "T" else M, 1) + pad, g)
    B = _colmajor(N if tb == "T" else K, K if tb == "T" else N,
                  max(N if tb == "T" else K, 1) + pad, g)
    C0 = _colmajor(M, N, M + pad, g)
    C = C0.clone().t().contiguous().t() if pad == 0 else _colmajor(M, N, M + pad, g)
    C.copy_(C0)
    alpha, beta = -1.25, 0.5
    gpu_case(f"dgemm {ta}{tb} M={M} N={N} K={K} pad={pad}")
    ddcsvd.dgemm_torch(ta, tb, A, B, C, alpha, beta)
    torch.cuda.synchronize()
    opA = A.t() if ta == "T" else A
    opB = B.t() if tb == "T" else B
    ref = alpha * (opA.cpu() @ opB.cpu()) + beta * C0.cpu()
    amax = float(A.abs().max()) if A.numel() else 0.0
    bmax = float(B.abs().max()) if B.numel() else 0.0
    bound = 8 * EPS * max(K, 1) * amax * bmax * abs(alpha) + EPS * abs(beta) * float(C0.abs().max())
    err = float((C.cpu() - ref).abs().max())
    assert err <= bound + 1e-300, f"max error {err:.3e} > {bound:.3e}"


def test_dgemm_trailing_update_shape(ddcsvd):
    """The trailing-update call of the bidiagonalization: C -= [U X] [Y V]^T,
    K = 64, on a sub-block of a matrix with leading dimension 8200."""
    g = np.random.default_rng(5)
    lda = 8200
    big = torch.from_numpy(g.standard_normal((1100, lda))).cuda().t()  # lda x 1100 view
    UX = torch.from_numpy(g.standard_normal((64, lda))).cuda().t()
    YV = torch.from_numpy(g.standard_normal((64, 1104))).cuda().t()
    r0, m, n = 32, 8192, 1100
    before = big.clone()
    C = big[r0:m, r0:n]
    C0 = C.clone()
    gpu_case("dgemm trailing update 8160x1068 K=64 lda 8200")
    ddcsvd.dgemm_torch("N", "T", UX[r0:m, :], YV[r0:n, :], C, -1.0, 1.0)
    torch.cuda.synchronize()
    ref = C0.cpu() - UX[r0:m, :].cpu() @ YV[r0:n, :].cpu().t()
    err = float((C.cpu() - ref).abs().max())
    assert err <= 8 * EPS * 64 * 25.0, err
    # the rows above r0, the padding rows and the columns left of r0 are untouched
    assert torch.equal(big[:r0, :], before[:r0, :])
    assert torch.equal(big[m:, :], before[m:, :])
    assert torch.equal(big[:, :r0], before[:, :r0])
