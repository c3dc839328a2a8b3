"""The measured configurations at their full sizes, through svd_gpu_dev.

  C2  plain FP64, 8192 x 4096 and 8192 x 8192, entries N(0,1), seed 2:
      sigma against LAPACK (numpy, values only, on the host) within
      1e-12 sigma_max; residual ||A - U S V^T||_F / ||A||_F and
      ||U^T U - I||_F, ||V^T V - I||_F -- the reconstruction check that the
      reference's driver compiles out (test-whole-svd.c:81-96).
  C3  plain FP64, 8192 x 8192 A = Q diag(sigma) P^T, Haar Q and P, sigma_i =
      10^(-6 i / 8191): every |sigma_hat_i - sigma_i| / sigma_max <= 1e-12
      against the exact values.

Bounds: residual <= 1e-12 sqrt(n), orthogonality <= 2e-13 n (the bounds of
tests/test_gpu_svd.py); products are formed on the GPU in FP64.
"""
import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _svd(ddcsvd, A):
    s, U, V = ddcsvd.svd_torch(A)
    torch.cuda.synchronize()
    return s, U, V


def _quality(A, s, U, V):
    n = s.numel()
    eye = torch.eye(n, dtype=torch.float64, device="cuda")
    res = float(torch.linalg.norm(A - (U * s) @ V.T) / torch.linalg.norm(A))
    ou = float(torch.linalg.norm(U.T @ U - eye))
    ov = float(torch.linalg.norm(V.T @ V - eye))
    return res, ou, ov


@pytest.mark.timeout(400)
@pytest.mark.parametrize("m,n", [(8192, 4096), (8192, 8192)])
def test_c2_gaussian(ddcsvd, m, n):
    Ah = gen.normal_matrix(m, n, 2)
    A = ddcsvd.col_major(torch.from_numpy(Ah).cuda())
    gpu_case(f"C2 svd_gpu_dev {m}x{n} N(0,1) seed 2")
    s, U, V = _svd(ddcsvd, A)
    res, ou, ov = _quality(A, s, U, V)
    print(f"[C2 {m}x{n}] residual {res:.2e} orth U {ou:.2e} orth V {ov:.2e}", flush=True)
    assert bool(torch.all(s[1:] >= s[:-1]))
    assert res <= 1e-12 * np.sqrt(n)
    assert ou <= 2e-13 * n and ov <= 2e-13 * n
    del U, V
    exact = np.sort(np.linalg.svd(Ah, compute_uv=False))
    err = float(np.max(np.abs(s.cpu().numpy() - exact)) / exact[-1])
    print(f"[C2 {m}x{n}] max |sigma - sigma_LAPACK| / sigma_max = {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.timeout(400)
def test_c3_graded_plain(ddcsvd):
    n = 8192
    sig = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    qs = []
    for k in range(2):
        g = torch.from_numpy(gen.normal_matrix(n, n, 3 + k)).cuda()
        q, r = torch.linalg.qr(g)
        qs.append(q * torch.sign(torch.diagonal(r)))
        del g, r
    A = ddcsvd.col_major((qs[0] * torch.from_numpy(sig).cuda()) @ qs[1].T)
    del qs
    gpu_case("C3 plain svd_gpu_dev 8192^2 graded 1..1e-6")
    s, U, V = _svd(ddcsvd, A)
    res, ou, ov = _quality(A, s, U, V)
    exact = np.sort(sig)
    err = np.abs(s.cpu().numpy() - exact)
    print(f"[C3 plain] residual {res:.2e} orth U {ou:.2e} orth V {ov:.2e} "
          f"max |dsigma|/sigma_max {err.max():.2e} max rel {np.max(err / exact):.2e}", flush=True)
    assert err.max() <= 1e-12
    assert res <= 1e-12 * np.sqrt(n)
    assert ou <= 2e-13 * n and ov <= 2e-13 * n
