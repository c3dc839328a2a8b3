"""GPU tests of the SVD derivative: svd_gpu_backward through the C-ABI and the
torch.autograd harness around it.

Checks:
  analytic   against torch's own FP64 SVD backward on the CPU (an independent
             implementation of the same derivative), signs of the CPU factors
             aligned to the GPU ones:  ||gA - gA_ref||_F / ||gA_ref||_F <= 1e-10
  finite     central differences along 64 directions with i.i.d. N(0,1)
  differences entries, step h = 1e-6, forward SVDs on the GPU with the signs of
             U and V fixed to the unperturbed ones.  The fourth-order central
             stencil (f(-2h), f(-h), f(h), f(2h)) must agree to 1e-8 relative
             (norm over the 64 directions); the two-point stencil, whose h^2
             term alone is about 1e-6 here, is reported beside it.
  gap guard  repeated and near-repeated singular values give finite gradients;
             a singular-value-only loss gives U diag(gs) V^T exactly.
"""
import ctypes

import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda:0")


def _to_gpu(a):
    return torch.from_numpy(np.asfortranarray(a)).to(_dev())


def _ref_grad(A, U_gpu, gU, gs, gV):
    """torch's CPU FP64 SVD backward of L = <gU, U> + <gs, sigma> + <gV, V>
    (thin factors, sigma ascending, columns signed like U_gpu)."""
    At = torch.from_numpy(np.array(A)).requires_grad_(True)
    Ut, St, Vht = torch.linalg.svd(At, full_matrices=False)
    Ut, St, Vt = Ut.flip(1), St.flip(0), Vht.flip(0).t()
    sgn = torch.sign((Ut.detach() * torch.from_numpy(U_gpu)).sum(0))
    sgn[sgn == 0] = 1.0
    L = St @ torch.from_numpy(gs)
    if gU is not None:
        L = L + (Ut * sgn * torch.from_numpy(gU)).sum()
    if gV is not None:
        L = L + (Vt * sgn * torch.from_numpy(gV)).sum()
    L.backward()
    return At.grad.numpy()


def _gpu_grad(ddcsvd, A, gU, gs, gV, guard=1e-13):
    s, U, V = ddcsvd.svd_torch(_to_gpu(A))
    gA = ddcsvd.svd_backward_torch(U, s, V, None if gU is None else _to_gpu(gU),
                                   _to_gpu(gs), None if gV is None else _to_gpu(gV), guard)
    torch.cuda.synchronize()
    return s.cpu().numpy(), U.cpu().numpy(), V.cpu().numpy(), gA.cpu().numpy()


def _cotangents(m, n, seed):
    mn = min(m, n)
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((m, mn)), rng.standard_normal(mn),
            rng.standard_normal((n, mn)))


@pytest.mark.parametrize("shape", [(64, 48), (300, 200), (130, 130), (48, 80), (257, 33)])
def test_backward_matches_torch_reference(ddcsvd, shape):
    m, n = shape
    mn = min(m, n)
    gpu_case(f"adjoint vs torch reference {m}x{n}")
    A = gen.graded_matrix(m, n, np.linspace(1.0, 2.0, mn), seed=3 + m)
    gU, gs, gV = _cotangents(m, n, 7)
    _, U, _, gA = _gpu_grad(ddcsvd, A, gU, gs, gV)
    ref = _ref_grad(A, U, gU, gs, gV)
    err = np.linalg.norm(gA - ref) / np.linalg.norm(ref)
    assert err <= 1e-10, f"relative gradient error {err:.3e}"


def test_backward_gaussian_matches_torch_reference(ddcsvd):
    m, n = 400, 240
    gpu_case(f"adjoint gaussian {m}x{n}")
    A = gen.normal_matrix(m, n, seed=21)
    gU, gs, gV = _cotangents(m, n, 8)
    s, U, _, gA = _gpu_grad(ddcsvd, A, gU, gs, gV)
    ref = _ref_grad(A, U, gU, gs, gV)
    # both sides amplify the factors' rounding by 1 / (smallest relative gap)
    relgap = np.min(np.diff(s)) / s[-1]
    err = np.linalg.norm(gA - ref) / np.linalg.norm(ref)
    assert err <= max(1e-10, 1e-14 / relgap), f"relative gradient error {err:.3e}"


def test_backward_sigma_only_and_null_cotangents(ddcsvd):
    m, n = 96, 64
    gpu_case("adjoint null cotangents")
    A = gen.uniform_matrix(m, n, seed=4)
    gs = np.random.default_rng(1).standard_normal(n)
    s, U, V, gA = _gpu_grad(ddcsvd, A, None, gs, None)
    np.testing.assert_allclose(gA, (U * gs) @ V.T, rtol=0, atol=1e-13 * np.abs(gs).max() * n)
    st, Ut, Vt = ddcsvd.svd_torch(_to_gpu(A))
    zero = ddcsvd.svd_backward_torch(Ut, st, Vt, None, None, None)
    assert torch.count_nonzero(zero).item() == 0


def test_gap_guard_repeated_singular_values(ddcsvd):
    m, n = 80, 40
    gpu_case("adjoint gap guard")
    sig = np.linspace(1.0, 2.0, n)
    sig[10] = sig[11]             # exact repeat
    sig[20] = sig[21] * (1 + 1e-15)  # below the guard
    A = gen.graded_matrix(m, n, sig, seed=9)
    gU, gs, gV = _cotangents(m, n, 2)
    s, U, V, gA = _gpu_grad(ddcsvd, A, gU, gs, gV, guard=1e-10)
    assert np.all(np.isfinite(gA))
    # the forward factors stay orthonormal through the exact repeat
    assert np.linalg.norm(U.T @ U - np.eye(n)) <= 2e-13 * n
    assert np.linalg.norm(V.T @ V - np.eye(n)) <= 2e-13 * n
    # d / delta^2 stays below 1 / delta in size, delta = guard * s_max^2
    bound = (1.0 / (1e-10 * s[-1] ** 2)) * (np.linalg.norm(gU) + np.linalg.norm(gV)) * 4
    assert np.linalg.norm(gA) <= bound
    # singular-value-only loss: exact and independent of the guard
    _, U2, V2, gA2 = _gpu_grad(ddcsvd, A, None, gs, None, guard=0.0)
    np.testing.assert_allclose(gA2, (U2 * gs) @ V2.T, rtol=0, atol=1e-13 * n)


def test_forward_is_deterministic(ddcsvd):
    gpu_case("svd determinism 700x300")
    A = _to_gpu(gen.normal_matrix(700, 300, seed=5))
    r1 = ddcsvd.svd_torch(A)
    r2 = ddcsvd.svd_torch(A)
    for a, b in zip(r1, r2):
        assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(200, 120), (90, 150)])
def test_autograd_harness_recompute_matches_stored(ddcsvd, shape):
    import ddcsvd.autograd as ag
    m, n = shape
    mn = min(m, n)
    gpu_case(f"autograd harness {m}x{n}")
    A = gen.graded_matrix(m, n, np.linspace(1.0, 2.0, mn), seed=12)
    gU, gs, gV = _cotangents(m, n, 5)
    grads = []
    for recompute in (False, True):
        At = _to_gpu(A).requires_grad_(True)
        s, U, V = ag.svd(At, recompute=recompute)
        L = (s * _to_gpu(gs)).sum() + (U * _to_gpu(gU)).sum() + (V * _to_gpu(gV)).sum()
        L.backward()
        grads.append(At.grad.detach().cpu().numpy())
        Uc = U.detach().cpu().numpy()
    assert np.array_equal(grads[0], grads[1])
    ref = _ref_grad(A, Uc, gU, gs, gV)
    err = np.linalg.norm(grads[0] - ref) / np.linalg.norm(ref)
    assert err <= 1e-10, f"relative gradient error {err:.3e}"
    # only some outputs used
    At = _to_gpu(A).requires_grad_(True)
    s, U, V = ag.svd(At)
    s.sum().backward()
    Uc, Vc = U.detach().cpu().numpy(), V.detach().cpu().numpy()
    np.testing.assert_allclose(At.grad.cpu().numpy(), Uc @ Vc.T, rtol=0, atol=1e-12)


def test_finite_differences_2048x1024(ddcsvd):
    """A = Q diag(linspace(1, 2, 1024)) P^T, cotangents i.i.d. N(0,1)."""
    m, n = 2048, 1024
    gpu_case(f"adjoint finite differences {m}x{n}")
    A = gen.graded_matrix(m, n, np.linspace(1.0, 2.0, n), seed=31)
    gU, gs, gV = _cotangents(m, n, 77)
    s0, U0, V0, gA = _gpu_grad(ddcsvd, A, gU, gs, gV)
    ref = _ref_grad(A, U0, gU, gs, gV)
    err_ref = np.linalg.norm(gA - ref) / np.linalg.norm(ref)
    assert err_ref <= 1e-10, f"relative error vs torch reference {err_ref:.3e}"

    Ad, U0d = _to_gpu(A), _to_gpu(U0)
    gUd, gsd, gVd, gAd = _to_gpu(gU), _to_gpu(gs), _to_gpu(gV), _to_gpu(gA)

    def loss(Ap):
        s, U, V = ddcsvd.svd_torch(Ap)
        sg = torch.sign((U * U0d).sum(0))
        return ((U * sg * gUd).sum() + (s * gsd).sum() + (V * sg * gVd).sum()).item()

    h = 1e-6
    g = torch.Generator().manual_seed(2024)
    an, fd2, fd4 = [], [], []
    for k in range(64):
        E = torch.randn((m, n), generator=g, dtype=torch.float64).to(_dev())
        fm2, fm1, fp1, fp2 = (loss(Ad + c * h * E) for c in (-2, -1, 1, 2))
        an.append(torch.sum(gAd * E).item())
        fd2.append((fp1 - fm1) / (2 * h))
        fd4.append((fm2 - 8 * fm1 + 8 * fp1 - fp2) / (12 * h))
    an, fd2, fd4 = map(np.array, (an, fd2, fd4))
    e4 = np.linalg.norm(fd4 - an) / np.linalg.norm(an)
    e2 = np.linalg.norm(fd2 - an) / np.linalg.norm(an)
    print(f"[adjoint-fd] {m}x{n} 64 directions h={h}: 4th-order rel err {e4:.3e}, "
          f"2-point rel err {e2:.3e}, vs torch reference {err_ref:.3e}", flush=True)
    assert e4 <= 1e-8, f"finite-difference relative error {e4:.3e}"
    assert e2 <= 1e-5, f"two-point finite-difference relative error {e2:.3e}"


def test_backward_rejects_bad_arguments(ddcsvd):
    gpu_case("adjoint argument checks")
    lib = ddcsvd.load()
    U = torch.zeros((8, 4), dtype=torch.float64, device=_dev())
    s = torch.ones(4, dtype=torch.float64, device=_dev())
    V = torch.zeros((4, 4), dtype=torch.float64, device=_dev())
    gA = torch.zeros((8, 4), dtype=torch.float64, device=_dev())
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rc = lib.svd_gpu_backward(8, 4, p(U), 4, p(s), p(V), 4, None, 0, p(s), None, 0, p(gA), 8,
                              1e-13, None)
    assert rc == ddcsvd.E_ARG  # ldu < m
    rc = lib.svd_gpu_backward(0, 4, p(U), 8, p(s), p(V), 4, None, 0, p(s), None, 0, p(gA), 8,
                              1e-13, None)
    assert rc == ddcsvd.E_ARG
    rc = lib.svd_gpu_backward(8, 4, p(U), 8, p(s), p(V), 4, None, 0, p(s), None, 0, p(gA), 8,
                              -1.0, None)
    assert rc == ddcsvd.E_ARG
