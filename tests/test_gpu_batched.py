"""GPU tests of the batched SVD (svd_gpu_batched_*): one persistent launch
over many small matrices of uneven shapes, replayed from a HIP graph.

Per matrix, against LAPACK (numpy, FP64), with eps the unit roundoff of the
working precision, mt = max(m, n), mn = min(m, n):
  |sigma - sigma_LAPACK| / sigma_max   <= 8 eps sqrt(mt mn)   (FP64: or 1e-13)
  ||A - U S V^T||_F / ||A||_F          <= 8 eps sqrt(mt mn)   (FP64: or 1e-13)
  ||U^T U - I||_F, ||V^T V - I||_F     <= 4 eps sqrt(mt) mn   (FP64: or 1e-13 mn)
(one-sided Jacobi stops when every pair of columns is orthogonal to
sqrt(mt) eps, so the Frobenius orthogonality error grows with mn).
"""
import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _tol(dtype, mt, mn):
    """(sigma, residual, orthogonality) bounds for an mt x mn problem."""
    eps = np.finfo(np.float64 if dtype == torch.float64 else np.float32).eps / 2
    mn4 = max(mn, 4)
    ts = 8 * eps * np.sqrt(mt * mn4)
    to = 4 * eps * np.sqrt(mt) * mn4
    if dtype == torch.float64:
        return max(ts, 1e-13), max(ts, 1e-13), max(to, 1e-13 * mn4)
    return ts, ts, to


def _mats(ms, ns, seed):
    return [gen.uniform_matrix(int(m), int(n), seed=seed * 100003 + k) for k, (m, n) in
            enumerate(zip(ms, ns))]


def _check_batch(b, mats, S, U, V, dtype, info=None, sample=None):
    torch.cuda.synchronize()
    ks = range(b.count) if sample is None else sample
    worst = [0.0, 0.0, 0.0]
    bad = []
    for k in ks:
        A = mats[k]
        s, u, v = b.unpack(k, S, U, V)
        sl = np.sort(np.linalg.svd(A, compute_uv=False))
        smax = max(sl[-1], 1e-300)
        mn = len(sl)
        es = np.max(np.abs(s - sl)) / smax
        er = np.linalg.norm(A - (u * s) @ v.T) / max(np.linalg.norm(A), 1e-300)
        eo = max(np.linalg.norm(u.T @ u - np.eye(mn)), np.linalg.norm(v.T @ v - np.eye(mn)))
        ts, tr, to = _tol(dtype, max(A.shape), mn)
        if not np.all(np.diff(s) >= 0):
            bad.append(f"matrix {k}: sigma not ascending")
        worst = [max(worst[0], es / ts), max(worst[1], er / tr), max(worst[2], eo / to)]
        if es > ts or er > tr or eo > to:
            bad.append(f"matrix {k} {A.shape}: sigma {es:.2e} residual {er:.2e} orth {eo:.2e}")
    print(f"[batched-check] {dtype} worst error / bound: sigma {worst[0]:.2f} "
          f"residual {worst[1]:.2f} orth {worst[2]:.2f}", flush=True)
    assert not bad, "; ".join(bad[:5])
    if info is not None:
        inf = info.cpu().numpy()
        assert np.all(inf[list(ks)] >= 1), "a matrix did not converge"
    return worst


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_uneven_shapes(ddcsvd, dtype):
    from ddcsvd.batched import BatchedSvd
    ms, ns = gen.batch_shapes(300, seed=4)
    gpu_case(f"batched 300 uneven {dtype}")
    mats = _mats(ms, ns, 1)
    b = BatchedSvd(ms, ns, dtype)
    A = b.pack(mats)
    S, U, V = b.outputs()
    info = torch.zeros(b.count, dtype=torch.int32, device="cuda")
    b.run(A, S, U, V, info)
    w = _check_batch(b, mats, S, U, V, dtype, info)
    print(f"[batched] {dtype} config {b.config()}", flush=True)
    b.close()


def test_batched_edge_shapes(ddcsvd):
    from ddcsvd.batched import BatchedSvd
    shapes = [(1, 1), (1, 7), (7, 1), (2, 2), (3, 200), (256, 256), (256, 1), (1, 256),
              (64, 64), (65, 64), (128, 129), (192, 191), (17, 256), (100, 3), (5, 5)]
    ms = [s[0] for s in shapes]
    ns = [s[1] for s in shapes]
    gpu_case("batched edge shapes")
    mats = _mats(ms, ns, 2)
    b = BatchedSvd(ms, ns, torch.float64)
    A = b.pack(mats)
    S, U, V = b.outputs()
    info = torch.zeros(b.count, dtype=torch.int32, device="cuda")
    b.run(A, S, U, V, info)
    _check_batch(b, mats, S, U, V, torch.float64, info)
    b.close()


def test_batched_every_variant(ddcsvd):
    from ddcsvd.batched import BatchedSvd
    ms, ns = gen.batch_shapes(120, seed=9)
    gpu_case("batched variants")
    mats = _mats(ms, ns, 3)
    b = BatchedSvd(ms, ns, torch.float64)
    A = b.pack(mats)
    nvar = b.config()["variants"]
    for var in range(nvar):
        for wgs, coop in ((1, 257), (2, 32), (3, 96)):
            S, U, V = b.outputs()
            b.set_config(var, wgs, coop)
            b.run(A, S, U, V)
            _check_batch(b, mats, S, U, V, torch.float64)
    b.close()


def test_batched_graph_replay(ddcsvd):
    from ddcsvd.batched import BatchedSvd
    ms, ns = gen.batch_shapes(200, seed=5)
    gpu_case("batched graph replay")
    mats = _mats(ms, ns, 4)
    b = BatchedSvd(ms, ns, torch.float64)
    A = b.pack(mats)
    S, U, V = b.outputs()
    b.run(A, S, U, V)
    torch.cuda.synchronize()
    ref = [t.clone() for t in (S, U, V)]
    for rep in range(3):
        for t in (S, U, V):
            t.fill_(float("nan"))  # eager work between the replays
        torch.cuda.synchronize()
        b.run(A, S, U, V)
        torch.cuda.synchronize()
        for a, r in zip((S, U, V), ref):
            assert torch.equal(a, r), f"replay {rep} differs"
    # new buffers: a new graph
    S2, U2, V2 = b.outputs()
    b.run(A, S2, U2, V2)
    torch.cuda.synchronize()
    for a, r in zip((S2, U2, V2), ref):
        assert torch.equal(a, r)
    b.close()


def test_batched_rank_deficient(ddcsvd):
    from ddcsvd.batched import BatchedSvd
    gpu_case("batched rank deficient")
    z = np.zeros((30, 12))
    r1 = np.outer(np.arange(1, 41), np.arange(1, 9)).astype(float)
    dup = gen.uniform_matrix(50, 10, seed=8)
    dup[:, 5] = dup[:, 2]
    wide0 = np.zeros((6, 20))
    mats = [z, r1, dup, wide0]
    b = BatchedSvd([a.shape[0] for a in mats], [a.shape[1] for a in mats], torch.float64)
    A = b.pack(mats)
    S, U, V = b.outputs()
    b.run(A, S, U, V)
    torch.cuda.synchronize()
    for k, a in enumerate(mats):
        s, u, v = b.unpack(k, S, U, V)
        mn = min(a.shape)
        assert np.linalg.norm(a - (u * s) @ v.T) <= 1e-13 * max(np.linalg.norm(a), 1.0)
        assert np.linalg.norm(u.T @ u - np.eye(mn)) <= 1e-12
        assert np.linalg.norm(v.T @ v - np.eye(mn)) <= 1e-12
    b.close()


def test_batched_rejects_bad_shapes(ddcsvd):
    from ddcsvd.batched import BatchedSvd
    gpu_case("batched argument checks")
    with pytest.raises(ddcsvd.SvdGpuError):
        BatchedSvd([300], [10])
    with pytest.raises(ddcsvd.SvdGpuError):
        BatchedSvd([10, 0], [10, 5])
