"""svd_gpu_batched_noalloc against svd_gpu_batched_run: the no-alloc entry
point runs the same kernel at a fixed configuration; a plan forced to that
configuration must give bit-identical results.  Two calls in a row, the
second with a smaller batch, then a batch over the scratch limit (rejected).
"""
import numpy as np
import pytest

from conftest import gpu_case
from ddcsvd import gen

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NOALLOC_CONFIG = (1, 2, 32)  # variant, workgroups per CU, coop_min_n


def _batch(count, seed, mhi):
    ms, ns = gen.batch_shapes(count, seed, mlo=8, mhi=mhi, nlo=4)
    mats = [gen.uniform_matrix(int(m), int(n), seed=seed * 7919 + k)
            for k, (m, n) in enumerate(zip(ms, ns))]
    return ms, ns, mats


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_noalloc_matches_planned_run(ddcsvd, dtype):
    from ddcsvd.batched import BatchedSvd
    for call, (count, seed, mhi) in enumerate(((300, 11, 96), (41, 12, 64))):
        ms, ns, mats = _batch(count, seed, mhi)
        b = BatchedSvd(ms, ns, dtype=dtype)
        b.set_config(*NOALLOC_CONFIG)
        A = b.pack(mats)
        S1, U1, V1 = b.outputs()
        S2, U2, V2 = b.outputs()
        info1 = torch.zeros(count, dtype=torch.int32, device="cuda")
        info2 = torch.full((count,), -7, dtype=torch.int32, device="cuda")
        gpu_case(f"batched planned run, call {call}, {count} matrices {dtype}")
        b.run(A, S1, U1, V1, info1)
        torch.cuda.synchronize()
        gpu_case(f"batched no-alloc run, call {call}, {count} matrices {dtype}")
        b.run_noalloc(A, S2, U2, V2, info2)
        torch.cuda.synchronize()
        assert torch.equal(S1, S2) and torch.equal(U1, U2) and torch.equal(V1, V2)
        assert torch.equal(info1, info2)
        # and the values against LAPACK
        eps = np.finfo(np.float64 if dtype == torch.float64 else np.float32).eps
        for k in range(0, count, 7):
            s, _, _ = b.unpack(k, S2, U2, V2)
            sl = np.sort(np.linalg.svd(mats[k], compute_uv=False))
            assert np.max(np.abs(s - sl)) <= 64 * eps * max(sl[-1], 1e-300) * max(mats[k].shape)
        b.close()


def test_noalloc_rejects_oversized_batch(ddcsvd):
    from ddcsvd.batched import BatchedSvd
    ms = np.full(200, 256)
    ns = np.full(200, 256)
    b = BatchedSvd(ms, ns, dtype=torch.float64)  # 200 x (256^2 + 256^2) x 8 B > 32 MiB
    A = torch.zeros(int(b.offA[-1]), dtype=torch.float64, device="cuda")
    S, U, V = b.outputs()
    gpu_case("batched no-alloc oversized batch (rejected on the host)")
    with pytest.raises(ddcsvd.SvdGpuError):
        b.run_noalloc(A, S, U, V)
    assert "no-alloc limit" in ddcsvd.last_error()
    b.close()
