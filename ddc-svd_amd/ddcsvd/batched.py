"""Batched SVD of many small matrices through libddcsvd's svd_gpu_batched_*.

    ms, ns = shapes
    b = BatchedSvd(ms, ns, dtype=torch.float64)
    A = b.pack([...])                       # one flat CUDA buffer, column-major blocks
    S, U, V = b.outputs()
    b.run(A, S, U, V)                       # first call tunes and captures a graph
    sigma_k, U_k, V_k = b.unpack(k, S, U, V)

Every call goes to the HIP library (one persistent launch per run); there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, load

F64, F32 = 0, 1


def _bind(lib):
    if getattr(lib, "_batched_bound", False):
        return lib
    i, v, pll = ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)
    pi = ctypes.POINTER(ctypes.c_int)
    lib.svd_gpu_batched_plan.argtypes = [ctypes.POINTER(v), i, pi, pi, i]
    lib.svd_gpu_batched_plan.restype = i
    lib.svd_gpu_batched_offsets.argtypes = [v, pll, pll, pll, pll]
    lib.svd_gpu_batched_offsets.restype = i
    lib.svd_gpu_batched_run.argtypes = [v, v, v, v, v, v, v]
    lib.svd_gpu_batched_run.restype = i
    lib.svd_gpu_batched_tune.argtypes = [v, v, v, v, v, v]
    lib.svd_gpu_batched_tune.restype = i
    lib.svd_gpu_batched_config.argtypes = [v, pi, pi, pi, pi, pi]
    lib.svd_gpu_batched_config.restype = i
    lib.svd_gpu_batched_set_config.argtypes = [v, i, i, i]
    lib.svd_gpu_batched_set_config.restype = i
    lib.svd_gpu_batched_destroy.argtypes = [v]
    lib.svd_gpu_batched_destroy.restype = None
    lib.svd_gpu_batched_noalloc.argtypes = [i, pi, pi, i, v, v, v, v, v, v]
    lib.svd_gpu_batched_noalloc.restype = i
    lib._batched_bound = True
    return lib


class BatchedSvd:
    """A plan for one batch of shapes (m_k, n_k), 1 <= m_k, n_k <= 256."""

    def __init__(self, ms, ns, dtype=None):
        import torch
        self.torch = torch
        self.dtype = torch.float64 if dtype is None else dtype
        if self.dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be torch.float64 or torch.float32")
        self.lib = _bind(load())
        self.ms = np.ascontiguousarray(ms, dtype=np.int32)
        self.ns = np.ascontiguousarray(ns, dtype=np.int32)
        self.count = len(self.ms)
        self.plan = ctypes.c_void_p()
        pi = ctypes.POINTER(ctypes.c_int)
        _check(self.lib.svd_gpu_batched_plan(
            ctypes.byref(self.plan), self.count, self.ms.ctypes.data_as(pi),
            self.ns.ctypes.data_as(pi), F64 if self.dtype == torch.float64 else F32))
        offs = [np.zeros(self.count + 1, dtype=np.int64) for _ in range(4)]
        pll = ctypes.POINTER(ctypes.c_longlong)
        _check(self.lib.svd_gpu_batched_offsets(self.plan, *[o.ctypes.data_as(pll) for o in offs]))
        self.offA, self.offS, self.offU, self.offV = offs

    def close(self):
        if self.plan:
            self.lib.svd_gpu_batched_destroy(self.plan)
            self.plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _empty(self, n):
        return self.torch.empty(int(n), dtype=self.dtype, device="cuda")

    def outputs(self):
        """Fresh (sigma, U, V) buffers of the packed sizes."""
        return (self._empty(self.offS[-1]), self._empty(self.offU[-1]),
                self._empty(self.offV[-1]))

    def pack(self, mats):
        """One flat CUDA buffer holding the matrices column-major, back to back."""
        flat = np.empty(int(self.offA[-1]), dtype=np.float64)
        for k, a in enumerate(mats):
            a = np.asarray(a, dtype=np.float64)
            assert a.shape == (self.ms[k], self.ns[k])
            flat[self.offA[k]:self.offA[k + 1]] = a.ravel(order="F")
        return self.torch.from_numpy(flat).to(device="cuda", dtype=self.dtype)

    def unpack(self, k, S, U, V):
        """(sigma, U, V) of matrix k as host float64 arrays."""
        m, n = int(self.ms[k]), int(self.ns[k])
        mn = min(m, n)
        s = S[self.offS[k]:self.offS[k + 1]].double().cpu().numpy()
        u = U[self.offU[k]:self.offU[k + 1]].double().cpu().numpy().reshape(mn, m).T
        v = V[self.offV[k]:self.offV[k + 1]].double().cpu().numpy().reshape(mn, n).T
        return s, u, v

    def _ptrs(self, A, S, U, V):
        for t in (A, S, U, V):
            if t.dtype != self.dtype or not t.is_cuda or not t.is_contiguous():
                raise ValueError("buffers must be contiguous CUDA tensors of the plan's dtype")
        if (A.numel() < self.offA[-1] or S.numel() < self.offS[-1] or U.numel() < self.offU[-1]
                or V.numel() < self.offV[-1]):
            raise ValueError("buffer too small for the batch")
        return [ctypes.c_void_p(t.data_ptr()) for t in (A, S, U, V)]

    def run(self, A, S, U, V, info=None, stream=None):
        """Enqueue the batch on `stream` (default: torch's current stream)."""
        if stream is None:
            stream = self.torch.cuda.current_stream().cuda_stream
        ip = None if info is None else ctypes.c_void_p(info.data_ptr())
        _check(self.lib.svd_gpu_batched_run(self.plan, *self._ptrs(A, S, U, V), ip,
                                            ctypes.c_void_p(stream)))

    def run_noalloc(self, A, S, U, V, info=None, stream=None):
        """The same batch through svd_gpu_batched_noalloc (no plan buffers, no
        device allocation; fixed launch configuration)."""
        if stream is None:
            stream = self.torch.cuda.current_stream().cuda_stream
        ip = None if info is None else ctypes.c_void_p(info.data_ptr())
        pi = ctypes.POINTER(ctypes.c_int)
        _check(self.lib.svd_gpu_batched_noalloc(
            self.count, self.ms.ctypes.data_as(pi), self.ns.ctypes.data_as(pi),
            F64 if self.dtype == self.torch.float64 else F32, *self._ptrs(A, S, U, V), ip,
            ctypes.c_void_p(stream)))

    def tune(self, A, S, U, V, stream=None):
        if stream is None:
            stream = self.torch.cuda.current_stream().cuda_stream
        _check(self.lib.svd_gpu_batched_tune(self.plan, *self._ptrs(A, S, U, V),
                                             ctypes.c_void_p(stream)))

    def config(self) -> dict:
        vals = [ctypes.c_int() for _ in range(5)]
        nvar = self.lib.svd_gpu_batched_config(self.plan, *[ctypes.byref(x) for x in vals])
        return {"variant": vals[0].value, "workgroups_per_cu": vals[1].value,
                "coop_min_n": vals[2].value, "block_le128_rows": vals[3].value,
                "block_gt128_rows": vals[4].value, "variants": nvar}

    def set_config(self, variant: int, workgroups_per_cu: int, coop_min_n: int = 96):
        _check(self.lib.svd_gpu_batched_set_config(self.plan, variant, workgroups_per_cu,
                                                   coop_min_n))
