"""Seeded synthetic matrices shared by the tests, the bench and the CPU baseline.

Counter-based splitmix64: entry k (column-major index) of the stream with
seed s is mix64(s + (k+1) * 0x9E3779B97F4A7C15), so every entry is
independent and numpy, C and device code produce the same bits for the
uniform draws.  Normal draws use Box-Muller on two uniform streams.

The literal input of the reference's smoke driver (test-whole-svd.c:38,69-73:
square m x m, U[1,4) from glibc rand() with the default seed 1) is available
as ``glibc_rand_matrix``.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(x: np.ndarray) -> np.ndarray:
    x = (x ^ (x >> np.uint64(30))) * _M1
    x = (x ^ (x >> np.uint64(27))) * _M2
    return x ^ (x >> np.uint64(31))


def uniform01(count: int, seed: int, stream: int = 0) -> np.ndarray:
    """``count`` doubles in [0, 1) with 53 random bits each."""
    base = np.uint64((int(seed) * 0x632BE59BD9B4E019 + int(stream) * 0x2545F4914F6CDD1D) & (2**64 - 1))
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = _mix64(base + k * _GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_matrix(m: int, n: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """m x n Fortran-ordered matrix, entries i.i.d. U[lo, hi)."""
    u = uniform01(m * n, seed)
    return np.asfortranarray((lo + (hi - lo) * u).reshape(n, m).T)


def normal_matrix(m: int, n: int, seed: int) -> np.ndarray:
    """m x n Fortran-ordered matrix, entries i.i.d. N(0, 1) (Box-Muller)."""
    u1 = uniform01(m * n, seed, stream=1)
    u2 = uniform01(m * n, seed, stream=2)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    return np.asfortranarray(z.reshape(n, m).T)


def haar_orthogonal(n: int, seed: int) -> np.ndarray:
    """Haar-random n x n orthogonal matrix: QR of a seeded N(0,1) matrix with
    the columns of Q multiplied by sign(diag R)."""
    g = normal_matrix(n, n, seed)
    q, r = np.linalg.qr(g)
    s = np.sign(np.diag(r))
    s[s == 0] = 1.0
    return np.asfortranarray(q * s)


def graded_matrix(m: int, n: int, sigma: np.ndarray, seed: int) -> np.ndarray:
    """A = Q diag(sigma) P^T with Haar Q (m x m) and P (n x n), seeds seed and seed+1."""
    q = haar_orthogonal(m, seed)[:, : len(sigma)]
    p = haar_orthogonal(n, seed + 1)[:, : len(sigma)]
    return np.asfortranarray((q * sigma) @ p.T)


def glibc_rand_matrix(m: int, n: int, lo: float = 1.0, hi: float = 4.0, seed: int = 1) -> np.ndarray:
    """The reference driver's input: lo + rand()/(RAND_MAX+1) * (hi-lo), drawn
    in column-major order from glibc rand() after srand(seed)
    (test-whole-svd.c:18-24, :69-73; the driver never seeds, i.e. seed 1)."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.rand.restype = ctypes.c_int
    libc.srand(ctypes.c_uint(seed))
    rand_max = 2147483647.0
    vals = np.empty(m * n, dtype=np.float64)
    for k in range(m * n):
        vals[k] = lo + (libc.rand() / (1.0 * rand_max + 1)) * (hi - lo)
    return np.asfortranarray(vals.reshape(n, m).T)


def batch_shapes(count: int, seed: int, mlo: int = 16, mhi: int = 256, nlo: int = 8):
    """Shapes of the batched workload: m uniform in [mlo, mhi], n uniform in
    [nlo, m] (both inclusive), drawn from the same counter stream."""
    u = uniform01(2 * count, seed, stream=7).reshape(count, 2)
    ms = mlo + np.floor(u[:, 0] * (mhi - mlo + 1)).astype(np.int64)
    ns = nlo + np.floor(u[:, 1] * (ms - nlo + 1)).astype(np.int64)
    return ms, ns
