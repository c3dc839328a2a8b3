"""Writes tests/golden/ref_fixtures.npz from the reference's own code.

Runs the reference host routines compiled from their sources
(oracle/_ref/libddcref.so, built by `make -C oracle ref` where the reference
sources exist) on seeded inputs and stores inputs/outputs as small fixtures,
so the tests can pin results without the reference present.

    python tests/golden/gen_golden.py

Cases:
  c1lit_*   the literal `test-whole-svd 512 256` input: 512 x 512 U[1,4) from
            glibc rand() with the default seed (test-whole-svd.c:38,69-73);
            the reference's sigma, alpha, beta
  c1_*      512 x 256 U(-1,1), seed 1: sigma, alpha, beta, 4 edge columns of U
  full64_*  64 x 64 U(-1,1) seed 3: sigma, U, V, alpha, beta, A after bidiag
  full96_*  96 x 64 U(-1,1) seed 4: the same (the reference's V for m != n is
            stored too: it carries the multV indexing defect)
  dcN_*     value solver on random N x (N+1) bidiagonals, N in 1,2,3,5,17,64
  twN_*     twisted right vectors for N in 8, 64
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "ddc-svd_amd"))

import pyoracle as po  # noqa: E402
from ddcsvd import gen  # noqa: E402


def main() -> None:
    if po.ref_lib() is None:
        raise SystemExit("oracle/_ref/libddcref.so is not built (make -C oracle ref)")
    out = {}
    A = gen.glibc_rand_matrix(512, 512, 1.0, 4.0, seed=1)
    r = po.ref_svd(A)
    out.update(c1lit_sigma=r["sigma"], c1lit_alpha=r["alpha"], c1lit_beta=r["beta"])

    A = gen.uniform_matrix(512, 256, 1)
    r = po.ref_svd(A)
    out.update(c1_sigma=r["sigma"], c1_alpha=r["alpha"], c1_beta=r["beta"],
               c1_U_edge=np.hstack([r["U"][:, :4], r["U"][:, -4:]]))

    for tag, (m, n, seed) in (("full64", (64, 64, 3)), ("full96", (96, 64, 4))):
        A = gen.uniform_matrix(m, n, seed)
        r = po.ref_svd(A)
        out.update({f"{tag}_sigma": r["sigma"], f"{tag}_U": r["U"], f"{tag}_V": r["V"],
                    f"{tag}_alpha": r["alpha"], f"{tag}_beta": r["beta"],
                    f"{tag}_Amod": r["A_mod"]})

    rng_seed = 11
    for N in (1, 2, 3, 5, 17, 64):
        u = gen.uniform01(2 * N, rng_seed + N)
        b1 = 2.0 * u[:N] - 1.0
        b2 = 2.0 * u[N:] - 1.0
        out[f"dc{N}_b1"] = b1
        out[f"dc{N}_b2"] = b2
        out[f"dc{N}_sigma"] = po.ref_dc_values(b1, b2)

    lib = po.ref_lib()
    for N in (8, 64):
        u = gen.uniform01(2 * N, 100 + N)
        a = 2.0 * u[:N] - 1.0
        b = 2.0 * u[N:] - 1.0
        b[N - 1] = 0.0
        sig = po.ref_dc_values(a, b)
        X = np.zeros(N * N)
        lib.ref_right_vectors(N, N, po._ptr(a), po._ptr(b), po._ptr(sig), po._ptr(X))
        out[f"tw{N}_a"] = a
        out[f"tw{N}_b"] = b
        out[f"tw{N}_sigma"] = sig
        out[f"tw{N}_X"] = X.reshape(N, N)  # row i = vector i

    np.savez_compressed(os.path.join(HERE, "ref_fixtures.npz"), **out)
    print(f"wrote {len(out)} arrays to ref_fixtures.npz")


if __name__ == "__main__":
    main()
