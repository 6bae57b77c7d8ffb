"""Write the LAPACK golden vectors used by the oracle and GPU parity tests.

Every expected output here comes from LAPACK's own RFP routines (scipy.linalg.lapack,
scipy-openblas), called exactly as the reference's tests call them
(python_tests/test_vs_lapack.py:32-120, python_tests/test_stfttr_strttf.py:44-60):

  rfp_perm.npz   — ?trttf of A[i, j] = i*n + j + 1 (float64, exact), i.e. the
                   RFP index permutation, for the 19 sizes of
                   python_tests/test_stfttr_strttf.py:25 and all 4 transr x uplo.
  chain_f32.npz  — ssfrk(A A^T, trans='N') + ssfrk(I, beta=1) -> C, spftrf -> F,
                   spftrs(B, nrhs=3) -> X for the 15 (n, k) pairs of
                   python_tests/test_vs_lapack.py:133-152, all 4 variants, seeds as
                   there (42 + case, B seed + 1000).
  chain_f64.npz  — the same with dsfrk / dpftrf / dpftrs for a subset of sizes.
  info.npz       — LAPACK ?pftrf info for crafted non-SPD RFP matrices.

Run:  python tests/golden/make_golden.py   (writes next to this file)
"""
import math
import os

import numpy as np
import scipy.linalg.lapack as lapack

HERE = os.path.dirname(os.path.abspath(__file__))
N_PERM = [1, 2, 3, 4, 5, 6, 7, 8, 15, 16, 17, 32, 33, 63, 64, 65, 128, 129, 257]
CHAIN = [(1, 1), (2, 2), (3, 3), (4, 4), (5, 5), (8, 6), (9, 7), (32, 16), (33, 17),
         (64, 32), (65, 33), (128, 64), (129, 65), (256, 128), (257, 129)]
CHAIN64 = [(1, 1), (2, 2), (3, 3), (5, 5), (8, 6), (9, 7), (33, 17), (64, 32), (129, 65)]
VARIANTS = [(t, u) for t in "NT" for u in "LU"]


def perm():
    out = {}
    for n in N_PERM:
        A = np.arange(1, n * n + 1, dtype=np.float64).reshape(n, n)
        for t, u in VARIANTS:
            arf, info = lapack.dtrttf(A, transr=t.encode(), uplo=u.encode())
            assert info == 0
            out[f"n{n}_{t}{u}"] = arf.astype(np.int32)  # values are exact integers
    np.savez_compressed(os.path.join(HERE, "rfp_perm.npz"), **out)


def chain(dtype, sizes, fname):
    sfrk = lapack.ssfrk if dtype == np.float32 else lapack.dsfrk
    pftrf = lapack.spftrf if dtype == np.float32 else lapack.dpftrf
    pftrs = lapack.spftrs if dtype == np.float32 else lapack.dpftrs
    out = {}
    seed = 42
    for n, k in sizes:
        for t, u in VARIANTS:
            rng = np.random.default_rng(seed)
            A = (rng.standard_normal((n, k)) / math.sqrt(k)).astype(dtype)
            C = np.zeros(n * (n + 1) // 2, dtype=dtype)
            C = sfrk(n, k, 1.0, np.asfortranarray(A), 0.0, C, transr=t.encode(), uplo=u.encode(), trans=b"N")
            C = sfrk(n, n, 1.0, np.asfortranarray(np.eye(n, dtype=dtype)), 1.0, C,
                     transr=t.encode(), uplo=u.encode(), trans=b"N")
            F, info = pftrf(n, C.copy(), transr=t.encode(), uplo=u.encode())
            assert info == 0
            rng = np.random.default_rng(seed + 1000)
            B = rng.standard_normal((n, 3)).astype(dtype)
            X, info = pftrs(n, F, B.copy(), transr=t.encode(), uplo=u.encode())
            assert info == 0
            key = f"n{n}_k{k}_{t}{u}"
            out[key + "_A"] = A
            out[key + "_C"] = C
            out[key + "_F"] = F
            out[key + "_B"] = B
            out[key + "_X"] = X
            seed += 1
    np.savez_compressed(os.path.join(HERE, fname), **out)


def info_cases():
    """Symmetric matrices whose leading minor of order p fails first, packed
    with ?trttf; LAPACK ?pftrf gives the expected info for each variant."""
    out = {}
    rng = np.random.default_rng(7)
    for n in [1, 2, 3, 5, 8, 9, 33, 64, 65, 131]:
        for p in sorted({1, max(1, n // 2), max(1, n // 2 + 1), n}):
            G = rng.standard_normal((n, n))
            S = G @ G.T / n + np.eye(n)
            # make the p-th leading minor singular-negative: shift the (p-1, p-1) entry
            S[p - 1, p - 1] -= 10.0 * n + S[p - 1, p - 1]
            for dt, tag in ((np.float32, "s"), (np.float64, "d")):
                Sd = S.astype(dt)
                for t, u in VARIANTS:
                    arf, _ = (lapack.strttf if dt == np.float32 else lapack.dtrttf)(
                        Sd, transr=t.encode(), uplo=u.encode())
                    _, info = (lapack.spftrf if dt == np.float32 else lapack.dpftrf)(
                        n, arf.copy(), transr=t.encode(), uplo=u.encode())
                    key = f"{tag}_n{n}_p{p}_{t}{u}"
                    out[key + "_arf"] = arf
                    out[key + "_info"] = np.array(info, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "info.npz"), **out)


if __name__ == "__main__":
    perm()
    chain(np.float32, CHAIN, "chain_f32.npz")
    chain(np.float64, CHAIN64, "chain_f64.npz")
    info_cases()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
