"""curfpSpftrsNoAlloc (the allocation-free spftrs) through the C ABI against
curfpSpftrs on the same factor and right-hand sides: bit-identical results
(the same kernels), on two calls in a row, the second smaller, plus the size
limits and a residual check of the second solve in float64 on the host."""
import ctypes
import math
import os

import numpy as np
import pytest
import torch

from oracle import rfp_oracle as O

pytestmark = pytest.mark.gpu

OPS = {"N": 0, "T": 1}
FILLS = {"L": 0, "U": 1}


@pytest.fixture(scope="module")
def abi(cuda):
    import curfp_amd

    lib = ctypes.CDLL(os.path.join(os.path.dirname(curfp_amd.__file__), "libcurfp_amd.so"))
    h = ctypes.c_void_p()
    assert lib.curfpCreate(ctypes.byref(h)) == 0
    yield curfp_amd, lib, h
    assert lib.curfpDestroy(h) == 0


def _factor(curfp, n, k, t, u, seed):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((n, k)) / math.sqrt(k)).astype(np.float32)
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(torch.from_numpy(A).cuda(), C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    assert curfp.spftrf(C, transr=t, uplo=u) == 0
    return A, C


def _solve_both(lib, h, C, B, n, nrhs, t, u):
    """B: (nrhs, n) contiguous = the n x nrhs column-major block, ldb = n."""
    X_ref, X_na = B.clone(), B.clone()
    torch.cuda.synchronize()
    p = ctypes.c_void_p
    assert lib.curfpSpftrs(h, OPS[t], FILLS[u], n, nrhs, p(C.data_ptr()), p(X_ref.data_ptr()), n) == 0
    assert lib.curfpSpftrsNoAlloc(h, OPS[t], FILLS[u], n, nrhs, p(C.data_ptr()), p(X_na.data_ptr()), n) == 0
    assert lib.curfpSynchronize(h) == 0
    return X_ref, X_na


def test_no_alloc_matches_spftrs(abi):
    curfp, lib, h = abi
    # first call: the largest size the fixed scratch holds
    n, nrhs, t, u = 16384, 64, "T", "U"
    _, C = _factor(curfp, n, 256, t, u, seed=61)
    B = torch.from_numpy(np.random.default_rng(62).standard_normal((nrhs, n)).astype(np.float32)).cuda()
    X_ref, X_na = _solve_both(lib, h, C, B, n, nrhs, t, u)
    assert bool(torch.isfinite(X_na).all())
    assert torch.equal(X_na, X_ref)
    del C
    # second call, smaller and another layout: the array still holds the first call's contents
    n, nrhs, t, u = 3001, 17, "N", "L"
    A, C = _factor(curfp, n, 200, t, u, seed=63)
    B = torch.from_numpy(np.random.default_rng(64).standard_normal((nrhs, n)).astype(np.float32)).cuda()
    X_ref, X_na = _solve_both(lib, h, C, B, n, nrhs, t, u)
    assert torch.equal(X_na, X_ref)
    S = A.astype(np.float64) @ A.astype(np.float64).T + np.eye(n)
    r = O.dense_relative_residual(S, X_na.cpu().numpy().T, B.cpu().numpy().T)
    assert r <= 1e-6, r


def test_no_alloc_limits(abi):
    curfp, lib, h = abi
    p = ctypes.c_void_p
    dummy = torch.zeros(16, device="cuda")
    ptr = p(dummy.data_ptr())
    assert lib.curfpSpftrsNoAlloc(h, 1, 1, 16385, 1, ptr, ptr, 16385) == 3
    assert lib.curfpSpftrsNoAlloc(h, 1, 1, 64, 4097, ptr, ptr, 64) == 3
    assert lib.curfpSpftrsNoAlloc(h, 1, 1, 64, 1, ptr, ptr, 63) == 3
    assert lib.curfpSpftrsNoAlloc(h, 1, 1, 0, 1, None, None, 1) == 0
    assert lib.curfpSpftrsNoAlloc(None, 1, 1, 4, 1, ptr, ptr, 4) == 1
