"""GPU tests of the handle's failure and stream behaviour:

* a stream switch orders the handle's shared workspace and sweep counters
  (curfpSetStream makes the new stream wait for the old one);
* a substitution sweep whose bounded wait runs out reports
  CURFP_STATUS_EXECUTION_FAILED at the next synchronising call
  (curfpSynchronize, spftrf), after which the handle works again.  The
  failure is forced from the host with a wait limit of zero polls, so every
  workgroup stops waiting at once and every kernel exits normally.
"""
import math

import numpy as np
import pytest
import torch

from oracle import rfp_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def curfp(cuda):
    import curfp_amd

    return curfp_amd


def _factor(curfp, n, k, seed, t="T", u="U"):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((n, k)) / math.sqrt(k)).astype(np.float32)
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(torch.from_numpy(A).cuda(), C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    assert curfp.spftrf(C, transr=t, uplo=u) == 0
    S = A.astype(np.float64) @ A.astype(np.float64).T + np.eye(n)
    return C, S


def test_stream_switch_orders_shared_state(curfp):
    """Two spftrs calls on the same handle, the second issued on another
    stream while the first is still queued: both sweeps use the handle's
    ticket counter, progress words and look-back products, so without the
    fence they would overlap.  Both results must equal the one-stream result
    bit for bit, and meet the residual bound."""
    n, nrhs = 8192, 64
    C, S = _factor(curfp, n, 256, seed=41)
    B = torch.from_numpy(np.random.default_rng(42).standard_normal((n, nrhs)).astype(np.float32)).cuda()
    Xref = B.clone()
    curfp.spftrs(C, Xref)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    X1, X2 = B.clone(), B.clone()
    Ys = [B.clone() for _ in range(3)]
    torch.cuda.synchronize()
    try:
        curfp.set_stream(s1)
        for Y in Ys:                    # keep s1 busy when s2's work is queued
            curfp.spftrs(C, Y)
        curfp.spftrs(C, X1)
        curfp.set_stream(s2)
        curfp.spftrs(C, X2)
        curfp.synchronize()
        s1.synchronize()
    finally:
        curfp.set_stream(None)
    torch.cuda.synchronize()
    assert torch.equal(X1, Xref) and torch.equal(X2, Xref)
    assert O.dense_relative_residual(S, X2.cpu().numpy(), B.cpu().numpy()) <= 1e-6


def test_sweep_timeout_is_reported_and_cleared(curfp):
    """Wait limit 0: every sweep workgroup that finds an earlier block
    unfinished gives up at once and writes NaN.  curfpSynchronize returns
    EXECUTION_FAILED (RuntimeError status=4) once; a later solve with the
    default limit is exact again."""
    n, nrhs = 8192, 64
    C, S = _factor(curfp, n, 256, seed=43)
    B = torch.from_numpy(np.random.default_rng(44).standard_normal((n, nrhs)).astype(np.float32)).cuda()
    Xref = B.clone()
    curfp.spftrs(C, Xref)
    curfp.synchronize()
    h = curfp._get_handle()
    X = B.clone()
    try:
        h.set_sweep_wait_limit(0)
        curfp.spftrs(C, X)
        with pytest.raises(RuntimeError, match="status=4"):
            curfp.synchronize()
    finally:
        h.set_sweep_wait_limit(1 << 20)
    assert bool(torch.isnan(X).any())
    curfp.synchronize()                # reported once, then cleared
    X = B.clone()
    curfp.spftrs(C, X)
    curfp.synchronize()
    assert torch.equal(X, Xref)
    with pytest.raises(RuntimeError):
        h.set_sweep_wait_limit(-1)


def test_spftrf_reports_an_earlier_sweep_failure(curfp):
    """spftrf synchronises for info anyway and reports a failed earlier sweep
    (EXECUTION_FAILED); the next spftrf on a fresh matrix succeeds."""
    n, nrhs = 4096, 32
    C, S = _factor(curfp, n, 128, seed=45)
    B = torch.from_numpy(np.random.default_rng(46).standard_normal((n, nrhs)).astype(np.float32)).cuda()
    h = curfp._get_handle()
    try:
        h.set_sweep_wait_limit(0)
        curfp.spftrs(C, B.clone())
    finally:
        h.set_sweep_wait_limit(1 << 20)
    rng = np.random.default_rng(47)
    A = (rng.standard_normal((n, 128)) / math.sqrt(128)).astype(np.float32)
    C2 = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(torch.from_numpy(A).cuda(), C2)
    curfp.add_to_diagonal(C2, 1.0)
    C3 = C2.clone()
    with pytest.raises(RuntimeError, match="status=4"):
        curfp.spftrf(C2)
    assert curfp.spftrf(C3) == 0
    X = B.clone()
    curfp.spftrs(C3, X)
    S3 = A.astype(np.float64) @ A.astype(np.float64).T + np.eye(n)
    assert O.dense_relative_residual(S3, X.cpu().numpy(), B.cpu().numpy()) <= 1e-6


def test_sweep_wait_limit_raw_api(curfp):
    """The C entry points through ctypes: a negative limit is rejected
    (INVALID_VALUE), curfpSynchronize on a healthy handle succeeds, and a null
    handle is NOT_INITIALIZED."""
    import ctypes
    import os

    lib = ctypes.CDLL(os.path.join(os.path.dirname(curfp.__file__), "libcurfp_amd.so"))
    h = ctypes.c_void_p()
    assert lib.curfpCreate(ctypes.byref(h)) == 0
    try:
        assert lib.curfpSetSweepWaitLimit(h, -1) == 3
        assert lib.curfpSetSweepWaitLimit(h, 1000) == 0
        assert lib.curfpSynchronize(h) == 0
        assert lib.curfpSynchronize(None) == 1
        assert lib.curfpSetSweepWaitLimit(None, 5) == 1
    finally:
        assert lib.curfpDestroy(h) == 0
