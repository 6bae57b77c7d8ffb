"""GPU tests of the differentiable chain (curfp_amd.autograd.krr_solve):
kernel-ridge loss L = sum(W * X), (A A^T + lam I) X = Y, gradients w.r.t.
A, lam and Y checked against FP64 central finite differences (step 1e-6),
against the closed form, and against dense torch.linalg.cholesky /
cholesky_solve autograd (float64)."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ag(cuda):
    from curfp_amd import autograd

    return autograd


def problem(n, k, nrhs, seed, dtype=torch.float64):
    g = torch.Generator().manual_seed(seed)
    A = (torch.randn(n, k, generator=g, dtype=torch.float64) / math.sqrt(k)).to(dtype).cuda()
    Y = torch.randn(n, nrhs, generator=g, dtype=torch.float64).to(dtype).cuda()
    W = torch.randn(n, nrhs, generator=g, dtype=torch.float64).to(dtype).cuda()
    return A, Y, W


@pytest.mark.parametrize("t,u", [("T", "U"), ("N", "L"), ("T", "L"), ("N", "U")])
def test_finite_differences_fp64(ag, t, u):
    """n=512, k=64 (the differentiation workload's FD size), FP64, step 1e-6."""
    n, k, nrhs, lam0 = 512, 64, 16, 0.1
    A, Y, W = problem(n, k, nrhs, seed=3)
    A.requires_grad_(True)
    Y.requires_grad_(True)
    lam = torch.tensor(lam0, dtype=torch.float64, device="cuda", requires_grad=True)
    L = (W * ag.krr_solve(A, lam, Y, transr=t, uplo=u)).sum()
    gA, glam, gY = torch.autograd.grad(L, (A, lam, Y))

    def loss(Av, lv, Yv):
        with torch.no_grad():
            return float((W * ag.krr_solve(Av, lv, Yv, transr=t, uplo=u)).sum())

    h = 1e-6
    rng = np.random.default_rng(0)
    for _ in range(6):
        i, p = int(rng.integers(n)), int(rng.integers(k))
        Ap, Am = A.detach().clone(), A.detach().clone()
        Ap[i, p] += h
        Am[i, p] -= h
        fd = (loss(Ap, lam0, Y.detach()) - loss(Am, lam0, Y.detach())) / (2 * h)
        assert abs(fd - gA[i, p].item()) <= 1e-5 * max(1.0, abs(fd)), (i, p, fd, gA[i, p].item())
    fd = (loss(A.detach(), lam0 + h, Y.detach()) - loss(A.detach(), lam0 - h, Y.detach())) / (2 * h)
    assert abs(fd - glam.item()) <= 1e-5 * max(1.0, abs(fd))
    for _ in range(4):
        i, c = int(rng.integers(n)), int(rng.integers(nrhs))
        Yp, Ym = Y.detach().clone(), Y.detach().clone()
        Yp[i, c] += h
        Ym[i, c] -= h
        fd = (loss(A.detach(), lam0, Yp) - loss(A.detach(), lam0, Ym)) / (2 * h)
        assert abs(fd - gY[i, c].item()) <= 1e-5 * max(1.0, abs(fd))


@pytest.mark.parametrize("n,dtype", [(8192, torch.float64), (4096, torch.float32)])
def test_against_dense_torch_autograd(ag, n, dtype):
    """Dense reference: torch.linalg.cholesky + cholesky_solve autograd in float64."""
    k, nrhs, lam0 = 512, 16, 0.1
    A, Y, W = problem(n, k, nrhs, seed=4, dtype=dtype)
    A.requires_grad_(True)
    Y.requires_grad_(True)
    lam = torch.tensor(lam0, dtype=dtype, device="cuda", requires_grad=True)
    L = (W * ag.krr_solve(A, lam, Y)).sum()
    gA, glam, gY = torch.autograd.grad(L, (A, lam, Y))

    A64 = A.detach().double().requires_grad_(True)
    Y64 = Y.detach().double().requires_grad_(True)
    lam64 = torch.tensor(lam0, dtype=torch.float64, device="cuda", requires_grad=True)
    C = A64 @ A64.t() + lam64 * torch.eye(n, dtype=torch.float64, device="cuda")
    Lc = torch.linalg.cholesky(C)
    X = torch.cholesky_solve(Y64, Lc)
    Lref = (W.double() * X).sum()
    rA, rlam, rY = torch.autograd.grad(Lref, (A64, lam64, Y64))
    tol = 1e-9 if dtype == torch.float64 else 2e-3
    assert (gA.double() - rA).norm() <= tol * rA.norm()
    assert abs(glam.double().item() - rlam.item()) <= tol * max(1.0, abs(rlam.item()))
    assert (gY.double() - rY).norm() <= tol * rY.norm()


def test_recompute_matches_stored_factor(ag):
    n, k, nrhs = 1500, 100, 8
    A, Y, W = problem(n, k, nrhs, seed=5, dtype=torch.float32)
    grads = []
    for rec in (False, True):
        Ar = A.clone().requires_grad_(True)
        Yr = Y.clone().requires_grad_(True)
        L = (W * ag.krr_solve(Ar, 0.5, Yr, recompute=rec)).sum()
        grads.append(torch.autograd.grad(L, (Ar, Yr)))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])


def test_gemm_nt_matches_fp64(cuda):
    import curfp_amd as curfp

    for dt in (torch.float32, torch.float64):
        A = torch.randn(300, 70, dtype=dt, device="cuda")
        B = torch.randn(50, 70, dtype=dt, device="cuda")
        C = curfp.gemm_nt(A, B)
        ref = A.double().cpu() @ B.double().cpu().t()
        tol = 1e-5 if dt == torch.float32 else 1e-13
        assert (C.double().cpu() - ref).abs().max() <= tol * ref.abs().max()
        # transposed views (unit stride along rows)
        C2 = curfp.gemm_nt(A.t(), A.t())
        ref2 = A.double().cpu().t() @ A.double().cpu()
        assert (C2.double().cpu() - ref2).abs().max() <= tol * ref2.abs().max() * 10
