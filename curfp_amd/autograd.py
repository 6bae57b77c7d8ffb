"""Differentiable RFP chain: X = (alpha*A@A.T + lam*I)^{-1} Y.

``krr_solve(A, lam, Y)`` runs ssfrk -> add_to_diagonal -> spftrf -> spftrs
(FP32 or FP64, by A's dtype) and carries a hand-written backward:

* adjoint solve  Ybar = C^{-1} Xbar   — spftrs with the same RFP factor;
* Cholesky adjoint.  For the solve X = L^{-T} L^{-1} Y the factor's adjoint
  is Lbar = -(Ybar X^T + X Ybar^T) L, and the Cholesky backward
  C̄ = L^{-T} Φ(L^T Lbar) L^{-1} (Φ: lower triangle with halved diagonal),
  symmetrised, collapses to  C̄ = -sym(Ybar X^T) — a symmetric matrix of rank
  at most 2*nrhs.  It is therefore never formed (neither densely nor in RFP);
* the symmetric multiply  Abar = alpha (C̄ + C̄^T) A = -alpha (Ybar (X^T A) +
  X (Ybar^T A)) is four thin GEMMs on the library's MFMA GEMM
  (curfpSgemmStrided / curfpDgemmStrided), and lambar = tr(C̄) = -<Ybar, X>.

The factor computed by the forward is kept for the backward (it is the
forward's own RFP buffer, not a second packed matrix); with
``recompute=True`` it is released after the forward and recomputed in the
backward instead, for problems where one packed matrix is all the memory
there is.
"""
from __future__ import annotations

import torch

import curfp_amd as _c


def _factor(A: torch.Tensor, lam: float, alpha: float, transr: str, uplo: str) -> torch.Tensor:
    n = A.shape[0]
    f64 = A.dtype == torch.float64
    C = torch.empty(n * (n + 1) // 2, dtype=A.dtype, device=A.device)
    (_c.dsfrk if f64 else _c.ssfrk)(A, C, alpha=alpha, transr=transr, uplo=uplo, trans="T")
    _c.add_to_diagonal(C, lam, transr=transr, uplo=uplo)
    (_c.dpftrf if f64 else _c.spftrf)(C, transr=transr, uplo=uplo, check=True)
    return C


def _solve(C: torch.Tensor, B: torch.Tensor, transr: str, uplo: str) -> torch.Tensor:
    X = B.contiguous().clone()
    (_c.dpftrs if C.dtype == torch.float64 else _c.spftrs)(C, X, transr=transr, uplo=uplo)
    return X


class KRRSolve(torch.autograd.Function):
    """Autograd function behind :func:`krr_solve`."""

    @staticmethod
    def forward(ctx, A, lam, Y, alpha, transr, uplo, recompute):
        A = A.contiguous()
        lam_f = float(lam)
        C = _factor(A, lam_f, alpha, transr, uplo)
        X = _solve(C, Y, transr, uplo)
        ctx.alpha, ctx.lam, ctx.transr, ctx.uplo = alpha, lam_f, transr, uplo
        ctx.lam_is_tensor = torch.is_tensor(lam)
        if recompute:
            ctx.save_for_backward(A, X)
            ctx.has_factor = False
        else:
            ctx.save_for_backward(A, X, C)
            ctx.has_factor = True
        return X

    @staticmethod
    def backward(ctx, Xbar):
        if ctx.has_factor:
            A, X, C = ctx.saved_tensors
        else:
            A, X = ctx.saved_tensors
            C = _factor(A, ctx.lam, ctx.alpha, ctx.transr, ctx.uplo)
        squeeze = X.ndim == 1
        X2 = X.unsqueeze(1) if squeeze else X
        Xb = Xbar.unsqueeze(1) if squeeze else Xbar
        # adjoint solve (C is symmetric): Ybar = C^{-1} Xbar
        Ybar = _solve(C, Xb.to(X.dtype), ctx.transr, ctx.uplo)
        Abar = lam_bar = None
        if ctx.needs_input_grad[0]:
            At = A.t()
            XtA = _c.gemm_nt(X2.t(), At)          # (nrhs, k) = X^T A
            YtA = _c.gemm_nt(Ybar.t(), At)        # (nrhs, k) = Ybar^T A
            Abar = _c.gemm_nt(Ybar, XtA.t(), alpha=-ctx.alpha)
            Abar = _c.gemm_nt(X2, YtA.t(), Abar, alpha=-ctx.alpha, beta=1.0)
        if ctx.needs_input_grad[1]:
            lam_bar = -(Ybar * X2).sum()
            if not ctx.lam_is_tensor:
                lam_bar = None
        Ybar_out = Ybar[:, 0] if squeeze else Ybar
        return Abar, lam_bar, Ybar_out, None, None, None, None


def krr_solve(A: torch.Tensor, lam, Y: torch.Tensor, *, alpha: float = 1.0, transr: str = "T",
              uplo: str = "U", recompute: bool = False) -> torch.Tensor:
    """X solving (alpha*A@A.T + lam*I) X = Y, differentiable in A, lam and Y.

    A: (n, k) float32 or float64 device tensor (row-major, the ssfrk default
    trans='T').  lam: float or 0-d tensor.  Y: (n, nrhs) or (n,), same dtype.
    """
    if A.dtype not in (torch.float32, torch.float64) or Y.dtype != A.dtype:
        raise ValueError("A and Y must share a float32 or float64 dtype")
    if Y.shape[0] != A.shape[0]:
        raise ValueError(f"Y has {Y.shape[0]} rows, A has {A.shape[0]}")
    return KRRSolve.apply(A, lam, Y, float(alpha), transr, uplo, bool(recompute))
