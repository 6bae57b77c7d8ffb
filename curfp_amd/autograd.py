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


def _factor(A: torch.Tensor, lam: float, alpha: float, transr: str, uplo: str, math: str) -> torch.Tensor:
    n = A.shape[0]
    f64 = A.dtype == torch.float64
    C = torch.empty(n * (n + 1) // 2, dtype=A.dtype, device=A.device)
    if f64:
        _c.dsfrk(A, C, alpha=alpha, transr=transr, uplo=uplo, trans="T")
    elif math == "bf16":
        # BF16 MFMA with FP32 accumulation; exact when A's entries are BF16 values
        _c.ssfrk(A.to(torch.bfloat16), C, alpha=alpha, transr=transr, uplo=uplo, trans="T")
    else:
        with _c.math_mode(math):
            _c.ssfrk(A, C, alpha=alpha, transr=transr, uplo=uplo, trans="T")
    _c.add_to_diagonal(C, lam, transr=transr, uplo=uplo)
    if f64:
        _c.dpftrf(C, transr=transr, uplo=uplo, check=True)
    else:
        with _c.math_mode(math):
            _c.spftrf(C, transr=transr, uplo=uplo, check=True)
    return C


def _solve(C: torch.Tensor, B: torch.Tensor, A: torch.Tensor, lam: float, alpha: float, transr: str, uplo: str,
           refine_tol, max_iter: int, strict: bool = True):
    """X = C^{-1} B with the factor C; with refine_tol, FP64 refinement
    against the generator.  Returns X, or None when the refinement misses
    refine_tol and strict is False."""
    if refine_tol is None:
        X = B.to(C.dtype).contiguous().clone()
        (_c.dpftrs if C.dtype == torch.float64 else _c.spftrs)(C, X, transr=transr, uplo=uplo)
        return X
    from .refine import refine_solve

    X, info = refine_solve(C, A, lam, B, alpha=alpha, tol=refine_tol, max_iter=max_iter, transr=transr, uplo=uplo)
    if not info["rel_residual"] <= refine_tol:
        if not strict:
            return None
        raise RuntimeError(f"iterative refinement did not reach {refine_tol:g}: {info}")
    return X


def _factor_solve(A, lam, alpha, transr, uplo, math, Y, refine_tol, max_iter):
    """Factor and solve with the precision policy of the BF16 mode: a BF16
    factor that is not positive definite, or whose refined solve (classical
    steps, then flexible GMRES) misses refine_tol, is replaced by the
    split-BF16 (FP32-accurate) factor.  Returns (factor, X, math used)."""
    try:
        C = _factor(A, lam, alpha, transr, uplo, math)
    except torch.linalg.LinAlgError:
        if math != "bf16":
            raise
        math = "bf16x3"
        C = _factor(A, lam, alpha, transr, uplo, math)
    X = _solve(C, Y, A, lam, alpha, transr, uplo, refine_tol, max_iter, strict=math != "bf16")
    if X is None:
        del C
        math = "bf16x3"
        C = _factor(A, lam, alpha, transr, uplo, math)
        X = _solve(C, Y, A, lam, alpha, transr, uplo, refine_tol, max_iter)
    return C, X, math


class KRRSolve(torch.autograd.Function):
    """Autograd function behind :func:`krr_solve`."""

    @staticmethod
    def forward(ctx, A, lam, Y, alpha, transr, uplo, recompute, math, refine_tol, max_iter):
        A = A.contiguous()
        lam_f = float(lam)
        C, X, math = _factor_solve(A, lam_f, alpha, transr, uplo, math, Y, refine_tol, max_iter)
        ctx.alpha, ctx.lam, ctx.transr, ctx.uplo = alpha, lam_f, transr, uplo
        ctx.math, ctx.refine_tol, ctx.max_iter = math, refine_tol, max_iter
        ctx.lam_is_tensor = torch.is_tensor(lam)
        if recompute:
            del C
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
            C = _factor(A, ctx.lam, ctx.alpha, ctx.transr, ctx.uplo, ctx.math)
        squeeze = X.ndim == 1
        X2 = X.unsqueeze(1) if squeeze else X
        Xb = Xbar.unsqueeze(1) if squeeze else Xbar
        # adjoint solve (C is symmetric): Ybar = C^{-1} Xbar
        Ybar = _solve(C, Xb, A, ctx.lam, ctx.alpha, ctx.transr, ctx.uplo, ctx.refine_tol, ctx.max_iter)
        del C
        Abar = lam_bar = None
        if ctx.needs_input_grad[0]:
            Ag = A if A.dtype == X2.dtype else A.to(X2.dtype)
            At = Ag.t()
            XtA = _c.gemm_nt(X2.t(), At)          # (nrhs, k) = X^T A
            YtA = _c.gemm_nt(Ybar.t(), At)        # (nrhs, k) = Ybar^T A
            Abar = _c.gemm_nt(Ybar, XtA.t(), alpha=-ctx.alpha)
            Abar = _c.gemm_nt(X2, YtA.t(), Abar, alpha=-ctx.alpha, beta=1.0).to(A.dtype)
        if ctx.needs_input_grad[1] and ctx.lam_is_tensor:
            lam_bar = -(Ybar * X2).sum()
        Ybar_out = Ybar[:, 0] if squeeze else Ybar
        return Abar, lam_bar, Ybar_out, None, None, None, None, None, None, None


def krr_solve(A: torch.Tensor, lam, Y: torch.Tensor, *, alpha: float = 1.0, transr: str = "T",
              uplo: str = "U", recompute: bool = False, math: str = "fp32", refine_tol: float = None,
              max_iter: int = 30) -> torch.Tensor:
    """X solving (alpha*A@A.T + lam*I) X = Y, differentiable in A, lam and Y.

    A: (n, k) float32 or float64 device tensor (row-major, the ssfrk default
    trans='T').  lam: float or 0-d tensor.  Y: (n, nrhs) or (n,).
    math: arithmetic of ssfrk and the spftrf updates for a float32 A
    ('fp32', 'bf16': A rounded to BF16, or 'bf16x3'); a 'bf16' factor that
    fails (not positive definite, or refinement short of refine_tol) is
    redone in 'bf16x3', which the backward then reuses.  refine_tol: when set,
    both the forward solve and the adjoint solve use FP64 iterative
    refinement against the FP32 generator down to that relative residual, and
    X (and the gradients w.r.t. Y) are float64.
    recompute: release the factor after the forward and recompute it in the
    backward.
    """
    if A.dtype not in (torch.float32, torch.float64):
        raise ValueError("A must be float32 or float64")
    if A.dtype == torch.float64 and (math != "fp32" or refine_tol is not None):
        raise ValueError("BF16 math and refinement apply to a float32 A")
    if Y.shape[0] != A.shape[0]:
        raise ValueError(f"Y has {Y.shape[0]} rows, A has {A.shape[0]}")
    if refine_tol is None and Y.dtype != A.dtype:
        raise ValueError("Y must have A's dtype")
    return KRRSolve.apply(A, lam, Y, float(alpha), transr, uplo, bool(recompute), math, refine_tol, int(max_iter))
