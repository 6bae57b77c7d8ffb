"""Mixed-precision solve of (alpha*A@A.T + lam*I) X = B with FP64 refinement.

The factor comes from spftrf run in any math mode (FP32, BF16, split-BF16);
the refinement loop (curfpSpftrsRefine in include/curfp.h) forms the
residual B - C X in FP64 from the FP32 generator A on the GPU, solves for the
correction with the FP32 factor and accumulates it in FP64, until
||B - C X||_F <= tol * ||C||_F * ||X||_F.
"""
from __future__ import annotations

import torch

from . import (_get_handle, _op, _fill, _validate, _n_from_rfp, _spftrs_refine_C, _spftrs_C, _dsfmm_C, _dlansf_C,
               NORM_FRO, SIDE_LEFT)


def refine_solve(F: torch.Tensor, A: torch.Tensor, lam: float, B: torch.Tensor, *, alpha: float = 1.0,
                 tol: float = 1e-6, max_iter: int = 30, transr: str = "T", uplo: str = "U", trans: str = "T"):
    """Solve C X = B, C = alpha*op(A)@op(A).T + lam*I, from the RFP factor F.

    A: float32 generator, row-major (n, k) for trans='T' (the ssfrk default).
    B: (n, nrhs) or (n,) tensor of any float dtype.
    Returns (X, info) with X float64 (n, nrhs) and info a dict holding the
    number of FP32 solves and the final relative residual.
    """
    _validate(F, "F")
    _validate(A, "A")
    n = _n_from_rfp(F.numel())
    t = trans.upper()
    k = A.shape[1] if t == "T" else A.shape[0]
    lda = A.shape[1] if t == "T" else A.shape[0]
    if (A.shape[0] if t == "T" else A.shape[1]) != n:
        raise ValueError(f"generator A has shape {tuple(A.shape)}, expected n={n} rows")
    squeeze = B.ndim == 1
    B2 = B.unsqueeze(1) if squeeze else B
    if B2.shape[0] != n:
        raise ValueError(f"B has {B2.shape[0]} rows but n={n}")
    nrhs = B2.shape[1]
    # column-major FP64 n x nrhs: the transpose of a contiguous (nrhs, n) tensor
    Bc = B2.t().to(torch.float64).contiguous()
    Xc = torch.empty_like(Bc)
    iters, resid = _spftrs_refine_C(_get_handle(), _op(transr), _fill(uplo), _op(trans), n, k, nrhs, float(alpha),
                                    A.data_ptr(), lda, float(lam), F.data_ptr(), Bc.data_ptr(), n,
                                    Xc.data_ptr(), n, float(tol), int(max_iter))
    X = Xc.t()
    if squeeze:
        X = X[:, 0]
    return X, {"solves": iters, "rel_residual": resid}


def refine_solve_rfp(F: torch.Tensor, C: torch.Tensor, B: torch.Tensor, *, tol: float = 1e-6, max_iter: int = 30,
                     transr: str = "T", uplo: str = "U"):
    """Solve C X = B for a general SPD matrix given as an FP64 RFP array C,
    from its FP32 RFP factor F (spftrf of an FP32 copy of C, any math mode).

    Each step solves for the correction with F (FP32 substitution sweeps)
    and forms the residual B - C X with the FP64 symmetric multiply dsfmm on
    C, until ||B - C X||_F <= tol * ||C||_F * ||X||_F, max_iter solves, or a
    step that does not reduce the residual (whose iterate is then dropped).  B: (n, nrhs) or (n,).  Returns
    (X, info) with X float64 like B and info holding the number of FP32
    solves and the final relative residual.
    """
    _validate(F, "F")
    _validate(C, "C", torch.float64)
    n = _n_from_rfp(F.numel())
    if C.numel() != F.numel():
        raise ValueError("C and F must hold the same RFP size")
    squeeze = B.ndim == 1
    B2 = B.unsqueeze(1) if squeeze else B
    if B2.shape[0] != n:
        raise ValueError(f"B has {B2.shape[0]} rows but n={n}")
    nrhs = B2.shape[1]
    h, t, u = _get_handle(), _op(transr), _fill(uplo)
    # column-major n x nrhs blocks: contiguous (nrhs, n) tensors, leading dimension n
    Bc = B2.t().to(torch.float64).contiguous()
    X = torch.zeros_like(Bc)
    R = Bc.clone()
    normC = float(_dlansf_C(h, NORM_FRO, t, u, n, C.data_ptr()))
    solves, resid = 0, float("inf")
    for _ in range(max(1, max_iter)):
        D = R.to(torch.float32)
        _spftrs_C(h, t, u, n, nrhs, F.data_ptr(), D.data_ptr(), n)
        Xn = X + D.to(torch.float64)
        Rn = Bc.clone()
        _dsfmm_C(h, t, u, SIDE_LEFT, n, nrhs, -1.0, C.data_ptr(), Xn.data_ptr(), n, 1.0, Rn.data_ptr(), n)
        solves += 1
        nx = float(torch.linalg.vector_norm(Xn))
        rn = float(torch.linalg.vector_norm(Rn)) / (normC * nx) if nx > 0 else 0.0
        if rn >= resid:          # no progress: keep the previous iterate
            break
        X, R, resid = Xn, Rn, rn
        if resid <= tol:
            break
    Xo = X.t()
    if squeeze:
        Xo = Xo[:, 0]
    return Xo, {"solves": solves, "rel_residual": resid}
