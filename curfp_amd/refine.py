"""Mixed-precision solve of (alpha*A@A.T + lam*I) X = B with FP64 refinement.

The factor comes from spftrf run in any math mode (FP32, BF16, split-BF16);
the refinement loop (curfpSpftrsRefine in include/curfp.h) forms the
residual B - C X in FP64 from the FP32 generator A on the GPU, solves for the
correction with the FP32 factor and accumulates it in FP64, until
||B - C X||_F <= tol * ||C||_F * ||X||_F.
"""
from __future__ import annotations

import torch

from . import _get_handle, _op, _fill, _validate, _n_from_rfp, _spftrs_refine_C


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
