"""
CPU oracle for the RFP hot path — TEST INFRASTRUCTURE ONLY.

This module restates, on the CPU with numpy / scipy BLAS+LAPACK, the algorithm
of the reference's RFP routines so that the HIP kernels can be checked
against it.  Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may import it, and only as the checker: the product path
(``curfp_amd``) never calls it.

Geometry.  LAPACK's RFP layout (the reference's tables
src/curfp_ssfrk.cpp:52-133, src/curfp_spftrf.cpp:62-180,
src/curfp_spftrs.cpp:60-213; LAWN 199) is restated as three strided views of
the lower factor L (C = L L^T; for uplo = 'U', L = U^T):

    v11 : n1 x n1 lower block, v21 : n2 x n1, v22 : n2 x n2 lower block,
    element (i, j) of a block at  arf[off + i*rs + j*cs].

Functions and what pins each one (see tests/test_oracle.py):
  views / lower_index_map / trttf / tfttr / diag_indices
      — bit-exact against LAPACK ?trttf / ?tfttr golden vectors (all 8
        variants, n <= 257), the reference's hand layouts (n=2 case 5,
        n=3 case 1, n=6 LAPACK doc example), bijection and round trip.
  add_to_diagonal  — closed form fl(c + fl(value)) at the diagonal indices;
        LAPACK ssfrk(A=I, beta=1) golden vectors.
  sfrk   — reference src/curfp_ssfrk.cpp:135-212 (SYRK v11, GEMM v21, SYRK
        v22); pinned by LAPACK ?sfrk golden vectors, exact small-integer
        products, and the quick-return known answers of
        python_tests/test_ssfrk.py:12-111.
  pftrf  — reference src/curfp_spftrf.cpp:182-292 (POTRF v11, TRSM v21,
        SYRK v22, POTRF v22, info offset n1); pinned by LAPACK ?pftrf golden
        vectors, the hand Cholesky known answers of
        python_tests/test_spftrf.py:19-128, reconstruction L L^T = C, and the
        LAPACK info of crafted non-SPD matrices.
  pftrs  — reference src/curfp_spftrs.cpp:215-302 (4 TRSM + 2 GEMM); pinned
        by LAPACK ?pftrs golden vectors and the dense residual B - C X.
  sfmm   — reference src/curfp_ssfmm.cpp:107-306 (side L: alpha S B + beta C,
        side R: alpha B S + beta C with S the symmetric matrix); pinned against
        the dense product of an independently built symmetric matrix, the
        known answer of S = I, alpha = 0 / beta = 0 cases, and a planted
        one-triangle mistake.
  sfr2k  — reference src/curfp_ssfr2k.cpp:121-231 (per view: SYR2K on v11
        and v22, two GEMMs on v21); sfr / sfr2 — src/curfp_ssfr.cpp:99-172,
        src/curfp_ssfr2.cpp:94-189 (the k = 1, beta = 1 cases).  Pinned by
        sfr2k(A, A) = 2 sfrk(A) and the polarisation identity
        x y^T + y x^T = ((x+y)(x+y)^T - (x-y)(x-y)^T) / 2 against the
        LAPACK-pinned sfrk, by the dense product packed with trttf, the quick
        returns of python_tests/test_ssfr2k.py:81-102, and a planted
        halved-diagonal-block mistake.
  pftri  — reference src/curfp_spftri.cu:159-388 (inverse of the triangular
        factor, then L^-T L^-1); pinned against LAPACK ?pftri (scipy), an
        independent RFP routine, and C inv(C) = I.
  pfcon  — reference src/curfp_spfcon.cu:84-271 (LAPACK ?lacn2 estimate
        of ||A^-1||_1 with A^-1 = A^-T); pinned against LAPACK ?pocon
        (scipy) on the same factor, the exact 1-norm of the inverse (the
        estimate never exceeds it and equals it for diagonal matrices), the
        n = 1 known answer of python_tests/test_spfcon.py:107-117, and a
        planted forward-solve-only mistake.
  lansf  — reference src/curfp_slansf.cu:273-435 (max |a|, largest absolute
        column sum, Frobenius norm of the symmetric matrix); pinned by
        LAPACK ?lange on the dense matrix (an independent routine), the
        closed forms of c J + d I and of the identity, and the n = 1 known
        answer of python_tests/test_slansf.py:124-133.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import blas as _blas
from scipy.linalg import lapack as _lapack

__all__ = [
    "views", "lower_index_map", "trttf", "tfttr", "diag_indices", "add_to_diagonal",
    "rfp_to_dense_sym", "dense_sym_to_rfp", "sfrk", "sfr2k", "sfr", "sfr2", "pftrf", "pftrs", "pftri", "pfcon", "lansf", "sfmm", "chain_flops",
]


def views(n: int, transr: str = "T", uplo: str = "U"):
    """(n1, n2, ld, v11, v21, v22) with v = (off, rs, cs).

    n1/n2 split: src/curfp_ssfrk.cpp:176-182; offsets and leading dimensions:
    src/curfp_spftrf.cpp:62-180 (fill11/off11, trsm_b_off, off22 per case)."""
    lower = uplo.upper() == "L"
    tn = transr.upper() == "N"
    if n % 2:
        n1, n2 = (n - n // 2, n // 2) if lower else (n // 2, n - n // 2)
        if tn:
            ld = n
            o11, o21, o22 = (0, n1, n) if lower else (n2, 0, n1)
        elif lower:
            ld, o11, o21, o22 = n1, 0, n1 * n1, 1
        else:
            ld, o11, o21, o22 = n2, n2 * n2, 0, n1 * n2
    else:
        k = n // 2
        n1 = n2 = k
        if tn:
            ld = n + 1
            o11, o21, o22 = (1, k + 1, 0) if lower else (k + 1, 0, k)
        else:
            ld = k
            o11, o21, o22 = (k, k * (k + 1), 0) if lower else (k * (k + 1), 0, k * k)
    # A block stored with fill 'U' (or a TRSM that solves the transposed
    # system) holds the transpose of the lower factor's block: row view.
    row11, row22, row21 = (not tn), tn, (tn != lower)

    def v(o, row):
        return (o, ld, 1) if row else (o, 1, ld)

    return n1, n2, ld, v(o11, row11), v(o21, row21), v(o22, row22)


def _block_index(v, rows, cols):
    off, rs, cs = v
    i = np.arange(rows, dtype=np.int64)[:, None]
    j = np.arange(cols, dtype=np.int64)[None, :]
    return off + i * rs + j * cs


def lower_index_map(n: int, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """(n, n) int64: arf index of logical lower element (i, j), -1 above."""
    n1, n2, _, v11, v21, v22 = views(n, transr, uplo)
    M = np.full((n, n), -1, dtype=np.int64)
    M[:n1, :n1] = _block_index(v11, n1, n1)
    M[n1:, :n1] = _block_index(v21, n2, n1)
    M[n1:, n1:] = _block_index(v22, n2, n2)
    M[np.triu_indices(n, 1)] = -1
    return M


def _tri(n):
    return np.tril_indices(n)


def trttf(A: np.ndarray, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """Dense triangle -> RFP (LAPACK ?trttf; reference src/curfp_strttf.cu:46-415).

    A is indexed logically (A[i, j]); uplo = 'L' reads the lower triangle,
    'U' the upper."""
    n = A.shape[0]
    M = lower_index_map(n, transr, uplo)
    i, j = _tri(n)
    arf = np.empty(n * (n + 1) // 2, dtype=A.dtype)
    arf[M[i, j]] = A[i, j] if uplo.upper() == "L" else A[j, i]
    return arf


def tfttr(arf: np.ndarray, n: int, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """RFP -> dense triangle, other triangle zero (reference src/curfp_stfttr.cu:24-128)."""
    M = lower_index_map(n, transr, uplo)
    i, j = _tri(n)
    A = np.zeros((n, n), dtype=arf.dtype)
    if uplo.upper() == "L":
        A[i, j] = arf[M[i, j]]
    else:
        A[j, i] = arf[M[i, j]]
    return A


def diag_indices(n: int, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """Flat RFP indices of the diagonal (reference curfp/__init__.py:194-269)."""
    n1, n2, ld, v11, _, v22 = views(n, transr, uplo)
    return np.concatenate([v11[0] + np.arange(n1, dtype=np.int64) * (ld + 1),
                           v22[0] + np.arange(n2, dtype=np.int64) * (ld + 1)])


def add_to_diagonal(arf: np.ndarray, value: float, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """arf[diag] += value in arf's dtype (reference curfp/__init__.py:272-296)."""
    n = int((np.sqrt(8 * arf.size + 1) - 1) // 2)
    out = arf.copy()
    d = diag_indices(n, transr, uplo)
    out[d] = out[d] + arf.dtype.type(value)
    return out


def rfp_to_dense_sym(arf: np.ndarray, n: int, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """Full symmetric dense matrix from an (unfactored) RFP array."""
    M = lower_index_map(n, transr, uplo)
    i, j = _tri(n)
    S = np.zeros((n, n), dtype=arf.dtype)
    S[i, j] = arf[M[i, j]]
    S[j, i] = arf[M[i, j]]
    return S


def rfp_lower_factor(arf: np.ndarray, n: int, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """Dense lower factor L (C = L L^T) from a factored RFP array."""
    M = lower_index_map(n, transr, uplo)
    i, j = _tri(n)
    L = np.zeros((n, n), dtype=arf.dtype)
    L[i, j] = arf[M[i, j]]
    return L


def dense_sym_to_rfp(S: np.ndarray, transr: str = "T", uplo: str = "U") -> np.ndarray:
    n = S.shape[0]
    M = lower_index_map(n, transr, uplo)
    i, j = _tri(n)
    arf = np.empty(n * (n + 1) // 2, dtype=S.dtype)
    arf[M[i, j]] = S[i, j]
    return arf


# --- block access -----------------------------------------------------------

def _view(arf, v, rows, cols):
    """Strided numpy view of a block (no copy)."""
    off, rs, cs = v
    es = arf.itemsize
    return np.lib.stride_tricks.as_strided(arf[off:], shape=(rows, cols), strides=(rs * es, cs * es))


def _get(arf, v, rows, cols):
    return np.array(_view(arf, v, rows, cols), copy=True)


def _put(arf, v, rows, cols, X, lower=False):
    W = _view(arf, v, rows, cols)
    if not lower:
        W[...] = X
        return
    step = 1024
    for r0 in range(0, rows, step):
        r1 = min(rows, r0 + step)
        W[r0:r1, :r0] = X[r0:r1, :r0]
        blk = W[r0:r1, r0:r1]
        ii, jj = np.tril_indices(r1 - r0)
        blk[ii, jj] = X[r0:r1, r0:r1][ii, jj]


def _op_rows(A: np.ndarray, trans: str, n: int, k: int) -> np.ndarray:
    """op(A) as an n x k array: trans='N' -> A (n x k col-major buffer given as
    its logical (n, k) numpy array); trans='T' -> A is k x n (row-major (n, k))."""
    return A if trans.upper() == "N" else A.T


def sfrk(n, k, alpha, A, beta, C, transr="T", uplo="U", trans="N", exact=True):
    """C := alpha op(A) op(A)^T + beta C in RFP (reference src/curfp_ssfrk.cpp:135-212).

    ``A`` is the logical matrix LAPACK sees: (n, k) for trans='N', (k, n) for
    trans='T'.  With exact=True products are accumulated in float64 and
    rounded once to C's dtype; exact=False multiplies in C's dtype (BLAS
    ?gemm, the arithmetic LAPACK ?sfrk itself uses — the CPU timing mode).
    beta == 0 never reads C; alpha == 0 never reads A (quick returns of
    :154-161)."""
    C = np.array(C, copy=True)
    if n == 0:
        return C
    if (alpha == 0 or k == 0) and beta == 1:
        return C
    if alpha == 0 and beta == 0:
        C[:] = 0
        return C
    n1, n2, _, v11, v21, v22 = views(n, transr, uplo)
    acc = np.float64 if exact else C.dtype
    if alpha == 0 or k == 0:
        G = np.zeros((n, max(k, 1)), dtype=acc)
    else:
        G = np.asarray(A, dtype=acc)
        G = G if trans.upper() == "N" else G.T
    A1, A2 = G[:n1], G[n1:]
    for v, X, Y, r, c, lo in ((v11, A1, A1, n1, n1, True), (v21, A2, A1, n2, n1, False),
                              (v22, A2, A2, n2, n2, True)):
        if r == 0 or c == 0:
            continue
        P = alpha * (X @ Y.T)
        if beta != 0:
            P = P + beta * _get(C, v, r, c).astype(acc)
        _put(C, v, r, c, P.astype(C.dtype), lower=lo)
    return C


def sfr2k(n, k, alpha, A, B, beta, C, transr="T", uplo="U", trans="N"):
    """C := alpha op(A) op(B)^T + alpha op(B) op(A)^T + beta C in RFP
    (reference src/curfp_ssfr2k.cpp:121-231), float64 products rounded once
    to C's dtype.  ``A``, ``B`` are the logical matrices: (n, k) for
    trans='N', (k, n) for trans='T'.  Quick returns as sfrk (:137-148)."""
    C = np.array(C, copy=True)
    if n == 0 or ((alpha == 0 or k == 0) and beta == 1):
        return C
    if (alpha == 0 or k == 0) and beta == 0:
        C[:] = 0
        return C
    n1, n2, _, v11, v21, v22 = views(n, transr, uplo)
    if alpha == 0 or k == 0:
        X = Y = np.zeros((n, 1))
    else:
        X = np.asarray(A, dtype=np.float64)
        Y = np.asarray(B, dtype=np.float64)
        if trans.upper() != "N":
            X, Y = X.T, Y.T
    X1, X2, Y1, Y2 = X[:n1], X[n1:], Y[:n1], Y[n1:]
    for v, Xr, Xc, Yr, Yc, r, c, lo in ((v11, X1, X1, Y1, Y1, n1, n1, True), (v21, X2, X1, Y2, Y1, n2, n1, False),
                                        (v22, X2, X2, Y2, Y2, n2, n2, True)):
        if r == 0 or c == 0:
            continue
        P = alpha * (Xr @ Yc.T + Yr @ Xc.T)
        if beta != 0:
            P = P + beta * _get(C, v, r, c).astype(np.float64)
        _put(C, v, r, c, P.astype(C.dtype), lower=lo)
    return C


def sfr(n, alpha, x, C, transr="T", uplo="U"):
    """C := alpha x x^T + C (reference src/curfp_ssfr.cpp:99-172)."""
    x = np.asarray(x, dtype=np.float64).reshape(n, 1)
    return sfr2k(n, 1, alpha / 2.0, x, x, 1.0, C, transr, uplo, "N")


def sfr2(n, alpha, x, y, C, transr="T", uplo="U"):
    """C := alpha (x y^T + y x^T) + C (reference src/curfp_ssfr2.cpp:94-189)."""
    return sfr2k(n, 1, alpha, np.asarray(x, dtype=np.float64).reshape(n, 1),
                 np.asarray(y, dtype=np.float64).reshape(n, 1), 1.0, C, transr, uplo, "N")


def _potrf_lower(T: np.ndarray):
    """LAPACK ?potrf('L') on a dense block; returns (L, info)."""
    fn = _lapack.spotrf if T.dtype == np.float32 else _lapack.dpotrf
    L, info = fn(np.asfortranarray(T), lower=1, clean=1)
    return np.tril(L), int(info)


def pftrf(n, arf, transr="T", uplo="U"):
    """RFP Cholesky (reference src/curfp_spftrf.cpp:182-292); returns (arf, info).

    POTRF(v11) -> TRSM (v21 := v21 L11^{-T}) -> SYRK (v22 -= v21 v21^T) ->
    POTRF(v22); a failure in v22 reports info + n1 (:282)."""
    arf = np.array(arf, copy=True)
    if n == 0:
        return arf, 0
    n1, n2, _, v11, v21, v22 = views(n, transr, uplo)
    dt = arf.dtype
    if n1 > 0:
        T11 = _get(arf, v11, n1, n1)
        L11, info = _potrf_lower(np.tril(T11) + np.tril(T11, -1).T)
        if info:
            return arf, info
        _put(arf, v11, n1, n1, L11, lower=True)
    if n1 > 0 and n2 > 0:
        S = _get(arf, v21, n2, n1)
        trsm = _blas.strsm if dt == np.float32 else _blas.dtrsm
        # X L11^T = S  ->  side='R', lower, trans
        X = trsm(1.0, np.asfortranarray(L11), np.asfortranarray(S), side=1, lower=1, trans_a=1)
        _put(arf, v21, n2, n1, X)
        T22 = _get(arf, v22, n2, n2)
        syrk = _blas.ssyrk if dt == np.float32 else _blas.dsyrk
        U = syrk(-1.0, np.asfortranarray(X), beta=1.0, c=np.asfortranarray(T22), lower=1)
        _put(arf, v22, n2, n2, U, lower=True)
    if n2 > 0:
        T22 = _get(arf, v22, n2, n2)
        L22, info = _potrf_lower(np.tril(T22) + np.tril(T22, -1).T)
        if info:
            return arf, info + n1
        _put(arf, v22, n2, n2, L22, lower=True)
    return arf, 0


def pftrs(n, arf, B, transr="T", uplo="U"):
    """Solve C X = B with the RFP factor (reference src/curfp_spftrs.cpp:215-302).

    Forward: L11 y1 = b1, b2 -= L21 y1, L22 y2 = b2; backward: L22^T x2 = y2,
    b1 -= L21^T x2, L11^T x1 = b1.  B is (n, nrhs)."""
    B = np.array(B, copy=True)
    if n == 0 or B.size == 0:
        return B
    squeeze = B.ndim == 1
    if squeeze:
        B = B[:, None]
    n1, n2, _, v11, v21, v22 = views(n, transr, uplo)
    dt = arf.dtype
    trsm = _blas.strsm if dt == np.float32 else _blas.dtrsm
    gemm = _blas.sgemm if dt == np.float32 else _blas.dgemm
    L11 = np.asfortranarray(np.tril(_get(arf, v11, n1, n1))) if n1 else None
    L22 = np.asfortranarray(np.tril(_get(arf, v22, n2, n2))) if n2 else None
    L21 = np.asfortranarray(_get(arf, v21, n2, n1)) if n1 and n2 else None
    b1 = np.asfortranarray(B[:n1].astype(dt))
    b2 = np.asfortranarray(B[n1:].astype(dt))
    if n1:
        b1 = trsm(1.0, L11, b1, side=0, lower=1, trans_a=0)
    if L21 is not None:
        b2 = gemm(-1.0, L21, b1, beta=1.0, c=b2)
    if n2:
        b2 = trsm(1.0, L22, b2, side=0, lower=1, trans_a=0)
        b2 = trsm(1.0, L22, b2, side=0, lower=1, trans_a=1)
    if L21 is not None:
        b1 = gemm(-1.0, L21, b2, beta=1.0, c=b1, trans_a=1)
    if n1:
        b1 = trsm(1.0, L11, b1, side=0, lower=1, trans_a=1)
    X = np.concatenate([b1, b2], axis=0).astype(dt)
    return X[:, 0] if squeeze else X


def pfcon(n: int, arf: np.ndarray, anorm: float, transr: str = "T", uplo: str = "U", itmax: int = 5) -> float:
    """Reciprocal condition estimate 1 / (est ||A^-1||_1 * anorm) from the RFP
    Cholesky factor (reference src/curfp_spfcon.cu:84-271): LAPACK ?lacn2's
    Hager-Higham iteration with A^-T = A^-1 and the reference's decisions —
    start at inv(A) 1/n, then inv(A) sign(x), then unit vectors e_j at the
    largest |x_j| until the sign pattern repeats, the largest entry stays put
    or itmax iterations have run, and finally max(est, 2 |inv(A) v|_1 / (3n))
    for v_i = (-1)^i (1 + i/(n-1)).  Products inv(A) x are two float64
    triangular solves with the dense factor."""
    if n == 0 or anorm == 0:
        return 0.0
    L = rfp_lower_factor(np.asarray(arf, dtype=np.float64), n, transr, uplo)

    def solve(v):
        y, _ = _lapack.dtrtrs(L, v, lower=1)
        z, _ = _lapack.dtrtrs(L, y, lower=1, trans=1)
        return z

    x = solve(np.full(n, 1.0 / n))
    est = float(np.abs(x).sum())
    if n > 1:
        isgn = np.where(x >= 0, 1.0, -1.0)
        x = solve(isgn.copy())
        j, it = int(np.argmax(np.abs(x))), 2
        while True:
            x = solve(np.eye(n)[j])
            est = float(np.abs(x).sum())
            if np.all(isgn * x >= 0):          # sign pattern repeated
                break
            isgn = np.where(x >= 0, 1.0, -1.0)
            x = solve(isgn.copy())
            jn = int(np.argmax(np.abs(x)))
            if abs(x[j]) == abs(x[jn]) or it >= itmax:
                break
            j, it = jn, it + 1
        v = np.where(np.arange(n) % 2 == 0, 1.0, -1.0) * (1.0 + np.arange(n) / (n - 1))
        est = max(est, 2.0 * float(np.abs(solve(v)).sum()) / (3.0 * n))
    return (1.0 / est) / anorm if est > 0 else 0.0


def pftri(n: int, arf: np.ndarray, transr: str = "T", uplo: str = "U") -> np.ndarray:
    """inv(C) in RFP from the RFP Cholesky factor (reference
    src/curfp_spftri.cu:159-388): X = inv(L) (LAPACK ?trtri on the dense
    factor), inv(C) = X^T X, packed back into the same layout; float64."""
    if n == 0:
        return np.zeros(0)
    L = rfp_lower_factor(np.asarray(arf, dtype=np.float64), n, transr, uplo)
    X, info = _lapack.dtrtri(L, lower=1)
    if info != 0:
        raise np.linalg.LinAlgError(f"singular factor (trtri info {info})")
    X = np.tril(X)
    return dense_sym_to_rfp(X.T @ X, transr, uplo)


def lansf(arf: np.ndarray, n: int, norm: str, transr: str = "T", uplo: str = "U") -> float:
    """Norm of the symmetric matrix stored in RFP (reference
    src/curfp_slansf.cu:273-435): 'M' max |a_ij|, '1' (= 'O' = 'I') the
    largest absolute column sum, 'F' (= 'E') Frobenius; evaluated in float64
    on the dense symmetric matrix."""
    if n == 0:
        return 0.0
    S = np.abs(rfp_to_dense_sym(np.asarray(arf, dtype=np.float64), n, transr, uplo))
    key = norm.upper()
    if key in ("M", "MAX"):
        return float(S.max())
    if key in ("1", "O", "I"):
        return float(S.sum(axis=0).max())
    if key in ("F", "E", "FRO"):
        return float(np.sqrt((S * S).sum()))
    raise ValueError(f"unknown norm {norm!r}")


def sfmm(arf: np.ndarray, nA: int, B: np.ndarray, C: np.ndarray, alpha: float, beta: float, transr: str = "T",
         uplo: str = "U", side: str = "L") -> np.ndarray:
    """Symmetric RFP matrix times a dense block (reference
    src/curfp_ssfmm.cpp:107-306), float64: side 'L' alpha S B + beta C,
    side 'R' alpha B S + beta C.  alpha == 0 does not read S or B; beta == 0
    does not read C (BLAS conventions)."""
    B = np.asarray(B, dtype=np.float64)
    out = np.zeros(B.shape) if beta == 0 else beta * np.asarray(C, dtype=np.float64)
    if alpha == 0 or nA == 0:
        return out
    S = rfp_to_dense_sym(np.asarray(arf, dtype=np.float64), nA, transr, uplo)
    return out + alpha * (S @ B if side.upper() == "L" else B @ S)


def chain_flops(n: int, k: int, nrhs: int) -> dict:
    """Algorithmic flop counts of ssfrk, spftrf, spftrs (LAPACK conventions)."""
    f_sfrk = float(n) * (n + 1) * k
    f_pftrf = n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0
    f_pftrs = 2.0 * n * n * nrhs
    return {"ssfrk": f_sfrk, "spftrf": f_pftrf, "spftrs": f_pftrs, "chain": f_sfrk + f_pftrf + f_pftrs}


def dense_relative_residual(S: np.ndarray, X: np.ndarray, B: np.ndarray) -> float:
    """||B - S X||_F / (||S||_F ||X||_F), evaluated in float64."""
    S = S.astype(np.float64)
    X = X.astype(np.float64)
    B = B.astype(np.float64)
    return float(np.linalg.norm(B - S @ X) / (np.linalg.norm(S) * np.linalg.norm(X)))

