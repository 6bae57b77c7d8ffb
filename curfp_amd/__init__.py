"""
curfp_amd — RFP (Rectangular Full Packed) routines for AMD Instinct MI355X.

The Python surface mirrors the reference package ``curfp`` (curfp/__init__.py)
for the RFP hot path, so ``import curfp_amd as curfp`` is a drop-in there:

    import torch, curfp_amd as curfp
    n, k, nrhs = 4096, 128, 10
    A = torch.randn(n, k, device="cuda") / k**0.5
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    B = torch.randn(n, nrhs, device="cuda")
    curfp.ssfrk(A, C)                 # C = A @ A.T in RFP format
    curfp.add_to_diagonal(C, 1.0)     # C += I
    curfp.spftrf(C)                   # Cholesky, in place
    curfp.spftrs(C, B)                # B := C^{-1} B

Every call goes through the C ABI in include/curfp.h (libcurfp_amd.so) to
hand-written gfx950 HIP kernels; there is no CPU or PyTorch fallback: when the
native library is missing the import fails.

Beyond the reference: ``set_math_mode`` (BF16 / split-BF16 MFMA for ssfrk and
the spftrf updates), ``ssfrk`` on a bfloat16 ``A``, the FP64 generator
residual ``gen_residual`` and the iterative-refinement solver in
``curfp_amd.refine``, and the autograd chain in ``curfp_amd.autograd``.
"""
from __future__ import annotations

import math

import torch

try:
    from ._curfp_C import (  # noqa: F401
        Handle as _HandleBase,
        OP_N,
        OP_T,
        FILL_LOWER,
        FILL_UPPER,
        MATH_FP32,
        MATH_BF16,
        MATH_BF16X3,
        NORM_MAX,
        NORM_ONE,
        NORM_FRO,
        SIDE_LEFT,
        SIDE_RIGHT,
        KERNEL_CLASSES,
        ssfrk as _ssfrk_C,
        ssfrk_bf16 as _ssfrk_bf16_C,
        spftrf as _spftrf_C,
        spftrs as _spftrs_C,
        spftrs_rm as _spftrs_rm_C,
        strttf as _strttf_C,
        stfttr as _stfttr_C,
        sadd_to_diagonal as _sadd_diag_C,
        sgen_residual as _sgen_residual_C,
        spftrs_refine as _spftrs_refine_C,
        sgemm_strided as _sgemm_strided_C,
        dgemm_strided as _dgemm_strided_C,
        dsfrk as _dsfrk_C,
        dpftrf as _dpftrf_C,
        dpftrs as _dpftrs_C,
        dpftrs_rm as _dpftrs_rm_C,
        dstrttf as _dstrttf_C,
        dstfttr as _dstfttr_C,
        dadd_to_diagonal as _dadd_diag_C,
        slansf as _slansf_C,
        dlansf as _dlansf_C,
        ssfmm as _ssfmm_C,
        dsfmm as _dsfmm_C,
        spftri as _spftri_C,
        dpftri as _dpftri_C,
        ssfmv as _ssfmv_C,
        dsfmv as _dsfmv_C,
        spfcon as _spfcon_C,
        dpfcon as _dpfcon_C,
        ssfr as _ssfr_C,
        dsfr as _dsfr_C,
        ssfr2 as _ssfr2_C,
        dsfr2 as _dsfr2_C,
        ssfr2k as _ssfr2k_C,
        dsfr2k as _dsfr2k_C,
    )
except ImportError as exc:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "curfp_amd native extension is not built; run `python curfp_amd/_build.py` "
        "(hipcc --offload-arch=gfx950) first"
    ) from exc

# norm strings accepted by slansf / dlansf (curfp/__init__.py:114-123)
_NORM_MAP = {"M": NORM_MAX, "MAX": NORM_MAX, "1": NORM_ONE, "O": NORM_ONE, "I": NORM_ONE,
             "F": NORM_FRO, "E": NORM_FRO, "FRO": NORM_FRO}


def _norm(s: str) -> int:
    try:
        return _NORM_MAP[s.upper()]
    except KeyError:
        raise ValueError(f"norm must be 'M'/'1'/'F' (or 'O'/'I'/'E'), got {s!r}") from None

# ---------------------------------------------------------------------------
# Global handle cache (one handle per device) — curfp/__init__.py:88-106
# ---------------------------------------------------------------------------
_handles: dict = {}


def _get_handle() -> "Handle":
    dev = torch.cuda.current_device()
    h = _handles.get(dev)
    if h is None:
        h = Handle()
        _handles[dev] = h
    return h


def set_stream(stream) -> None:
    """Bind curfp operations on the current device to ``stream`` (None = default)."""
    h = _get_handle()
    h.set_stream_ptr(0 if stream is None else stream.cuda_stream)


_MATH_MAP = {"fp32": MATH_FP32, "bf16": MATH_BF16, "bf16x3": MATH_BF16X3}


def synchronize() -> None:
    """Wait for the current device's curfp stream; raises RuntimeError
    ("curfp error: status=4") if an earlier asynchronous call failed on the
    device (a substitution sweep that gave up waiting and wrote NaN)."""
    _get_handle().synchronize()


def set_math_mode(mode: str) -> None:
    """Arithmetic of ssfrk and the spftrf updates on the current device's handle:
    'fp32' (default), 'bf16' or 'bf16x3' (split-BF16, FP32-accurate products)."""
    try:
        m = _MATH_MAP[mode.lower()]
    except KeyError:
        raise ValueError(f"math mode must be one of {sorted(_MATH_MAP)}, got {mode!r}")
    _get_handle().set_math_mode(m)


def get_math_mode() -> str:
    inv = {v: k for k, v in _MATH_MAP.items()}
    return inv[_get_handle().get_math_mode()]


def set_profiling(enable: bool) -> None:
    """Bracket every library kernel with HIP events (and reset the totals)."""
    _get_handle().set_profiling(bool(enable))


def workspace_bytes() -> int:
    """Device memory held by the current device's handle (workspace + BF16 shadow)."""
    return int(_get_handle().workspace_bytes())


def kernel_stats(reset: bool = False) -> dict:
    """Per kernel class: launches, algorithmic flops / bytes, summed ms."""
    h = _get_handle()
    return {name: h.kernel_stats(cls, reset) for name, cls in KERNEL_CLASSES.items()}


class math_mode:
    """Context manager: ``with curfp.math_mode('bf16'): ...``."""

    def __init__(self, mode: str):
        self.mode = mode
        self.prev = None

    def __enter__(self):
        self.prev = get_math_mode()
        set_math_mode(self.mode)
        return self

    def __exit__(self, *exc):
        set_math_mode(self.prev)
        return False


# ---------------------------------------------------------------------------
# String -> enum helpers — curfp/__init__.py:112-154
# ---------------------------------------------------------------------------
_OP_MAP = {"N": OP_N, "T": OP_T}
_FILL_MAP = {"L": FILL_LOWER, "U": FILL_UPPER}


def _op(s: str) -> int:
    try:
        return _OP_MAP[s.upper()]
    except (KeyError, AttributeError):
        raise ValueError(f"transr/trans must be 'N' or 'T', got {s!r}")


def _fill(s: str) -> int:
    try:
        return _FILL_MAP[s.upper()]
    except (KeyError, AttributeError):
        raise ValueError(f"uplo must be 'L' or 'U', got {s!r}")


def _n_from_rfp(numel: int) -> int:
    """Recover n from numel == n*(n+1)//2 (curfp/__init__.py:160-167), exactly."""
    n = (math.isqrt(8 * numel + 1) - 1) // 2
    if n * (n + 1) // 2 != numel:
        raise ValueError(f"Tensor has {numel} elements which is not a valid RFP size n*(n+1)//2")
    return n


def _validate(t: torch.Tensor, name: str, dtype=torch.float32) -> None:
    """curfp/__init__.py:173-188."""
    if not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.dtype != dtype:
        name64 = "float64" if dtype == torch.float64 else "float32"
        raise ValueError(f"{name} must be {name64} (got {t.dtype})")


def _validate64(t: torch.Tensor, name: str) -> None:
    _validate(t, name, torch.float64)


# ---------------------------------------------------------------------------
# RFP geometry (same 8-variant layout as LAPACK; curfp/__init__.py:194-269)
# ---------------------------------------------------------------------------
def rfp_blocks(n: int, transr: str = "T", uplo: str = "U"):
    """(n1, n2, ld, v11, v21, v22): the three blocks of the lower factor as
    strided views (offset, row stride, col stride) into the RFP array."""
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

    def view(o, row):
        return (o, ld, 1) if row else (o, 1, ld)

    return n1, n2, ld, view(o11, not tn), view(o21, tn != lower), view(o22, tn)


def rfp_diag_indices(n: int, transr: str = "T", uplo: str = "U", device=None) -> torch.Tensor:
    """Flat RFP indices of the n diagonal elements (int64), all 8 variants."""
    if device is None:
        device = torch.cuda.current_device()
    n1, n2, ld, v11, _, v22 = rfp_blocks(n, transr, uplo)
    d1 = v11[0] + torch.arange(n1, device=device, dtype=torch.int64) * (ld + 1)
    d2 = v22[0] + torch.arange(n2, device=device, dtype=torch.int64) * (ld + 1)
    return torch.cat([d1, d2])


def add_to_diagonal(C: torch.Tensor, value: float, transr: str = "T", uplo: str = "U", n: int = None) -> None:
    """C[diag] += value (curfp/__init__.py:272-296), as one HIP kernel."""
    if not C.is_cuda:
        raise ValueError("C must be a CUDA tensor")
    if not C.is_contiguous():
        raise ValueError("C must be contiguous")
    if C.dtype not in (torch.float32, torch.float64):
        raise ValueError(f"C must be float32 or float64 (got {C.dtype})")
    if n is None:
        n = _n_from_rfp(C.numel())
    fn = _sadd_diag_C if C.dtype == torch.float32 else _dadd_diag_C
    fn(_get_handle(), _op(transr), _fill(uplo), n, float(value), C.data_ptr())


class Handle(_HandleBase):
    """curfp library handle (stream, math mode, cached workspace)."""

    def set_stream(self, stream) -> None:
        self.set_stream_ptr(stream.cuda_stream)


# ---------------------------------------------------------------------------
# Raw API — explicit handle, integer enums (curfp/__init__.py:323-405)
# ---------------------------------------------------------------------------
def ssfrk_raw(handle, transr, uplo, trans, n, k, alpha, A, lda, beta, C) -> None:
    _validate(C, "C")
    if A.dtype == torch.bfloat16:
        if not A.is_cuda or not A.is_contiguous():
            raise ValueError("A must be a contiguous CUDA tensor")
        _ssfrk_bf16_C(handle, transr, uplo, trans, n, k, float(alpha), A.data_ptr(), lda, float(beta), C.data_ptr())
        return
    _validate(A, "A")
    _ssfrk_C(handle, transr, uplo, trans, n, k, float(alpha), A.data_ptr(), lda, float(beta), C.data_ptr())


def spftrf_raw(handle, transr, uplo, n, A) -> int:
    _validate(A, "A")
    return _spftrf_C(handle, transr, uplo, n, A.data_ptr())


def spftrs_raw(handle, transr, uplo, n, nrhs, A, B, ldb) -> None:
    """B is (n x nrhs) column-major with leading dimension ldb (C ABI layout)."""
    _validate(A, "A")
    _validate(B, "B")
    _spftrs_C(handle, transr, uplo, n, nrhs, A.data_ptr(), B.data_ptr(), ldb)


def strttf_raw(handle, transr, uplo, n, A, lda, arf) -> None:
    _validate(A, "A")
    _validate(arf, "arf")
    _strttf_C(handle, transr, uplo, n, A.data_ptr(), lda, arf.data_ptr())


def stfttr_raw(handle, transr, uplo, n, arf, A, lda) -> None:
    _validate(arf, "arf")
    _validate(A, "A")
    _stfttr_C(handle, transr, uplo, n, arf.data_ptr(), A.data_ptr(), lda)


def dsfrk_raw(handle, transr, uplo, trans, n, k, alpha, A, lda, beta, C) -> None:
    _validate64(A, "A")
    _validate64(C, "C")
    _dsfrk_C(handle, transr, uplo, trans, n, k, float(alpha), A.data_ptr(), lda, float(beta), C.data_ptr())


def dpftrf_raw(handle, transr, uplo, n, A) -> int:
    _validate64(A, "A")
    return _dpftrf_C(handle, transr, uplo, n, A.data_ptr())


def dpftrs_raw(handle, transr, uplo, n, nrhs, A, B, ldb) -> None:
    _validate64(A, "A")
    _validate64(B, "B")
    _dpftrs_C(handle, transr, uplo, n, nrhs, A.data_ptr(), B.data_ptr(), ldb)


def dstrttf_raw(handle, transr, uplo, n, A, lda, arf) -> None:
    _validate64(A, "A")
    _validate64(arf, "arf")
    _dstrttf_C(handle, transr, uplo, n, A.data_ptr(), lda, arf.data_ptr())


def dstfttr_raw(handle, transr, uplo, n, arf, A, lda) -> None:
    _validate64(arf, "arf")
    _validate64(A, "A")
    _dstfttr_C(handle, transr, uplo, n, arf.data_ptr(), A.data_ptr(), lda)


# ---------------------------------------------------------------------------
# High-level API — global handle, string parameters, shape inference
# ---------------------------------------------------------------------------
def _sfrk(A, C, alpha, beta, transr, uplo, trans, n, k, lda, f64):
    if f64:
        _validate64(A, "A")
        _validate64(C, "C")
    elif A.dtype == torch.bfloat16:
        _validate(C, "C")
        if not A.is_cuda or not A.is_contiguous():
            raise ValueError("A must be a contiguous CUDA tensor")
    else:
        _validate(A, "A")
        _validate(C, "C")
    if n is None:
        n = _n_from_rfp(C.numel())
    t = trans.upper()
    if k is None:
        k = A.shape[1] if t == "T" else A.shape[0]
    if lda is None:
        lda = A.shape[1] if t == "T" else A.shape[0]
    if f64:
        fn = _dsfrk_C
    else:
        fn = _ssfrk_bf16_C if A.dtype == torch.bfloat16 else _ssfrk_C
    fn(_get_handle(), _op(transr), _fill(uplo), _op(trans), n, k, float(alpha), A.data_ptr(), lda,
       float(beta), C.data_ptr())


def ssfrk(A, C, *, alpha=1.0, beta=0.0, transr="T", uplo="U", trans="T", n=None, k=None, lda=None) -> None:
    """C := alpha*A@A.T + beta*C into RFP storage (curfp/__init__.py:410-465).

    trans='T': A is row-major (n, k); trans='N': A is column-major n x k
    (i.e. a row-major (k, n) tensor).  A may be bfloat16 (BF16 MFMA with FP32
    accumulation; C stays float32)."""
    _sfrk(A, C, alpha, beta, transr, uplo, trans, n, k, lda, False)


def dsfrk(A, C, *, alpha=1.0, beta=0.0, transr="T", uplo="U", trans="T", n=None, k=None, lda=None) -> None:
    """FP64 ssfrk (curfp/__init__.py:1439-1474)."""
    _sfrk(A, C, alpha, beta, transr, uplo, trans, n, k, lda, True)


def _pftrf(C, n, transr, uplo, check, f64):
    (_validate64 if f64 else _validate)(C, "C")
    if n is None:
        n = _n_from_rfp(C.numel())
    fn = _dpftrf_C if f64 else _spftrf_C
    info = fn(_get_handle(), _op(transr), _fill(uplo), n, C.data_ptr())
    if check and info != 0:
        raise torch.linalg.LinAlgError(
            f"{'dpftrf' if f64 else 'spftrf'}: matrix is not positive definite "
            f"(leading minor of order {info} is not positive definite)"
        )
    return info


def spftrf(C, *, n=None, transr="T", uplo="U", check=True) -> int:
    """In-place RFP Cholesky (curfp/__init__.py:468-501); returns info."""
    return _pftrf(C, n, transr, uplo, check, False)


def dpftrf(C, *, n=None, transr="T", uplo="U", check=True) -> int:
    """FP64 spftrf (curfp/__init__.py:1511-1529)."""
    return _pftrf(C, n, transr, uplo, check, True)


def _pftrs(C, B, n, nrhs, transr, uplo, f64):
    v = _validate64 if f64 else _validate
    v(C, "C")
    v(B, "B")
    if n is None:
        n = _n_from_rfp(C.numel())
    squeezed = B.ndim == 1
    Bv = B.unsqueeze(1) if squeezed else B
    if nrhs is None:
        nrhs = Bv.shape[1]
    if Bv.shape[0] != n:
        raise ValueError(f"B has {Bv.shape[0]} rows but n={n}")
    # B is (n, nrhs) row-major: solved in place through the row-major entry
    # point, no transposed copies (the reference copies twice, :548 and :562).
    fn = _dpftrs_rm_C if f64 else _spftrs_rm_C
    fn(_get_handle(), _op(transr), _fill(uplo), n, nrhs, C.data_ptr(), Bv.data_ptr(), max(1, Bv.shape[1]))


def spftrs(C, B, *, n=None, nrhs=None, transr="T", uplo="U") -> None:
    """Solve C X = B in place on B, shape (n, nrhs) or (n,) (curfp/__init__.py:504-565)."""
    _pftrs(C, B, n, nrhs, transr, uplo, False)


def dpftrs(C, B, *, n=None, nrhs=None, transr="T", uplo="U") -> None:
    """FP64 spftrs (curfp/__init__.py:1547-1581)."""
    _pftrs(C, B, n, nrhs, transr, uplo, True)


def _trttf(A, transr, uplo, n, f64):
    (_validate64 if f64 else _validate)(A, "A")
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"A must be a square 2-D tensor, got shape {tuple(A.shape)}")
    if n is None:
        n = A.shape[0]
    arf = torch.empty(n * (n + 1) // 2, dtype=A.dtype, device=A.device)
    fn = _dstrttf_C if f64 else _strttf_C
    fn(_get_handle(), _op(transr), _fill(uplo), n, A.data_ptr(), A.shape[1], arf.data_ptr())
    return arf


def _tfttr(arf, transr, uplo, n, f64):
    (_validate64 if f64 else _validate)(arf, "arf")
    if n is None:
        n = _n_from_rfp(arf.numel())
    A = torch.zeros(n, n, dtype=arf.dtype, device=arf.device)
    fn = _dstfttr_C if f64 else _stfttr_C
    fn(_get_handle(), _op(transr), _fill(uplo), n, arf.data_ptr(), A.data_ptr(), n)
    return A


def strttf(A, *, transr="T", uplo="U", n=None) -> torch.Tensor:
    """Dense (n, n) triangle -> new RFP tensor (curfp/__init__.py:719-750)."""
    return _trttf(A, transr, uplo, n, False)


def stfttr(arf, *, transr="T", uplo="U", n=None) -> torch.Tensor:
    """RFP -> new dense (n, n) tensor, other triangle zero (curfp/__init__.py:753-781)."""
    return _tfttr(arf, transr, uplo, n, False)


def dstrttf(A, *, transr="T", uplo="U", n=None) -> torch.Tensor:
    """FP64 strttf (curfp/__init__.py:1347-1375)."""
    return _trttf(A, transr, uplo, n, True)


def dstfttr(arf, *, transr="T", uplo="U", n=None) -> torch.Tensor:
    """FP64 stfttr (curfp/__init__.py:1378-1403)."""
    return _tfttr(arf, transr, uplo, n, True)


def gemm_nt(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor = None, *, alpha: float = 1.0,
            beta: float = 0.0) -> torch.Tensor:
    """C = alpha * A @ B.T + beta * C on the library's MFMA GEMM (float32 or
    float64; any strides with a unit stride in one direction)."""
    if A.dtype != B.dtype or A.dtype not in (torch.float32, torch.float64):
        raise ValueError("A and B must be both float32 or both float64")
    if A.ndim != 2 or B.ndim != 2 or A.shape[1] != B.shape[1]:
        raise ValueError(f"shape mismatch {tuple(A.shape)} x {tuple(B.shape)}^T")
    M, K = A.shape
    N = B.shape[0]
    if C is None:
        C = torch.empty(M, N, dtype=A.dtype, device=A.device)
        beta = 0.0
    if tuple(C.shape) != (M, N) or C.dtype != A.dtype:
        raise ValueError("C has the wrong shape or dtype")
    fn = _sgemm_strided_C if A.dtype == torch.float32 else _dgemm_strided_C
    fn(_get_handle(), M, N, K, float(alpha), A.data_ptr(), A.stride(0), A.stride(1), B.data_ptr(), B.stride(0),
       B.stride(1), float(beta), C.data_ptr(), C.stride(0), C.stride(1))
    return C


def gen_residual(A, X, B, *, lam: float, alpha: float = 1.0, trans: str = "T", R=None) -> torch.Tensor:
    """FP64 residual R = B - (alpha*op(A)@op(A).T + lam*I) @ X from the FP32 generator A.

    A: float32, row-major (n, k) for trans='T'.  X, B (and R if given):
    float64 (n, nrhs) tensors stored column-major, i.e. ``t.stride() == (1, n)``
    (the transpose of a contiguous (nrhs, n) tensor)."""
    _validate(A, "A")
    t = trans.upper()
    n = A.shape[0] if t == "T" else A.shape[1]
    k = A.shape[1] if t == "T" else A.shape[0]
    lda = A.shape[1] if t == "T" else A.shape[0]
    nrhs = X.shape[1]
    if R is None:
        R = torch.empty_strided((n, nrhs), (1, n), dtype=torch.float64, device=A.device)
    for v, nm in ((X, "X"), (B, "B"), (R, "R")):
        if (v.dtype != torch.float64 or not v.is_cuda or tuple(v.shape) != (n, nrhs)
                or (nrhs > 1 and v.stride() != (1, n)) or (nrhs == 1 and v.stride()[0] != 1)):
            raise ValueError(f"{nm} must be a float64 CUDA (n, nrhs) column-major tensor")
    _sgen_residual_C(_get_handle(), _op(trans), n, k, nrhs, float(alpha), A.data_ptr(), lda, float(lam),
                     X.data_ptr(), n, B.data_ptr(), n, R.data_ptr(), n)
    return R


# ---------------------------------------------------------------------------
# Symmetric matrix-vector multiply — curfp/__init__.py:568-670
# ---------------------------------------------------------------------------
def ssfmv_raw(handle: Handle, transr: int, uplo: int, n: int, alpha: float, arf: torch.Tensor, x: torch.Tensor,
              incx: int, beta: float, y: torch.Tensor, incy: int) -> None:
    """Low-level ssfmv: y := alpha A x + beta y (BLAS increments incx, incy)."""
    _validate(arf, "arf")
    _validate_strided(x, "x", torch.float32)
    _validate_strided(y, "y", torch.float32)
    _ssfmv_C(handle, transr, uplo, n, float(alpha), arf.data_ptr(), x.data_ptr(), incx, float(beta), y.data_ptr(),
             incy)


def dsfmv_raw(handle: Handle, transr: int, uplo: int, n: int, alpha: float, arf: torch.Tensor, x: torch.Tensor,
              incx: int, beta: float, y: torch.Tensor, incy: int) -> None:
    """FP64 twin of ssfmv_raw."""
    _validate(arf, "arf", torch.float64)
    _validate_strided(x, "x", torch.float64)
    _validate_strided(y, "y", torch.float64)
    _dsfmv_C(handle, transr, uplo, n, float(alpha), arf.data_ptr(), x.data_ptr(), incx, float(beta), y.data_ptr(),
             incy)


def _sfmv(fn, dtype, arf, x, y, alpha, beta, transr, uplo, n):
    _validate(arf, "arf", dtype)
    _vec(x, "x", dtype)
    if n is None:
        n = _n_from_rfp(arf.numel())
    if y is None:
        y = torch.zeros(n, dtype=dtype, device=arf.device)
    _vec(y, "y", dtype)
    if x.numel() != n:
        raise ValueError(f"x has {x.numel()} elements but n={n}")
    if y.numel() != n:
        raise ValueError(f"y has {y.numel()} elements but n={n}")
    fn(_get_handle(), _op(transr), _fill(uplo), n, float(alpha), arf.data_ptr(), x.data_ptr(), 1, float(beta),
       y.data_ptr(), 1)
    return y


def ssfmv(arf: torch.Tensor, x: torch.Tensor, y: torch.Tensor = None, *, alpha: float = 1.0, beta: float = 0.0,
          transr: str = "T", uplo: str = "U", n: int = None) -> torch.Tensor:
    """y := alpha A x + beta y with A symmetric in RFP; y is created (zeros)
    when not given and returned."""
    return _sfmv(_ssfmv_C, torch.float32, arf, x, y, alpha, beta, transr, uplo, n)


def dsfmv(arf: torch.Tensor, x: torch.Tensor, y: torch.Tensor = None, *, alpha: float = 1.0, beta: float = 0.0,
          transr: str = "T", uplo: str = "U", n: int = None) -> torch.Tensor:
    """FP64 twin of ssfmv."""
    return _sfmv(_dsfmv_C, torch.float64, arf, x, y, alpha, beta, transr, uplo, n)


# ---------------------------------------------------------------------------
# Reciprocal condition estimate — curfp/__init__.py:844-955, 1739-1800
# ---------------------------------------------------------------------------
def spfcon_raw(handle: Handle, transr: int, uplo: int, n: int, C: torch.Tensor, anorm: float) -> float:
    """Low-level spfcon: rcond = 1 / (||A^-1||_1 anorm) estimated from the RFP Cholesky factor C."""
    _validate(C, "C")
    return float(_spfcon_C(handle, transr, uplo, n, C.data_ptr(), float(anorm)))


def dpfcon_raw(handle: Handle, transr: int, uplo: int, n: int, C: torch.Tensor, anorm: float) -> float:
    """FP64 twin of spfcon_raw."""
    _validate(C, "C", torch.float64)
    return float(_dpfcon_C(handle, transr, uplo, n, C.data_ptr(), float(anorm)))


def spfcon(C: torch.Tensor, anorm: float, *, transr: str = "T", uplo: str = "U", n: int = None) -> float:
    """Estimate the reciprocal condition number 1 / (||A^-1||_1 anorm) from the
    RFP Cholesky factor C (from spftrf, same transr / uplo); anorm is the
    1-norm of A before factorisation (``slansf(C, '1')`` then).  Hager-Higham
    estimator (LAPACK ?lacn2); 0.0 when anorm == 0."""
    _validate(C, "C")
    if n is None:
        n = _n_from_rfp(C.numel())
    return float(_spfcon_C(_get_handle(), _op(transr), _fill(uplo), n, C.data_ptr(), float(anorm)))


def dpfcon(C: torch.Tensor, anorm: float, *, transr: str = "T", uplo: str = "U", n: int = None) -> float:
    """FP64 twin of spfcon."""
    _validate(C, "C", torch.float64)
    if n is None:
        n = _n_from_rfp(C.numel())
    return float(_dpfcon_C(_get_handle(), _op(transr), _fill(uplo), n, C.data_ptr(), float(anorm)))


def _pftrf_rcond(lansf, pftrf, pfcon, C, transr, uplo, n, check, name):
    if n is None:
        n = _n_from_rfp(C.numel())
    h, t, u = _get_handle(), _op(transr), _fill(uplo)
    anorm = float(lansf(h, NORM_ONE, t, u, n, C.data_ptr()))
    info = pftrf(h, t, u, n, C.data_ptr())
    if check and info != 0:
        raise torch.linalg.LinAlgError(f"{name}: matrix is not positive definite "
                                       f"(leading minor of order {info} is not positive definite)")
    if info != 0 or anorm == 0.0:
        return 0.0
    return float(pfcon(h, t, u, n, C.data_ptr(), anorm))


def spftrf_rcond(C: torch.Tensor, *, transr: str = "T", uplo: str = "U", n: int = None,
                 check: bool = True) -> float:
    """slansf('1') + spftrf + spfcon: C is overwritten with its Cholesky factor
    and the estimated reciprocal condition number is returned (0.0 if C is not
    positive definite and check is False; LinAlgError if check is True)."""
    _validate(C, "C")
    return _pftrf_rcond(_slansf_C, _spftrf_C, _spfcon_C, C, transr, uplo, n, check, "spftrf_rcond")


def dpftrf_rcond(C: torch.Tensor, *, transr: str = "T", uplo: str = "U", n: int = None,
                 check: bool = True) -> float:
    """FP64 twin of spftrf_rcond."""
    _validate(C, "C", torch.float64)
    return _pftrf_rcond(_dlansf_C, _dpftrf_C, _dpfcon_C, C, transr, uplo, n, check, "dpftrf_rcond")


# ---------------------------------------------------------------------------
# Rank-1 / rank-2 / rank-2k updates — curfp/__init__.py:966-1200
# ---------------------------------------------------------------------------
def _vec(t: torch.Tensor, name: str, dtype) -> None:
    if not (t.is_cuda and t.dtype == dtype and t.ndim == 1 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous 1-D {dtype} CUDA tensor")


def ssfr_raw(handle: Handle, transr: int, uplo: int, n: int, alpha: float, x: torch.Tensor, incx: int,
             arf: torch.Tensor) -> None:
    """Low-level ssfr: arf := alpha x x^T + arf (x with BLAS increment incx)."""
    _validate_strided(x, "x", torch.float32)
    _validate(arf, "arf")
    _ssfr_C(handle, transr, uplo, n, float(alpha), x.data_ptr(), incx, arf.data_ptr())


def ssfr(arf: torch.Tensor, x: torch.Tensor, *, alpha: float = 1.0, transr: str = "T", uplo: str = "U",
         n: int = None) -> None:
    """Symmetric rank-1 update in RFP: arf := alpha x x^T + arf."""
    _validate(arf, "arf")
    _vec(x, "x", torch.float32)
    if n is None:
        n = _n_from_rfp(arf.numel())
    _ssfr_C(_get_handle(), _op(transr), _fill(uplo), n, float(alpha), x.data_ptr(), 1, arf.data_ptr())


def dsfr(arf: torch.Tensor, x: torch.Tensor, *, alpha: float = 1.0, transr: str = "T", uplo: str = "U",
         n: int = None) -> None:
    """FP64 twin of ssfr."""
    _validate(arf, "arf", torch.float64)
    _vec(x, "x", torch.float64)
    if n is None:
        n = _n_from_rfp(arf.numel())
    _dsfr_C(_get_handle(), _op(transr), _fill(uplo), n, float(alpha), x.data_ptr(), 1, arf.data_ptr())


def ssfr2_raw(handle: Handle, transr: int, uplo: int, n: int, alpha: float, x: torch.Tensor, incx: int,
              y: torch.Tensor, incy: int, arf: torch.Tensor) -> None:
    """Low-level ssfr2: arf := alpha (x y^T + y x^T) + arf."""
    _validate_strided(x, "x", torch.float32)
    _validate_strided(y, "y", torch.float32)
    _validate(arf, "arf")
    _ssfr2_C(handle, transr, uplo, n, float(alpha), x.data_ptr(), incx, y.data_ptr(), incy, arf.data_ptr())


def ssfr2(arf: torch.Tensor, x: torch.Tensor, y: torch.Tensor, *, alpha: float = 1.0, transr: str = "T",
          uplo: str = "U", n: int = None) -> None:
    """Symmetric rank-2 update in RFP: arf := alpha (x y^T + y x^T) + arf."""
    _validate(arf, "arf")
    _vec(x, "x", torch.float32)
    _vec(y, "y", torch.float32)
    if n is None:
        n = _n_from_rfp(arf.numel())
    _ssfr2_C(_get_handle(), _op(transr), _fill(uplo), n, float(alpha), x.data_ptr(), 1, y.data_ptr(), 1,
             arf.data_ptr())


def dsfr2(arf: torch.Tensor, x: torch.Tensor, y: torch.Tensor, *, alpha: float = 1.0, transr: str = "T",
          uplo: str = "U", n: int = None) -> None:
    """FP64 twin of ssfr2."""
    _validate(arf, "arf", torch.float64)
    _vec(x, "x", torch.float64)
    _vec(y, "y", torch.float64)
    if n is None:
        n = _n_from_rfp(arf.numel())
    _dsfr2_C(_get_handle(), _op(transr), _fill(uplo), n, float(alpha), x.data_ptr(), 1, y.data_ptr(), 1,
             arf.data_ptr())


def _sfr2k(fn, dtype, A, B, C, alpha, beta, transr, uplo, trans, n, k, lda, ldb):
    _validate(A, "A", dtype)
    _validate(B, "B", dtype)
    _validate(C, "C", dtype)
    if n is None:
        n = _n_from_rfp(C.numel())
    t = trans.upper()
    if A.ndim != 2 or B.shape != A.shape:
        raise ValueError("A and B must be 2-D tensors of the same shape")
    # trans='T': row-major (n, k) = column-major k x n (lda = k); trans='N': row-major (k, n) = column-major
    # n x k (lda = n; the reference infers lda = A.shape[0] here, which its C entry rejects unless k >= n)
    if k is None:
        k = A.shape[1] if t == "T" else A.shape[0]
    if lda is None:
        lda = A.shape[1]
    if ldb is None:
        ldb = B.shape[1]
    if n > 0 and k > 0:
        for X, ld, name in ((A, lda, "A"), (B, ldb, "B")):
            need = (n - 1) * ld + k if t == "T" else (k - 1) * ld + n
            if X.numel() < need:
                raise ValueError(f"{name} holds {X.numel()} elements; n={n}, k={k}, ld={ld} needs {need}")
    fn(_get_handle(), _op(transr), _fill(uplo), _op(trans), n, k, float(alpha), A.data_ptr(), lda, B.data_ptr(),
       ldb, float(beta), C.data_ptr())


def ssfr2k(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, *, alpha: float = 1.0, beta: float = 0.0,
           transr: str = "T", uplo: str = "U", trans: str = "T", n: int = None, k: int = None, lda: int = None,
           ldb: int = None) -> None:
    """Symmetric rank-2k update in RFP: C := alpha op(A) op(B)^T + alpha op(B) op(A)^T + beta C
    (trans='T': A, B row-major (n, k), i.e. column-major k x n; trans='N': A, B row-major (k, n), i.e.
    column-major n x k).  Honours the handle's math mode like ssfrk."""
    _sfr2k(_ssfr2k_C, torch.float32, A, B, C, alpha, beta, transr, uplo, trans, n, k, lda, ldb)


def dsfr2k(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, *, alpha: float = 1.0, beta: float = 0.0,
           transr: str = "T", uplo: str = "U", trans: str = "T", n: int = None, k: int = None, lda: int = None,
           ldb: int = None) -> None:
    """FP64 twin of ssfr2k."""
    _sfr2k(_dsfr2k_C, torch.float64, A, B, C, alpha, beta, transr, uplo, trans, n, k, lda, ldb)


def dsfr_raw(handle: Handle, transr: int, uplo: int, n: int, alpha: float, x: torch.Tensor, incx: int,
             arf: torch.Tensor) -> None:
    """FP64 twin of ssfr_raw."""
    _validate_strided(x, "x", torch.float64)
    _validate(arf, "arf", torch.float64)
    _dsfr_C(handle, transr, uplo, n, float(alpha), x.data_ptr(), incx, arf.data_ptr())


def dsfr2_raw(handle: Handle, transr: int, uplo: int, n: int, alpha: float, x: torch.Tensor, incx: int,
              y: torch.Tensor, incy: int, arf: torch.Tensor) -> None:
    """FP64 twin of ssfr2_raw."""
    _validate_strided(x, "x", torch.float64)
    _validate_strided(y, "y", torch.float64)
    _validate(arf, "arf", torch.float64)
    _dsfr2_C(handle, transr, uplo, n, float(alpha), x.data_ptr(), incx, y.data_ptr(), incy, arf.data_ptr())


def dsfr2k_raw(handle: Handle, transr: int, uplo: int, trans: int, n: int, k: int, alpha: float,
               A: torch.Tensor, lda: int, B: torch.Tensor, ldb: int, beta: float, C: torch.Tensor) -> None:
    """FP64 twin of ssfr2k_raw."""
    _validate_strided(A, "A", torch.float64)
    _validate_strided(B, "B", torch.float64)
    _validate(C, "C", torch.float64)
    _dsfr2k_C(handle, transr, uplo, trans, n, k, float(alpha), A.data_ptr(), lda, B.data_ptr(), ldb, float(beta),
              C.data_ptr())


def ssfr2k_raw(handle: Handle, transr: int, uplo: int, trans: int, n: int, k: int, alpha: float,
               A: torch.Tensor, lda: int, B: torch.Tensor, ldb: int, beta: float, C: torch.Tensor) -> None:
    """Low-level ssfr2k: integer enums, all dimensions explicit."""
    _validate_strided(A, "A", torch.float32)
    _validate_strided(B, "B", torch.float32)
    _validate(C, "C")
    _ssfr2k_C(handle, transr, uplo, trans, n, k, float(alpha), A.data_ptr(), lda, B.data_ptr(), ldb, float(beta),
              C.data_ptr())


# ---------------------------------------------------------------------------
# Inverse from the RFP Cholesky factor — curfp/__init__.py:392-404, :661-690
# ---------------------------------------------------------------------------
def spftri_raw(handle: Handle, transr: int, uplo: int, n: int, A: torch.Tensor) -> None:
    """Low-level: overwrite A (the RFP Cholesky factor from spftrf) with inv(C)."""
    _validate(A, "A")
    _spftri_C(handle, transr, uplo, n, A.data_ptr())


def spftri(C: torch.Tensor, *, n: int = None, transr: str = "T", uplo: str = "U") -> None:
    """inv(C) in place from the RFP Cholesky factor produced by spftrf (same
    transr / uplo), as L^-T L^-1 in the same RFP layout."""
    _validate(C, "C")
    if n is None:
        n = _n_from_rfp(C.numel())
    _spftri_C(_get_handle(), _op(transr), _fill(uplo), n, C.data_ptr())


def dpftri_raw(handle: Handle, transr: int, uplo: int, n: int, A: torch.Tensor) -> None:
    """FP64 twin of spftri_raw."""
    _validate(A, "A", torch.float64)
    _dpftri_C(handle, transr, uplo, n, A.data_ptr())


def dpftri(C: torch.Tensor, *, n: int = None, transr: str = "T", uplo: str = "U") -> None:
    """FP64 twin of spftri."""
    _validate(C, "C", torch.float64)
    if n is None:
        n = _n_from_rfp(C.numel())
    _dpftri_C(_get_handle(), _op(transr), _fill(uplo), n, C.data_ptr())


# ---------------------------------------------------------------------------
# Symmetric RFP matrix times a dense block — curfp/__init__.py:1203-1360
# ---------------------------------------------------------------------------
_SIDE_MAP = {"L": SIDE_LEFT, "R": SIDE_RIGHT}


def _side(s: str) -> int:
    try:
        return _SIDE_MAP[s.upper()]
    except KeyError:
        raise ValueError(f"side must be 'L' or 'R', got {s!r}") from None


def ssfmm_raw(handle: Handle, transr: int, uplo: int, side: int, m: int, n: int, alpha: float,
              arf: torch.Tensor, B: torch.Tensor, ldb: int, beta: float, C: torch.Tensor, ldc: int) -> None:
    """Low-level ssfmm: integer enums, B and C column-major m x n with leading
    dimensions ldb, ldc (side L: C = alpha A B + beta C, A m x m; side R:
    C = alpha B A + beta C, A n x n)."""
    _validate(arf, "arf")
    _validate_strided(B, "B", torch.float32)
    _validate_strided(C, "C", torch.float32)
    _ssfmm_C(handle, transr, uplo, side, m, n, float(alpha), arf.data_ptr(), B.data_ptr(), ldb, float(beta),
             C.data_ptr(), ldc)


def dsfmm_raw(handle: Handle, transr: int, uplo: int, side: int, m: int, n: int, alpha: float,
              arf: torch.Tensor, B: torch.Tensor, ldb: int, beta: float, C: torch.Tensor, ldc: int) -> None:
    """FP64 twin of ssfmm_raw."""
    _validate(arf, "arf", torch.float64)
    _validate_strided(B, "B", torch.float64)
    _validate_strided(C, "C", torch.float64)
    _dsfmm_C(handle, transr, uplo, side, m, n, float(alpha), arf.data_ptr(), B.data_ptr(), ldb, float(beta),
             C.data_ptr(), ldc)


def _validate_strided(t: torch.Tensor, name: str, dtype) -> None:
    """A dense operand addressed through an explicit leading dimension: CUDA and
    of the right dtype (a transposed view is fine)."""
    if not t.is_cuda or t.dtype != dtype:
        raise ValueError(f"{name} must be a {dtype} CUDA tensor")


def _sfmm(fn, dtype, arf, B, C, alpha, beta, transr, uplo, side, n_A):
    if not (arf.is_cuda and arf.is_contiguous() and arf.dtype == dtype):
        raise ValueError(f"arf must be a contiguous {dtype} CUDA tensor")
    if B.ndim != 2 or not B.is_cuda or B.dtype != dtype:
        raise ValueError(f"B must be a 2-D {dtype} CUDA tensor, got {tuple(B.shape)}")
    m, n = B.shape
    if n_A is None:
        n_A = _n_from_rfp(arf.numel())
    sd = side.upper()
    if sd == "L" and n_A != m:
        raise ValueError(f"side='L': A is {n_A}x{n_A} but B has {m} rows")
    if sd == "R" and n_A != n:
        raise ValueError(f"side='R': A is {n_A}x{n_A} but B has {n} cols")
    if C is None:
        C = torch.zeros(n, m, dtype=dtype, device=arf.device).t()     # column-major m x n
    if C.shape != (m, n) or not C.is_cuda or C.dtype != dtype:
        raise ValueError(f"C must be a {m}x{n} {dtype} CUDA tensor, got {tuple(C.shape)}")

    def colmajor(t):
        return t.stride(0) == 1 and (t.shape[1] <= 1 or t.stride(1) >= max(t.shape[0], 1))

    def rowmajor(t):
        return t.stride(1) == 1 and (t.shape[0] <= 1 or t.stride(0) >= max(t.shape[1], 1))

    h = _get_handle()
    if colmajor(B) and colmajor(C):
        fn(h, _op(transr), _fill(uplo), _side(sd), m, n, float(alpha), arf.data_ptr(), B.data_ptr(),
           max(B.stride(1), m, 1), float(beta), C.data_ptr(), max(C.stride(1), m, 1))
    elif rowmajor(B) and rowmajor(C):
        # row-major m x n is column-major n x m holding the transpose: C^T = alpha B^T A + beta C^T
        # (or alpha A B^T + beta C^T), i.e. the other side
        fn(h, _op(transr), _fill(uplo), _side("R" if sd == "L" else "L"), n, m, float(alpha), arf.data_ptr(),
           B.data_ptr(), max(B.stride(0), n, 1), float(beta), C.data_ptr(), max(C.stride(0), n, 1))
    else:
        raise ValueError("B and C must both be column-major (F-contiguous) or both row-major")
    return C


def ssfmm(arf: torch.Tensor, B: torch.Tensor, C: torch.Tensor = None, *, alpha: float = 1.0, beta: float = 0.0,
          transr: str = "T", uplo: str = "U", side: str = "L", n_A: int = None) -> torch.Tensor:
    """Symmetric RFP matrix A times a dense block (curfp/__init__.py:1203-1302):
    side='L': C := alpha*A@B + beta*C (A m x m); side='R': C := alpha*B@A + beta*C
    (A n x n).  B, C float32 (m, n), both column-major or both row-major (the
    row-major case runs as the transposed product); C is created (zeros,
    column-major) when not given and returned."""
    return _sfmm(_ssfmm_C, torch.float32, arf, B, C, alpha, beta, transr, uplo, side, n_A)


def dsfmm(arf: torch.Tensor, B: torch.Tensor, C: torch.Tensor = None, *, alpha: float = 1.0, beta: float = 0.0,
          transr: str = "T", uplo: str = "U", side: str = "L", n_A: int = None) -> torch.Tensor:
    """FP64 twin of ssfmm."""
    return _sfmm(_dsfmm_C, torch.float64, arf, B, C, alpha, beta, transr, uplo, side, n_A)


# ---------------------------------------------------------------------------
# Norms of a symmetric RFP matrix — curfp/__init__.py:784-841 (slansf),
# :1700-1720 (dlansf)
# ---------------------------------------------------------------------------
def slansf_raw(handle: Handle, norm: int, transr: int, uplo: int, n: int, C: torch.Tensor) -> float:
    """Low-level norm of the symmetric matrix stored in C (RFP, float32):
    integer norm / transr / uplo codes, explicit n."""
    _validate(C, "C")
    return float(_slansf_C(handle, norm, transr, uplo, n, C.data_ptr()))


def slansf(C: torch.Tensor, norm: str = "1", *, transr: str = "T", uplo: str = "U", n: int = None) -> float:
    """Norm of a symmetric matrix stored in RFP format (float32 tensor of
    n*(n+1)/2 elements; the matrix itself, not its Cholesky factor):
    'M'/'max' max |a_ij|; '1'/'O'/'I' largest absolute column sum (the
    1-norm, equal to the infinity norm); 'F'/'E'/'fro' Frobenius norm.
    Sizes are 64-bit (n >= 65536 works; the reference's cuBLAS reductions
    stop at n(n+1)/2 = INT_MAX)."""
    _validate(C, "C")
    if n is None:
        n = _n_from_rfp(C.numel())
    return float(_slansf_C(_get_handle(), _norm(norm), _op(transr), _fill(uplo), n, C.data_ptr()))


def dlansf_raw(handle: Handle, norm: int, transr: int, uplo: int, n: int, C: torch.Tensor) -> float:
    """FP64 twin of slansf_raw."""
    _validate(C, "C", torch.float64)
    return float(_dlansf_C(handle, norm, transr, uplo, n, C.data_ptr()))


def dlansf(C: torch.Tensor, norm: str = "1", *, transr: str = "T", uplo: str = "U", n: int = None) -> float:
    """FP64 twin of slansf."""
    _validate(C, "C", torch.float64)
    if n is None:
        n = _n_from_rfp(C.numel())
    return float(_dlansf_C(_get_handle(), _norm(norm), _op(transr), _fill(uplo), n, C.data_ptr()))
