"""GPU tests of the rank updates ssfr / ssfr2 / ssfr2k and their FP64 twins
through the C ABI, against the oracle (float64 products; pinned in
test_oracle.py against the LAPACK-pinned sfrk), following
python_tests/test_ssfr.py, test_ssfr2.py and test_ssfr2k.py: all four
transr x uplo layouts of odd and even n, alpha = 0 (x, A, B not read),
beta = 0 (C not read) and beta = 1 pass-through, positive, strided and
negative increments, trans = N and T with padded leading dimensions, the
raw API, and k = 1000 / n = 1025 across the 128-row tiles."""
import numpy as np
import pytest
import torch

from oracle import rfp_oracle as O

pytestmark = pytest.mark.gpu

VARIANTS = [(t, u) for t in "NT" for u in "LU"]
N_VALS = [1, 2, 3, 4, 5, 6, 7, 8, 15, 16, 17, 32, 33, 64, 65, 128, 129, 300]


@pytest.fixture(scope="module")
def curfp(cuda):
    import curfp_amd

    return curfp_amd


def sym_rfp(n, rng, t, u, dtype):
    A = rng.standard_normal((n, n))
    return O.trttf((A + A.T) / 2.0, t, u).astype(dtype)


def blas_vec(x: np.ndarray, inc: int, dtype, rng) -> np.ndarray:
    """BLAS storage of x with increment inc (inc < 0: x[0] at the end), the
    gaps filled with NaN so a read between elements shows."""
    n = len(x)
    buf = np.full(1 + (n - 1) * abs(inc), np.nan)
    idx = np.arange(n) * abs(inc)
    buf[idx if inc > 0 else idx[::-1]] = x
    return buf.astype(dtype)


def err_bound(n, k, scale, dtype):
    u = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
    return 8.0 * (k + 2) * u * scale * np.sqrt(n * (n + 1) / 2) + 1e-30


@pytest.mark.parametrize("n", N_VALS)
@pytest.mark.parametrize("t,u", VARIANTS)
def test_ssfr_and_ssfr2_match_oracle(curfp, n, t, u):
    rng = np.random.default_rng(n * 100 + ord(t) + ord(u))
    alpha = float(rng.uniform(0.5, 2.5))
    C0 = sym_rfp(n, rng, t, u, np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    C = torch.from_numpy(C0).cuda()
    curfp.ssfr(C, torch.from_numpy(x).cuda(), alpha=alpha, transr=t, uplo=u)
    ref = O.sfr(n, alpha, x, C0.astype(np.float64), t, u)
    scale = np.abs(ref).max() + 1
    assert np.abs(C.cpu().numpy() - ref).max() <= 4 * 2.0 ** -24 * scale
    C = torch.from_numpy(C0).cuda()
    curfp.ssfr2(C, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), alpha=-alpha, transr=t, uplo=u)
    ref = O.sfr2(n, -alpha, x, y, C0.astype(np.float64), t, u)
    scale = np.abs(ref).max() + 1
    assert np.abs(C.cpu().numpy() - ref).max() <= 4 * 2.0 ** -24 * scale


@pytest.mark.parametrize("incx,incy", [(1, 1), (3, 2), (-1, 1), (2, -3), (-2, -1)])
@pytest.mark.parametrize("t,u", [("N", "L"), ("T", "U")])
def test_sfr_increments(curfp, incx, incy, t, u):
    """BLAS increments through the raw API: element i of x at x[i*incx]
    (incx > 0) or x[(n-1-i)*|incx|] (incx < 0)."""
    n = 37
    rng = np.random.default_rng(abs(incx) * 7 + abs(incy))
    for dtype, raw1, raw2 in ((np.float32, curfp.ssfr_raw, curfp.ssfr2_raw),
                              (np.float64, curfp.dsfr_raw, curfp.dsfr2_raw)):
        C0 = sym_rfp(n, rng, t, u, dtype)
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        xb = torch.from_numpy(blas_vec(x, incx, dtype, rng)).cuda()
        yb = torch.from_numpy(blas_vec(y, incy, dtype, rng)).cuda()
        h, T, U = curfp._get_handle(), curfp._op(t), curfp._fill(u)
        tol = 8 * (2.0 ** -24 if dtype == np.float32 else 2.0 ** -53)
        C = torch.from_numpy(C0).cuda()
        raw1(h, T, U, n, 1.5, xb, incx, C)
        ref = O.sfr(n, 1.5, x.astype(dtype), C0.astype(np.float64), t, u)
        assert np.abs(C.cpu().numpy() - ref).max() <= tol * (np.abs(ref).max() + 1)
        C = torch.from_numpy(C0).cuda()
        raw2(h, T, U, n, 0.5, xb, incx, yb, incy, C)
        ref = O.sfr2(n, 0.5, x.astype(dtype), y.astype(dtype), C0.astype(np.float64), t, u)
        assert np.abs(C.cpu().numpy() - ref).max() <= tol * (np.abs(ref).max() + 1)


def test_sfr_alpha0_n0_and_errors(curfp):
    """python_tests/test_ssfr.py:73-89: alpha = 0 leaves C unchanged and does
    not read x; n = 0 is a no-op; incx = 0 is an invalid value."""
    n = 20
    rng = np.random.default_rng(2)
    C0 = sym_rfp(n, rng, "T", "U", np.float32)
    C = torch.from_numpy(C0).cuda()
    xn = torch.full((n,), float("nan"), device="cuda")
    curfp.ssfr(C, xn, alpha=0.0)
    curfp.ssfr2(C, xn, xn, alpha=0.0)
    assert np.array_equal(C.cpu().numpy(), C0)
    curfp.ssfr(torch.zeros(0, device="cuda"), torch.zeros(0, device="cuda"), n=0)
    with pytest.raises(RuntimeError):
        curfp.ssfr_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, n, 1.0, xn, 0, C)
    with pytest.raises(RuntimeError):
        curfp.ssfr2_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, n, 1.0, xn, 1, xn, 0, C)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 17, 33, 64, 65, 127, 128, 129, 257])
@pytest.mark.parametrize("k", [1, 4, 16, 33])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_ssfr2k_trans_t_matches_oracle(curfp, n, k, t, u):
    """python_tests/test_ssfr2k.py:34-78: row-major (n, k) A and B, trans = T."""
    rng = np.random.default_rng(n * 1000 + k * 10 + ord(u) + ord(t))
    alpha, beta = float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.0, 1.5))
    A = rng.standard_normal((n, k)).astype(np.float32)
    B = rng.standard_normal((n, k)).astype(np.float32)
    C0 = sym_rfp(n, rng, t, u, np.float32)
    C = torch.from_numpy(C0).cuda()
    curfp.ssfr2k(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), C, alpha=alpha, beta=beta,
                 transr=t, uplo=u, trans="T")
    ref = O.sfr2k(n, k, alpha, A.T, B.T, beta, C0.astype(np.float64), t, u, "T")
    got = C.cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-4)       # the reference's tolerance
    assert np.abs(got - ref).max() <= 8 * (k + 2) * 2.0 ** -24 * (np.abs(ref).max() + 2 * alpha * k)


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_sfr2k_padded_leading_dims_f32_f64(curfp, trans, t, u):
    """Column-major operands with lda / ldb beyond the minimum (padding
    filled with NaN), both trans, FP32 and FP64, beta = 0 with a NaN C."""
    n, k = 131, 19
    rng = np.random.default_rng(ord(trans) + ord(t) + ord(u))
    rows, cols = (n, k) if trans == "N" else (k, n)
    lda, ldb = rows + 3, rows + 5
    for dtype, raw in ((np.float32, curfp.ssfr2k_raw), (np.float64, curfp.dsfr2k_raw)):
        A = rng.standard_normal((rows, cols))
        B = rng.standard_normal((rows, cols))
        Ab = np.full((cols, lda), np.nan)
        Ab[:, :rows] = A.T
        Bb = np.full((cols, ldb), np.nan)
        Bb[:, :rows] = B.T
        Ad, Bd = torch.from_numpy(Ab.astype(dtype)).cuda(), torch.from_numpy(Bb.astype(dtype)).cuda()
        C0 = sym_rfp(n, rng, t, u, dtype)
        u_ = 2.0 ** -24 if dtype == np.float32 else 2.0 ** -53
        h = curfp._get_handle()
        for beta, Cin in ((0.75, C0), (0.0, np.full_like(C0, np.nan))):
            C = torch.from_numpy(Cin.copy()).cuda()
            raw(h, curfp._op(t), curfp._fill(u), curfp._op(trans), n, k, -1.25, Ad, lda, Bd, ldb, beta, C)
            ref = O.sfr2k(n, k, -1.25, A.astype(dtype), B.astype(dtype), beta, C0.astype(np.float64), t, u, trans)
            got = C.cpu().numpy()
            assert np.isfinite(got).all()
            assert np.abs(got - ref).max() <= 8 * (k + 2) * u_ * (np.abs(ref).max() + 2.5 * k)


@pytest.mark.parametrize("t,u", VARIANTS)
def test_ssfr2k_passthrough_and_zero(curfp, t, u):
    """python_tests/test_ssfr2k.py:82-102: alpha = 0, beta = 1 leaves C
    unchanged; alpha = 0, beta = 0 zeroes it; neither reads A or B."""
    n = 17
    rng = np.random.default_rng(5)
    C0 = sym_rfp(n, rng, t, u, np.float32)
    A = torch.full((n, 4), float("nan"), device="cuda")
    C = torch.from_numpy(C0).cuda()
    curfp.ssfr2k(A, A, C, alpha=0.0, beta=1.0, transr=t, uplo=u)
    assert np.array_equal(C.cpu().numpy(), C0)
    curfp.ssfr2k(A, A, C, alpha=0.0, beta=0.5, transr=t, uplo=u)
    assert np.array_equal(C.cpu().numpy(), 0.5 * C0)
    curfp.ssfr2k(A, A, C, alpha=0.0, beta=0.0, transr=t, uplo=u)
    assert not C.cpu().numpy().any()


def test_ssfr2k_raw_api_and_python_trans_n(curfp):
    """python_tests/test_ssfr2k.py:105-140 (raw call = Python call), then
    trans = N from row-major (k, n) tensors."""
    n, k = 200, 24
    rng = np.random.default_rng(9)
    A = torch.from_numpy(rng.standard_normal((n, k)).astype(np.float32)).cuda()
    B = torch.from_numpy(rng.standard_normal((n, k)).astype(np.float32)).cuda()
    C0 = torch.from_numpy(sym_rfp(n, rng, "T", "U", np.float32)).cuda()
    C1, C2 = C0.clone(), C0.clone()
    curfp.ssfr2k(A, B, C1, alpha=1.3, beta=0.7)
    curfp.ssfr2k_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, curfp.OP_T, n, k, 1.3, A, k, B, k, 0.7, C2)
    assert torch.equal(C1, C2)
    C3 = C0.clone()
    curfp.ssfr2k(A.t().contiguous(), B.t().contiguous(), C3, alpha=1.3, beta=0.7, trans="N")
    assert torch.allclose(C1, C3, rtol=1e-5, atol=1e-5)
    with pytest.raises(ValueError):
        curfp.ssfr2k(A[:, :5], B, C1)
    with pytest.raises(ValueError):
        curfp.ssfr2k(A[: n - 1].contiguous(), B[: n - 1].contiguous(), C1)


@pytest.mark.parametrize("t,u", [("N", "L"), ("T", "U")])
def test_ssfr2k_large_k_and_math_modes(curfp, t, u):
    """n = 1025, k = 1000 over several 128 x 128 tiles; BF16x3 within FP32
    accuracy, BF16 within BF16 accuracy (|A|, |B| entries ~ 1)."""
    n, k = 1025, 1000
    rng = np.random.default_rng(31)
    A = rng.standard_normal((n, k)).astype(np.float32)
    B = rng.standard_normal((n, k)).astype(np.float32)
    C0 = sym_rfp(n, rng, t, u, np.float32)
    ref = O.sfr2k(n, k, 0.5, A.T, B.T, 1.0, C0.astype(np.float64), t, u, "T")
    Ad, Bd = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    for mode, tol in (("fp32", 1e-5), ("bf16x3", 1e-5), ("bf16", 2e-2)):
        C = torch.from_numpy(C0).cuda()
        with curfp.math_mode(mode):
            curfp.ssfr2k(Ad, Bd, C, alpha=0.5, beta=1.0, transr=t, uplo=u)
        err = np.abs(C.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err <= tol, (mode, err)


@pytest.mark.parametrize("n", [1, 6, 65, 130])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_dsfr_family_matches_oracle(curfp, n, t, u):
    rng = np.random.default_rng(77 + n)
    C0 = sym_rfp(n, rng, t, u, np.float64)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    tol = 16 * 2.0 ** -53
    C = torch.from_numpy(C0).cuda()
    curfp.dsfr(C, torch.from_numpy(x).cuda(), alpha=0.3, transr=t, uplo=u)
    ref = O.sfr(n, 0.3, x, C0, t, u)
    assert np.abs(C.cpu().numpy() - ref).max() <= tol * (np.abs(ref).max() + 1)
    C = torch.from_numpy(C0).cuda()
    curfp.dsfr2(C, torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), alpha=0.3, transr=t, uplo=u)
    ref = O.sfr2(n, 0.3, x, y, C0, t, u)
    assert np.abs(C.cpu().numpy() - ref).max() <= tol * (np.abs(ref).max() + 1)
    k = 7
    A, B = rng.standard_normal((n, k)), rng.standard_normal((n, k))
    C = torch.from_numpy(C0).cuda()
    curfp.dsfr2k(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), C, alpha=-0.6, beta=0.4, transr=t, uplo=u)
    ref = O.sfr2k(n, k, -0.6, A.T, B.T, 0.4, C0, t, u, "T")
    assert np.abs(C.cpu().numpy() - ref).max() <= 8 * (k + 2) * 2.0 ** -53 * (np.abs(ref).max() + 2 * k)


@pytest.mark.parametrize("trans,k", [("T", 520), ("N", 77), ("T", 33)])
@pytest.mark.parametrize("t,u", [("N", "L"), ("T", "U")])
def test_ssfr2k_bf16_row_blocks(curfp, trans, k, t, u):
    """n >= 2048 in the BF16 modes: op(A), op(B) converted once to BF16 row
    blocks (k padded to a multiple of 8), the updates on the BF16 row-block
    kernels.  BF16x3 within FP32 accuracy; BF16 equal to the float64 product
    of the BF16-rounded operands to FP32 accumulation accuracy."""
    n = 4100
    rng = np.random.default_rng(k + ord(trans))
    shape = (n, k) if trans == "T" else (k, n)      # row-major tensors
    A = rng.standard_normal(shape).astype(np.float32)
    B = rng.standard_normal(shape).astype(np.float32)
    C0 = sym_rfp(n, rng, t, u, np.float32)
    At, Bt = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    # the oracle takes the logical matrices: (k, n) for trans T, (n, k) for N — A.T of either tensor
    ref = O.sfr2k(n, k, 0.5, A.T, B.T, 0.75, C0.astype(np.float64), t, u, trans)
    C = torch.from_numpy(C0).cuda()
    with curfp.math_mode("bf16x3"):
        curfp.ssfr2k(At, Bt, C, alpha=0.5, beta=0.75, transr=t, uplo=u, trans=trans)
    err = np.abs(C.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err <= 1e-5, err
    bf = lambda X: torch.from_numpy(X).to(torch.bfloat16).double().numpy()
    ref16 = O.sfr2k(n, k, 0.5, bf(A).T, bf(B).T, 0.75, C0.astype(np.float64), t, u, trans)
    C = torch.from_numpy(C0).cuda()
    with curfp.math_mode("bf16"):
        curfp.ssfr2k(At, Bt, C, alpha=0.5, beta=0.75, transr=t, uplo=u, trans=trans)
    err = np.abs(C.cpu().numpy() - ref16).max() / np.abs(ref16).max()
    assert err <= 1e-5, err
