"""GPU tests of ssfmv / dsfmv (y := alpha A x + beta y, A symmetric in RFP)
through the C ABI against the oracle's symmetric multiply with a single
column (float64; pinned in test_oracle.py), following
python_tests/test_ssfmv.py: all four transr x uplo layouts of odd and even n,
beta accumulation, alpha = 0 / beta = 0 (operands not read), the raw API with
positive, strided and negative increments, n across the 128-row blocks, FP64."""
import numpy as np
import pytest
import torch

from oracle import rfp_oracle as O

pytestmark = pytest.mark.gpu

VARIANTS = [(t, u) for t in "NT" for u in "LU"]


@pytest.fixture(scope="module")
def curfp(cuda):
    import curfp_amd

    return curfp_amd


def sym(n, rng):
    A = rng.standard_normal((n, n))
    return (A + A.T) / 2.0


def ref_mv(arf, n, x, y, alpha, beta, t, u):
    return O.sfmm(arf, n, np.asarray(x, np.float64)[:, None], None if y is None else np.asarray(y)[:, None],
                  alpha, beta, t, u, "L")[:, 0]


def bound(n, S, x, y, alpha, beta, eps):
    return 4.0 * n * eps * (abs(alpha) * np.abs(S).sum(axis=1) @ np.abs(x) + abs(beta) * np.abs(y).sum()) + 1e-30


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 15, 16, 17, 32, 33, 64, 65, 127, 128, 129, 300, 1025])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_ssfmv_matches_oracle(curfp, n, t, u):
    """python_tests/test_ssfmv.py:40-98 (alpha random, beta = 0 then beta accumulating)."""
    rng = np.random.default_rng(n * 10 + ord(t) + ord(u))
    S = sym(n, rng).astype(np.float32)
    arf = O.trttf(S, t, u).astype(np.float32)
    x = rng.standard_normal(n).astype(np.float32)
    y0 = rng.standard_normal(n).astype(np.float32)
    alpha, beta = float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.1, 1.5))
    A, xd = torch.from_numpy(arf).cuda(), torch.from_numpy(x).cuda()
    y = curfp.ssfmv(A, xd, alpha=alpha, transr=t, uplo=u)
    ref = ref_mv(arf, n, x, None, alpha, 0.0, t, u)
    assert np.abs(y.cpu().numpy() - ref).sum() <= bound(n, S, x, np.zeros(n), alpha, 0.0, 2.0 ** -24)
    yd = torch.from_numpy(y0).cuda()
    out = curfp.ssfmv(A, xd, yd, alpha=alpha, beta=beta, transr=t, uplo=u)
    assert out.data_ptr() == yd.data_ptr()
    ref = ref_mv(arf, n, x, y0, alpha, beta, t, u)
    assert np.abs(yd.cpu().numpy() - ref).sum() <= bound(n, S, x, y0, alpha, beta, 2.0 ** -24)


@pytest.mark.parametrize("incx,incy", [(1, 1), (2, 3), (-1, 1), (3, -2), (-2, -1)])
def test_sfmv_increments(curfp, incx, incy):
    n = 131
    rng = np.random.default_rng(abs(incx) * 5 + abs(incy))
    for dtype, raw, eps in ((np.float32, curfp.ssfmv_raw, 2.0 ** -24), (np.float64, curfp.dsfmv_raw, 2.0 ** -53)):
        S = sym(n, rng).astype(dtype)
        arf = O.trttf(S, "N", "L").astype(dtype)
        x, y0 = rng.standard_normal(n).astype(dtype), rng.standard_normal(n).astype(dtype)
        xb = np.full(1 + (n - 1) * abs(incx), np.nan, dtype)
        xb[(np.arange(n) * abs(incx))[:: 1 if incx > 0 else -1]] = x
        yb = np.full(1 + (n - 1) * abs(incy), 7.0, dtype)
        yidx = (np.arange(n) * abs(incy))[:: 1 if incy > 0 else -1]
        yb[yidx] = y0
        yd = torch.from_numpy(yb).cuda()
        raw(curfp._get_handle(), curfp.OP_N, curfp.FILL_LOWER, n, 1.5, torch.from_numpy(arf).cuda(),
            torch.from_numpy(xb).cuda(), incx, -0.5, yd, incy)
        got = yd.cpu().numpy()
        ref = ref_mv(arf, n, x, y0, 1.5, -0.5, "N", "L")
        assert np.abs(got[yidx] - ref).sum() <= bound(n, S, x, y0, 1.5, -0.5, eps)
        gaps = np.ones(len(yb), bool)
        gaps[yidx] = False
        assert (got[gaps] == 7.0).all()            # the gaps of y are not written


def test_ssfmv_alpha0_beta0_raw_api_and_errors(curfp):
    """python_tests/test_ssfmv.py:101-125 (raw = Python call); alpha = 0
    reads neither A nor x; beta = 0 does not read y; incx = 0 is invalid."""
    n = 200
    rng = np.random.default_rng(3)
    nan_arf = torch.full((n * (n + 1) // 2,), float("nan"), device="cuda")
    nan_x = torch.full((n,), float("nan"), device="cuda")
    y0 = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
    y = y0.clone()
    curfp.ssfmv(nan_arf, nan_x, y, alpha=0.0, beta=2.0)
    assert torch.equal(y, 2.0 * y0)
    S = sym(n, rng).astype(np.float32)
    A = torch.from_numpy(O.trttf(S, "T", "U").astype(np.float32)).cuda()
    x = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
    y = torch.full((n,), float("nan"), device="cuda")
    curfp.ssfmv(A, x, y, alpha=1.0, beta=0.0)
    assert torch.isfinite(y).all()
    y2 = torch.zeros(n, device="cuda")
    curfp.ssfmv_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, n, 1.0, A, x, 1, 0.0, y2, 1)
    assert torch.equal(y, y2)
    with pytest.raises(RuntimeError):
        curfp.ssfmv_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, n, 1.0, A, x, 0, 0.0, y2, 1)
    with pytest.raises(ValueError):
        curfp.ssfmv(A, x[:-1])


@pytest.mark.parametrize("n", [1, 6, 64, 65, 257])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_dsfmv_matches_oracle(curfp, n, t, u):
    rng = np.random.default_rng(40 + n)
    S = sym(n, rng)
    arf = O.trttf(S, t, u)
    x, y0 = rng.standard_normal(n), rng.standard_normal(n)
    yd = torch.from_numpy(y0.copy()).cuda()
    curfp.dsfmv(torch.from_numpy(arf).cuda(), torch.from_numpy(x).cuda(), yd, alpha=0.7, beta=-1.2, transr=t,
                uplo=u)
    ref = ref_mv(arf, n, x, y0, 0.7, -1.2, t, u)
    assert np.abs(yd.cpu().numpy() - ref).sum() <= bound(n, S, x, y0, 0.7, -1.2, 2.0 ** -53)


@pytest.mark.parametrize("n,t,u", [(65537, "T", "U"), (65536, "N", "L")])
def test_ssfmv_64bit_sizes_closed_form(curfp, n, t, u):
    """n(n+1)/2 above 2^31 entries: A = c J + d I, x_i = 1 + i % 3, so
    y_i = c sum(x) + d x_i (every partial sum exact in FP32)."""
    c, d = 0.5, 2.0
    nt = n * (n + 1) // 2
    assert nt > 2 ** 31
    C = torch.full((nt,), c, device="cuda")
    curfp.add_to_diagonal(C, d, transr=t, uplo=u)
    xi = 1.0 + (np.arange(n) % 3)
    x = torch.from_numpy(xi.astype(np.float32)).cuda()
    y = curfp.ssfmv(C, x, transr=t, uplo=u).cpu().numpy().astype(np.float64)
    ref = c * xi.sum() + d * xi
    assert np.abs(y - ref).max() <= 1e-6 * np.abs(ref).max(), np.abs(y - ref).max()
    del C
    torch.cuda.empty_cache()
