"""GPU tests of ssfmm / dsfmm (symmetric RFP matrix times a dense block)
against the oracle (float64 dense product; the oracle is pinned in
test_oracle.py), following python_tests/test_ssfmm.py: both sides, all four
transr x uplo layouts, the reference's sizes plus sizes around the 128-row
blocks and every workgroup width (1..100 columns), alpha = 0 (A and B not
read) and beta = 0 (C not read), the row-major Python path, FP64."""
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


def sym(n, rng, dtype):
    A = rng.standard_normal((n, n))
    return ((A + A.T) / 2.0).astype(dtype)


def colmajor(x: np.ndarray, dtype) -> torch.Tensor:
    """(m, n) column-major CUDA tensor holding x."""
    return torch.from_numpy(np.ascontiguousarray(x.T.astype(dtype))).cuda().t()


def run_raw(curfp, fn, t, u, side, nA, nrhs, alpha, arf, B, beta, C, dtype):
    """The C-ABI call with column-major B, C (side L: m = nA, n = nrhs; side R:
    m = nrhs, n = nA)."""
    m, n = B.shape
    Bt, Ct = colmajor(B, dtype), colmajor(C, dtype)
    fn(curfp._get_handle(), curfp._op(t), curfp._fill(u), curfp._side(side), m, n, alpha,
       torch.from_numpy(arf.astype(dtype)).cuda(), Bt, m, beta, Ct, m)
    return Ct.cpu().numpy().astype(np.float64)


def bound(nA, S, B, alpha, beta, C, u):
    """normwise: c * nA * u * (|alpha| |S| |B| + |beta| |C|)"""
    return 4.0 * nA * u * (abs(alpha) * np.linalg.norm(np.abs(S)) * np.linalg.norm(np.abs(B))
                           + abs(beta) * np.linalg.norm(C)) + 1e-30


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 16, 17, 33, 64])
@pytest.mark.parametrize("nrhs", [1, 4, 8])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_ssfmm_reference_cases(curfp, n, nrhs, side, t, u):
    rng = np.random.default_rng(n * 1000 + nrhs * 10 + ord(side) + ord(u))
    alpha, beta = float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.0, 1.5))
    S = sym(n, rng, np.float32)
    shape = (n, nrhs) if side == "L" else (nrhs, n)
    B = rng.standard_normal(shape).astype(np.float32)
    C = rng.standard_normal(shape).astype(np.float32)
    arf = O.trttf(S, t, u).astype(np.float32)
    got = run_raw(curfp, curfp.ssfmm_raw, t, u, side, n, nrhs, alpha, arf, B, beta, C, np.float32)
    ref = O.sfmm(arf, n, B, C, alpha, beta, t, u, side)
    np.testing.assert_allclose(got, ref, rtol=1e-3, atol=1e-3)      # the reference's tolerance
    assert np.linalg.norm(got - ref) <= bound(n, S, B, alpha, beta, C, 2.0 ** -24)


@pytest.mark.parametrize("n,nrhs", [(127, 1), (128, 16), (129, 17), (255, 32), (300, 33), (513, 64),
                                    (1000, 100), (1025, 7)])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_ssfmm_block_edges_and_widths(curfp, n, nrhs, side, t, u):
    rng = np.random.default_rng(n + nrhs)
    S = sym(n, rng, np.float32)
    shape = (n, nrhs) if side == "L" else (nrhs, n)
    B = rng.standard_normal(shape).astype(np.float32)
    C = rng.standard_normal(shape).astype(np.float32)
    arf = O.trttf(S, t, u).astype(np.float32)
    got = run_raw(curfp, curfp.ssfmm_raw, t, u, side, n, nrhs, -0.75, arf, B, 0.5, C, np.float32)
    ref = O.sfmm(arf, n, B, C, -0.75, 0.5, t, u, side)
    assert np.linalg.norm(got - ref) <= bound(n, S, B, -0.75, 0.5, C, 2.0 ** -24)


@pytest.mark.parametrize("side", ["L", "R"])
def test_ssfmm_alpha0_beta0_do_not_read(curfp, side):
    n, nrhs = 200, 9
    rng = np.random.default_rng(5)
    shape = (n, nrhs) if side == "L" else (nrhs, n)
    arf_nan = np.full(n * (n + 1) // 2, np.nan, np.float32)
    B_nan = np.full(shape, np.nan, np.float32)
    C = rng.standard_normal(shape).astype(np.float32)
    got = run_raw(curfp, curfp.ssfmm_raw, "T", "U", side, n, nrhs, 0.0, arf_nan, B_nan, 2.0, C, np.float32)
    assert np.array_equal(got, 2.0 * C.astype(np.float64))
    got = run_raw(curfp, curfp.ssfmm_raw, "T", "U", side, n, nrhs, 0.0, arf_nan, B_nan, 0.0, C, np.float32)
    assert not got.any()
    S = sym(n, rng, np.float32)
    arf = O.trttf(S, "T", "U").astype(np.float32)
    B = rng.standard_normal(shape).astype(np.float32)
    got = run_raw(curfp, curfp.ssfmm_raw, "T", "U", side, n, nrhs, 1.0, arf, B, 0.0,
                  np.full(shape, np.nan, np.float32), np.float32)
    ref = O.sfmm(arf, n, B, None, 1.0, 0.0, "T", "U", side)
    assert np.linalg.norm(got - ref) <= bound(n, S, B, 1.0, 0.0, np.zeros(shape), 2.0 ** -24)


@pytest.mark.parametrize("side", ["L", "R"])
def test_ssfmm_python_api_both_layouts(curfp, side):
    n, nrhs = 333, 20
    rng = np.random.default_rng(11)
    S = sym(n, rng, np.float32)
    arf = torch.from_numpy(O.trttf(S, "N", "L").astype(np.float32)).cuda()
    shape = (n, nrhs) if side == "L" else (nrhs, n)
    B = rng.standard_normal(shape).astype(np.float32)
    C = rng.standard_normal(shape).astype(np.float32)
    ref = O.sfmm(arf.cpu().numpy(), n, B, C, 1.25, 0.75, "N", "L", side)
    for make in (colmajor, lambda x, d: torch.from_numpy(np.ascontiguousarray(x.astype(d))).cuda()):
        Bt, Ct = make(B, np.float32), make(C, np.float32)
        out = curfp.ssfmm(arf, Bt, Ct, alpha=1.25, beta=0.75, transr="N", uplo="L", side=side)
        assert out.data_ptr() == Ct.data_ptr()
        got = out.cpu().numpy().astype(np.float64)
        assert np.linalg.norm(got - ref) <= bound(n, S, B, 1.25, 0.75, C, 2.0 ** -24)
    fresh = curfp.ssfmm(arf, colmajor(B, np.float32), transr="N", uplo="L", side=side)
    ref0 = O.sfmm(arf.cpu().numpy(), n, B, None, 1.0, 0.0, "N", "L", side)
    assert np.linalg.norm(fresh.cpu().numpy() - ref0) <= bound(n, S, B, 1.0, 0.0, np.zeros(shape), 2.0 ** -24)
    with pytest.raises(ValueError):
        curfp.ssfmm(arf, colmajor(B[:-1] if side == "L" else B[:, :-1], np.float32), side=side)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (63, 5), (130, 70)])
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_dsfmm_matches_oracle(curfp, n, nrhs, side, t, u):
    rng = np.random.default_rng(3 * n + nrhs)
    S = sym(n, rng, np.float64)
    shape = (n, nrhs) if side == "L" else (nrhs, n)
    B = rng.standard_normal(shape)
    C = rng.standard_normal(shape)
    arf = O.trttf(S, t, u)
    got = run_raw(curfp, curfp.dsfmm_raw, t, u, side, n, nrhs, 0.9, arf, B, -1.1, C, np.float64)
    ref = O.sfmm(arf, n, B, C, 0.9, -1.1, t, u, side)
    assert np.linalg.norm(got - ref) <= bound(n, S, B, 0.9, -1.1, C, 2.0 ** -53)
