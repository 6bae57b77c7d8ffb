"""GPU tests of spftri / dpftri (inverse of an SPD matrix from its RFP
Cholesky factor) through the C ABI: all four transr x uplo layouts of odd
and even n across the 128-row leaves (FP32) / 64-row leaves (FP64), against
the oracle (LAPACK ?trtri on the same factor, float64; pinned against LAPACK
?pftri in test_oracle.py) and C inv(C) = I; at n = 8192 the round trip
inv(C) (C x) = x with ssfmm for both products."""
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


def spd(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, max(n // 2, 1))) / np.sqrt(max(n // 2, 1))
    return X @ X.T + np.eye(n)          # eigenvalues in [1, ~4]: inv(C) well conditioned


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 127, 128, 129, 255, 256, 257, 383, 600, 1025])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_spftri_matches_oracle(curfp, n, t, u):
    S = spd(n, n)
    C = torch.from_numpy(O.trttf(S, t, u).astype(np.float32)).cuda()
    assert curfp.spftrf(C, transr=t, uplo=u) == 0
    F = C.cpu().numpy().astype(np.float64)
    curfp.spftri(C, transr=t, uplo=u)
    got = C.cpu().numpy().astype(np.float64)
    ref = O.pftri(n, F, t, u)
    assert np.linalg.norm(got - ref) <= 2e-6 * np.sqrt(n) * np.linalg.norm(ref), np.linalg.norm(got - ref)
    Sinv = O.rfp_to_dense_sym(got, n, t, u)
    assert np.linalg.norm(S @ Sinv - np.eye(n)) <= 1e-4 * np.sqrt(n)


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 129, 300])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_dpftri_matches_oracle(curfp, n, t, u):
    S = spd(n, 50 + n)
    C = torch.from_numpy(O.trttf(S, t, u)).cuda()
    assert curfp.dpftrf(C, transr=t, uplo=u) == 0
    F = C.cpu().numpy()
    curfp.dpftri(C, transr=t, uplo=u)
    got = C.cpu().numpy()
    ref = O.pftri(n, F, t, u)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.sqrt(n) * np.linalg.norm(ref)
    assert np.linalg.norm(S @ O.rfp_to_dense_sym(got, n, t, u) - np.eye(n)) <= 1e-12 * np.sqrt(n)


def test_spftri_raw_api_and_n0(curfp):
    curfp.spftri(torch.zeros(0, device="cuda"), n=0)
    n = 200
    S = spd(n, 3)
    C1 = torch.from_numpy(O.trttf(S, "T", "U").astype(np.float32)).cuda()
    assert curfp.spftrf(C1) == 0
    C2 = C1.clone()
    curfp.spftri(C1)
    curfp.spftri_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, n, C2)
    assert torch.equal(C1, C2)


@pytest.mark.parametrize("t,u", [("T", "U"), ("N", "L")])
def test_spftri_round_trip_large(curfp, t, u):
    """n = 8192: y = C x, z = inv(C) y with ssfmm on the original and the
    inverted RFP matrices; z = x to FP32 accuracy (C well conditioned)."""
    n, k, nrhs = 8192, 256, 4
    rng = np.random.default_rng(17)
    A = torch.from_numpy((rng.standard_normal((n, k)) / np.sqrt(k)).astype(np.float32)).cuda()
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(A, C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    Cinv = C.clone()
    assert curfp.spftrf(Cinv, transr=t, uplo=u) == 0
    curfp.spftri(Cinv, transr=t, uplo=u)
    x = torch.from_numpy(rng.standard_normal((nrhs, n)).astype(np.float32)).cuda().t()   # column-major
    y = curfp.ssfmm(C, x, transr=t, uplo=u)
    z = curfp.ssfmm(Cinv, y, transr=t, uplo=u)
    err = float(torch.linalg.norm(z - x) / torch.linalg.norm(x))
    assert err <= 1e-4, err
