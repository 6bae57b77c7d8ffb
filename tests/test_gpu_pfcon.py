"""GPU tests of spfcon / dpfcon and spftrf_rcond / dpftrf_rcond (reciprocal
condition estimate from the RFP Cholesky factor) through the C ABI, following
python_tests/test_spfcon.py: all four transr x uplo layouts, well- and
ill-conditioned SPD matrices against LAPACK ?pocon within the reference's
factor of 3, against the oracle (the same estimator in float64 on the GPU's
own factor; pinned against LAPACK ?pocon in test_oracle.py), the n = 1 and
anorm = 0 answers, the raw API, n across the 128-row (FP32) / 64-row (FP64)
leaves, and the combined factor-and-estimate call."""
import numpy as np
import pytest
import torch

from oracle import rfp_oracle as O

pytestmark = pytest.mark.gpu

VARIANTS = [(t, u) for t in "NT" for u in "LU"]
FACTOR = 3.0


@pytest.fixture(scope="module")
def curfp(cuda):
    import curfp_amd

    return curfp_amd


def make_spd(n, rng, cond):
    if n == 1:
        return np.array([[1.0]])
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.linspace(1.0, cond, n)) @ Q.T
    return (A + A.T) / 2.0


def lapack_rcond(A):
    from scipy.linalg import lapack

    r, info = lapack.dpocon(np.linalg.cholesky(A).T, np.linalg.norm(A, 1))
    assert info == 0
    return float(r)


def run(curfp, A, t, u, dtype):
    n = A.shape[0]
    C = torch.from_numpy(O.trttf(A, t, u).astype(dtype)).cuda()
    fac, con = (curfp.spftrf, curfp.spfcon) if dtype == np.float32 else (curfp.dpftrf, curfp.dpfcon)
    assert fac(C, transr=t, uplo=u) == 0
    anorm = float(np.linalg.norm(A.astype(dtype).astype(np.float64), 1))
    return con(C, anorm, transr=t, uplo=u), C.cpu().numpy().astype(np.float64), anorm


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 16, 32, 64, 128, 129, 256, 300, 1000])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_spfcon_well_conditioned(curfp, n, t, u):
    """python_tests/test_spfcon.py:71-84."""
    A = make_spd(n, np.random.default_rng(n), 10.0)
    got, F, anorm = run(curfp, A, t, u, np.float32)
    ref = lapack_rcond(A)
    assert 1.0 / FACTOR <= got / ref <= FACTOR, (got, ref)
    orc = O.pfcon(n, F, anorm, t, u)
    assert abs(got - orc) <= 1e-4 * orc, (got, orc)


@pytest.mark.parametrize("n", [8, 32, 64, 200])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_spfcon_ill_conditioned(curfp, n, t, u):
    """python_tests/test_spfcon.py:87-101 (cond 1e4)."""
    A = make_spd(n, np.random.default_rng(99 + n), 1e4)
    got, F, anorm = run(curfp, A, t, u, np.float32)
    ref = lapack_rcond(A)
    assert 1.0 / FACTOR <= got / ref <= FACTOR, (got, ref)
    orc = O.pfcon(n, F, anorm, t, u)
    assert abs(got - orc) <= 1e-2 * orc, (got, orc)


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 200])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_dpfcon_matches_oracle(curfp, n, t, u):
    A = make_spd(n, np.random.default_rng(7 * n), 1e3)
    got, F, anorm = run(curfp, A, t, u, np.float64)
    assert abs(got - O.pfcon(n, F, anorm, t, u)) <= 1e-10 * got
    assert 1.0 / FACTOR <= got / lapack_rcond(A) <= FACTOR


def test_spfcon_n1_anorm0_raw_api(curfp):
    """python_tests/test_spfcon.py:107-150."""
    v = 4.0
    C = torch.tensor([v], device="cuda")
    assert curfp.spftrf(C) == 0
    assert abs(curfp.spfcon(C, v) - 1.0) < 1e-6
    assert curfp.spfcon(C, 0.0) == 0.0
    assert curfp.spfcon(torch.zeros(0, device="cuda"), 1.0, n=0) == 0.0
    n = 50
    A = make_spd(n, np.random.default_rng(4), 20.0)
    arf = torch.from_numpy(O.trttf(A, "T", "U").astype(np.float32)).cuda()
    assert curfp.spftrf(arf) == 0
    anorm = float(np.linalg.norm(A, 1))
    assert curfp.spfcon(arf, anorm) == curfp.spfcon_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, n,
                                                         arf, anorm)
    with pytest.raises(RuntimeError):
        curfp.spfcon_raw(curfp._get_handle(), 5, curfp.FILL_UPPER, n, arf, anorm)


@pytest.mark.parametrize("t,u", [("N", "L"), ("T", "U")])
def test_pftrf_rcond(curfp, t, u):
    """slansf + spftrf + spfcon in one call; a non-SPD matrix raises (or
    returns 0 with check=False)."""
    n = 300
    A = make_spd(n, np.random.default_rng(12), 50.0)
    C = torch.from_numpy(O.trttf(A, t, u).astype(np.float32)).cuda()
    got = curfp.spftrf_rcond(C, transr=t, uplo=u)
    assert 1.0 / FACTOR <= got / lapack_rcond(A) <= FACTOR
    D = torch.from_numpy(O.trttf(A, t, u)).cuda()
    got = curfp.dpftrf_rcond(D, transr=t, uplo=u)
    assert 1.0 / FACTOR <= got / lapack_rcond(A) <= FACTOR
    assert np.allclose(O.rfp_lower_factor(D.cpu().numpy(), n, t, u) @ O.rfp_lower_factor(D.cpu().numpy(), n, t, u).T,
                       A, atol=1e-10)
    bad = A.copy()
    bad[7, 7] = -1.0
    C = torch.from_numpy(O.trttf(bad, t, u).astype(np.float32)).cuda()
    with pytest.raises(torch.linalg.LinAlgError):
        curfp.spftrf_rcond(C, transr=t, uplo=u)
    C = torch.from_numpy(O.trttf(bad, t, u).astype(np.float32)).cuda()
    assert curfp.spftrf_rcond(C, transr=t, uplo=u, check=False) == 0.0


def test_spfcon_large(curfp):
    """n = 8192 on the chain's generator (A A^T / k + I): the estimate is a
    lower bound of the true ||A^-1||_1 within the reference's factor, checked
    through ||A^-1 y||_1 / ||y||_1 <= ||A^-1||_1 for the estimator's own
    kind of vector (inv(A) applied with spftrs)."""
    n, k = 8192, 256
    rng = np.random.default_rng(21)
    X = torch.from_numpy((rng.standard_normal((n, k)) / np.sqrt(k)).astype(np.float32)).cuda()
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(X, C)
    curfp.add_to_diagonal(C, 1.0)
    rc = curfp.spftrf_rcond(C)
    Xd = X.double().cpu().numpy()
    A = Xd @ Xd.T + np.eye(n)
    inv_norm = np.linalg.norm(np.linalg.inv(A), 1)
    exact = 1.0 / (inv_norm * np.linalg.norm(A, 1))
    assert exact * (1 - 1e-3) <= rc <= FACTOR * exact, (rc, exact)
