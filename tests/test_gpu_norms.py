"""GPU tests of slansf / dlansf (norms of a symmetric RFP matrix) against the
oracle (float64 on the dense matrix; the oracle itself is pinned against
LAPACK ?lange in test_oracle.py), following python_tests/test_slansf.py:
all four transr x uplo layouts of odd and even n, the three norms and their
aliases, n = 0 / 1, the raw API, and at n = 65536 / 65537 (n(n+1)/2 above
INT_MAX, where the reference's int-sized reductions give up) the closed
forms of c J + d I with a spike planted at the last packed entry."""
import math

import numpy as np
import pytest
import torch

from oracle import rfp_oracle as O

pytestmark = pytest.mark.gpu

VARIANTS = [(t, u) for t in "NT" for u in "LU"]
SIZES = [1, 2, 3, 4, 5, 8, 16, 32, 64, 127, 128, 129, 255, 256, 257, 383, 1000, 1025]


@pytest.fixture(scope="module")
def curfp(cuda):
    import curfp_amd

    return curfp_amd


def skewed_sym(n, seed, dtype):
    """Symmetric matrix whose stored-triangle column sums differ from its row
    sums, so a 1-norm that drops the mirrored half is caught."""
    rng = np.random.default_rng(seed)
    L = np.tril(rng.standard_normal((n, n)) * (1.0 + np.arange(n))[:, None])
    return (L + np.tril(L, -1).T).astype(dtype)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("t,u", VARIANTS)
def test_slansf_matches_oracle(curfp, n, t, u):
    S = skewed_sym(n, n, np.float32)
    arf = O.trttf(S, t, u).astype(np.float32)
    C = torch.from_numpy(arf).cuda()
    for norm in ("M", "1", "F"):
        ref = O.lansf(arf.astype(np.float64), n, norm, t, u)
        got = curfp.slansf(C, norm, transr=t, uplo=u)
        if norm == "M":
            assert got == np.float32(ref), (norm, got, ref)
        else:
            # FP32 partial sums of at most 128 terms, FP64 beyond: ~1e-6
            assert abs(got - ref) <= 2e-6 * ref, (norm, got, ref)


@pytest.mark.parametrize("n", [1, 5, 128, 129, 257, 1000])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_dlansf_matches_oracle(curfp, n, t, u):
    S = skewed_sym(n, 100 + n, np.float64)
    arf = O.trttf(S, t, u)
    C = torch.from_numpy(arf).cuda()
    for norm in ("M", "1", "F"):
        ref = O.lansf(arf, n, norm, t, u)
        got = curfp.dlansf(C, norm, transr=t, uplo=u)
        assert abs(got - ref) <= 1e-13 * ref, (norm, got, ref)


def test_slansf_aliases_n0_n1_and_raw_api(curfp):
    """python_tests/test_slansf.py:116-175."""
    for norm in ("M", "max", "1", "O", "I", "F", "E", "fro"):
        assert curfp.slansf(torch.zeros(0, device="cuda"), norm, n=0) == 0.0
        assert abs(curfp.slansf(torch.tensor([3.7], device="cuda"), norm) - 3.7) < 1e-6
        assert curfp.dlansf(torch.tensor([-2.5], device="cuda", dtype=torch.float64), norm) == 2.5
    n = 300
    arf = O.trttf(skewed_sym(n, 9, np.float32), "T", "U").astype(np.float32)
    C = torch.from_numpy(arf).cuda()
    one = curfp.slansf(C, "1")
    assert curfp.slansf(C, "O") == one and curfp.slansf(C, "I") == one
    fro = curfp.slansf(C, "F")
    assert curfp.slansf(C, "E") == fro and curfp.slansf(C, "fro") == fro
    h = curfp._get_handle()
    for s, code in (("M", curfp.NORM_MAX), ("1", curfp.NORM_ONE), ("F", curfp.NORM_FRO)):
        assert curfp.slansf_raw(h, code, curfp.OP_T, curfp.FILL_UPPER, n, C) == curfp.slansf(C, s)
    with pytest.raises(ValueError):
        curfp.slansf(C, "X")
    with pytest.raises(RuntimeError):
        curfp.slansf_raw(h, 7, curfp.OP_T, curfp.FILL_UPPER, n, C)


@pytest.mark.parametrize("n,t,u", [(65536, "N", "L"), (65537, "T", "U")])
def test_slansf_64bit_sizes_closed_form(curfp, n, t, u):
    """c J + d I: max = c + d, 1-norm = n c + d, Frobenius =
    sqrt(n (c + d)^2 + (n^2 - n) c^2); then a spike at the last packed entry
    (index n(n+1)/2 - 1 > 2^31) is the max."""
    c, d = 0.5, 2.0
    nt = n * (n + 1) // 2
    assert nt > 2 ** 31
    C = torch.full((nt,), c, device="cuda")
    curfp.add_to_diagonal(C, d, transr=t, uplo=u)
    assert curfp.slansf(C, "M", transr=t, uplo=u) == c + d
    one = curfp.slansf(C, "1", transr=t, uplo=u)
    assert abs(one - (n * c + d)) <= 1e-6 * (n * c + d), one
    fro_ref = math.sqrt(n * (c + d) ** 2 + (n * n - n) * c * c)
    fro = curfp.slansf(C, "F", transr=t, uplo=u)
    assert abs(fro - fro_ref) <= 1e-6 * fro_ref, fro
    C[nt - 1] = 1.0e4
    assert curfp.slansf(C, "M", transr=t, uplo=u) == 1.0e4
    del C
    torch.cuda.empty_cache()
