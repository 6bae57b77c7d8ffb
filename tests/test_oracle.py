"""CPU tests: pin the oracle against LAPACK golden vectors, the reference's
hand-written known answers and mathematical invariants.  No GPU needed."""
import math

import numpy as np
import pytest

from oracle import rfp_oracle as O

VARIANTS = [(t, u) for t in "NT" for u in "LU"]
N_PERM = [1, 2, 3, 4, 5, 6, 7, 8, 15, 16, 17, 32, 33, 63, 64, 65, 128, 129, 257]


# --------------------------------------------------------------------------- geometry

@pytest.mark.parametrize("n", N_PERM)
@pytest.mark.parametrize("t,u", VARIANTS)
def test_index_map_matches_lapack_trttf(golden, n, t, u):
    """strttf of A[i, j] = i*n + j + 1 through LAPACK is the index permutation."""
    perm = golden("rfp_perm.npz")[f"n{n}_{t}{u}"].astype(np.int64)
    A = np.arange(1, n * n + 1, dtype=np.int64).reshape(n, n)
    assert np.array_equal(O.trttf(A, t, u), perm)


@pytest.mark.parametrize("n", [1, 2, 5, 6, 17, 64, 65, 129])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_index_map_is_bijection_and_round_trips(n, t, u):
    M = O.lower_index_map(n, t, u)
    i, j = np.tril_indices(n)
    idx = M[i, j]
    assert np.array_equal(np.sort(idx), np.arange(n * (n + 1) // 2))
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    arf = O.trttf(A, t, u)
    back = O.tfttr(arf, n, t, u)
    tri = np.tril(A) if u == "L" else np.triu(A)
    assert np.array_equal(back, tri)


def test_hand_layouts_from_reference_tests():
    # n=2, TRANSR=N, UPLO=L (case 5): ARF = {A(1,1), A(0,0), A(1,0)}
    # (python_tests/test_spftrf.py:4-6, python_tests/test_ssfrk.py:84-111)
    M = O.lower_index_map(2, "N", "L")
    assert (M[1, 1], M[0, 0], M[1, 0]) == (0, 1, 2)
    # n=3, TRANSR=N, UPLO=L (case 1): {A00, A10, A20, A22, A11, A21}
    # (python_tests/test_spftrf.py:8-11)
    M = O.lower_index_map(3, "N", "L")
    assert [M[0, 0], M[1, 0], M[2, 0], M[2, 2], M[1, 1], M[2, 1]] == [0, 1, 2, 3, 4, 5]
    # LAPACK documentation example, n=6, TRANSR=T, UPLO=U (rows of the k x (n+1) array):
    #   03 13 23 33 00 01 02 / 04 14 24 34 44 11 12 / 05 15 25 35 45 55 22
    rows = [["03", "13", "23", "33", "00", "01", "02"],
            ["04", "14", "24", "34", "44", "11", "12"],
            ["05", "15", "25", "35", "45", "55", "22"]]
    A = np.arange(36).reshape(6, 6)
    arf = O.trttf(A, "T", "U")
    R = arf.reshape(7, 3).T  # column-major k x (n+1)
    for r in range(3):
        for c in range(7):
            i, j = int(rows[r][c][0]), int(rows[r][c][1])
            assert R[r, c] == A[i, j]


@pytest.mark.parametrize("n", [1, 2, 3, 6, 7, 64, 65])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_diag_indices_closed_form(n, t, u):
    M = O.lower_index_map(n, t, u)
    assert np.array_equal(O.diag_indices(n, t, u), M[np.arange(n), np.arange(n)])


def test_planted_geometry_mistake_is_caught(golden, monkeypatch):
    """A plausible slip (v21 read as a column view instead of a row view, i.e.
    a transposed off-diagonal block) must fail the golden comparison."""
    real = O.views

    def wrong(n, transr="T", uplo="U"):
        n1, n2, ld, v11, v21, v22 = real(n, transr, uplo)
        return n1, n2, ld, v11, (v21[0], v21[2], v21[1]), v22

    monkeypatch.setattr(O, "views", wrong)
    perm = golden("rfp_perm.npz")["n8_TU"].astype(np.int64)
    A = np.arange(1, 65, dtype=np.int64).reshape(8, 8)
    assert not np.array_equal(O.trttf(A, "T", "U"), perm)


# --------------------------------------------------------------------------- arithmetic

CHAIN = [(1, 1), (2, 2), (3, 3), (4, 4), (5, 5), (8, 6), (9, 7), (32, 16), (33, 17),
         (64, 32), (65, 33), (128, 64), (129, 65), (256, 128), (257, 129)]


@pytest.mark.parametrize("n,k", CHAIN)
@pytest.mark.parametrize("t,u", VARIANTS)
def test_oracle_chain_matches_lapack_f32(golden, n, k, t, u):
    g = golden("chain_f32.npz")
    key = f"n{n}_k{k}_{t}{u}"
    A, Cl, Fl, B, Xl = (g[key + s] for s in ("_A", "_C", "_F", "_B", "_X"))
    C = O.sfrk(n, k, 1.0, A, 0.0, np.zeros_like(Cl), t, u, "N")
    C = O.add_to_diagonal(C, 1.0, t, u)
    # the reference's own bar is 1e-3 (python_tests/test_vs_lapack.py:22-23)
    assert np.abs(C - Cl).max() <= 1e-5 * max(1.0, np.abs(Cl).max())
    F, info = O.pftrf(n, C, t, u)
    assert info == 0
    assert np.abs(F - Fl).max() <= 1e-4
    X = O.pftrs(n, F, B, t, u)
    assert np.abs(X - Xl).max() <= 1e-4 * max(1.0, np.abs(Xl).max())


@pytest.mark.parametrize("n,k", [(1, 1), (2, 2), (5, 5), (9, 7), (33, 17), (64, 32), (129, 65)])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_oracle_chain_matches_lapack_f64(golden, n, k, t, u):
    g = golden("chain_f64.npz")
    key = f"n{n}_k{k}_{t}{u}"
    A, Cl, Fl, B, Xl = (g[key + s] for s in ("_A", "_C", "_F", "_B", "_X"))
    C = O.add_to_diagonal(O.sfrk(n, k, 1.0, A, 0.0, np.zeros_like(Cl), t, u, "N"), 1.0, t, u)
    assert np.abs(C - Cl).max() <= 1e-13
    F, info = O.pftrf(n, C, t, u)
    assert info == 0 and np.abs(F - Fl).max() <= 1e-12
    X = O.pftrs(n, F, B, t, u)
    assert np.abs(X - Xl).max() <= 1e-11


@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("n,k", [(7, 3), (16, 5), (33, 9)])
def test_sfrk_exact_small_integers(t, u, n, k):
    """Small-integer A: every product and partial sum is exact in float32, so
    the packed result must equal trttf(A A^T) bit for bit."""
    rng = np.random.default_rng(n * 31 + k)
    A = rng.integers(-4, 5, size=(n, k)).astype(np.float32)
    C = O.sfrk(n, k, 1.0, A, 0.0, np.zeros(n * (n + 1) // 2, np.float32), t, u, "N")
    S = (A.astype(np.int64) @ A.T.astype(np.int64)).astype(np.float32)
    assert np.array_equal(C, O.dense_sym_to_rfp(S, t, u))


def test_sfrk_quick_returns():
    """python_tests/test_ssfrk.py:12-56: alpha=0, beta=1 passes C through;
    alpha=beta=0 zeroes C; neither reads A."""
    C = np.arange(15, dtype=np.float32)
    assert np.array_equal(O.sfrk(5, 3, 0.0, None, 1.0, C, "T", "U", "N"), C)
    assert np.array_equal(O.sfrk(5, 3, 0.0, None, 0.0, C, "T", "U", "N"), np.zeros(15, np.float32))


def test_sfrk_n2_rank1_known_answer():
    """python_tests/test_ssfrk.py:84-111 (case 5, alpha=2, beta=0.5)."""
    C = np.array([4.0, 1.0, 0.5], np.float32)  # {C11, C00, C10}
    A = np.array([[2.0], [3.0]], np.float32)
    out = O.sfrk(2, 1, 2.0, A, 0.5, C, "N", "L", "N")
    assert np.allclose(out, [2 * 9 + 2.0, 2 * 4 + 0.5, 2 * 6 + 0.25])


def test_pftrf_known_answers():
    """python_tests/test_spftrf.py:19-128."""
    for t, u in VARIANTS:
        F, info = O.pftrf(1, np.array([9.0], np.float32), t, u)
        assert info == 0 and abs(F[0] - 3.0) < 1e-6
    F, info = O.pftrf(2, np.array([3.0, 4.0, 2.0], np.float32), "N", "L")
    assert info == 0 and np.allclose(F, [math.sqrt(2.0), 2.0, 1.0], atol=1e-6)
    A = np.array([5.0, 1.0, 2.0, 6.0, 4.0, 0.0], np.float32)
    F, info = O.pftrf(3, A, "N", "L")
    l00 = math.sqrt(5)
    l10, l20 = 1 / l00, 2 / l00
    l11 = math.sqrt(4 - l10 ** 2)
    l21 = (0 - l20 * l10) / l11
    l22 = math.sqrt(6 - l20 ** 2 - l21 ** 2)
    assert info == 0 and np.allclose(F, [l00, l10, l20, l22, l11, l21], atol=1e-6)


def test_pftrf_info_matches_lapack(golden):
    g = golden("info.npz")
    keys = sorted(k[:-5] for k in g.files if k.endswith("_info"))
    for key in keys:
        arf = g[key + "_arf"]
        n = int(key.split("_")[1][1:])
        t, u = key[-2], key[-1]
        _, info = O.pftrf(n, arf, t, u)
        assert info == int(g[key + "_info"]), key


@pytest.mark.parametrize("t,u", VARIANTS)
def test_pftrf_reconstruction_and_residual(t, u):
    n, k = 97, 40
    rng = np.random.default_rng(3)
    A = rng.standard_normal((n, k))
    C = O.add_to_diagonal(O.sfrk(n, k, 1.0, A, 0.0, np.zeros(n * (n + 1) // 2), t, u, "N"), 0.5, t, u)
    S = O.rfp_to_dense_sym(C, n, t, u)
    F, info = O.pftrf(n, C, t, u)
    L = O.rfp_lower_factor(F, n, t, u)
    assert info == 0
    assert np.linalg.norm(L @ L.T - S) / np.linalg.norm(S) < 1e-14
    B = rng.standard_normal((n, 4))
    X = O.pftrs(n, F, B, t, u)
    assert O.dense_relative_residual(S, X, B) < 1e-15


def test_chain_flops_counts():
    f = O.chain_flops(65536, 1024, 64)
    assert abs(f["ssfrk"] - 4.3981e12) / 4.3981e12 < 1e-4
    assert abs(f["spftrf"] - 9.3827e13) / 9.3827e13 < 1e-4
    assert abs(f["spftrs"] - 5.4976e11) / 5.4976e11 < 1e-4


# --------------------------------------------------------------------------- norms

NORMS = ["M", "1", "F"]


def _sym_with_skewed_triangles(n, seed):
    """Symmetric S whose lower-triangle column sums differ from its row sums
    (large entries low in the matrix), so a norm that forgets the mirrored
    triangle is caught."""
    rng = np.random.default_rng(seed)
    L = np.tril(rng.standard_normal((n, n)) * (1.0 + np.arange(n))[:, None])
    return L + np.tril(L, -1).T


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 17, 64, 129, 257])
@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("norm", NORMS)
def test_lansf_matches_lapack_lange(n, t, u, norm):
    from scipy.linalg import lapack

    S = _sym_with_skewed_triangles(n, n)
    arf = O.trttf(S, t, u)
    ref = float(lapack.dlange(norm, S))
    assert abs(O.lansf(arf, n, norm, t, u) - ref) <= 1e-13 * ref


@pytest.mark.parametrize("t,u", VARIANTS)
def test_lansf_closed_forms(t, u):
    n, c, d = 37, 0.25, 3.0
    S = np.full((n, n), c) + d * np.eye(n)
    arf = O.trttf(S, t, u)
    assert O.lansf(arf, n, "M", t, u) == c + d
    assert abs(O.lansf(arf, n, "1", t, u) - (n * c + d)) < 1e-12
    assert abs(O.lansf(arf, n, "F", t, u) - math.sqrt(n * (c + d) ** 2 + (n * n - n) * c * c)) < 1e-12
    arf = O.trttf(np.eye(n), t, u)
    assert (O.lansf(arf, n, "M", t, u), O.lansf(arf, n, "1", t, u)) == (1.0, 1.0)
    assert abs(O.lansf(arf, n, "F", t, u) - math.sqrt(n)) < 1e-12


def test_lansf_n1_and_n0_known_answers():
    """python_tests/test_slansf.py:116-133: n = 0 gives 0; n = 1 gives |v|."""
    for norm in ("M", "1", "O", "I", "F", "E", "fro", "max"):
        assert O.lansf(np.zeros(0), 0, norm) == 0.0
        assert abs(O.lansf(np.array([3.7]), 1, norm) - 3.7) < 1e-15


def test_planted_norm_mistake_is_caught(monkeypatch):
    """Column sums of the stored triangle alone (the mirrored half dropped)
    must fail the LAPACK comparison."""
    from scipy.linalg import lapack

    n = 9
    S = _sym_with_skewed_triangles(n, 4)
    arf = O.trttf(S, "T", "U")
    real = O.rfp_to_dense_sym

    def lower_only(a, m, transr="T", uplo="U"):
        return np.tril(real(a, m, transr, uplo))

    monkeypatch.setattr(O, "rfp_to_dense_sym", lower_only)
    assert abs(O.lansf(arf, n, "1", "T", "U") - lapack.dlange("1", S)) > 1e-3


# --------------------------------------------------------------------------- symmetric multiply

@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 33, 64, 129])
@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("side", ["L", "R"])
def test_sfmm_matches_dense_product(n, t, u, side):
    rng = np.random.default_rng(7 * n + len(side))
    S = _sym_with_skewed_triangles(n, n + 1)
    arf = O.trttf(S, t, u)
    nrhs = 5
    B = rng.standard_normal((n, nrhs) if side == "L" else (nrhs, n))
    C = rng.standard_normal(B.shape)
    ref = 1.5 * (S @ B if side == "L" else B @ S) - 0.5 * C
    assert np.allclose(O.sfmm(arf, n, B, C, 1.5, -0.5, t, u, side), ref, rtol=1e-13, atol=1e-12)


def test_sfmm_known_answers():
    n = 6
    arf = O.trttf(np.eye(n), "N", "L")
    B = np.arange(12.0).reshape(n, 2)
    assert np.array_equal(O.sfmm(arf, n, B, None, 2.0, 0.0, "N", "L", "L"), 2.0 * B)
    C = np.full((n, 2), np.nan)
    assert np.array_equal(O.sfmm(arf, n, B, C, 1.0, 0.0, "N", "L", "L"), B)          # beta = 0: C not read
    Bn = np.full((n, 2), np.nan)
    C = np.ones((n, 2))
    assert np.array_equal(O.sfmm(arf * np.nan, n, Bn, C, 0.0, 3.0, "N", "L", "L"), 3.0 * C)   # alpha = 0


def test_planted_sfmm_mistake_is_caught(monkeypatch):
    """Multiplying by the stored triangle alone must fail the dense check."""
    n = 9
    S = _sym_with_skewed_triangles(n, 2)
    arf = O.trttf(S, "T", "U")
    B = np.ones((n, 3))
    real = O.rfp_to_dense_sym
    monkeypatch.setattr(O, "rfp_to_dense_sym", lambda a, m, transr="T", uplo="U": np.tril(real(a, m, transr, uplo)))
    assert not np.allclose(O.sfmm(arf, n, B, None, 1.0, 0.0, "T", "U", "L"), S @ B)


# --------------------------------------------------------------------------- inverse

@pytest.mark.parametrize("n", [1, 2, 3, 6, 17, 64, 129])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_pftri_matches_lapack_pftri(n, t, u):
    from scipy.linalg import lapack

    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    S = X @ X.T + n * np.eye(n)
    F, info = lapack.dpftrf(n, O.trttf(S, t, u), transr=t, uplo=u)
    assert info == 0
    ref, info = lapack.dpftri(n, F.copy(), transr=t, uplo=u)
    assert info == 0
    got = O.pftri(n, F, t, u)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.linalg.norm(S @ O.rfp_to_dense_sym(got, n, t, u) - np.eye(n)) < 1e-12


def test_planted_pftri_mistake_is_caught(monkeypatch):
    """Inverting the transposed factor (the wrong triangle) must fail the
    LAPACK comparison."""
    from scipy.linalg import lapack

    n = 7
    rng = np.random.default_rng(1)
    X = rng.standard_normal((n, n))
    S = X @ X.T + n * np.eye(n)
    F, _ = lapack.dpftrf(n, O.trttf(S, "N", "L"), transr="N", uplo="L")
    ref, _ = lapack.dpftri(n, F.copy(), transr="N", uplo="L")
    real = O.rfp_lower_factor
    monkeypatch.setattr(O, "rfp_lower_factor", lambda a, m, transr="T", uplo="U": real(a, m, transr, uplo).T)
    assert np.abs(O.pftri(n, F, "N", "L") - ref).max() > 1e-6


# --------------------------------------------------------------------------- rank-2k updates

@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("n,k,trans", [(1, 1, "N"), (6, 2, "T"), (7, 3, "N"), (33, 9, "T"), (64, 5, "N")])
def test_sfr2k_equals_twice_sfrk_for_equal_operands(t, u, n, k, trans):
    """Small-integer A: sfr2k(A, A) = 2 sfrk(A) bit for bit (sfrk is pinned by
    LAPACK ?sfrk golden vectors)."""
    rng = np.random.default_rng(n + 10 * k)
    A = rng.integers(-4, 5, size=(n, k) if trans == "N" else (k, n)).astype(np.float32)
    C0 = rng.integers(-3, 4, size=n * (n + 1) // 2).astype(np.float32)
    got = O.sfr2k(n, k, 1.5, A, A, 0.5, C0, t, u, trans)
    ref = O.sfrk(n, k, 3.0, A, 0.5, C0, t, u, trans)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("n", [1, 2, 5, 16, 17, 65])
def test_sfr2k_polarisation_against_sfrk(t, u, n):
    """X Y^T + Y X^T = ((X+Y)(X+Y)^T - (X-Y)(X-Y)^T) / 2, both sides in RFP."""
    k = 4
    rng = np.random.default_rng(3 * n)
    X, Y = rng.standard_normal((n, k)), rng.standard_normal((n, k))
    C0 = rng.standard_normal(n * (n + 1) // 2)
    got = O.sfr2k(n, k, -0.7, X, Y, 1.25, C0, t, u, "N")
    P = O.sfrk(n, k, -0.35, X + Y, 1.25, C0, t, u, "N")
    ref = O.sfrk(n, k, 0.35, X - Y, 1.0, P, t, u, "N")
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    dense = -0.7 * (X @ Y.T + Y @ X.T) + 1.25 * O.rfp_to_dense_sym(C0, n, t, u)
    assert np.allclose(got, O.trttf(dense, t, u), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("t,u", VARIANTS)
def test_sfr_and_sfr2_dense(t, u):
    n = 11
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    C0 = rng.standard_normal(n * (n + 1) // 2)
    S = O.rfp_to_dense_sym(C0, n, t, u)
    assert np.allclose(O.sfr(n, 2.5, x, C0, t, u), O.trttf(S + 2.5 * np.outer(x, x), t, u), atol=1e-12)
    assert np.allclose(O.sfr2(n, -1.5, x, y, C0, t, u),
                       O.trttf(S - 1.5 * (np.outer(x, y) + np.outer(y, x)), t, u), atol=1e-12)


def test_sfr2k_quick_returns():
    """python_tests/test_ssfr2k.py:81-102: alpha = 0, beta = 1 passes C
    through; alpha = beta = 0 zeroes C; neither reads A or B."""
    C = np.arange(15, dtype=np.float32)
    assert np.array_equal(O.sfr2k(5, 3, 0.0, None, None, 1.0, C), C)
    assert np.array_equal(O.sfr2k(5, 0, 2.0, None, None, 1.0, C), C)
    assert not O.sfr2k(5, 3, 0.0, None, None, 0.0, C).any()
    assert np.array_equal(O.sfr2k(5, 3, 0.0, None, None, 2.0, C), 2 * C)


def test_planted_sfr2k_mistake_is_caught(monkeypatch):
    """Halving the diagonal blocks (adding the symmetric part of only one of
    the two products there) must fail the dense check."""
    n, k = 9, 3
    rng = np.random.default_rng(8)
    X, Y = rng.standard_normal((n, k)), rng.standard_normal((n, k))
    C0 = np.zeros(n * (n + 1) // 2)
    real_put = O._put

    def put_half(arf, v, rows, cols, P, lower=False):
        return real_put(arf, v, rows, cols, P / 2 if lower else P, lower)

    monkeypatch.setattr(O, "_put", put_half)
    got = O.sfr2k(n, k, 1.0, X, Y, 0.0, C0, "T", "U", "N")
    assert not np.allclose(got, O.trttf(X @ Y.T + Y @ X.T, "T", "U"))


# --------------------------------------------------------------------------- condition estimate

def _spd_cond(n, seed, cond):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.linspace(1.0, cond, n)) @ Q.T
    return (A + A.T) / 2.0


@pytest.mark.parametrize("n", [2, 3, 8, 33, 128])
@pytest.mark.parametrize("cond", [10.0, 1e4])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_pfcon_matches_lapack_pocon(n, cond, t, u):
    """LAPACK ?pocon (scipy) on the same matrix; the estimate is a lower bound
    of ||A^-1||_1, so rcond is at least the exact value."""
    from scipy.linalg import lapack

    A = _spd_cond(n, 10 * n + int(cond), cond)
    anorm = np.linalg.norm(A, 1)
    F, info = lapack.dpftrf(n, O.trttf(A, t, u), transr=t, uplo=u)
    assert info == 0
    got = O.pfcon(n, F, anorm, t, u)
    ref, info = lapack.dpocon(np.linalg.cholesky(A).T, anorm)
    assert info == 0
    assert abs(got - ref) <= 1e-10 * ref
    exact = 1.0 / (np.linalg.norm(np.linalg.inv(A), 1) * anorm)
    assert got >= exact * (1 - 1e-12)


def test_pfcon_known_answers():
    """Diagonal A: the estimate is exact, rcond = min(d) / max(d); n = 1
    (python_tests/test_spfcon.py:107-117): rcond = 1; anorm = 0: 0."""
    d = np.array([3.0, 0.5, 7.0, 2.0, 9.0])
    arf = O.trttf(np.diag(np.sqrt(d)), "N", "L")
    assert abs(O.pfcon(5, arf, d.max(), "N", "L") - d.min() / d.max()) <= 1e-15
    assert abs(O.pfcon(1, np.array([np.sqrt(4.0)]), 4.0) - 1.0) <= 1e-15
    assert O.pfcon(5, arf, 0.0, "N", "L") == 0.0
    assert O.pfcon(0, np.zeros(0), 1.0) == 0.0


def test_planted_pfcon_mistake_is_caught(monkeypatch):
    """Applying inv(L) alone (a forgotten backward solve) must fail the
    LAPACK comparison."""
    from scipy.linalg import lapack

    n = 16
    A = _spd_cond(n, 3, 100.0)
    anorm = np.linalg.norm(A, 1)
    F, _ = lapack.dpftrf(n, O.trttf(A, "T", "U"), transr="T", uplo="U")
    ref, _ = lapack.dpocon(np.linalg.cholesky(A).T, anorm)
    real = O._lapack.dtrtrs

    def forward_only(a, b, lower=0, trans=0, **kw):
        return (b, 0) if trans else real(a, b, lower=lower, trans=trans, **kw)

    monkeypatch.setattr(O._lapack, "dtrtrs", forward_only)
    assert abs(O.pfcon(n, F, anorm, "T", "U") - ref) > 1e-3 * ref
