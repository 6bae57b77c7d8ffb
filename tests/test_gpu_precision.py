"""GPU tests of the precision modes and the FP64 iterative refinement:
BF16 / split-BF16 ssfrk and spftrf updates, the FP64 generator residual, and
refinement down to a relative residual of 1e-6 (the north-star target) and
below, checked against float64 dense arithmetic on the host."""
import math

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


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even to BF16, returned as float64 (via torch)."""
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(torch.bfloat16).double().numpy()


def problem(n, k, nrhs, seed, lam=1.0):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((n, k)) / math.sqrt(k)).astype(np.float32)
    B = rng.standard_normal((n, nrhs)).astype(np.float32)
    return A, B


@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("n,k", [(600, 130), (1001, 257)])
def test_ssfrk_bf16_input_is_exact_products(curfp, t, u, n, k):
    """A in BF16: every product is exact in FP32, so C matches the float64
    product of the BF16 values to FP32 accumulation accuracy."""
    A, _ = problem(n, k, 1, seed=n + k)
    Ab = torch.from_numpy(A).cuda().to(torch.bfloat16)
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(Ab, C, transr=t, uplo=u)
    A64 = Ab.double().cpu().numpy()
    S = A64 @ A64.T
    Cd = O.rfp_to_dense_sym(C.cpu().numpy().astype(np.float64), n, t, u)
    u32 = np.finfo(np.float32).eps
    assert np.linalg.norm(Cd - S) <= 2 * math.sqrt(k) * u32 * np.linalg.norm(S) + 1e-30


@pytest.mark.parametrize("n,k,ld,t,u,beta", [(8200, 136, 136, "T", "U", 0.0), (8193, 520, 520, "N", "L", 0.5),
                                              (4500, 1000, 1000, "T", "U", 0.0), (4097, 520, 520, "N", "L", 0.5),
                                              (4300, 1004, 1008, "T", "L", 0.0), (4100, 64, 64, "N", "U", -1.0),
                                              (3001, 1000, 1000, "N", "U", 0.25), (3333, 136, 136, "T", "L", 0.0)])
def test_ssfrk_bf16_input_large_tiles(curfp, n, k, ld, t, u, beta):
    """Sizes whose RFP blocks reach the BF16 tile kernels for larger problems:
    256 x 256 (blocks >= 4096 rows, n >= 8192) and 128 x 128 with direct-to-LDS
    loads when K % 8 == 0, the register-staged kernel otherwise (k = 1004
    inside rows of 1008); partial tiles at every edge, beta != 0.
    Products of BF16 values are exact in FP32, so C matches the float64
    result to FP32 accumulation accuracy."""
    rng = np.random.default_rng(n + k)
    A = (rng.standard_normal((n, ld)) / math.sqrt(k)).astype(np.float32)
    Ab = torch.from_numpy(A).cuda().to(torch.bfloat16)
    C0 = torch.from_numpy(rng.standard_normal(n * (n + 1) // 2).astype(np.float32)).cuda()
    C = C0.clone()
    curfp.ssfrk(Ab, C, alpha=0.75, beta=beta, transr=t, uplo=u, n=n, k=k, lda=ld)
    A64 = Ab[:, :k].double()
    S = (0.75 * (A64 @ A64.t())).cpu().numpy()
    if beta != 0.0:
        S += beta * O.rfp_to_dense_sym(C0.cpu().numpy().astype(np.float64), n, t, u)
    Cd = O.rfp_to_dense_sym(C.cpu().numpy().astype(np.float64), n, t, u)
    u32 = np.finfo(np.float32).eps
    assert np.linalg.norm(Cd - S) <= 4 * math.sqrt(k) * u32 * np.linalg.norm(S)


@pytest.mark.parametrize("n,k", [(777, 300), (3000, 300), (2500, 77)])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_ssfrk_math_modes(curfp, t, u, n, k):
    """FP32 A in the three math modes; from n = 2048 (k >= 32) the BF16 modes
    convert A once to BF16 row blocks (k = 77 padded to 80)."""
    A, _ = problem(n, k, 1, seed=11)
    Ad = torch.from_numpy(A).cuda()
    S = A.astype(np.float64) @ A.astype(np.float64).T
    Sb = bf16_round(A) @ bf16_round(A).T
    u32 = np.finfo(np.float32).eps
    res = {}
    for mode in ("fp32", "bf16", "bf16x3"):
        C = torch.empty(n * (n + 1) // 2, device="cuda")
        with curfp.math_mode(mode):
            curfp.ssfrk(Ad, C, transr=t, uplo=u)
        res[mode] = O.rfp_to_dense_sym(C.cpu().numpy().astype(np.float64), n, t, u)
    nS = np.linalg.norm(S)
    # BF16 mode = exact product of the BF16-rounded inputs
    assert np.linalg.norm(res["bf16"] - Sb) <= 4 * math.sqrt(k) * u32 * np.linalg.norm(Sb)
    # split-BF16 recovers ~FP32 accuracy; FP32 is FP32
    assert np.linalg.norm(res["bf16x3"] - S) <= 1e-5 * nS
    assert np.linalg.norm(res["fp32"] - S) <= 4 * math.sqrt(k) * u32 * nS
    assert np.linalg.norm(res["bf16"] - S) > np.linalg.norm(res["bf16x3"] - S)


@pytest.mark.parametrize("mode", ["bf16", "bf16x3"])
@pytest.mark.parametrize("n,t,u", [(2048, "T", "U"), (2047, "N", "L"), (1500, "N", "U"), (1501, "T", "L"),
                                   (6001, "T", "U"), (5000, "N", "L")])
def test_spftrf_math_modes_backward_error(curfp, mode, n, t, u):
    """Factor with BF16 / split-BF16 updates: ||L L^T - C|| / ||C|| at the
    precision of the mode (BF16 ~ 2^-8 scale, split-BF16 ~ FP32 scale)."""
    k = 256
    A, _ = problem(n, k, 1, seed=3)
    Ad = torch.from_numpy(A).cuda()
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(Ad, C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    S = O.rfp_to_dense_sym(C.cpu().numpy().astype(np.float64), n, t, u)
    with curfp.math_mode(mode):
        assert curfp.spftrf(C, transr=t, uplo=u) == 0
    L = O.rfp_lower_factor(C.cpu().numpy().astype(np.float64), n, t, u)
    err = np.linalg.norm(L @ L.T - S) / np.linalg.norm(S)
    bound = 2.0 ** -7 if mode == "bf16" else 1e-5
    assert err <= bound, err


def test_gen_residual_fp64(curfp):
    n, k, nrhs = 9000, 150, 7
    A, B = problem(n, k, nrhs, seed=5)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((n, nrhs))
    lam = 0.3
    Ad = torch.from_numpy(A).cuda()
    Xc = torch.from_numpy(X.T.copy()).cuda().t()
    Bc = torch.from_numpy(B.astype(np.float64).T.copy()).cuda().t()
    R = curfp.gen_residual(Ad, Xc, Bc, lam=lam).cpu().numpy()
    A64 = A.astype(np.float64)
    Rref = B.astype(np.float64) - (A64 @ (A64.T @ X) + lam * X)
    assert np.abs(R - Rref).max() <= 1e-12 * np.abs(Rref).max() + 1e-12


@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_refinement_reaches_target(curfp, mode, t, u):
    """BF16 factor + FP64 refinement reaches ||B - C X|| / (||C|| ||X||) <= 1e-6,
    verified independently in float64 on the host."""
    n, k, nrhs, lam = 2047 if t == "N" else 2048, 256, 16, 1.0
    A, B = problem(n, k, nrhs, seed=21)
    Ad = torch.from_numpy(A).cuda()
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    with curfp.math_mode(mode):
        curfp.ssfrk(Ad, C, transr=t, uplo=u)
        curfp.add_to_diagonal(C, lam, transr=t, uplo=u)
        assert curfp.spftrf(C, transr=t, uplo=u) == 0
    from curfp_amd import refine

    X, info = refine.refine_solve(C, Ad, lam, torch.from_numpy(B).cuda(), tol=1e-6, transr=t, uplo=u)
    S = A.astype(np.float64) @ A.astype(np.float64).T + lam * np.eye(n)
    r = O.dense_relative_residual(S, X.cpu().numpy(), B.astype(np.float64))
    assert info["rel_residual"] <= 1e-6 and r <= 1.5e-6, (info, r)
    assert abs(r - info["rel_residual"]) <= 0.05 * r + 1e-16
    if mode == "fp32":
        assert info["solves"] <= 2


def test_refinement_to_fp64_accuracy(curfp):
    """Driving the tolerance down shows the FP64 accumulation: the error
    against a float64 dense solve falls far below FP32 roundoff."""
    n, k, nrhs, lam = 1500, 200, 4, 0.5
    A, B = problem(n, k, nrhs, seed=8)
    Ad = torch.from_numpy(A).cuda()
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(Ad, C)
    curfp.add_to_diagonal(C, lam)
    curfp.spftrf(C)
    from curfp_amd import refine

    X, info = refine.refine_solve(C, Ad, lam, torch.from_numpy(B).cuda(), tol=1e-14, max_iter=20)
    S = A.astype(np.float64) @ A.astype(np.float64).T + lam * np.eye(n)
    Xref = np.linalg.solve(S, B.astype(np.float64))
    err = np.linalg.norm(X.cpu().numpy() - Xref) / np.linalg.norm(Xref)
    assert info["rel_residual"] <= 1e-14 and err <= 1e-12, (info, err)


def _mismatched_factor(curfp, n, k, nrhs, seed, t="T", u="U"):
    """The factor of C~ = A A^T + 0.01 I for the system C = A A^T + 0.05 I
    with rank(A A^T) = k < n: I - C~^{-1} C has the eigenvalue
    -(0.05 - 0.01)/0.01 = -4 on the null space of A^T, so classical
    refinement diverges, while C~^{-1} C has its spectrum in [1, 5]."""
    A, B = problem(n, k, nrhs, seed=seed)
    Ad = torch.from_numpy(A).cuda()
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(Ad, C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 0.01, transr=t, uplo=u)
    assert curfp.spftrf(C, transr=t, uplo=u) == 0
    S = A.astype(np.float64) @ A.astype(np.float64).T + 0.05 * np.eye(n)
    return A, B, Ad, C, S


@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("nrhs", [1, 4, 33])
def test_refinement_krylov_when_classical_diverges(curfp, t, u, nrhs):
    """Classical refinement diverges on this pair (see _mismatched_factor); the
    driver switches to flexible GMRES preconditioned by the factor and reaches
    1e-6 and 1e-12, verified in float64 on the host, with the reported residual
    matching the host's."""
    n, k = 1200 if t == "T" else 1201, 100
    A, B, Ad, C, S = _mismatched_factor(curfp, n, k, nrhs, seed=12 + nrhs, t=t, u=u)
    from curfp_amd import refine

    for tol in (1e-6, 1e-12):
        X, info = refine.refine_solve(C, Ad, 0.05, torch.from_numpy(B).cuda(), tol=tol, max_iter=60, transr=t,
                                      uplo=u)
        r = O.dense_relative_residual(S, X.cpu().numpy(), B.astype(np.float64))
        assert info["rel_residual"] <= tol and r <= 1.5 * tol, (tol, info, r)
        assert abs(r - info["rel_residual"]) <= 0.05 * r + 1e-16
        assert info["solves"] <= (30 if tol == 1e-6 else 50), info


def test_refinement_returns_best_iterate(curfp):
    """With too few solves to converge the driver returns its best iterate:
    the residual it reports is the host's residual of the X it returns, and
    no worse than that of the first solve alone."""
    n, k, nrhs = 1200, 100, 4
    A, B, Ad, C, S = _mismatched_factor(curfp, n, k, nrhs, seed=12)
    from curfp_amd import refine

    Bd = torch.from_numpy(B).cuda()
    X1, i1 = refine.refine_solve(C, Ad, 0.05, Bd, tol=1e-6, max_iter=1)
    r1 = O.dense_relative_residual(S, X1.cpu().numpy(), B.astype(np.float64))
    for max_iter in (2, 3, 4, 6):
        X, info = refine.refine_solve(C, Ad, 0.05, Bd, tol=1e-6, max_iter=max_iter)
        r = O.dense_relative_residual(S, X.cpu().numpy(), B.astype(np.float64))
        assert bool(torch.isfinite(X).all())
        assert info["solves"] <= max_iter
        assert abs(r - info["rel_residual"]) <= 0.05 * r + 1e-16, (max_iter, info, r)
        assert r <= r1 * (1 + 1e-9), (max_iter, r, r1)


def test_krr_solve_bf16_policy(curfp):
    """krr_solve with math='bf16' and a refinement tolerance at conditionings
    (kappa ~ 2.5e3 and 2.5e4) where classical refinement of a BF16 factor
    stalls or the BF16 factor is not even positive definite: the result
    reaches 1e-6 (flexible GMRES on the BF16 factor, or the split-BF16
    fallback), checked in float64 on the host, and no LinAlgError escapes."""
    from curfp_amd.autograd import krr_solve

    n, k, nrhs = 2048, 128, 8
    rng = np.random.default_rng(77)
    A = (rng.standard_normal((n, k)) / math.sqrt(k)).astype(np.float32)
    A = bf16_round(A).astype(np.float32)
    Y = rng.standard_normal((n, nrhs)).astype(np.float32)
    Ad, Yd = torch.from_numpy(A).cuda(), torch.from_numpy(Y).cuda()
    for lam in (1e-2, 1e-3):
        X = krr_solve(Ad, lam, Yd, math="bf16", refine_tol=1e-6, max_iter=60)
        S = A.astype(np.float64) @ A.astype(np.float64).T + lam * np.eye(n)
        r = O.dense_relative_residual(S, X.detach().cpu().numpy(), Y.astype(np.float64))
        assert r <= 1.5e-6, (lam, r)


@pytest.mark.parametrize("t,u", VARIANTS)
@pytest.mark.parametrize("nrhs", [1, 17, 64, 100])
def test_refinement_bf16_solves(curfp, t, u, nrhs):
    """Refinement under the BF16 math mode runs the sweeps' off-diagonal
    products in BF16.  Checked: (1) the first solve is a solve with the
    computed factor F = L L^T to BF16 backward error — ||F X0 - B|| /
    (||F|| ||X0||) within a small multiple of 2^-8 — and differs from the FP32
    sweep's X0 (so the BF16 products ran); (2) the refinement still reaches
    1e-6, verified in float64 on the host, within 4 solves."""
    n, k, lam = 1000 if t == "N" else 1153, 200, 1.0
    A, B = problem(n, k, nrhs, seed=33 + nrhs)
    Ad = torch.from_numpy(A).cuda()
    Bd = torch.from_numpy(B).cuda()
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    from curfp_amd import refine

    with curfp.math_mode("bf16"):
        curfp.ssfrk(Ad, C, transr=t, uplo=u)
        curfp.add_to_diagonal(C, lam, transr=t, uplo=u)
        assert curfp.spftrf(C, transr=t, uplo=u) == 0
        X0b, _ = refine.refine_solve(C, Ad, lam, Bd, tol=0.0, max_iter=1, transr=t, uplo=u)
        X, info = refine.refine_solve(C, Ad, lam, Bd, tol=1e-6, transr=t, uplo=u)
    X0f, _ = refine.refine_solve(C, Ad, lam, Bd, tol=0.0, max_iter=1, transr=t, uplo=u)
    L = O.rfp_lower_factor(C.cpu().numpy().astype(np.float64), n, t, u)
    F = L @ L.T
    x0b, x0f = X0b.cpu().numpy(), X0f.cpu().numpy()
    back = np.linalg.norm(F @ x0b - B) / (np.linalg.norm(F) * np.linalg.norm(x0b))
    assert back <= 4 * 2.0 ** -8, back
    d = np.linalg.norm(x0b - x0f) / np.linalg.norm(x0f)
    assert 1e-7 <= d <= 5e-2, d
    S = A.astype(np.float64) @ A.astype(np.float64).T + lam * np.eye(n)
    r = O.dense_relative_residual(S, X.cpu().numpy(), B.astype(np.float64))
    assert info["rel_residual"] <= 1e-6 and r <= 1.5e-6, (info, r)
    assert info["solves"] <= 4, info


@pytest.mark.parametrize("mode", ["fp32", "bf16"])
@pytest.mark.parametrize("t,u", [("N", "L"), ("T", "U")])
def test_refinement_general_rfp(curfp, mode, t, u):
    """refine_solve_rfp: a general SPD RFP matrix (no generator) in FP64, its
    FP32 factor in the given math mode, the residual from dsfmm; the relative
    residual is checked independently on the host in float64."""
    from curfp_amd.refine import refine_solve_rfp

    n, k, nrhs = 1500, 700, 3
    rng = np.random.default_rng(5)
    G = rng.standard_normal((n, k)) / math.sqrt(k)
    S = G @ G.T + 0.05 * np.eye(n)
    C64 = torch.from_numpy(O.trttf(S, t, u)).cuda()
    F = C64.to(torch.float32)
    with curfp.math_mode(mode):
        assert curfp.spftrf(F, transr=t, uplo=u) == 0
    B = rng.standard_normal((n, nrhs))
    for tol in (1e-6, 1e-12):
        X, info = refine_solve_rfp(F, C64, torch.from_numpy(B).cuda(), tol=tol, transr=t, uplo=u)
        Xh = X.cpu().numpy()
        res = np.linalg.norm(B - S @ Xh) / (np.linalg.norm(S) * np.linalg.norm(Xh))
        assert res <= tol, (tol, res, info)
        assert abs(res - info["rel_residual"]) <= 0.01 * res + 1e-17
    x, info = refine_solve_rfp(F, C64, torch.from_numpy(B[:, 0]).cuda(), tol=1e-10, transr=t, uplo=u)
    assert x.shape == (n,) and info["rel_residual"] <= 1e-10
