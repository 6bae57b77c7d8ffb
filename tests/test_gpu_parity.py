"""GPU parity tests: the HIP kernels (through the C ABI, via the _curfp_C
extension and ctypes) against the LAPACK golden vectors and the oracle.

Bars: pack / unpack / add_to_diagonal / info are bit-exact; floating-point
results meet the reference's 1e-3 max-abs bar (python_tests/test_vs_lapack.py:22-23)
and, much tighter, a normwise relative error of a few units of roundoff
times the problem size, stated per test."""
import ctypes
import math
import os

import numpy as np
import pytest
import torch

from oracle import rfp_oracle as O

pytestmark = pytest.mark.gpu

VARIANTS = [(t, u) for t in "NT" for u in "LU"]
N_PERM = [1, 2, 3, 4, 5, 6, 7, 8, 15, 16, 17, 32, 33, 63, 64, 65, 128, 129, 257]
CHAIN = [(1, 1), (2, 2), (3, 3), (4, 4), (5, 5), (8, 6), (9, 7), (32, 16), (33, 17),
         (64, 32), (65, 33), (128, 64), (129, 65), (256, 128), (257, 129)]
CHAIN64 = [(1, 1), (2, 2), (3, 3), (5, 5), (8, 6), (9, 7), (33, 17), (64, 32), (129, 65)]


@pytest.fixture(scope="module")
def curfp(cuda):
    import curfp_amd

    return curfp_amd


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


# ----------------------------------------------------------------- pack / unpack

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_pack_unpack_bit_exact_vs_lapack(curfp, golden, dtype):
    g = golden("rfp_perm.npz")
    pack = curfp.strttf if dtype == np.float32 else curfp.dstrttf
    unpack = curfp.stfttr if dtype == np.float32 else curfp.dstfttr
    for n in N_PERM:
        A = np.arange(1, n * n + 1, dtype=dtype).reshape(n, n)
        for t, u in VARIANTS:
            perm = g[f"n{n}_{t}{u}"].astype(dtype)
            arf = pack(dev(A), transr=t, uplo=u).cpu().numpy()
            assert np.array_equal(arf, perm), (n, t, u)
            back = unpack(dev(perm), transr=t, uplo=u).cpu().numpy()
            tri = np.tril(A) if u == "L" else np.triu(A)
            assert np.array_equal(back, tri), (n, t, u)


@pytest.mark.parametrize("n", [300, 301, 1000, 1025])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_pack_random_and_lda(curfp, n, t, u):
    """Random data, and a padded lda > n through the raw API."""
    rng = np.random.default_rng(n)
    lda = n + 7
    Ap = rng.standard_normal((n, lda)).astype(np.float32)
    h = curfp._get_handle()
    arf = torch.empty(n * (n + 1) // 2, dtype=torch.float32, device="cuda")
    curfp.strttf_raw(h, curfp._op(t), curfp._fill(u), n, dev(Ap), lda, arf)
    ref = O.trttf(Ap[:, :n], t, u)
    assert np.array_equal(arf.cpu().numpy(), ref)
    out = torch.full((n, lda), -7.0, dtype=torch.float32, device="cuda")
    curfp.stfttr_raw(h, curfp._op(t), curfp._fill(u), n, arf, out, lda)
    o = out.cpu().numpy()
    tri = np.tril(Ap[:, :n]) if u == "L" else np.triu(Ap[:, :n])
    mask = np.tril(np.ones((n, n), bool)) if u == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(o[:, :n][mask], tri[mask])
    assert np.all(o[:, :n][~mask] == -7.0)   # other triangle untouched
    assert np.all(o[:, n:] == -7.0)          # padding untouched


@pytest.mark.parametrize("t,u", VARIANTS)
def test_add_to_diagonal_bit_exact(curfp, t, u):
    for n in [1, 2, 7, 64, 65, 1001]:
        rng = np.random.default_rng(n)
        arf = rng.standard_normal(n * (n + 1) // 2).astype(np.float32)
        x = dev(arf)
        curfp.add_to_diagonal(x, 0.37, transr=t, uplo=u)
        assert np.array_equal(x.cpu().numpy(), O.add_to_diagonal(arf, 0.37, t, u))
        arf64 = arf.astype(np.float64)
        x = dev(arf64)
        curfp.add_to_diagonal(x, 1e-2, transr=t, uplo=u)
        assert np.array_equal(x.cpu().numpy(), O.add_to_diagonal(arf64, 1e-2, t, u))


# ----------------------------------------------------------------- chain vs LAPACK

def _chain_gpu(curfp, A, n, t, u, B, f64=False):
    dt = torch.float64 if f64 else torch.float32
    C = torch.empty(n * (n + 1) // 2, dtype=dt, device="cuda")
    (curfp.dsfrk if f64 else curfp.ssfrk)(dev(A), C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    Cn = C.cpu().numpy()
    info = (curfp.dpftrf if f64 else curfp.spftrf)(C, transr=t, uplo=u, check=False)
    F = C.cpu().numpy()
    X = dev(B)
    (curfp.dpftrs if f64 else curfp.spftrs)(C, X, transr=t, uplo=u)
    return Cn, info, F, X.cpu().numpy()


@pytest.mark.parametrize("n,k", CHAIN)
def test_chain_vs_lapack_f32(curfp, golden, n, k):
    g = golden("chain_f32.npz")
    for t, u in VARIANTS:
        key = f"n{n}_k{k}_{t}{u}"
        A, Cl, Fl, B, Xl = (g[key + s] for s in ("_A", "_C", "_F", "_B", "_X"))
        C, info, F, X = _chain_gpu(curfp, A, n, t, u, B)
        assert info == 0
        # reference bar (test_vs_lapack.py:96-120) and a tighter normwise one
        assert np.abs(C - Cl).max() < 1e-3 and np.abs(F - Fl).max() < 1e-3 and np.abs(X - Xl).max() < 1e-3, key
        u32 = np.finfo(np.float32).eps
        assert np.linalg.norm(C - Cl) <= 8 * k * u32 * np.linalg.norm(Cl), key
        assert np.linalg.norm(F - Fl) <= 8 * n * u32 * np.linalg.norm(Fl), key
        assert np.linalg.norm(X - Xl) <= 16 * n * u32 * np.linalg.norm(Xl), key


@pytest.mark.parametrize("n,k", CHAIN64)
def test_chain_vs_lapack_f64(curfp, golden, n, k):
    g = golden("chain_f64.npz")
    u64 = np.finfo(np.float64).eps
    for t, u in VARIANTS:
        key = f"n{n}_k{k}_{t}{u}"
        A, Cl, Fl, B, Xl = (g[key + s] for s in ("_A", "_C", "_F", "_B", "_X"))
        C, info, F, X = _chain_gpu(curfp, A, n, t, u, B, f64=True)
        assert info == 0
        assert np.linalg.norm(C - Cl) <= 8 * k * u64 * np.linalg.norm(Cl), key
        assert np.linalg.norm(F - Fl) <= 8 * n * u64 * np.linalg.norm(Fl), key
        assert np.linalg.norm(X - Xl) <= 16 * n * u64 * np.linalg.norm(Xl), key


# ----------------------------------------------------------------- known answers

def test_known_answers(curfp):
    """python_tests/test_spftrf.py:19-128 and test_ssfrk.py:84-111."""
    with curfp.Handle() as h:
        for t in (curfp.OP_N, curfp.OP_T):
            for u in (curfp.FILL_LOWER, curfp.FILL_UPPER):
                A = torch.tensor([9.0], device="cuda")
                assert curfp.spftrf_raw(h, t, u, 1, A) == 0
                assert abs(A.item() - 3.0) < 1e-6
        A = torch.tensor([3.0, 4.0, 2.0], device="cuda")
        assert curfp.spftrf_raw(h, curfp.OP_N, curfp.FILL_LOWER, 2, A) == 0
        assert np.allclose(A.cpu().numpy(), [math.sqrt(2), 2, 1], atol=1e-6)
        A = torch.tensor([5.0, 1.0, 2.0, 6.0, 4.0, 0.0], device="cuda")
        assert curfp.spftrf_raw(h, curfp.OP_N, curfp.FILL_LOWER, 3, A) == 0
        l00 = math.sqrt(5)
        l10, l20 = 1 / l00, 2 / l00
        l11 = math.sqrt(4 - l10 ** 2)
        l21 = -l20 * l10 / l11
        l22 = math.sqrt(6 - l20 ** 2 - l21 ** 2)
        assert np.allclose(A.cpu().numpy(), [l00, l10, l20, l22, l11, l21], atol=1e-6)
        C = torch.tensor([4.0, 1.0, 0.5], device="cuda")
        Av = torch.tensor([2.0, 3.0], device="cuda")
        curfp.ssfrk_raw(h, curfp.OP_N, curfp.FILL_LOWER, curfp.OP_N, 2, 1, 2.0, Av, 2, 0.5, C)
        assert np.allclose(C.cpu().numpy(), [20.0, 8.5, 12.25], atol=1e-5)


def test_ssfrk_never_reads_A_on_quick_paths(curfp):
    """tests/test_ssfrk.cu:39-101: alpha=0 with A = NULL, through the C ABI."""
    lib = ctypes.CDLL(os.path.join(os.path.dirname(curfp.__file__), "libcurfp_amd.so"))
    h = ctypes.c_void_p()
    assert lib.curfpCreate(ctypes.byref(h)) == 0
    C = torch.arange(15, dtype=torch.float32, device="cuda")
    zero, one, half = ctypes.c_float(0.0), ctypes.c_float(1.0), ctypes.c_float(0.5)
    cp = ctypes.c_void_p(C.data_ptr())
    assert lib.curfpSsfrk(h, 1, 1, 1, 5, 3, ctypes.byref(zero), None, 3, ctypes.byref(one), cp) == 0
    torch.cuda.synchronize()
    assert torch.equal(C.cpu(), torch.arange(15, dtype=torch.float32))
    assert lib.curfpSsfrk(h, 1, 1, 1, 5, 3, ctypes.byref(zero), None, 3, ctypes.byref(half), cp) == 0
    torch.cuda.synchronize()
    assert torch.equal(C.cpu(), torch.arange(15, dtype=torch.float32) * 0.5)
    assert lib.curfpSsfrk(h, 1, 1, 1, 5, 3, ctypes.byref(zero), None, 3, ctypes.byref(zero), cp) == 0
    torch.cuda.synchronize()
    assert torch.count_nonzero(C).item() == 0
    assert lib.curfpDestroy(h) == 0


def test_beta_zero_ignores_nan_in_C(curfp):
    n, k = 70, 9
    A = torch.randn(n, k, device="cuda")
    C = torch.full((n * (n + 1) // 2,), float("nan"), device="cuda")
    curfp.ssfrk(A, C, transr="N", uplo="L")
    assert torch.isfinite(C).all()


def test_pftrf_info_exact(curfp, golden):
    """Exact LAPACK info on crafted non-SPD matrices (the reference only pins info > 0)."""
    g = golden("info.npz")
    keys = sorted(k[:-5] for k in g.files if k.endswith("_info"))
    for key in keys:
        arf = g[key + "_arf"]
        n = int(key.split("_")[1][1:])
        t, u = key[-2], key[-1]
        f64 = key.startswith("d_")
        info = (curfp.dpftrf if f64 else curfp.spftrf)(dev(arf), transr=t, uplo=u, check=False)
        assert info == int(g[key + "_info"]), key
    with pytest.raises(torch.linalg.LinAlgError):
        curfp.spftrf(dev(-np.ones(3, np.float32)), check=True)


@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("n,t,u", [(16385, "N", "L"), (17000, "T", "U"), (20001, "T", "L"), (24576, "N", "U")])
def test_spftrf_tall_panels(curfp, mode, n, t, u):
    """Orders whose panel solves include panels of 8192 .. 12288 rows (one
    panel launch per 512 columns) and, at 24576, the two 12288-row halves of
    the top-level solve: ||L L^T - C||_F / ||C||_F formed in FP64 on the GPU,
    at the mode's precision."""
    g = torch.Generator(device="cuda").manual_seed(n)
    A = torch.randn(n, 256, device="cuda", generator=g) / 16.0
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(A, C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    T = curfp.stfttr(C, transr=t, uplo=u).double()
    S = T + T.T - torch.diag(torch.diagonal(T))
    del T
    with curfp.math_mode(mode):
        assert curfp.spftrf(C, transr=t, uplo=u) == 0
    F = curfp.stfttr(C, transr=t, uplo=u).double()
    L = F if u == "L" else F.T.contiguous()
    del F
    err = (torch.linalg.norm(L @ L.T - S) / torch.linalg.norm(S)).item()
    u32 = np.finfo(np.float32).eps
    bound = 2.0 ** -7 if mode == "bf16" else (1e-4 if mode == "bf16x3" else 4 * math.sqrt(n) * u32)
    assert err <= bound, (n, mode, err)


@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3", "fp64"])
def test_gemm_routes_agree(curfp, mode, monkeypatch):
    """The large plain GEMM updates go to hipBLASLt / rocBLAS; a handle created
    with CURFP_GEMM_NATIVE=1 keeps every GEMM on the hand-written kernels.
    Both factorisations of the same matrix (n = 20001, with updates of up to
    10001 x 10000 x 10000 and SYRKs split into library rectangles) meet the
    mode's backward-error bound and agree to within the condition number
    times it; each route repeats bit-identically."""
    n, t, u = 20001, "T", "U"
    f64 = mode == "fp64"
    dt = torch.float64 if f64 else torch.float32
    g = torch.Generator(device="cuda").manual_seed(5)
    A = (torch.randn(n, 256, device="cuda", generator=g) / 16.0).to(dt)
    C0 = torch.empty(n * (n + 1) // 2, device="cuda", dtype=dt)
    (curfp.dsfrk if f64 else curfp.ssfrk)(A, C0, transr=t, uplo=u)
    curfp.add_to_diagonal(C0, 1.0, transr=t, uplo=u)
    T = (curfp.dstfttr if f64 else curfp.stfttr)(C0, transr=t, uplo=u).double()
    S = T + T.T - torch.diag(torch.diagonal(T))
    del T
    math_code = {"fp32": curfp.MATH_FP32, "bf16": curfp.MATH_BF16, "bf16x3": curfp.MATH_BF16X3, "fp64": curfp.MATH_FP32}
    factors = {}
    for route in ("library", "native"):
        if route == "native":
            monkeypatch.setenv("CURFP_GEMM_NATIVE", "1")
        else:
            monkeypatch.delenv("CURFP_GEMM_NATIVE", raising=False)
        with curfp.Handle() as h:
            h.set_math_mode(math_code[mode])
            out = []
            for _ in range(2):
                C = C0.clone()
                fn = curfp.dpftrf_raw if f64 else curfp.spftrf_raw
                assert fn(h, curfp._op(t), curfp._fill(u), n, C) == 0
                torch.cuda.synchronize()
                out.append(C)
            assert torch.equal(out[0], out[1]), route
            factors[route] = out[0]
    u_ = np.finfo(np.float64 if f64 else np.float32).eps
    bound = 2.0 ** -7 if mode == "bf16" else (1e-4 if mode == "bf16x3" else 4 * math.sqrt(n) * u_)
    Ls = {}
    for route, C in factors.items():
        F = (curfp.dstfttr if f64 else curfp.stfttr)(C, transr=t, uplo=u).double()
        L = F.T.contiguous()
        err = (torch.linalg.norm(L @ L.T - S) / torch.linalg.norm(S)).item()
        assert err <= bound, (route, mode, err)
        Ls[route] = L
    diff = (torch.linalg.norm(Ls["library"] - Ls["native"]) / torch.linalg.norm(Ls["native"])).item()
    assert diff <= 200 * bound, (mode, diff)   # condition number of S ~ 80


@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("t,u", VARIANTS)
def test_spftrf_two_leaf_blocks(curfp, mode, t, u):
    """Orders whose recursion ends in diagonal blocks of 129 .. 256 rows (one
    launch for both leaves, the panel and the update between them in FP32 and
    BF16 mode; the four-launch recursion in split-BF16): backward error at the
    mode's precision, and exact LAPACK info for a failing pivot in the first
    leaf, on the leaf boundary, in the second leaf and in the second half.
    A diagonal entry set to -100 in an SPD matrix with eigenvalues >= 1 makes
    that column's pivot the first non-positive one."""
    rng = np.random.default_rng(11)
    u32 = np.finfo(np.float32).eps
    for n in [129, 200, 256, 257, 300, 511, 512, 700]:
        A = rng.standard_normal((n, 64)) / 8.0
        S = (A @ A.T + np.eye(n)).astype(np.float32)
        S64 = S.astype(np.float64)
        arf = dev(O.trttf(S, t, u))
        with curfp.math_mode(mode):
            assert curfp.spftrf(arf, transr=t, uplo=u, check=False) == 0, n
        L = O.rfp_lower_factor(arf.cpu().numpy().astype(np.float64), n, t, u)
        err = np.linalg.norm(L @ L.T - S64) / np.linalg.norm(S64)
        assert err <= (2.0 ** -7 if mode == "bf16" else 8 * math.sqrt(n) * u32), (n, err)
    for n, cols in [(300, [1, 100, 128, 129, 140, 150, 151, 250, 300]),
                    (512, [129, 200, 256, 257, 385, 512]), (384, [129, 192, 193, 320])]:
        A = rng.standard_normal((n, 64)) / 8.0
        S = (A @ A.T + np.eye(n)).astype(np.float32)
        for p in cols:
            Sp = S.copy()
            Sp[p - 1, p - 1] = -100.0
            with curfp.math_mode(mode):
                info = curfp.spftrf(dev(O.trttf(Sp, t, u)), transr=t, uplo=u, check=False)
            assert info == p, (n, p, info)


# ----------------------------------------------------------------- larger sizes vs oracle

def _spd_problem(n, k, nrhs, seed, dtype=np.float32, lam=1.0):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((n, k)) / math.sqrt(k)).astype(dtype)
    B = rng.standard_normal((n, nrhs)).astype(dtype)
    S = A.astype(np.float64) @ A.astype(np.float64).T + lam * np.eye(n)
    return A, B, S


@pytest.mark.parametrize("n,t,u", [(2048, "T", "U"), (2047, "N", "L"), (2047, "T", "U"), (2048, "N", "L"),
                                   (1500, "N", "U"), (1501, "T", "L"), (777, "N", "U"), (778, "T", "L")])
def test_chain_residual_f32(curfp, n, t, u):
    """Config 1 (n=2048, k=256, 64 rhs; and n=2047 N/L) plus the remaining
    variants: ssfrk vs the exact product, the factor's reconstruction and the
    solve's relative residual, against float64 dense arithmetic."""
    k, nrhs = 256, 64
    A, B, S = _spd_problem(n, k, nrhs, seed=0)
    C = torch.empty(n * (n + 1) // 2, dtype=torch.float32, device="cuda")
    curfp.ssfrk(dev(A), C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    Cd = O.rfp_to_dense_sym(C.cpu().numpy().astype(np.float64), n, t, u)
    u32 = np.finfo(np.float32).eps
    assert np.linalg.norm(Cd - S) <= 4 * k * u32 * np.linalg.norm(S)
    info = curfp.spftrf(C, transr=t, uplo=u)
    assert info == 0
    L = O.rfp_lower_factor(C.cpu().numpy().astype(np.float64), n, t, u)
    assert np.linalg.norm(L @ L.T - S) <= 4 * n * u32 * np.linalg.norm(S)
    X = dev(B)
    curfp.spftrs(C, X, transr=t, uplo=u)
    res = O.dense_relative_residual(S, X.cpu().numpy(), B)
    assert res <= 1e-6, res


@pytest.mark.parametrize("n,t,u", [(1024, "T", "U"), (1023, "N", "L"), (600, "N", "U"), (601, "T", "L")])
def test_chain_residual_f64(curfp, n, t, u):
    k, nrhs = 200, 16
    A, B, S = _spd_problem(n, k, nrhs, seed=1, dtype=np.float64)
    C = torch.empty(n * (n + 1) // 2, dtype=torch.float64, device="cuda")
    curfp.dsfrk(dev(A), C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    assert curfp.dpftrf(C, transr=t, uplo=u) == 0
    X = dev(B)
    curfp.dpftrs(C, X, transr=t, uplo=u)
    res = O.dense_relative_residual(S, X.cpu().numpy(), B)
    assert res <= 1e-14, res


@pytest.mark.parametrize("n", [1, 63, 64, 65, 511, 512, 513, 1000, 1100])
def test_dpftrs_block_geometry(curfp, n):
    """The FP64 substitution's 512-row block solves over 64-row leaves (one
    workgroup per 16 right-hand sides): sizes around the leaf and block
    edges, partial column groups, two layouts, row-major B (Python) and
    column-major B with ldb > n (C ABI), against float64 dense arithmetic;
    repeated solves are bit-identical."""
    k = 48
    for nrhs in (1, 15, 16, 17, 64, 100):
        A, B, S = _spd_problem(n, k, nrhs, seed=300 + n + nrhs, dtype=np.float64)
        for t, u in (("T", "U"), ("N", "L")):
            C = torch.empty(n * (n + 1) // 2, dtype=torch.float64, device="cuda")
            curfp.dsfrk(dev(A), C, transr=t, uplo=u)
            curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
            assert curfp.dpftrf(C, transr=t, uplo=u) == 0
            X = dev(B)
            curfp.dpftrs(C, X, transr=t, uplo=u)
            assert O.dense_relative_residual(S, X.cpu().numpy(), B) <= 1e-14, (nrhs, t, u)
            X2 = dev(B)
            curfp.dpftrs(C, X2, transr=t, uplo=u)
            assert torch.equal(X, X2), (nrhs, t, u)
            ldb = n + 3
            Bc = np.zeros((nrhs, ldb), np.float64)
            Bc[:, :n] = B.T
            Bt = dev(Bc)
            curfp.dpftrs_raw(curfp._get_handle(), curfp.OP_T if t == "T" else curfp.OP_N,
                             curfp.FILL_UPPER if u == "U" else curfp.FILL_LOWER, n, nrhs, C, Bt, ldb)
            assert O.dense_relative_residual(S, Bt.cpu().numpy()[:, :n].T, B) <= 1e-14, (nrhs, t, u)
            assert np.all(Bt.cpu().numpy()[:, n:] == 0)


def test_spftrs_1d_and_column_major_abi(curfp):
    n, k = 301, 40
    A, B, S = _spd_problem(n, k, 5, seed=5)
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(dev(A), C)
    curfp.add_to_diagonal(C, 1.0)
    curfp.spftrf(C)
    x = dev(B[:, 0].copy())
    curfp.spftrs(C, x)
    assert x.shape == (n,)
    assert O.dense_relative_residual(S, x.cpu().numpy()[:, None], B[:, :1]) < 1e-6
    # the reference C ABI layout: column-major B with ldb > n
    ldb = n + 3
    Bc = np.zeros((5, ldb), np.float32)
    Bc[:, :n] = B.T
    Bt = dev(Bc)
    curfp.spftrs_raw(curfp._get_handle(), curfp.OP_T, curfp.FILL_UPPER, n, 5, C, Bt, ldb)
    Xc = Bt.cpu().numpy()[:, :n].T
    assert O.dense_relative_residual(S, Xc, B) < 1e-6
    assert np.all(Bt.cpu().numpy()[:, n:] == 0)


@pytest.mark.parametrize("n,t,u", [(777, "T", "U"), (1025, "N", "L"), (1154, "N", "U"), (300, "T", "L")])
def test_spftrs_sweep_widths(curfp, n, t, u):
    """FP32 spftrs runs one substitution launch per direction over 128-row
    blocks, with 16/32/64 right-hand sides per workgroup and several column
    groups beyond 64.  Every width and group count, partial column groups,
    both B layouts and ragged last blocks (n1, n2 not multiples of 128) must
    solve to the FP32 residual bound; repeated calls (the launches' progress
    words are reused) must give bit-identical results; more than 4096
    right-hand sides take the recursive block substitution."""
    k = 96
    A, _, S = _spd_problem(n, k, 1, seed=11)
    C = torch.empty(n * (n + 1) // 2, dtype=torch.float32, device="cuda")
    curfp.ssfrk(dev(A), C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
    assert curfp.spftrf(C, transr=t, uplo=u) == 0
    rng = np.random.default_rng(n)
    widths = [1, 16, 17, 33, 64, 65, 200] + ([4100] if n == 300 else [])
    for nrhs in widths:
        B = rng.standard_normal((n, nrhs)).astype(np.float32)
        X = dev(B)
        curfp.spftrs(C, X, transr=t, uplo=u)
        res = O.dense_relative_residual(S, X.cpu().numpy(), B)
        assert res <= 1e-6, (nrhs, res)
        X2 = dev(B)
        curfp.spftrs(C, X2, transr=t, uplo=u)
        assert torch.equal(X, X2), nrhs
        if nrhs in (17, 65):
            # column-major B with ldb > n through the C ABI
            ldb = n + 5
            Bc = np.zeros((nrhs, ldb), np.float32)
            Bc[:, :n] = B.T
            Bt = dev(Bc)
            curfp.spftrs_raw(curfp._get_handle(), curfp.OP_T if t == "T" else curfp.OP_N,
                             curfp.FILL_UPPER if u == "U" else curfp.FILL_LOWER, n, nrhs, C, Bt, ldb)
            Xc = Bt.cpu().numpy()[:, :n].T
            assert O.dense_relative_residual(S, Xc, B) <= 1e-6, nrhs
            assert np.all(Bt.cpu().numpy()[:, n:] == 0)


@pytest.mark.parametrize("n", [1, 2, 5, 127, 128, 129, 255, 256, 257, 383, 513, 640])
def test_chain_small_and_boundary_sizes(curfp, n):
    """Sizes around the 128-row leaves and the 512-column panels, where the
    panel TRSM, the sweep's block geometry and the look-back products take
    their ragged paths: the factor against float64 reconstruction and the
    solve's residual, all 8 variants, row-major and (C ABI) column-major B."""
    k, nrhs = 24, 3
    A, B, S = _spd_problem(n, k, nrhs, seed=100 + n)
    u32 = np.finfo(np.float32).eps
    for t in "NT":
        for u in "LU":
            C = torch.empty(n * (n + 1) // 2, dtype=torch.float32, device="cuda")
            curfp.ssfrk(dev(A), C, transr=t, uplo=u)
            curfp.add_to_diagonal(C, 1.0, transr=t, uplo=u)
            assert curfp.spftrf(C, transr=t, uplo=u) == 0
            L = O.rfp_lower_factor(C.cpu().numpy().astype(np.float64), n, t, u)
            assert np.linalg.norm(L @ L.T - S) <= 8 * max(n, 16) * u32 * np.linalg.norm(S), (t, u)
            X = dev(B)
            curfp.spftrs(C, X, transr=t, uplo=u)
            assert O.dense_relative_residual(S, X.cpu().numpy(), B) <= 1e-6, (t, u)
            Bt = dev(np.ascontiguousarray(B.T))
            curfp.spftrs_raw(curfp._get_handle(), curfp.OP_T if t == "T" else curfp.OP_N,
                             curfp.FILL_UPPER if u == "U" else curfp.FILL_LOWER, n, nrhs, C, Bt, n)
            assert O.dense_relative_residual(S, Bt.cpu().numpy().T, B) <= 1e-6, (t, u)


def test_stream_binding(curfp):
    """Operations follow the stream bound with set_stream (test_spftrs.py:169-188)."""
    n, k = 513, 64
    A, B, S = _spd_problem(n, k, 4, seed=9)
    s = torch.cuda.Stream()
    curfp.set_stream(s)
    try:
        with torch.cuda.stream(s):
            Ad, X = dev(A), dev(B)
            C = torch.empty(n * (n + 1) // 2, device="cuda")
            curfp.ssfrk(Ad, C)
            curfp.add_to_diagonal(C, 1.0)
            curfp.spftrf(C)
            curfp.spftrs(C, X)
        s.synchronize()
    finally:
        curfp.set_stream(None)
    assert O.dense_relative_residual(S, X.cpu().numpy(), B) < 1e-6


def test_pack_round_trip_past_int32(curfp):
    """n = 46342: n*(n+1) exceeds 2^31 (the reference's `int nt` overflows);
    the round trip must stay bit-exact and sampled positions must match the
    closed-form index map."""
    n = 46342
    for t, u in [("T", "U"), ("N", "L")]:
        g = torch.Generator(device="cuda").manual_seed(3)
        A = torch.randn(n, n, device="cuda", generator=g)
        arf = curfp.strttf(A, transr=t, uplo=u)
        rng = np.random.default_rng(1)
        i = rng.integers(0, n, 4096)
        j = rng.integers(0, n, 4096)
        lo_i, lo_j = np.maximum(i, j), np.minimum(i, j)
        n1, n2, ld, v11, v21, v22 = O.views(n, t, u)
        idx = np.where(lo_j >= n1, v22[0] + (lo_i - n1) * v22[1] + (lo_j - n1) * v22[2],
                       np.where(lo_i >= n1, v21[0] + (lo_i - n1) * v21[1] + lo_j * v21[2],
                                v11[0] + lo_i * v11[1] + lo_j * v11[2]))
        di, dj = (lo_i, lo_j) if u == "L" else (lo_j, lo_i)
        got = arf[torch.from_numpy(idx).cuda()].cpu()
        want = A[torch.from_numpy(di).cuda(), torch.from_numpy(dj).cuda()].cpu()
        assert torch.equal(got, want)
        back = curfp.stfttr(arf, transr=t, uplo=u)
        tri = torch.tril(A) if u == "L" else torch.triu(A)
        assert torch.equal(back, tri)
        del A, arf, back, tri
        torch.cuda.empty_cache()
