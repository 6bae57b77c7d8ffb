"""The reference's benchmark workloads at full size on one MI355X (parity at scale).

  plain method, small: n = 2048 (T, U) and 2047 (N, L), k = 256, 64 rhs,
             FP32 against the FP64 d* chain on the same inputs;
  plain method at scale: all 8 storage variants at n = 65536 and
             65535 (k = 1024, lambda = 1, 64 rhs, FP32), each against the
             FP64 d* chain at the same n, the FP64 chain at n = 32768, and a
             bit-exact strttf/stfttr round trip at n = 32768;
  precision: n = 90000, k = 4096, BF16 A (and its FP32 copy),
             lambda = 1e-2, 128 rhs: BF16, split-BF16 and FP32 factors, each
             refined in FP64 to a relative residual <= 1e-6 (the BF16 factor
             through flexible GMRES once classical steps stall); forward error
             against an FP64 dpftrf/dpftrs solution at n = 16384;
  differentiation cost: forward + backward within 3x forward at
             n = 16384, k = 512, 16 rhs;
  combined: n = 100000, k = 2048, 128 rhs, BF16 A, lambda = 1e-2,
             split-BF16 factor + FP64 refinement, forward + backward with the
             factor recomputed instead of stored; gradient error against FP64
             on a 4096-row problem of the same distribution.

Residuals are ||B - C X||_F / (||C||_F ||X||_F) formed in FP64 on the GPU
from the generator A.  Each test prints its figures (pytest -s) and, when the
environment variable CURFP_SCALE_OUT names a directory, writes them to
scale_results.json there.
"""
import json
import math
import os
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

RESULTS = {}


@pytest.fixture(scope="module")
def curfp(cuda):
    import curfp_amd

    yield curfp_amd
    out = os.environ.get("CURFP_SCALE_OUT", "")
    if out and os.path.isdir(out) and RESULTS:
        with open(os.path.join(out, "scale_results.json"), "w") as f:
            json.dump(RESULTS, f, indent=1)


def record(key, **kw):
    RESULTS[key] = kw
    print(key, kw, flush=True)


def gen(n, k, nrhs, seed, bf16=False):
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(n, k, device="cuda", generator=g) / math.sqrt(k)
    if bf16:
        A = A.to(torch.bfloat16).float()   # BF16 data with an exact FP32 copy
    B = torch.randn(n, nrhs, device="cuda", generator=g)
    return A, B


def norm_C(A, lam):
    G = A.double().t() @ A.double()
    return math.sqrt(float((G * G).sum() + 2 * lam * torch.trace(G) + A.shape[0] * lam * lam))


def rel_resid(curfp, A, lam, X, B, nc):
    Xc = X.double().t().contiguous().t()
    Bc = B.double().t().contiguous().t()
    R = curfp.gen_residual(A, Xc, Bc, lam=lam)
    return float(R.norm() / (nc * X.double().norm()))


def timed(fn):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t


@pytest.mark.parametrize("n,t,u", [(2048, "T", "U"), (2047, "N", "L")])
def test_small_vs_fp64(curfp, n, t, u):
    """The small plain-method workload (k = 256, 64 rhs, lambda = 1): FP32
    chain against the FP64 d* chain on the same inputs."""
    k, nrhs, lam = 256, 64, 1.0
    A, B = gen(n, k, nrhs, seed=0)
    nc = norm_C(A, lam)
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    curfp.ssfrk(A, C, transr=t, uplo=u)
    curfp.add_to_diagonal(C, lam, transr=t, uplo=u)
    assert curfp.spftrf(C, transr=t, uplo=u) == 0
    X = B.clone()
    curfp.spftrs(C, X, transr=t, uplo=u)
    C64 = torch.empty(n * (n + 1) // 2, dtype=torch.float64, device="cuda")
    curfp.dsfrk(A.double(), C64, transr=t, uplo=u)
    curfp.add_to_diagonal(C64, lam, transr=t, uplo=u)
    assert curfp.dpftrf(C64, transr=t, uplo=u) == 0
    Xref = B.double()
    curfp.dpftrs(C64, Xref, transr=t, uplo=u)
    err = float((X.double() - Xref).norm() / Xref.norm())
    r = rel_resid(curfp, A, lam, X, B, nc)
    r64 = rel_resid(curfp, A, lam, Xref, B, nc)
    record(f"c1_n{n}_{t}{u}", rel_residual=r, forward_error_vs_fp64=err, fp64_rel_residual=r64)
    assert r <= 1e-6 and r64 <= 1e-14 and err <= 1e-5


@pytest.mark.parametrize("n", [65536, 65535])
def test_plain_all_variants_n65536(curfp, n):
    k, nrhs, lam = 1024, 64, 1.0
    A, B = gen(n, k, nrhs, seed=1)
    nc = norm_C(A, lam)
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    sols = {}
    for t in "NT":
        for u in "LU":
            X = B.clone()

            def run():
                curfp.ssfrk(A, C, transr=t, uplo=u)
                curfp.add_to_diagonal(C, lam, transr=t, uplo=u)
                info = curfp.spftrf(C, transr=t, uplo=u, check=False)
                curfp.spftrs(C, X, transr=t, uplo=u)
                return info

            info, sec = timed(run)
            r = rel_resid(curfp, A, lam, X, B, nc)
            sols[t + u] = X
            record(f"c2_n{n}_{t}{u}", seconds=sec, rel_residual=r, info=info)
            assert info == 0 and r <= 1e-6
    del C
    # forward error of every variant against the FP64 d* chain on the same inputs
    C64 = torch.empty(n * (n + 1) // 2, dtype=torch.float64, device="cuda")
    Xref = B.double()

    def run64():
        curfp.dsfrk(A.double(), C64)
        curfp.add_to_diagonal(C64, lam)
        info = curfp.dpftrf(C64, check=False)
        curfp.dpftrs(C64, Xref)
        return info

    info64, sec64 = timed(run64)
    assert info64 == 0
    r64 = rel_resid(curfp, A, lam, Xref, B, nc)
    errs = {v: float((X.double() - Xref).norm() / Xref.norm()) for v, X in sols.items()}
    record(f"c2_n{n}_vs_fp64", fp64_seconds=sec64, fp64_rel_residual=r64, forward_error_vs_fp64=errs)
    # kappa(C) ~ 5 here: the FP32 chain's forward error is a small multiple of n u
    assert r64 <= 1e-14 and max(errs.values()) <= 1e-4


def test_plain_fp64_and_pack_n32768(curfp):
    n, k, nrhs, lam = 32768, 1024, 64, 1.0
    A32, B32 = gen(n, k, nrhs, seed=1)
    A, B = A32.double(), B32.double()
    C = torch.empty(n * (n + 1) // 2, dtype=torch.float64, device="cuda")
    X = B.clone()

    def run():
        curfp.dsfrk(A, C)
        curfp.add_to_diagonal(C, lam)
        info = curfp.dpftrf(C, check=False)
        curfp.dpftrs(C, X)
        return info

    info, sec = timed(run)
    R = B - curfp.gemm_nt(A, curfp.gemm_nt(X.t(), A.t())) - lam * X
    r = float(R.norm() / (norm_C(A32, lam) * X.norm()))
    record("c2_fp64_n32768", seconds=sec, rel_residual=r, info=info)
    assert info == 0 and r <= 1e-14
    del C, X
    # strttf / stfttr round trip of a dense upper triangle, bit-exact
    g = torch.Generator(device="cuda").manual_seed(2)
    D = torch.randn(n, n, device="cuda", generator=g)
    for t in "NT":
        arf = curfp.strttf(D, transr=t, uplo="U")
        back = curfp.stfttr(arf, transr=t, uplo="U")
        assert torch.equal(back, torch.triu(D))
    record("c2_pack_roundtrip_n32768", bit_exact=True)


def test_precision_n90000_k4096(curfp):
    n, k, nrhs, lam = 90000, 4096, 128, 1e-2
    A, B = gen(n, k, nrhs, seed=2, bf16=True)
    nc = norm_C(A, lam)
    Abf = A.to(torch.bfloat16)
    C = torch.empty(n * (n + 1) // 2, device="cuda")
    from curfp_amd import refine

    res = {}
    for mode in ("bf16", "bf16x3", "fp32"):
        def factor():
            if mode == "fp32":
                curfp.ssfrk(A, C)
            else:
                curfp.ssfrk(Abf, C)
            curfp.add_to_diagonal(C, lam)
            with curfp.math_mode(mode):
                return curfp.spftrf(C, check=False)

        info, t_fac = timed(factor)
        if info != 0:
            res[mode] = {"info": info, "factor_seconds": t_fac}
            record(f"c3_n{n}_{mode}", **res[mode])
            continue
        (X, it), t_sol = timed(lambda: refine.refine_solve(C, A, lam, B, tol=1e-6, max_iter=40))
        r = rel_resid(curfp, A, lam, X, B, nc)
        res[mode] = {"factor_seconds": t_fac, "solve_seconds": t_sol, "seconds": t_fac + t_sol,
                     "solves": it["solves"], "rel_residual": r, "info": info}
        record(f"c3_n{n}_{mode}", **res[mode])
    del C
    for mode in ("bf16", "bf16x3", "fp32"):
        assert res[mode].get("rel_residual", 1.0) <= 1e-6, (mode, res[mode])

    # forward error against FP64 dpftrf / dpftrs at n = 16384, same distribution
    n2 = 16384
    A2, B2 = gen(n2, k, nrhs, seed=2, bf16=True)
    C64 = torch.empty(n2 * (n2 + 1) // 2, dtype=torch.float64, device="cuda")
    curfp.dsfrk(A2.double(), C64)
    curfp.add_to_diagonal(C64, lam)
    assert curfp.dpftrf(C64) == 0
    Xref = B2.double().clone()
    curfp.dpftrs(C64, Xref)
    del C64
    C2 = torch.empty(n2 * (n2 + 1) // 2, device="cuda")
    for mode in ("bf16", "bf16x3", "fp32"):
        curfp.ssfrk(A2 if mode == "fp32" else A2.to(torch.bfloat16), C2)
        curfp.add_to_diagonal(C2, lam)
        with curfp.math_mode(mode):
            info = curfp.spftrf(C2, check=False)
        if info != 0:
            record(f"c3_n{n2}_{mode}_vs_fp64", info=info)
            continue
        X, it = refine.refine_solve(C2, A2, lam, B2, tol=1e-6, max_iter=40)
        Xp = B2.clone()
        curfp.spftrs(C2, Xp)
        err = float((X - Xref).norm() / Xref.norm())
        err_plain = float((Xp.double() - Xref).norm() / Xref.norm())
        record(f"c3_n{n2}_{mode}_vs_fp64", forward_error_refined=err, forward_error_unrefined=err_plain,
               solves=it["solves"], rel_residual=it["rel_residual"])
        if it["rel_residual"] <= 1e-6:
            # kappa(C) ~ 3e3 here: a 1e-6 residual bounds the forward error to ~1e-2
            # refinement never makes the error worse (equal when one solve already met the tolerance)
            assert err <= 5e-2 and err <= err_plain * (1 + 1e-9)


def test_backward_cost_n16384(curfp):
    from curfp_amd.autograd import krr_solve

    n, k, nrhs, lam = 16384, 512, 16, 0.1
    g = torch.Generator(device="cuda").manual_seed(3)
    A = (torch.randn(n, k, device="cuda", generator=g) / math.sqrt(k)).requires_grad_(True)
    Y = torch.randn(n, nrhs, device="cuda", generator=g).requires_grad_(True)
    W = torch.randn(n, nrhs, device="cuda", generator=g)
    lt = torch.tensor(lam, device="cuda", requires_grad=True)
    for _ in range(2):  # warm-up, then timed
        (X, t_fwd) = timed(lambda: krr_solve(A, lt, Y))
        L = (W * X).sum()
        (_, t_bwd) = timed(lambda: L.backward())
    record("c4_n16384", forward_seconds=t_fwd, backward_seconds=t_bwd,
           fwd_plus_bwd_over_fwd=(t_fwd + t_bwd) / t_fwd)
    assert (t_fwd + t_bwd) <= 3.0 * t_fwd


def test_combined_n100000_k2048(curfp):
    from curfp_amd.autograd import krr_solve

    n, k, nrhs, lam = 100000, 2048, 128, 1e-2
    g = torch.Generator(device="cuda").manual_seed(4)
    A = (torch.randn(n, k, device="cuda", generator=g) / math.sqrt(k)).to(torch.bfloat16).float()
    Y = torch.randn(n, nrhs, device="cuda", generator=g)
    W = torch.randn(n, nrhs, device="cuda", generator=g)
    A.requires_grad_(True)
    lt = torch.tensor(lam, device="cuda", requires_grad=True)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    X = krr_solve(A, lt, Y, math="bf16x3", refine_tol=1e-6, recompute=True)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    L = (W.double() * X).sum()
    L.backward()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # tensors (torch allocator peak) + the library's workspace and BF16 shadow
    peak_gb = (torch.cuda.max_memory_allocated() + curfp.workspace_bytes()) / 1e9
    gA, glam = A.grad.detach().clone(), lt.grad.detach().clone()
    assert torch.isfinite(gA).all() and math.isfinite(float(glam))
    del X, L, A

    # gradient error against FP64 on a 4096-row problem of the same distribution
    m = 4096
    g2 = torch.Generator(device="cuda").manual_seed(5)
    As = (torch.randn(m, k, device="cuda", generator=g2) / math.sqrt(k)).to(torch.bfloat16).float()
    Ys = torch.randn(m, nrhs, device="cuda", generator=g2)
    Ws = torch.randn(m, nrhs, device="cuda", generator=g2)
    a1 = As.clone().requires_grad_(True)
    l1 = torch.tensor(lam, device="cuda", requires_grad=True)
    (Ws.double() * krr_solve(a1, l1, Ys, math="bf16x3", refine_tol=1e-6, recompute=True)).sum().backward()
    a2 = As.double().requires_grad_(True)
    l2 = torch.tensor(lam, dtype=torch.float64, device="cuda", requires_grad=True)
    (Ws.double() * krr_solve(a2, l2, Ys.double())).sum().backward()
    errA = float((a1.grad.double() - a2.grad).norm() / a2.grad.norm())
    errl = abs(float(l1.grad) - float(l2.grad)) / abs(float(l2.grad))
    # the same pipeline refined to 1e-12 converges to the FP64 gradients
    a3 = As.clone().requires_grad_(True)
    l3 = torch.tensor(lam, device="cuda", requires_grad=True)
    (Ws.double() * krr_solve(a3, l3, Ys, math="bf16x3", refine_tol=1e-12, recompute=True)).sum().backward()
    errA_t = float((a3.grad.double() - a2.grad).norm() / a2.grad.norm())
    errl_t = abs(float(l3.grad) - float(l2.grad)) / abs(float(l2.grad))
    record("c5_n100000", forward_seconds=t_fwd, wall_seconds=wall, peak_hbm_gb=peak_gb,
           grad_rel_error_A_n4096=errA, grad_rel_error_lambda_n4096=errl,
           grad_rel_error_A_n4096_tol1e12=errA_t, grad_rel_error_lambda_n4096_tol1e12=errl_t)
    # tol 1e-6 at kappa(C) ~ 6e2 bounds the gradient error to ~1e-3 scale
    assert errA <= 1e-2 and errl <= 1e-2
    assert errA_t <= 1e-6 and errl_t <= 1e-6
