"""Mixed-precision LU (FP32 / BF16 factors) with FP64 iterative refinement on the GPU.

Bar: normwise backward error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
<= 1e-13, the FP64 accuracy the reference's LUFactorization delivers; the reference's own mixed
algorithms (OpenBLAS32MixedLUFactorization, src/openblas.jl:461-543) stop at ~1e-5
(test/Core/test_mixed_precision.jl:29-31).
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from conftest import backward_error

pytestmark = pytest.mark.gpu


def conditioned(n, cond, seed):
    from linearsolve_jl_amd import problems as P
    return P.conditioned(n, cond, seed=seed)


@pytest.mark.parametrize("precision,refinement", [("fp32", "ir"), ("fp32", "auto"), ("bf16", "auto"),
                                                  ("bf16", "gmres-ir")])
@pytest.mark.parametrize("n", [100, 700, 1500])
def test_random_reaches_fp64(lsx, precision, refinement, n):
    print(f"mixed {precision} {refinement} n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, 3))
    alg = lsx.MI355XMixedLUFactorization(precision=precision, refinement=refinement, tol=1e-13, max_iters=60)
    sol = lsx.solve(lsx.LinearProblem(A, B), alg)
    assert sol.retcode == lsx.ReturnCode.Success, sol.stats
    assert backward_error(A, sol.u, B) <= 1e-13
    assert sol.stats["backward_error"] <= 1e-13


@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_conditioned(lsx, cond):
    # A = Q1 diag(s) Q2^T.  FP32 factors refine to FP64 by IR up to cond 1e6 (u32 * cond < 1).
    # BF16 factors (two-piece BF16 products, u ~ 2^-17) refine by IR while u * cond is small and
    # precondition GMRES-IR beyond; where that stalls the 'auto' mode escalates to FP32 factors,
    # and must still reach 1e-13.
    n = 768
    A = conditioned(n, cond, seed=1)
    b = np.random.default_rng(2).standard_normal(n)
    for precision in ("fp32", "bf16"):
        alg = lsx.MI355XMixedLUFactorization(precision=precision, refinement="auto", tol=1e-13, max_iters=60)
        sol = lsx.solve(lsx.LinearProblem(A, b), alg)
        print(f"cond={cond:.0e} {precision}: {sol.stats}", flush=True)
        assert sol.retcode == lsx.ReturnCode.Success
        assert backward_error(A, sol.u, b) <= 1e-13
        if cond <= 1e4:
            np.testing.assert_allclose(sol.u, np.linalg.solve(A, b), rtol=0, atol=1e-9 * cond * np.abs(sol.u).max())
        if precision == "bf16" and cond <= 1e2:
            assert sol.stats["escalated"] == 0


def test_reference_mixed_cases(lsx):
    # test/Core/test_mixed_precision.jl: rand(100,100) + 5I, b = rand(100)
    rng = np.random.default_rng(0)
    A = rng.random((100, 100)) + 5 * np.eye(100)
    b = rng.random(100)
    sol = lsx.solve(lsx.LinearProblem(A, b), lsx.MI355XMixedLUFactorization())
    x64 = np.linalg.solve(A, b)
    assert np.linalg.norm(sol.u - x64) / np.linalg.norm(x64) < 1e-12
    # test/Core/resolve.jl:78-107 known answer
    A0 = np.array([[1.0, 2.0], [3.0, 4.0]])
    sol = lsx.solve(lsx.LinearProblem(A0.T @ A0, np.array([1.0, 2.0])), lsx.MI355XMixedLUFactorization())
    # backward error 1e-13 on a cond ~ 1e3 matrix bounds the forward error by ~1e-10
    np.testing.assert_allclose(sol.u, [-2.0, 1.5], rtol=1e-10)


def test_singular_is_failure(lsx):
    A = np.ones((2, 2))
    sol = lsx.solve(lsx.LinearProblem(A, np.ones(2)), lsx.MI355XMixedLUFactorization())
    assert sol.retcode == lsx.ReturnCode.Failure


def test_pivots_match_lapack_fp32(lsx):
    # FP32 factors of a matrix with decisive pivot margins pick LAPACK's pivots
    import scipy.linalg.lapack as lapack
    rng = np.random.default_rng(5)
    n = 100  # magnitudes 2^-49 .. 2^50 stay inside the FP32 range
    mags = rng.permutation(2.0 ** np.arange(1, n + 1) / 2.0 ** (n / 2))
    A = np.diag(mags) @ (np.eye(n) + 0.01 * rng.standard_normal((n, n)))
    A = A[rng.permutation(n), :]
    h = lsx._LUHandle(lsx.Context.get(0), n, lsx.LSX_F32)
    assert h.factor(A) == 0
    _, ipiv = h.factors()
    np.testing.assert_array_equal(ipiv, lapack.dgetrf(A)[1] + 1)


@pytest.mark.parametrize("cond", [1e2, 1e4, 1e6])
def test_conditioned_full_size(lsx, cond):
    """The mixed-precision workload at its full size: n = 16384, A = Q1 diag(s) Q2^T with s
    log-spaced from 1 to 1/cond, factors in FP32 and in BF16, refined to a normwise backward error
    <= 1e-13; GMRES-IR at cond 1e6.  The forward error is checked against LAPACK's FP64 solve of
    the same system (scipy's OpenBLAS dgetrf/dgetrs, the library LUFactorization calls): two
    solutions with backward errors <= 1e-13 of a system whose smallest singular value is 1/cond
    differ by at most about cond * 1e-13 relative."""
    import scipy.linalg as sla
    import torch
    from linearsolve_jl_amd import problems as P
    n = 16384
    print(f"mixed full size n={n} cond={cond:.0e}: generating", flush=True)
    A = P.conditioned(n, cond, device="cuda")  # row-major (n, n) tensor of A
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    Ah = A.cpu().numpy()
    bh = b.cpu().numpy()
    lu, piv = sla.lu_factor(Ah, check_finite=False)
    x64 = sla.lu_solve((lu, piv), bh, check_finite=False)
    del lu
    nA = float(torch.linalg.matrix_norm(A, ord=float("inf")))

    def berr(x):
        r = b - A @ x
        return float(r.abs().max() / (nA * x.abs().max() + b.abs().max()))

    runs = [("fp32", "gmres-ir" if cond >= 1e6 else "auto"), ("bf16", "auto")]
    for precision, refinement in runs:
        print(f"  {precision} {refinement}", flush=True)
        alg = lsx.MI355XMixedLUFactorization(precision=precision, refinement=refinement, tol=1e-13, max_iters=100)
        sol = lsx.solve(lsx.LinearProblem(A, b), alg)
        st = sol.stats
        print(f"  {precision}: {st}", flush=True)
        assert sol.retcode == lsx.ReturnCode.Success, st
        x = sol.u
        assert berr(x) <= 1e-13
        assert st["backward_error"] <= 1e-13
        fwd = np.linalg.norm(x.cpu().numpy() - x64) / np.linalg.norm(x64)
        print(f"  {precision}: forward error vs LAPACK {fwd:.2e}", flush=True)
        assert fwd <= cond * 1e-12
        if refinement == "gmres-ir":
            assert st["mode_used"] == 1  # GMRES-IR, as the workload asks at cond 1e6
    del A, Ah
    torch.cuda.empty_cache()


def test_beyond_tournament_blocks(lsx):
    # n = 17000: the first block columns have more rows than one tournament launch holds (16384)
    # and take the exact panels; the rest tournament pivoting; plain streams (no CU partition).
    # Refined to a backward error <= 1e-13 with FP32 and BF16 factors.
    n = 17000
    rng = np.random.default_rng(17)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    for precision in ("fp32", "bf16"):
        sol = lsx.solve(lsx.LinearProblem(A, b), lsx.MI355XMixedLUFactorization(precision=precision, tol=1e-13))
        print(f"n={n} {precision}: {sol.stats}", flush=True)
        assert sol.retcode == lsx.ReturnCode.Success
        assert backward_error(A, sol.u, b) <= 1e-13


@pytest.mark.parametrize("n", [200, 1000, 2500])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_factor_scaled_residual(lsx, precision, n):
    """The low-precision factors satisfy the scaled-residual test of the reference's
    test/Core/blocked_lufact.jl:11-28, ||PA - LU||_1 / (||L||_1 ||U||_1 n u) < 20, with u the unit
    roundoff the factorization promises: FP32's 2^-24 for FP32 factors (their trailing products
    come from three BF16 pieces per operand, which must stay below that), 2^-16 for BF16 factors
    (two pieces per operand: 3 2^-18 per product).  A dropped piece would leave FP32 factors at
    about 2^-16 and BF16 factors at about 2^-8."""
    rng = np.random.default_rng(n + (0 if precision == "fp32" else 1))
    A = rng.standard_normal((n, n))
    dt = lsx.LSX_F32 if precision == "fp32" else lsx.LSX_BF16
    h = lsx._LUHandle(lsx.Context.get(0), n, dt)
    assert h.factor(A) == 0
    F, ipiv = h.factors()
    Lf = np.tril(F, -1) + np.eye(n)
    Uf = np.triu(F)
    PA = A.copy()
    for k in range(n):  # LAPACK swap sequence
        p = ipiv[k] - 1
        if p != k:
            PA[[k, p], :] = PA[[p, k], :]
    u = 2.0 ** -24 if precision == "fp32" else 2.0 ** -16
    res = np.abs(PA - Lf @ Uf).sum(axis=0).max() / (np.abs(Lf).sum(axis=0).max() * np.abs(Uf).sum(axis=0).max() * n * u)
    # and against LAPACK's FP32 sgetrf of the same matrix: the relative factorization error
    # ||PA - LU||_F / ||A||_F of FP32 factors must be of sgetrf's size (products of two BF16 pieces
    # per operand instead of three would leave it about 2^7 times larger)
    e_dev = np.linalg.norm(PA - Lf @ Uf) / np.linalg.norm(A)
    lu32, piv32, info32 = lapack.sgetrf(A.astype(np.float32))
    P32 = A.copy()
    for k in range(n):
        p = piv32[k]
        if p != k:
            P32[[k, p], :] = P32[[p, k], :]
    lu32 = lu32.astype(np.float64)
    e_ref = np.linalg.norm(P32 - (np.tril(lu32, -1) + np.eye(n)) @ np.triu(lu32)) / np.linalg.norm(A)
    print(f"{precision} n={n}: scaled residual {res:.3e}, ||PA - LU||_F / ||A||_F {e_dev:.3e} "
          f"(LAPACK sgetrf {e_ref:.3e})", flush=True)
    assert res < 20.0
    if precision == "fp32":
        assert e_dev <= 4.0 * e_ref
    # the pivots form a permutation (every row chosen once)
    perm = np.arange(n)
    for k in range(n):
        p = ipiv[k] - 1
        assert k <= p < n
        perm[[k, p]] = perm[[p, k]]
    assert sorted(perm) == list(range(n))
