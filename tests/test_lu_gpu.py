"""FP64 LU (getrf + getrs) on the GPU through the C-ABI, against the oracle and LAPACK.

Parity bar: pivots bit-exact (LAPACK dgetrf semantics: first index of max |a|), factors equal to
the oracle's up to rounding (the blocked update orders differ), backward error <= 1e-13.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from conftest import backward_error
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def factor(lsx, A):
    n = A.shape[0]
    h = lsx._LUHandle(lsx.Context.get(0), n, lsx.LSX_F64)
    info = h.factor(A)
    F, ipiv = h.factors()
    return h, info, F, ipiv


@pytest.mark.parametrize("n", [1, 2, 5, 8, 17, 31, 32, 33, 64, 100, 127, 128, 129, 200, 255, 256, 257,
                               300, 511, 512, 513, 700, 1023, 1024, 1025, 1500, 2048, 2049, 3000])
def test_getrf_matches_oracle_and_lapack(lsx, n):
    print(f"getrf n={n}", flush=True)
    rng = np.random.default_rng(1000 + n)
    A = rng.standard_normal((n, n))
    h, info, F, ipiv = factor(lsx, A)
    Fo, po, io = O.dgetrf(A)
    lu, piv, il = lapack.dgetrf(A)
    assert info == io == il == 0
    np.testing.assert_array_equal(ipiv, po)
    np.testing.assert_array_equal(ipiv, piv + 1)
    growth = max(1.0, np.abs(Fo).max())
    assert np.abs(F - Fo).max() <= 1e-12 * n * growth
    b = rng.standard_normal(n)
    x = np.empty(n)
    h.solve(b, x)
    assert backward_error(A, x, b) <= 1e-13
    B = rng.standard_normal((n, 16))
    X = np.empty((n, 16), order="F")
    h.solve(B, X)
    assert backward_error(A, X, B) <= 1e-13
    Xo = O.dgetrs(Fo, po, B)
    assert np.abs(X - Xo).max() <= 1e-9 * max(1.0, np.abs(Xo).max())
    if n in (33, 300, 1025):
        B2 = rng.standard_normal((n, 37))
        X2 = np.empty((n, 37), order="F")
        h.solve(B2, X2)
        assert backward_error(A, X2, B2) <= 1e-13


def test_permutation_matrix_exact(lsx):
    # test/Core/blocked_lufact.jl:163-173: permutation input reproduces LAPACK exactly
    for n in (16, 65, 200, 1100):
        p = np.random.default_rng(n).permutation(n)
        A = np.eye(n)[p, :]
        h, info, F, ipiv = factor(lsx, A)
        lu, piv, il = lapack.dgetrf(A)
        assert info == 0
        np.testing.assert_array_equal(ipiv, piv + 1)
        np.testing.assert_array_equal(F, lu)


def test_wilkinson_growth(lsx):
    # test/Core/blocked_lufact.jl:175-184: ipiv = 1:n and U[n,n] = 2^(n-1)
    for n in (24, 53):
        A = np.eye(n) - np.tril(np.ones((n, n)), -1)
        A[:, n - 1] = 1.0
        h, info, F, ipiv = factor(lsx, A)
        np.testing.assert_array_equal(ipiv, np.arange(1, n + 1))
        assert F[n - 1, n - 1] == 2.0 ** (n - 1)


def test_decisive_margins_match_lapack(lsx):
    # test/Core/blocked_lufact.jl:148-161
    rng = np.random.default_rng(7)
    for n in (17, 64, 129, 300):
        for trial in range(3):
            mags = rng.permutation(2.0 ** np.arange(1, n + 1))
            A = np.diag(mags) @ (np.eye(n) + 0.01 * rng.standard_normal((n, n)))
            A = A[rng.permutation(n), :]
            h, info, F, ipiv = factor(lsx, A)
            lu, piv, il = lapack.dgetrf(A)
            np.testing.assert_array_equal(ipiv, piv + 1)


def test_singular_info_matches_lapack(lsx):
    # test/Core/blocked_lufact.jl:186-205
    rng = np.random.default_rng(11)
    for n in (10, 50, 130, 1200):
        for zc in (1, 4, n):
            A = rng.standard_normal((n, n))
            A[:, zc - 1] = 0.0
            h, info, F, ipiv = factor(lsx, A)
            lu, piv, il = lapack.dgetrf(A)
            Fo, po, io = O.dgetrf(A)
            assert info > 0 and info == il == io
            np.testing.assert_array_equal(ipiv, po)
    h, info, F, ipiv = factor(lsx, np.zeros((50, 50)))
    assert info == 1


def test_nan_propagates_without_error(lsx):
    A = np.random.default_rng(3).standard_normal((30, 30))
    A[1, 1] = np.nan
    h, info, F, ipiv = factor(lsx, A)
    assert np.isnan(F).any()


def test_interface_known_answers(lsx):
    # test/Core/resolve.jl:78-107: A = [1 2;3 4]'[1 2;3 4], b = [1,2] -> u = [-2, 1.5]
    A0 = np.array([[1.0, 2.0], [3.0, 4.0]])
    A = A0.T @ A0
    sol = lsx.solve(lsx.LinearProblem(A, np.array([1.0, 2.0])), lsx.MI355XLUFactorization())
    assert sol.retcode == lsx.ReturnCode.Success
    np.testing.assert_allclose(sol.u, [-2.0, 1.5], rtol=1e-12)
    # test/Core/retcodes.jl:15-52
    ok = lsx.solve(lsx.LinearProblem(np.array([[2.0, 1.0], [-1.0, 1.0]]), np.ones(2)),
                   lsx.MI355XLUFactorization())
    assert ok.retcode == lsx.ReturnCode.Success
    bad = lsx.solve(lsx.LinearProblem(np.array([[1.0, 1.0], [1.0, 1.0]]), np.ones(2)),
                    lsx.MI355XLUFactorization())
    assert bad.retcode == lsx.ReturnCode.Failure
    # identity and I/2 (test/Core/basictests.jl:30-41)
    for s in (1.0, 0.5):
        b = np.arange(1.0, 11.0)
        sol = lsx.solve(lsx.LinearProblem(s * np.eye(10), b), lsx.MI355XLUFactorization())
        np.testing.assert_allclose(sol.u, b / s, rtol=1e-15)


def test_cache_reuse_new_b_then_new_A(lsx):
    # test_interface (test/Core/basictests.jl:64-91): solve, reuse with new b, change A, change A and b
    rng = np.random.default_rng(5)
    n = 300
    A1, A2 = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    b1, b2 = rng.standard_normal(n), rng.standard_normal(n)
    cache = lsx.init(lsx.LinearProblem(A1, b1), lsx.MI355XLUFactorization())
    sol = lsx.solve_(cache)
    assert backward_error(A1, sol.u, b1) <= 1e-13
    handle = cache.cacheval
    cache.b = b2
    sol = lsx.solve_(cache)
    assert cache.cacheval is handle and not cache.isfresh
    assert backward_error(A1, sol.u, b2) <= 1e-13
    cache.A = A2
    assert cache.isfresh
    sol = lsx.solve_(cache)
    assert backward_error(A2, sol.u, b2) <= 1e-13
    cache.A, cache.b = A1, b1
    sol = lsx.solve_(cache)
    assert backward_error(A1, sol.u, b1) <= 1e-13
    # matrix right-hand side (test/Core/batch.jl:9-49)
    B = rng.standard_normal((n, 5))
    sol = lsx.solve(lsx.LinearProblem(A1, B), lsx.MI355XLUFactorization())
    np.testing.assert_allclose(sol.u, np.linalg.solve(A1, B), rtol=1e-9, atol=1e-11)


def test_singular_then_recovery(lsx):
    # test/Core/lu_refactorization.jl:83-110: Failure leaves the cache fresh, a good A recovers
    rng = np.random.default_rng(9)
    n = 64
    A = rng.standard_normal((n, n))
    S = A.copy()
    S[:, 3] = 0.0
    b = rng.standard_normal(n)
    cache = lsx.init(lsx.LinearProblem(S, b), lsx.MI355XLUFactorization())
    assert lsx.solve_(cache).retcode == lsx.ReturnCode.Failure
    assert cache.isfresh
    cache.A = A
    sol = lsx.solve_(cache)
    assert sol.retcode == lsx.ReturnCode.Success
    assert backward_error(A, sol.u, b) <= 1e-13


def test_config1_uniform_1024_two_rhs(lsx):
    # Plain dense LU, n = 1024, A and b iid U(0,1), second b reusing the cached factors
    rng = np.random.default_rng(np.random.SeedSequence(20261))
    n = 1024
    A = rng.random((n, n))
    b1, b2 = rng.random(n), rng.random(n)
    cache = lsx.init(lsx.LinearProblem(A, b1), lsx.MI355XLUFactorization())
    s1 = lsx.solve_(cache)
    F, ipiv = cache.cacheval.factors()
    lu, piv, _ = lapack.dgetrf(A)
    np.testing.assert_array_equal(ipiv, piv + 1)
    assert backward_error(A, s1.u, b1) <= 1e-13
    cache.b = b2
    s2 = lsx.solve_(cache)
    assert backward_error(A, s2.u, b2) <= 1e-13
    x_ref = lapack.dgetrs(lu, piv, b2)[0]
    np.testing.assert_allclose(s2.u, x_ref, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("n", [8192, 16384])
def test_config2_full_size_pivots_bit_exact(lsx, n):
    # Plain FP64 LU at scale: A iid N(0,1), 16 right-hand sides; pivots bit-exact vs LAPACK
    print(f"config2 n={n}", flush=True)
    rng = np.random.default_rng(np.random.SeedSequence([2, n]))
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, 16))
    h, info, F, ipiv = factor(lsx, A)
    assert info == 0
    X = np.empty((n, 16), order="F")
    h.solve(B, X)
    be = backward_error(A, X, B)
    print(f"n={n} backward error {be:.3e}", flush=True)
    assert be <= 1e-13
    del F
    lu, piv, il = lapack.dgetrf(A, overwrite_a=True)
    mism = np.nonzero(ipiv != piv + 1)[0]
    assert mism.size == 0, f"pivot mismatch at columns {mism[:10]}"


def test_beyond_the_partitioned_streams(lsx):
    # n = 17000: more than 32 panel workgroups, so the panel and the updates share plain streams
    # (no CU partition), and the last block is partial (17000 = 132 * 128 + 104).  Pivots
    # bit-exact vs LAPACK, backward error <= 1e-13.
    n = 17000
    print(f"n={n}", flush=True)
    rng = np.random.default_rng(np.random.SeedSequence([5, n]))
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, 2))
    h, info, F, ipiv = factor(lsx, A)
    assert info == 0
    X = np.empty((n, 2), order="F")
    h.solve(B, X)
    assert backward_error(A, X, B) <= 1e-13
    del F
    lu, piv, il = lapack.dgetrf(A, overwrite_a=True)
    mism = np.nonzero(ipiv != piv + 1)[0]
    assert mism.size == 0, f"pivot mismatch at columns {mism[:10]}"


# ---- adjoint solves: solve!(cache; adjoint = true) reuses the factorization (getrs 'T')

@pytest.mark.parametrize("n", [1, 7, 64, 65, 300, 1000, 2049])
def test_adjoint_solve_matches_lapack_getrs_t(lsx, n):
    # A^T x = b with the resident factors: backward error of the transposed system <= 1e-13, and
    # the solution equal (to rounding) to LAPACK dgetrs with trans = 'T' on the same factors
    print(f"adjoint n={n}", flush=True)
    rng = np.random.default_rng(3000 + n)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, 3))
    h, info, F, ipiv = factor(lsx, A)
    assert info == 0
    X = np.empty((n, 3), order="F")
    h.solve_adjoint(np.asfortranarray(B), X)
    assert backward_error(A.T, X, B) <= 1e-13
    Xl, linfo = lapack.dgetrs(F, ipiv - 1, B, trans=1)
    assert linfo == 0
    np.testing.assert_allclose(X, Xl, rtol=0, atol=1e-10 * np.abs(Xl).max())
    # one right-hand side (the vector case)
    x = np.empty(n)
    h.solve_adjoint(B[:, 0].copy(), x)
    assert backward_error(A.T, x, B[:, 0]) <= 1e-13


def test_reference_adjoint_cache_cases(lsx):
    # test/Core/adjoint.jl:446-498 (LUFactorization): a primal solve, then cache.b = c and an
    # adjoint solve from the same factorization; both directions from one cache; a fresh cache
    # factorizes first
    rng = np.random.default_rng(92)
    m = 60
    A = rng.random((m, m)) + m * np.eye(m)
    b = rng.random(m)
    c = rng.random(m)
    cache = lsx.init(lsx.LinearProblem(A, b), lsx.MI355XLUFactorization())
    np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(A, b), rtol=1e-12)
    cache.b = c
    sol = lsx.solve_(cache, adjoint=True)
    assert sol.retcode == lsx.ReturnCode.Success
    np.testing.assert_allclose(sol.u, np.linalg.solve(A.T, c), rtol=1e-12)
    np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(A, c), rtol=1e-12)  # and back again
    A2 = rng.random((m, m)) + m * np.eye(m)
    cache2 = lsx.init(lsx.LinearProblem(A2, b), lsx.MI355XLUFactorization())
    np.testing.assert_allclose(lsx.solve_(cache2, adjoint=True).u, np.linalg.solve(A2.T, b), rtol=1e-12)


def test_adjoint_fallbacks(lsx):
    # algorithms without factorization reuse solve the problem with adjoint(A): the mixed LU
    # refactorizes A^T, SimpleGMRES solves with A^T and its preconditioners swapped
    import scipy.sparse as sp
    from linearsolve_jl_amd import problems as P
    rng = np.random.default_rng(5)
    n = 200
    A = rng.standard_normal((n, n)) + 0.5 * n ** 0.5 * np.eye(n)
    b = rng.standard_normal(n)
    cache = lsx.init(lsx.LinearProblem(A, b), lsx.MI355XMixedLUFactorization(precision="fp32"))
    x = lsx.solve_(cache, adjoint=True).u
    assert backward_error(A.T, x, b) <= 1e-13
    rp, ci, v = P.convection_diffusion(30, 4.0)
    S = sp.csr_matrix((v, ci, rp), shape=(900, 900))
    bb = rng.standard_normal(900)
    cache = lsx.init(lsx.LinearProblem(lsx.CSRMatrix(rp, ci, v, (900, 900)), bb), lsx.MI355XSimpleGMRES(memory=30),
                     Pl=np.full(900, 4.0), reltol=1e-10, maxiters=5000)
    x = lsx.solve_(cache, adjoint=True).u
    assert np.linalg.norm(S.T @ x - bb) <= 1e-8 * np.linalg.norm(bb)


def test_default_solver_routes_to_device_algorithms(lsx):
    # solve(prob) with no algorithm: dense -> MI355XLUFactorization, CSR -> MI355XKrylovJL_GMRES,
    # BlockDiagonal -> MI355XBatchedLUFactorization, each answering to its own bar
    import scipy.sparse as sp
    from linearsolve_jl_amd import problems as P
    rng = np.random.default_rng(77)
    A = rng.standard_normal((300, 300))
    b = rng.standard_normal(300)
    sol = lsx.solve(lsx.LinearProblem(A, b))
    assert isinstance(sol.cache.alg, lsx.MI355XLUFactorization)
    assert sol.retcode == lsx.ReturnCode.Success and backward_error(A, sol.u, b) <= 1e-13
    rp, ci, v = P.convection_diffusion(20, 2.0)
    S = lsx.CSRMatrix(rp, ci, v, (400, 400))
    bs = rng.standard_normal(400)
    sol = lsx.solve(lsx.LinearProblem(S, bs), Pl=np.full(400, 4.0), reltol=1e-10)
    assert isinstance(sol.cache.alg, lsx.MI355XKrylovJL_GMRES) and sol.retcode == lsx.ReturnCode.Success
    Ssp = sp.csr_matrix((v, ci, rp), shape=(400, 400))
    assert np.linalg.norm(Ssp @ sol.u - bs) <= 1e-8 * np.linalg.norm(bs)
    blocks = [P.dense_block(n, seed=n)[0] for n in (5, 17, 40)]
    Bd = lsx.BlockDiagonal(blocks)
    bb = rng.standard_normal(62)
    sol = lsx.solve(lsx.LinearProblem(Bd, bb))
    assert isinstance(sol.cache.alg, lsx.MI355XBatchedLUFactorization) and sol.retcode == lsx.ReturnCode.Success
    off = 0
    for B in blocks:
        n = B.shape[0]
        assert backward_error(B, sol.u[off:off + n], bb[off:off + n]) <= 1e-13
        off += n


def test_refactorization_reuses_device_handle(lsx):
    # test/Core/lu_refactorization.jl:28-81: repeated refactorize + solve cycles on one cache are
    # correct, and after warm-up a refactorization reuses the device factor / pivot buffers (the
    # same handle) instead of allocating; resize_ marks the cache fresh and keeps the handle
    rng = np.random.default_rng(51)
    n = 51
    A1 = rng.random((n, n)) + n * np.eye(n)
    A2 = rng.random((n, n)) + n * np.eye(n)
    b = rng.random(n)
    cache = lsx.init(lsx.LinearProblem(A1.copy(), b.copy()), lsx.MI355XLUFactorization(), alias_A=True)
    np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(A1, b), rtol=1e-12)
    handle = cache.cacheval.h.value
    for Ak in (A2, A1, A2):
        cache.A = Ak.copy()
        sol = lsx.solve_(cache)
        assert sol.retcode == lsx.ReturnCode.Success
        np.testing.assert_allclose(sol.u, np.linalg.solve(Ak, b), rtol=1e-12)
        assert cache.cacheval.h.value == handle
    lsx.resize_(cache, n)
    assert cache.isfresh
    np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(A2, b), rtol=1e-12)
    assert cache.cacheval.h.value == handle
    # matrix right-hand side (test/Core/batch.jl:9-49): every column solved
    B = rng.random((n, 4))
    cache.b = B
    cache.u = np.zeros((n, 4))
    np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(A2, B), rtol=1e-12)


@pytest.mark.parametrize("alg", ["lu", "mixed"])
def test_torch_inputs_wait_for_torch_stream(lsx, alg):
    """The library reads device arrays on its own non-blocking streams, which do not wait for
    torch's.  The Python layer synchronizes torch's stream before each call, so an input that
    torch is still producing (here A * 1.0 and its column-major copy, queued behind long products)
    is read complete, not half-written."""
    import torch
    n = 2048
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    torch.cuda.synchronize()
    for rep in range(3):
        busy = torch.randn(4096, 4096, dtype=torch.float64, device="cuda", generator=g) * 1e-3
        for _ in range(6):
            busy = busy @ busy  # keeps torch's stream busy for tens of milliseconds
        Ai = A * 1.0            # queued behind the products
        bi = b * 1.0
        a = lsx.MI355XLUFactorization() if alg == "lu" else lsx.MI355XMixedLUFactorization(precision="fp32")
        sol = lsx.solve(lsx.LinearProblem(Ai, bi), a)
        assert sol.retcode == lsx.ReturnCode.Success
        x = sol.u
        r = b - A @ x
        be = float(r.abs().max() / (A.abs().sum(dim=1).max() * x.abs().max() + b.abs().max()))
        assert be <= 1e-13, (rep, be)
        del busy
