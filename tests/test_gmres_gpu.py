"""SimpleGMRES on the GPU through the C-ABI, against the oracle (a restatement of
src/simplegmres.jl) and the reference's closed-form tests.  Dot products are summed in a different
order than the reference's BLAS, so parity is tolerance-based."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def problems(lsx):
    from linearsolve_jl_amd import problems as P
    return P


def test_reference_6x6_resolve(lsx):
    # test/Core/basictests.jl:651-684: rtol 1e-13 vs A \ b, then re-solve with b mutated / replaced /
    # unchanged
    nr = 6
    Ar = np.array([[10.0 + i if i == j else 0.3 * (i + j) for j in range(1, nr + 1)] for i in range(1, nr + 1)])
    br = np.arange(1.0, nr + 1)
    br2 = np.array([0.3 * i + 1 for i in range(1, nr + 1)])
    for restart in (True, False):
        for mode in ("mutated", "replaced", "unchanged"):
            A = lsx.CSRMatrix.from_scipy(sp.csr_matrix(Ar))
            cache = lsx.init(lsx.LinearProblem(A, br.copy()), lsx.MI355XSimpleGMRES(restart=restart),
                             abstol=1e-12, reltol=1e-12)
            np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(Ar, br), rtol=1e-13)
            if mode == "mutated":
                cache.b[:] = br2
                expected = br2
            elif mode == "replaced":
                cache.b = br2.copy()
                expected = br2
            else:
                expected = br
            np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(Ar, expected), rtol=1e-13)


def test_reference_6x6_preconditioned(lsx):
    # test/Core/basictests.jl:686-710: Pl, Pr, both; success and rtol 1e-8
    npr = 6
    Ap = np.array([[10.0 + 5 * i if i == j else 0.3 * (i + j) for j in range(1, npr + 1)] for i in range(1, npr + 1)])
    bp = np.arange(1.0, npr + 1)
    d = np.diag(Ap)
    for restart in (True, False):
        for Pl, Pr in ((d, None), (None, d), (np.sqrt(d), np.sqrt(d))):
            A = lsx.CSRMatrix.from_scipy(sp.csr_matrix(Ap))
            sol = lsx.solve(lsx.LinearProblem(A, bp.copy()), lsx.MI355XSimpleGMRES(restart=restart), Pl=Pl, Pr=Pr,
                            abstol=1e-12, reltol=1e-12, maxiters=100)
            assert sol.retcode == lsx.ReturnCode.Success
            np.testing.assert_allclose(sol.u, np.linalg.solve(Ap, bp), rtol=1e-8)
            ref = O.gmres(A.rowptr, A.colind, A.vals, bp, memory=20, restart=restart, abstol=1e-12, reltol=1e-12,
                          maxiters=100, pl=Pl, pr=Pr)
            assert ref["retcode"] == int(sol.retcode)
            assert ref["iters"] == sol.iters


@pytest.mark.parametrize("m,pe", [(8, 0.0), (16, 1.0), (32, 5.0), (64, 3.0), (100, 7.0), (128, 2.0)])
def test_convdiff_matches_oracle(lsx, m, pe):
    print(f"gmres convdiff m={m} pe={pe}", flush=True)
    P = problems(lsx)
    rp, ci, v = P.convection_diffusion(m, pe)
    N = m * m
    b = P.rng_for(7, m).standard_normal(N)
    d = np.full(N, 4.0)
    maxit = 3000
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    sol = lsx.solve(lsx.LinearProblem(A, b), lsx.MI355XSimpleGMRES(restart=True, memory=30), Pl=d, reltol=1e-8,
                    maxiters=maxit)
    ref = O.gmres(rp, ci, v, b, memory=30, restart=True, reltol=1e-8, maxiters=maxit, pl=d)
    assert int(sol.retcode) == ref["retcode"]
    assert abs(sol.iters - ref["iters"]) <= 1
    S = sp.csr_matrix((v, ci, rp), shape=(N, N))
    if ref["retcode"] == O.RET_SUCCESS:
        # converged: true preconditioned residual within the tolerance the solver stopped at
        eps_ = np.sqrt(np.finfo(float).eps) + 1e-8 * np.linalg.norm(b / d)
        assert np.linalg.norm((b - S @ sol.u) / d) <= 1.5 * eps_
        np.testing.assert_allclose(sol.u, ref["u"], rtol=0, atol=1e-5 * np.abs(ref["u"]).max())
    else:
        np.testing.assert_allclose(sol.resid, ref["resid"], rtol=1e-3)


@pytest.mark.parametrize("m", [256, 1024])
def test_large_grid_residual_estimate_and_determinism(lsx, m):
    # Full-size property (grid side up to 1024, N = 1,048,576): the Givens estimate rNorm equals the
    # true preconditioned residual ||Pl \ (b - A x)|| (exact in exact arithmetic), and two solves are
    # bitwise identical.
    print(f"gmres large m={m}", flush=True)
    P = problems(lsx)
    rp, ci, v = P.convection_diffusion(m, 4.0)
    N = m * m
    b = P.rng_for(8, m).standard_normal(N)
    d = np.full(N, 4.0)
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    alg = lsx.MI355XSimpleGMRES(restart=True, memory=30)
    s1 = lsx.solve(lsx.LinearProblem(A, b), alg, Pl=d, reltol=1e-8, maxiters=120)
    s2 = lsx.solve(lsx.LinearProblem(A, b), alg, Pl=d, reltol=1e-8, maxiters=120)
    assert s1.iters == 120 and s1.retcode == lsx.ReturnCode.MaxIters
    np.testing.assert_array_equal(s1.u, s2.u)
    S = sp.csr_matrix((v, ci, rp), shape=(N, N))
    true = np.linalg.norm((b - S @ s1.u) / d)
    assert abs(true - s1.resid) <= 1e-6 * np.linalg.norm(b / d)
    assert s1.resid < np.linalg.norm(b / d)


def test_spmv_matches_scipy(lsx):
    P = problems(lsx)
    rp, ci, v = P.convection_diffusion(50, 3.0)
    N = 2500
    x = np.random.default_rng(1).standard_normal(N)
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    h = lsx._CSRHandle(lsx.Context.get(0), A)
    y = h.spmv(x)
    np.testing.assert_allclose(y, sp.csr_matrix((v, ci, rp), shape=(N, N)) @ x, rtol=1e-14, atol=1e-14)


def test_zero_rhs_and_maxiters(lsx):
    P = problems(lsx)
    rp, ci, v = P.convection_diffusion(20, 1.0)
    N = 400
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    sol = lsx.solve(lsx.LinearProblem(A, np.zeros(N)), lsx.MI355XSimpleGMRES())
    assert sol.retcode == lsx.ReturnCode.Success and sol.iters == 0 and not sol.u.any()
    b = np.ones(N)
    sol = lsx.solve(lsx.LinearProblem(A, b), lsx.MI355XSimpleGMRES(memory=5), reltol=1e-14, maxiters=7)
    ref = O.gmres(rp, ci, v, b, memory=5, restart=True, reltol=1e-14, maxiters=7)
    assert sol.retcode == lsx.ReturnCode.MaxIters and sol.iters == 7 == ref["iters"]
    np.testing.assert_allclose(sol.u, ref["u"], rtol=1e-10, atol=1e-12)


def test_matrix_rhs_rejected(lsx):
    # src/common.jl:614-625
    A = lsx.CSRMatrix.from_scipy(sp.identity(4, format="csr"))
    with pytest.raises(ValueError):
        lsx.init(lsx.LinearProblem(A, np.ones((4, 2))), lsx.MI355XSimpleGMRES())


# ---- batched layout (SimpleGMRES(blocksize > 0), src/simplegmres.jl:190-373)

def test_reference_6x6_blocksize_2(lsx):
    # test/Core/basictests.jl:651-710 with blocksize = 2: the matrices keep only their 2 x 2
    # diagonal blocks; resolve with b mutated / replaced / unchanged at rtol 1e-13, and the
    # preconditioned cases at rtol 1e-8 with Success
    nr = 6
    bs = 2
    mask = np.array([[(i - 1) // bs == (j - 1) // bs for j in range(1, nr + 1)] for i in range(1, nr + 1)])
    Ar = np.array([[10.0 + i if i == j else 0.3 * (i + j) for j in range(1, nr + 1)] for i in range(1, nr + 1)]) * mask
    br = np.arange(1.0, nr + 1)
    br2 = np.array([0.3 * i + 1 for i in range(1, nr + 1)])
    for restart in (True, False):
        for mode in ("mutated", "replaced", "unchanged"):
            A = lsx.CSRMatrix.from_scipy(sp.csr_matrix(Ar))
            cache = lsx.init(lsx.LinearProblem(A, br.copy()), lsx.MI355XSimpleGMRES(restart=restart, blocksize=bs),
                             abstol=1e-12, reltol=1e-12)
            np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(Ar, br), rtol=1e-13)
            if mode == "mutated":
                cache.b[:] = br2
                expected = br2
            elif mode == "replaced":
                cache.b = br2.copy()
                expected = br2
            else:
                expected = br
            np.testing.assert_allclose(lsx.solve_(cache).u, np.linalg.solve(Ar, expected), rtol=1e-13)
    Ap = np.array([[10.0 + 5 * i if i == j else 0.3 * (i + j) for j in range(1, nr + 1)] for i in range(1, nr + 1)]) * mask
    d = np.diag(Ap)
    for restart in (True, False):
        for Pl, Pr in ((d, None), (None, d), (np.sqrt(d), np.sqrt(d))):
            A = lsx.CSRMatrix.from_scipy(sp.csr_matrix(Ap))
            sol = lsx.solve(lsx.LinearProblem(A, br.copy()), lsx.MI355XSimpleGMRES(restart=restart, blocksize=bs),
                            Pl=Pl, Pr=Pr, abstol=1e-12, reltol=1e-12, maxiters=100)
            assert sol.retcode == lsx.ReturnCode.Success
            np.testing.assert_allclose(sol.u, np.linalg.solve(Ap, br), rtol=1e-8)


@pytest.mark.parametrize("bs,nblocks", [(16, 40), (100, 12), (512, 4), (1024, 3)])
def test_batched_blocks_match_oracle_per_block(lsx, bs, nblocks):
    # every block is its own SimpleGMRES: each block's result equals the oracle's SimpleGMRES on
    # that block alone with the batch's tolerance eps = abstol + reltol ||Pl \ b||_2 (same
    # retcode, iterations within 1, solution within 1e-6 relative); the batch reports its slowest
    # block's iterations
    print(f"batched gmres bs={bs} blocks={nblocks}", flush=True)
    P = problems(lsx)
    rng = np.random.default_rng(bs)
    m = int(round(bs ** 0.5))
    blocks = []
    for k in range(nblocks):
        if m * m == bs:
            rp, ci, v = P.convection_diffusion(m, 1.0 + k % 5)
            blocks.append(sp.csr_matrix((v, ci, rp), shape=(bs, bs)))
        else:
            B = sp.random(bs, bs, density=4.0 / bs, random_state=int(rng.integers(1 << 30))) + 4.0 * sp.eye(bs)
            blocks.append(sp.csr_matrix(B))
    S = sp.block_diag(blocks, format="csr")
    S.sort_indices()
    N = bs * nblocks
    b = rng.standard_normal(N)
    d = S.diagonal()
    reltol, maxit = 1e-8, 2000
    A = lsx.CSRMatrix.from_scipy(S)
    sol = lsx.solve(lsx.LinearProblem(A, b), lsx.MI355XSimpleGMRES(memory=20, blocksize=bs), Pl=d, reltol=reltol,
                    maxiters=maxit)
    eps = np.sqrt(np.finfo(float).eps) + reltol * np.linalg.norm(b / d)
    worst_iters = 0
    for k in range(nblocks):
        sl = slice(k * bs, (k + 1) * bs)
        Bk = blocks[k]
        Bk.sort_indices()
        ref = O.gmres(Bk.indptr, Bk.indices, Bk.data, b[sl], memory=20, restart=True, abstol=eps, reltol=0.0,
                      maxiters=maxit, pl=d[sl])
        assert ref["retcode"] == O.RET_SUCCESS
        worst_iters = max(worst_iters, ref["iters"])
        np.testing.assert_allclose(sol.u[sl], ref["u"], rtol=0, atol=1e-6 * np.abs(ref["u"]).max())
        assert np.linalg.norm((b[sl] - Bk @ sol.u[sl]) / d[sl]) <= 1.5 * eps
    assert sol.retcode == lsx.ReturnCode.Success
    assert abs(sol.iters - worst_iters) <= 1


def test_batched_rejects_bad_layouts(lsx):
    # blocksize must divide N (simplegmres.jl:422); entries outside the diagonal blocks are an error
    S = sp.csr_matrix(np.diag(np.arange(1.0, 7.0)) + np.diag(np.ones(5), 1))
    A = lsx.CSRMatrix.from_scipy(S)
    with pytest.raises(ValueError):
        lsx.solve(lsx.LinearProblem(A, np.ones(6)), lsx.MI355XSimpleGMRES(blocksize=4))
    with pytest.raises(lsx.LSXError):
        lsx.solve(lsx.LinearProblem(A, np.ones(6)), lsx.MI355XSimpleGMRES(blocksize=2))


def test_malformed_csr_is_rejected(lsx):
    # the structure check (rowptr bounds and monotonicity, column range) runs before any solver
    # kernel reads the matrix: bad input is an error, not an out-of-range access
    good_rp = np.array([0, 2, 3, 5], dtype=np.int32)
    good_ci = np.array([0, 1, 1, 0, 2], dtype=np.int32)
    vals = np.ones(5)
    cases = [
        (np.array([0, 2, 3, 9], dtype=np.int32), good_ci),      # rowptr[N] != nnz
        (np.array([0, 3, 2, 5], dtype=np.int32), good_ci),      # decreasing row
        (np.array([1, 2, 3, 5], dtype=np.int32), good_ci),      # rowptr[0] != 0
        (good_rp, np.array([0, 1, 1, 0, 7], dtype=np.int32)),   # column out of range
        (good_rp, np.array([0, -1, 1, 0, 2], dtype=np.int32)),  # negative column
    ]
    for rp, ci in cases:
        with pytest.raises(lsx.LSXError):
            lsx._CSRHandle(lsx.Context.get(0), lsx.CSRMatrix(rp, ci, vals, (3, 3)))
    lsx._CSRHandle(lsx.Context.get(0), lsx.CSRMatrix(good_rp, good_ci, vals, (3, 3)))
    # the serving lane checks a request's CSR the same way: the bad request fails alone
    srv = lsx.Server(hbm_budget=1 << 30)
    bad = srv.submit_gmres(lsx.CSRMatrix(cases[3][0], cases[3][1], vals, (3, 3)), np.ones(3), reltol=1e-8, maxiters=50)
    good = srv.submit_gmres(lsx.CSRMatrix(good_rp, good_ci, vals + np.array([3.0, 0, 3.0, 0, 3.0]), (3, 3)),
                            np.ones(3), reltol=1e-10, maxiters=50)
    srv.drain()
    rb, rg = srv.poll(bad), srv.poll(good)
    from linearsolve_jl_amd import _abi
    assert rb["status"] == _abi.LSX_EINVAL
    assert rg["status"] == 0 and rg["retcode"] == lsx.ReturnCode.Success
    srv.close()


# ---- KrylovJL_GMRES (src/iterative_wrappers.jl:200-228) on the device GMRES

def _csr(lsx, M):
    return lsx.CSRMatrix.from_scipy(sp.csr_matrix(M))


def test_krylovjl_gmres_reference_interface(lsx):
    # test/Core/basictests.jl:712-735 ("GMRES" and "GMRES_prec", gmres_restart = 5) through
    # test_interface (:64-91, real systems prob1 / prob2 of :30-57, cache_kwargs of :62) and
    # test_tolerance_update (:93-104) on prob5
    n = 8
    rng = np.random.default_rng(712)
    A1, A2 = np.eye(n), np.eye(n) / 2
    b1, b2 = rng.random(n), rng.random(n)
    kw = dict(abstol=1e-8, reltol=1e-8, maxiters=30)
    precs = lambda A, p=None: (1.0 / A.diagonal(), None)  # (Diagonal(inv.(diag(A))), I)
    for alg in (lsx.MI355XKrylovJL_GMRES(gmres_restart=5),
                lsx.MI355XKrylovJL_GMRES(gmres_restart=5, precs=precs, ldiv=False)):
        close = lambda x, y: np.linalg.norm(x - y) <= 1.5e-8 * max(np.linalg.norm(x), np.linalg.norm(y))
        sol = lsx.solve(lsx.LinearProblem(_csr(lsx, A1), b1.copy()), alg, **kw)
        assert close(A1 @ sol.u, b1)
        sol = lsx.solve(lsx.LinearProblem(_csr(lsx, A2), b2.copy()), alg, **kw)
        assert close(A2 @ sol.u, b2)
        cache = lsx.init(lsx.LinearProblem(_csr(lsx, A1), b1.copy()), alg, **kw)
        assert close(A1 @ lsx.solve_(cache).u, b1)
        cache.A = _csr(lsx, A2)
        assert close(A2 @ lsx.solve_(cache).u, b1)
        cache.A = _csr(lsx, A2)
        cache.b = b2.copy()
        assert close(A2 @ lsx.solve_(cache).u, b2)
        # test_tolerance_update: A5 = A5_' A5_, A5_ = I - 0.01 Tridiagonal(ones) + 0.5 e1 e8'
        A5_ = np.eye(n) - 0.01 * (np.eye(n) + np.eye(n, k=1) + np.eye(n, k=-1))
        A5_[0, 7] += 0.5
        A5 = A5_.T @ A5_
        u5 = np.ones(n)
        cache = lsx.init(lsx.LinearProblem(_csr(lsx, A5), A5 @ u5), alg)
        cache.reltol, cache.abstol = 1e-2, 1e-8
        u1 = lsx.solve_(cache).u.copy()
        cache.reltol, cache.abstol = 1e-8, 1e-8
        u2 = lsx.solve_(cache).u
        assert np.linalg.norm(u2 - u5) < np.linalg.norm(u1 - u5)


def test_krylovjl_gmres_reference_issues(lsx):
    # test/Core/basictests.jl:1112-1119 (issue 347: a 4 x 4 system with default options agrees with
    # LU) and :1121-1142 (issue 869: the memory keyword sizes the workspace; memory = 30 solves a
    # sprand(100, 100, 0.1) + 10I system to ||Ax - b|| < 1e-6)
    rng = np.random.default_rng(347)
    A = rng.random((4, 4))
    b = rng.random(4)
    sol = lsx.solve(lsx.LinearProblem(_csr(lsx, A), b.copy()), lsx.MI355XKrylovJL_GMRES())
    x = np.linalg.solve(A, b)
    assert np.linalg.norm(sol.u - x) <= 1.5e-8 * np.linalg.norm(x)
    S = sp.random(100, 100, density=0.1, random_state=869, format="csr") + 10 * sp.identity(100, format="csr")
    b = np.random.default_rng(869).random(100)
    sol = lsx.solve_(lsx.init(lsx.LinearProblem(lsx.CSRMatrix.from_scipy(S), b.copy()),
                              lsx.MI355XKrylovJL_GMRES(memory=30)))
    assert sol.retcode == lsx.ReturnCode.Success
    assert np.linalg.norm(S @ sol.u - b) < 1e-6
    sol2 = lsx.solve_(lsx.init(lsx.LinearProblem(lsx.CSRMatrix.from_scipy(S), b.copy()),
                               lsx.MI355XKrylovJL_GMRES(memory=10)))
    assert isinstance(sol2.u, np.ndarray) and sol2.u.shape == (100,)


def test_krylovjl_gmres_resize(lsx):
    # test/Core/resize.jl:5-48,70-72: solve at n = 3, resize to 6, solve again (atol 1e-6,
    # maxiters = 100)
    rng = np.random.default_rng(70)
    A0 = rng.random((3, 3)) + 5 * np.eye(3)
    cache = lsx.init(lsx.LinearProblem(_csr(lsx, A0), rng.random(3)), lsx.MI355XKrylovJL_GMRES(), maxiters=100)
    sol1 = lsx.solve_(cache)
    assert sol1.retcode == lsx.ReturnCode.Success and len(sol1.u) == 3
    lsx.resize_(cache, 6)
    assert cache.isfresh
    A1 = rng.random((6, 6)) + 5 * np.eye(6)
    b1 = rng.random(6)
    expected = np.linalg.solve(A1, b1)
    cache.A = _csr(lsx, A1)
    cache.b = b1
    cache.u = np.zeros(6)
    sol2 = lsx.solve_(cache)
    assert sol2.retcode == lsx.ReturnCode.Success and len(sol2.u) == 6
    np.testing.assert_allclose(sol2.u, expected, atol=1e-6)


def test_krylovjl_gmres_options_map_to_the_device_solve(lsx):
    # gmres_restart = 0: no restart, memory = min(20, n) (src/iterative_wrappers.jl:211-214): the
    # same solve as SimpleGMRES(restart = false, memory = 20), and the oracle's; gmres_restart = k:
    # SimpleGMRES(restart = true, memory = k).  Without restart the basis grows to maxiters
    # (here 400 vectors: beyond the 128 whose bookkeeping the LDS holds at N = 400).
    P = problems(lsx)
    rp, ci, v = P.convection_diffusion(20, 2.0)
    N = 400
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    b = np.random.default_rng(3).standard_normal(N)
    d = np.full(N, 4.0)
    kw = dict(Pl=d, abstol=0.0, reltol=1e-10)
    k0 = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XKrylovJL_GMRES(), maxiters=100, **kw)
    s0 = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XSimpleGMRES(restart=False, memory=20), maxiters=100, **kw)
    np.testing.assert_array_equal(k0.u, s0.u)
    assert k0.iters == s0.iters and k0.retcode == s0.retcode
    ref = O.gmres(rp, ci, v, b, memory=20, restart=False, abstol=0.0, reltol=1e-10, maxiters=100, pl=d)
    assert ref["iters"] == k0.iters and ref["retcode"] == int(k0.retcode)
    np.testing.assert_allclose(k0.u, ref["u"], rtol=1e-9, atol=1e-12)
    k5 = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XKrylovJL_GMRES(gmres_restart=7), maxiters=300, **kw)
    s5 = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XSimpleGMRES(restart=True, memory=7), maxiters=300, **kw)
    np.testing.assert_array_equal(k5.u, s5.u)
    assert k5.iters == s5.iters
    big = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XKrylovJL_GMRES(), maxiters=400, Pl=d, abstol=0.0,
                    reltol=1e-13)
    ref = O.gmres(rp, ci, v, b, memory=20, restart=False, abstol=0.0, reltol=1e-13, maxiters=400, pl=d)
    assert big.iters == ref["iters"] and int(big.retcode) == ref["retcode"] and big.stats is None
    np.testing.assert_allclose(big.u, ref["u"], rtol=1e-8, atol=1e-12)


def test_max_memory_bounds_the_unrestarted_basis(lsx):
    # lsx_gmres_max_memory: 2048 vectors whatever N (the bookkeeping moves from LDS to HBM where
    # the LDS cannot hold it); a solve without restart at that bound runs, one vector more is
    # refused with LSX_EINVAL, and the batched layout (bookkeeping in LDS) refuses dimensions the
    # LDS cannot hold
    import ctypes
    from linearsolve_jl_amd import _abi
    P = problems(lsx)
    L = _abi.lib()
    for m in (8, 45, 90):
        rp, ci, v = P.convection_diffusion(m, 1.0)
        N = m * m
        h = lsx._CSRHandle(lsx.Context.get(0), lsx.CSRMatrix(rp, ci, v, (N, N)))
        lim = ctypes.c_int32()
        assert L.lsx_gmres_max_memory(h.h, ctypes.byref(lim)) == 0
        assert lim.value == 2048, (m, lim.value)
        b = np.ones(N)
        x = np.empty(N)
        st = _abi.GmresStats()
        rc = L.lsx_gmres_solve(h.h, b.ctypes.data, x.ctypes.data, lim.value, 0, 0.0, 1e-8, lim.value,
                               None, None, ctypes.byref(st))
        assert rc == 0 and st.retcode == 1 and st.iters <= lim.value
        S = sp.csr_matrix((v, ci, rp), shape=(N, N))
        assert np.linalg.norm(b - S @ x) <= 1.01e-8 * np.linalg.norm(b)
        rc = L.lsx_gmres_solve(h.h, b.ctypes.data, x.ctypes.data, lim.value, 0, 0.0, 1e-8, lim.value + 1,
                               None, None, ctypes.byref(st))
        assert rc == _abi.LSX_EINVAL
    rp, ci, v = P.convection_diffusion(64, 1.0)
    N = 64 * 64
    h = lsx._CSRHandle(lsx.Context.get(0), lsx.CSRMatrix(rp, ci, v, (N, N)))
    b = np.ones(N)
    x = np.empty(N)
    st = _abi.GmresStats()
    rc = L.lsx_gmres_solve_batched(h.h, N, b.ctypes.data, x.ctypes.data, 1000, 0, 0.0, 1e-8, 1000, None, None,
                                   ctypes.byref(st))
    assert rc == _abi.LSX_EINVAL


def test_unrestarted_beyond_lds_matches_oracle(lsx):
    # N = 4096: the LDS beside each workgroup's rows holds the bookkeeping of about 70 vectors; this
    # solve needs 218 without restart, so R, the Givens pairs and z live in HBM.  Iterates,
    # iteration count and retcode are the oracle's SimpleGMRES(restart = false)
    # (src/simplegmres.jl:222-235, 500-606).
    P = problems(lsx)
    m = 64
    rp, ci, v = P.convection_diffusion(m, 2.0)
    N = m * m
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    b = np.random.default_rng(3).standard_normal(N)
    d = np.full(N, 4.0)
    sol = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XSimpleGMRES(restart=False, memory=20), Pl=d,
                    abstol=0.0, reltol=1e-10, maxiters=600)
    ref = O.gmres(rp, ci, v, b, memory=20, restart=False, abstol=0.0, reltol=1e-10, maxiters=600, pl=d)
    print(f"unrestarted N = {N}: {sol.iters} iterations (oracle {ref['iters']})", flush=True)
    assert ref["iters"] > 100
    assert sol.iters == ref["iters"] and int(sol.retcode) == ref["retcode"] and sol.stats is None
    np.testing.assert_allclose(sol.u, ref["u"], rtol=1e-8, atol=1e-10 * np.abs(ref["u"]).max())


def test_unrestarted_beyond_the_device_bound_warns(lsx):
    # maxiters above 2048 without restart: the solve restarts every 2048 iterations and says so
    P = problems(lsx)
    rp, ci, v = P.convection_diffusion(20, 2.0)
    N = 400
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    b = np.random.default_rng(4).standard_normal(N)
    with pytest.warns(RuntimeWarning, match="restarts every 2048"):
        sol = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XSimpleGMRES(restart=False), maxiters=3000,
                        abstol=0.0, reltol=1e-10)
    assert sol.stats == {"krylov_restart": 2048}
    assert sol.retcode == lsx.ReturnCode.Success
    ref = O.gmres(rp, ci, v, b, memory=2048, restart=True, abstol=0.0, reltol=1e-10, maxiters=3000)
    assert sol.iters == ref["iters"]
    np.testing.assert_allclose(sol.u, ref["u"], rtol=1e-8, atol=1e-12)
