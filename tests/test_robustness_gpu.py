"""Error paths and rarely taken branches on the GPU:

- the Infeasible termination of SimpleGMRES (inconsistent back-substitution,
  src/simplegmres.jl:576-585, retcode order :603-606) against the oracle;
- the post-solve residual check of LUFactorization(residualsafety = true)
  (src/factorization.jl:127-156, test/Core/default_algs.jl:487-530, test/Core/batch.jl:52-57);
- bounded device-side waits: with a wait bound of one clock tick every wait that has to wait gives
  up, the call returns LSX_ETIMEOUT promptly (the other waits of the call give up at once), the
  kernels end normally, and the handle works again once the default bound is back.
"""
import time

import numpy as np
import pytest
import scipy.linalg.lapack as lapack
import scipy.sparse as sp

from conftest import backward_error
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LSX_ETIMEOUT = -5

INCONSISTENT = [
    # nilpotent: A e2 = e1, A e1 = 0; b = e2 is not in range(A)
    (np.array([[0.0, 1.0], [0.0, 0.0]]), np.array([0.0, 1.0])),
    (np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]]), np.array([1.0, 0.0, 1.0])),
    (np.diag([1.0, 2.0, 0.0]), np.array([1.0, 1.0, 1.0])),
]


@pytest.mark.parametrize("case", range(len(INCONSISTENT)))
@pytest.mark.parametrize("restart", [True, False])
def test_gmres_infeasible_matches_oracle(lsx, case, restart):
    Ad, b = INCONSISTENT[case]
    print(f"gmres infeasible case={case} restart={restart}", flush=True)
    S = sp.csr_matrix(Ad)
    S.sort_indices()
    A = lsx.CSRMatrix.from_scipy(S)
    sol = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XSimpleGMRES(restart=restart, memory=20), maxiters=10)
    ref = O.gmres(A.rowptr, A.colind, A.vals, b, memory=20, restart=restart, maxiters=10)
    assert ref["retcode"] == O.RET_INFEASIBLE
    assert sol.retcode == lsx.ReturnCode.Infeasible
    assert sol.iters == ref["iters"]
    # y_k is zeroed where |R_kk| <= btol, so the iterate does not depend on rounding
    np.testing.assert_allclose(np.asarray(sol.u), ref["u"], rtol=1e-12, atol=1e-15)
    # The returned resid is |zeta_{k+1}| = |s_k zeta_k|, and at this breakdown the last Givens
    # rotation is taken between R_kk and Hbis, which are both rounding noise (about 1e-16 in
    # cases 1 and 2), so s_k is any value in [0, 1] that the summation order leaves.  What is fixed
    # is the bound |zeta_{k+1}| <= |zeta_k|: the estimate after iters - 1 steps.
    prev = O.gmres(A.rowptr, A.colind, A.vals, b, memory=20, restart=restart, maxiters=ref["iters"] - 1)
    assert 0.0 <= sol.resid <= prev["resid"] * (1 + 1e-12)
    if case == 0:  # exact zeros throughout: the rotation is the identity and the estimate is exact
        assert sol.resid == ref["resid"]
    # KrylovJL_GMRES maps an inconsistent system to Failure (src/iterative_wrappers.jl:733-745)
    solk = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XKrylovJL_GMRES(gmres_restart=20), maxiters=10)
    assert solk.retcode == lsx.ReturnCode.Failure


def near_singular(n=20, seed=42):
    """test/Core/default_algs.jl:434-441: Q diag(1e-15, 1, ..., 1) Q^T plus 1e-16 noise."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    U = Q * np.concatenate([[1e-15], np.ones(n - 1)])
    A = U @ U.T
    A += 1e-16 * rng.standard_normal((n, n))
    return A, rng.standard_normal(n)


def test_residualsafety_near_singular_fails(lsx):
    A, b = near_singular()
    sol = lsx.solve(lsx.LinearProblem(A.copy(), b.copy()), lsx.MI355XLUFactorization(residualsafety=True))
    assert sol.retcode == lsx.ReturnCode.APosterioriSafetyFailure
    # without the check the garbage solve is reported as Success (default_algs.jl:518-523)
    sol0 = lsx.solve(lsx.LinearProblem(A.copy(), b.copy()), lsx.MI355XLUFactorization())
    assert sol0.retcode == lsx.ReturnCode.Success


def test_residualsafety_well_conditioned_succeeds(lsx):
    # default_algs.jl:525-530
    A = np.array([[4.0, 1.0], [1.0, 3.0]])
    b = np.array([1.0, 2.0])
    sol = lsx.solve(lsx.LinearProblem(A.copy(), b.copy()), lsx.MI355XLUFactorization(residualsafety=True))
    assert sol.retcode == lsx.ReturnCode.Success
    np.testing.assert_allclose(sol.u, np.linalg.solve(A, b), rtol=1e-14)


def test_residualsafety_matrix_rhs(lsx):
    # test/Core/batch.jl:52-57: A = rand(n, n) + n I, B = rand(n, k)
    rng = np.random.default_rng(3)
    n, k = 50, 3
    A = rng.random((n, n)) + n * np.eye(n)
    B = rng.random((n, k))
    sol = lsx.solve(lsx.LinearProblem(A.copy(), B.copy()), lsx.MI355XLUFactorization(residualsafety=True))
    assert sol.retcode == lsx.ReturnCode.Success
    np.testing.assert_allclose(sol.u, np.linalg.solve(A, B), rtol=1e-12)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (63, 1), (64, 2), (1000, 1), (1000, 17), (4099, 16)])
def test_residual_norm_matches_numpy(lsx, n, nrhs):
    print(f"residual norm n={n} nrhs={nrhs}", flush=True)
    rng = np.random.default_rng(n + 7 * nrhs)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs))
    X = rng.standard_normal((n, nrhs))
    h = lsx._LUHandle(lsx.Context.get(0), n, lsx.LSX_F64)
    h.factor(A)
    rn, bn = h.residual_norm(A, B, X)
    np.testing.assert_allclose(rn, np.linalg.norm(B - A @ X), rtol=1e-12)
    np.testing.assert_allclose(bn, np.linalg.norm(B), rtol=1e-13)
    # device-resident A, B, X give the same bits
    import torch
    dA = torch.from_numpy(A).cuda()
    rn2, bn2 = h.residual_norm(dA, torch.from_numpy(B).cuda(), torch.from_numpy(X).cuda())
    assert rn2 == rn and bn2 == bn


def _expect_timeout(fn):
    t0 = time.time()
    with pytest.raises(Exception) as ei:
        fn()
    dt = time.time() - t0
    assert getattr(ei.value, "code", None) == LSX_ETIMEOUT, ei.value
    return dt


def test_wait_bound_gmres_gives_up_promptly(lsx):
    from linearsolve_jl_amd import problems as P
    ctx = lsx.Context.get(0)
    m = 100  # N = 10000 rows: two workgroups, so every all-reduce waits on the other one
    rp, ci, v = P.convection_diffusion(m, 3.0)
    N = m * m
    b = P.rng_for(11, m).standard_normal(N)
    d = np.full(N, 4.0)
    A = lsx.CSRMatrix(rp, ci, v, (N, N))
    cache = lsx.init(lsx.LinearProblem(A, b.copy()), lsx.MI355XSimpleGMRES(restart=True, memory=30), Pl=d,
                     reltol=1e-8, maxiters=3000)
    print("gmres wait bound 1e-9 s", flush=True)
    ctx.set_wait_bound(1e-9)
    try:
        dt = _expect_timeout(lambda: lsx.solve_(cache))
    finally:
        ctx.set_wait_bound(0.0)
    assert dt < 10.0, dt
    # the same handle solves correctly again
    sol = lsx.solve_(cache)
    ref = lsx.solve(lsx.LinearProblem(A, b.copy()), lsx.MI355XSimpleGMRES(restart=True, memory=30), Pl=d,
                    reltol=1e-8, maxiters=3000)
    assert sol.retcode == ref.retcode == lsx.ReturnCode.Success
    assert sol.iters == ref.iters
    np.testing.assert_array_equal(np.asarray(sol.u), np.asarray(ref.u))


def test_wait_bound_lu_gives_up_promptly(lsx):
    ctx = lsx.Context.get(0)
    n = 2048  # four panel workgroups and 32 triangular-solve blocks
    rng = np.random.default_rng(77)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal(n)
    h = lsx._LUHandle(ctx, n, lsx.LSX_F64)
    print("getrf wait bound 1e-9 s", flush=True)
    ctx.set_wait_bound(1e-9)
    try:
        dt = _expect_timeout(lambda: h.factor(A))
    finally:
        ctx.set_wait_bound(0.0)
    assert dt < 10.0, dt
    assert h.factor(A) == 0
    F, ipiv = h.factors()
    _, piv, _ = lapack.dgetrf(A)
    np.testing.assert_array_equal(ipiv, piv + 1)
    x = np.empty(n)
    print("getrs wait bound 1e-9 s", flush=True)
    ctx.set_wait_bound(1e-9)
    try:
        dt = _expect_timeout(lambda: h.solve(b, x))
    finally:
        ctx.set_wait_bound(0.0)
    assert dt < 10.0, dt
    h.solve(b, x)
    assert backward_error(A, x, b) <= 1e-13


def test_serve_timeout_reports_failure(lsx):
    from linearsolve_jl_amd import problems as P
    ctx = lsx.Context.get(0)
    srv = lsx.Server(ctx, hbm_budget=4 << 30, nstreams=2)
    try:
        m = 100
        rp, ci, v = P.convection_diffusion(m, 1.0)
        N = m * m
        b = P.rng_for(12, m).standard_normal(N)
        A = lsx.CSRMatrix(rp, ci, v, (N, N))
        Ad = np.random.default_rng(5).standard_normal((1024, 1024))
        bd = np.random.default_rng(6).standard_normal(1024)
        print("serve wait bound 1e-9 s", flush=True)
        ctx.set_wait_bound(1e-9)
        try:
            tg = srv.submit_gmres(A, b, pl=np.full(N, 4.0), memory=30, reltol=1e-8, maxiters=3000)
            td = srv.submit_dense(Ad, bd)
            srv.drain()
        finally:
            ctx.set_wait_bound(0.0)
        for t in (tg, td):
            r = srv.poll(t)
            assert r["status"] == LSX_ETIMEOUT, r
            assert r["retcode"] == lsx.ReturnCode.Failure
        # the server keeps serving
        tg = srv.submit_gmres(A, b, pl=np.full(N, 4.0), memory=30, reltol=1e-8, maxiters=3000)
        td = srv.submit_dense(Ad, bd)
        srv.drain()
        rg, rd = srv.poll(tg), srv.poll(td)
        assert rg["status"] == 0 and rg["retcode"] == lsx.ReturnCode.Success
        assert rd["status"] == 0 and rd["retcode"] == lsx.ReturnCode.Success
        assert backward_error(Ad, rd["x"], bd) <= 1e-13
    finally:
        srv.close()
