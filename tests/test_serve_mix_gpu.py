"""A slice of the serving workload, drawn from its own distributions: Poisson arrivals at
2000 requests/s; 70 % dense systems with n log-uniform in [16, 4096], A = I - g J
(J iid N(0,1)/sqrt(n), g ~ U[1e-3, 1e-1]), b iid N(0,1); 30 % SimpleGMRES(restart, memory 30,
reltol 1e-8, default abstol, Jacobi Pl) on 5-point convection-diffusion grids with side
m ~ U{64..1024} and Peclet ~ U[0, 10], maxiters = N (the reference's default, src/common.jl:764),
so every solve runs to its stopping rule.

Each result is checked against the reference's arithmetic:
  - dense: backward error <= 1e-13 and the solution against LAPACK's dgetrf/dgetrs (scipy's
    OpenBLAS); systems above 512 rows run on the dense lanes, whose result must equal the plain
    LUFactorization solve bit for bit, and that solve's pivots equal LAPACK's;
  - SimpleGMRES: grids with side <= 128 against the oracle's SimpleGMRES (retcode, iterations
    within 1, solution); every grid by its true preconditioned residual
    ||Pl \\ (b - A x)||_2 <= 1.5 eps, eps = abstol + reltol ||Pl \\ b||_2 (src/simplegmres.jl:110-129);
  - a sample of requests served again alone gives the same bits as in the mix.
"""
import time

import numpy as np
import pytest
import scipy.linalg.lapack as lapack
import scipy.sparse as sp

from conftest import backward_error
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RATE = 2000.0


def draw(count, seed=5):
    from linearsolve_jl_amd import problems as P
    g = P.rng_for(5, count, seed)
    arrivals = np.cumsum(g.exponential(1.0 / RATE, size=count))
    reqs = []
    for i in range(count):
        if g.random() < 0.7:
            n = int(round(np.exp(g.uniform(np.log(16), np.log(4096)))))
            gam = g.uniform(1e-3, 1e-1)
            A = -gam * g.standard_normal((n, n)) / np.sqrt(n)
            A[np.diag_indices(n)] += 1.0
            reqs.append(("dense", np.asfortranarray(A), g.standard_normal(n)))
        else:
            m = int(g.integers(64, 1024 + 1))
            pe = float(g.uniform(0.0, 10.0))
            rp, ci, v = P.convection_diffusion(m, pe)
            reqs.append(("gmres", (m, pe, rp, ci, v), g.standard_normal(m * m)))
    return arrivals, reqs


def submit(lsx, srv, req):
    kind, A, b = req
    if kind == "dense":
        return srv.submit_dense(A, b)
    m, pe, rp, ci, v = A
    N = m * m
    return srv.submit_gmres(lsx.CSRMatrix(rp, ci, v, (N, N)), b, pl=np.full(N, 4.0), memory=30, reltol=1e-8,
                            maxiters=N)


def wait_all(srv, tickets, timeout):
    out = {}
    t0 = time.time()
    while len(out) < len(tickets):
        for t in tickets:
            if t not in out:
                r = srv.poll(t)
                if r is not None:
                    out[t] = r
        assert time.time() - t0 < timeout, f"{len(tickets) - len(out)} serving results did not arrive"
        time.sleep(0.002)
    return out


def test_serving_mix_slice(lsx):
    count = 60
    arrivals, reqs = draw(count)
    kinds = [r[0] for r in reqs]
    print(f"serving mix slice: {kinds.count('dense')} dense, {kinds.count('gmres')} SimpleGMRES "
          f"(grid sides {sorted(r[1][0] for r in reqs if r[0] == 'gmres')})", flush=True)
    srv = lsx.Server(hbm_budget=48 << 30, nstreams=4)
    t0 = time.time()
    tickets = []
    for at, req in zip(arrivals, reqs):
        dt = at - (time.time() - t0)
        if dt > 0:
            time.sleep(dt)
        tickets.append(submit(lsx, srv, req))
    res = wait_all(srv, tickets, timeout=600.0)
    print(f"served {count} requests in {time.time() - t0:.1f} s", flush=True)
    srv.close()
    eps_def = np.sqrt(np.finfo(float).eps)
    for i, (req, t) in enumerate(zip(reqs, tickets)):
        kind, A, b = req
        r = res[t]
        assert r["status"] == 0, (i, r)
        assert r["retcode"] == lsx.ReturnCode.Success, (i, kind, r["retcode"], r["iters"])
        x = r["x"]
        if kind == "dense":
            n = A.shape[0]
            assert backward_error(A, x, b) <= 1e-13, i
            lu, piv, info = lapack.dgetrf(A)
            xl, _ = lapack.dgetrs(lu, piv, b)
            assert info == 0
            # A = I - g J with g <= 0.1 is well conditioned: the two solutions agree to rounding
            np.testing.assert_allclose(x, xl, rtol=0, atol=1e-13 * np.abs(xl).max())
            if n > 512:  # dense lanes: the plain LUFactorization kernels, bit for bit
                h = lsx._LUHandle(lsx.Context.get(0), n, lsx.LSX_F64)
                assert h.factor(A) == 0
                _, ipiv = h.factors()
                np.testing.assert_array_equal(ipiv, piv + 1)
                xp = np.empty(n)
                h.solve(b, xp)
                np.testing.assert_array_equal(x, xp)
        else:
            m, pe, rp, ci, v = A
            N = m * m
            S = sp.csr_matrix((v, ci, rp), shape=(N, N))
            eps_ = eps_def + 1e-8 * np.linalg.norm(b / 4.0)
            assert np.linalg.norm((b - S @ x) / 4.0) <= 1.5 * eps_, (i, m, pe)
            if m <= 128:
                ref = O.gmres(rp, ci, v, b, memory=30, restart=True, reltol=1e-8, maxiters=N, pl=np.full(N, 4.0))
                assert ref["retcode"] == O.RET_SUCCESS
                assert abs(r["iters"] - ref["iters"]) <= 1, (r["iters"], ref["iters"])
                np.testing.assert_allclose(x, ref["u"], rtol=1e-5, atol=1e-5 * np.abs(ref["u"]).max())
    # alone == in the mix: the two smallest grids and four dense systems of both lanes
    gm = sorted((i for i, q in enumerate(reqs) if q[0] == "gmres"), key=lambda i: reqs[i][1][0])[:2]
    dn = [i for i, q in enumerate(reqs) if q[0] == "dense"]
    dn = sorted(dn, key=lambda i: reqs[i][1].shape[0])
    sample = gm + dn[:2] + dn[-2:]
    for i in sample:
        srv1 = lsx.Server(hbm_budget=8 << 30, nstreams=2)
        t1 = submit(lsx, srv1, reqs[i])
        r1 = wait_all(srv1, [t1], timeout=120.0)[t1]
        srv1.close()
        np.testing.assert_array_equal(r1["x"], res[tickets[i]]["x"], err_msg=f"request {i} alone vs in the mix")
