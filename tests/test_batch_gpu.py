"""Batched small dense LU on the GPU (one persistent launch) against per-block LAPACK / the oracle.

Reference: BlockDiagonalFactorization's blockwise lu!/ldiv! (ext/LinearSolveBlockDiagonalsExt.jl)
and its tests (test/Core/basictests.jl:1168-1195).  Per block: pivots bit-exact, backward error
<= 1e-13, info equal to LAPACK's.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from conftest import backward_error
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run_batch(lsx, blocks, b):
    A = lsx.BlockDiagonal(blocks)
    alg = lsx.MI355XBatchedLUFactorization()
    sol = lsx.solve(lsx.LinearProblem(A, b), alg)
    return sol, alg


def test_reference_block_sizes(lsx):
    rng = np.random.default_rng(0)
    for bsizes in ([3, 3, 3, 3], [2, 3, 4]):
        blocks = [rng.standard_normal((s, s)) + s * np.eye(s) for s in bsizes]
        b = rng.standard_normal(sum(bsizes))
        sol, alg = run_batch(lsx, blocks, b)
        assert sol.retcode == lsx.ReturnCode.Success
        D = lsx.BlockDiagonal(blocks).to_dense()
        np.testing.assert_allclose(sol.u, np.linalg.solve(D, b), rtol=1e-12, atol=1e-13)


def test_random_sizes_pivots_bit_exact(lsx):
    from linearsolve_jl_amd import problems as P
    rng = np.random.default_rng(1)
    sizes = list(rng.integers(1, 513, size=150)) + [1, 2, 15, 16, 17, 31, 32, 33, 255, 256, 511, 512]
    blocks = [rng.standard_normal((int(s), int(s))) for s in sizes]
    b = rng.standard_normal(int(sum(sizes)))
    sol, alg = run_batch(lsx, blocks, b)
    assert sol.retcode == lsx.ReturnCode.Success
    off = 0
    for i, Bk in enumerate(blocks):
        n = Bk.shape[0]
        lu, piv, info = lapack.dgetrf(Bk)
        assert info == 0
        np.testing.assert_array_equal(alg.last_ipiv[off:off + n], piv + 1, err_msg=f"block {i} n={n}")
        assert backward_error(Bk, sol.u[off:off + n], b[off:off + n]) <= 1e-13, f"block {i} n={n}"
        off += n


def test_singular_block_info(lsx):
    rng = np.random.default_rng(2)
    blocks = [rng.standard_normal((40, 40)) for _ in range(3)]
    blocks[1][:, 5] = 0.0
    b = rng.standard_normal(120)
    sol, alg = run_batch(lsx, blocks, b)
    assert sol.retcode == lsx.ReturnCode.Failure
    assert alg.last_info[0] == 0 and alg.last_info[2] == 0
    assert alg.last_info[1] == lapack.dgetrf(blocks[1])[2] == O.dgetrf(blocks[1])[2]


def test_config4_full_batch(lsx):
    # 20,000 systems, n ~ U{8..512}, A = I - g J: every block solved to backward error <= 1e-13,
    # and a sample of blocks has LAPACK's pivots
    from linearsolve_jl_amd import problems as P
    print("batched config: 20000 systems", flush=True)
    sizes, Ap, bp = P.batched_systems()
    A = lsx.BlockDiagonal.from_packed(sizes, Ap)
    alg = lsx.MI355XBatchedLUFactorization()
    sol = lsx.solve(lsx.LinearProblem(A, bp), alg)
    assert sol.retcode == lsx.ReturnCode.Success
    ao = vo = 0
    worst = 0.0
    rng = np.random.default_rng(3)
    sample = set(rng.choice(len(sizes), size=200, replace=False).tolist())
    for i, n in enumerate(sizes):
        n = int(n)
        Ai = Ap[ao:ao + n * n].reshape((n, n), order="F")
        xi = sol.u[vo:vo + n]
        bi = bp[vo:vo + n]
        r = bi - Ai @ xi
        be = np.abs(r).max() / (np.abs(Ai).sum(axis=1).max() * np.abs(xi).max() + np.abs(bi).max())
        worst = max(worst, be)
        if i in sample:
            piv = lapack.dgetrf(Ai)[1]
            np.testing.assert_array_equal(alg.last_ipiv[vo:vo + n], piv + 1)
        ao += n * n
        vo += n
    print(f"worst backward error {worst:.3e}", flush=True)
    assert worst <= 1e-13


def test_fp32_refined_blocks(lsx):
    # FP32 factorization + FP64 refinement: every block reaches the refinement target in a few steps
    rng = np.random.default_rng(5)
    sizes = list(rng.integers(1, 513, size=120)) + [1, 2, 8, 9, 16, 17, 64, 65, 511, 512]
    blocks = []
    for s in sizes:
        s = int(s)
        J = rng.standard_normal((s, s)) / np.sqrt(s)
        blocks.append(np.eye(s) - rng.uniform(1e-3, 1e-1) * J if s % 2 else J + 2.0 * np.eye(s))
    b = rng.standard_normal(int(sum(sizes)))
    alg = lsx.MI355XBatchedLUFactorization(precision="fp32", tol=1e-14)
    sol = lsx.solve(lsx.LinearProblem(lsx.BlockDiagonal(blocks), b), alg)
    assert sol.retcode == lsx.ReturnCode.Success
    assert np.all(alg.last_iters >= 1) and np.all(alg.last_iters <= 10)
    off = 0
    for i, Bk in enumerate(blocks):
        n = Bk.shape[0]
        be = backward_error(Bk, sol.u[off:off + n], b[off:off + n])
        # the kernel's own FP64 residual met 1e-14; a residual recomputed in another summation
        # order may differ in the last bits
        assert be <= 2e-14, f"block {i} n={n}: {be}"
        assert alg.last_backward_error[i] <= 1e-14
        off += n


@pytest.mark.parametrize("bw", [None, 8, 16, 32])
def test_fp32_refined_pivoting_blocks(lsx, bw):
    # FP32 factors of systems that pivot in almost every column (iid N(0,1) blocks), sizes on both
    # sides of 256 rows and of the panel widths, with the tile table and with each panel width:
    # the interchanges applied in memory must leave factors that refine every block to the target
    rng = np.random.default_rng(55)
    sizes = list(rng.integers(1, 513, size=60)) + [1, 7, 8, 9, 15, 16, 17, 31, 32, 33, 255, 256, 257, 511, 512]
    blocks = [rng.standard_normal((int(s), int(s))) for s in sizes]
    b = rng.standard_normal(int(sum(sizes)))
    alg = lsx.MI355XBatchedLUFactorization(precision="fp32", tol=1e-14, tiles="auto" if bw is None else [(512, bw)])
    sol = lsx.solve(lsx.LinearProblem(lsx.BlockDiagonal(blocks), b), alg)
    assert sol.retcode == lsx.ReturnCode.Success
    assert np.all(alg.last_info == 0) and np.all(alg.last_iters <= 10)
    off = 0
    for i, Bk in enumerate(blocks):
        n = Bk.shape[0]
        assert backward_error(Bk, sol.u[off:off + n], b[off:off + n]) <= 2e-14, f"block {i} n={n}"
        assert alg.last_backward_error[i] <= 1e-14
        off += n


def test_fp32_singular_block(lsx):
    rng = np.random.default_rng(6)
    blocks = [rng.standard_normal((30, 30)) + 5 * np.eye(30) for _ in range(3)]
    blocks[2][:, 7] = 0.0
    b = rng.standard_normal(90)
    alg = lsx.MI355XBatchedLUFactorization(precision="fp32")
    sol = lsx.solve(lsx.LinearProblem(lsx.BlockDiagonal(blocks), b), alg)
    assert sol.retcode == lsx.ReturnCode.Failure
    assert alg.last_info[0] == 0 and alg.last_info[1] == 0 and alg.last_info[2] > 0


@pytest.mark.parametrize("bw", [8, 16, 32])
def test_tile_widths_keep_lapack_pivots(lsx, bw):
    # every panel width of the tile table gives LAPACK's pivots and a backward error <= 1e-13
    rng = np.random.default_rng(7)
    sizes = list(rng.integers(1, 513, size=60)) + [7, 8, 9, 15, 16, 17, 31, 32, 33, 63, 64, 65, 512]
    blocks = [rng.standard_normal((int(s), int(s))) for s in sizes]
    b = rng.standard_normal(int(sum(sizes)))
    alg = lsx.MI355XBatchedLUFactorization(tiles=[(512, bw)])
    sol = lsx.solve(lsx.LinearProblem(lsx.BlockDiagonal(blocks), b), alg)
    assert sol.retcode == lsx.ReturnCode.Success
    off = 0
    for i, Bk in enumerate(blocks):
        n = Bk.shape[0]
        np.testing.assert_array_equal(alg.last_ipiv[off:off + n], lapack.dgetrf(Bk)[1] + 1,
                                      err_msg=f"bw {bw} block {i} n={n}")
        assert backward_error(Bk, sol.u[off:off + n], b[off:off + n]) <= 1e-13
        off += n


def test_graph_replays_match_direct_launch(lsx, monkeypatch):
    # the captured launch replayed on changing inputs, with other work in between, gives the same
    # bits as the directly launched kernel on every replay
    import torch
    rng = np.random.default_rng(8)
    sizes = rng.integers(1, 400, size=300)
    ctx = lsx.Context.get(0)
    hg = lsx._BatchHandle(ctx, sizes, lsx.LSX_F64, tiles=None)
    monkeypatch.setenv("LSX_BATCH_NO_GRAPH", "1")
    hd = lsx._BatchHandle(ctx, sizes, lsx.LSX_F64, tiles=None)
    monkeypatch.delenv("LSX_BATCH_NO_GRAPH")
    tot = int((sizes * sizes).sum())
    for rep in range(4):
        A = rng.standard_normal(tot)
        b = rng.standard_normal(int(sizes.sum()))
        xg = np.empty_like(b)
        xd = np.empty_like(b)
        hg.factor_solve(A, b, xg)
        t = torch.randn(1 << 20, device="cuda").sum()   # unrelated device work between replays
        torch.cuda.synchronize()
        hd.factor_solve(A, b, xd)
        np.testing.assert_array_equal(xg, xd, err_msg=f"replay {rep}")
        assert float(t) == float(t)


def test_set_tiles_and_refinement_reject_bad_arguments(lsx):
    ctx = lsx.Context.get(0)
    h = lsx._BatchHandle(ctx, [4, 5], lsx.LSX_F32, tiles=None)
    with pytest.raises(lsx.LSXError):
        h.set_tiles([(512, 12)])
    with pytest.raises(lsx.LSXError):
        h.set_refinement(-1.0, 3)
