"""lsx_residual_norm_noalloc, the entry point that allocates nothing (its partial sums live in a
file-scope device array), against lsx_lu_residual_norm, the entry point it is a variant of:
the same two norms bit for bit, on two calls in a row (the second with a smaller input, so stale
scratch contents from the first call would show), and against numpy.  Also the host-side
refusals: an order above the scratch bound and host arrays."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LSX_EINVAL = -1
NOALLOC_MAX_N = 65536


def _noalloc(lsx, ctx, n, nrhs, A, B, X):
    """Calls the C-ABI directly with device (torch) arrays in column-major layout: A is n x n,
    B and X are n x nrhs (torch arrays of shape (nrhs, n), the transposes)."""
    from linearsolve_jl_amd import _abi
    assert tuple(A.shape) == (n, n) and tuple(B.shape) == (nrhs, n) and tuple(X.shape) == (nrhs, n)
    out = (ctypes.c_double * 2)()
    rc = _abi.lib().lsx_residual_norm_noalloc(ctx.handle, n, A.data_ptr(), n, nrhs, B.data_ptr(), n, X.data_ptr(), n,
                                              out)
    return rc, (out[0], out[1])


def _dev_colmajor(M):
    import torch
    # column-major n x k: the transpose of a C-contiguous k x n array
    return torch.from_numpy(np.ascontiguousarray(M.T)).cuda()


@pytest.mark.parametrize("sizes", [((3000, 20), (1000, 3))])
def test_noalloc_matches_residual_norm(lsx, sizes):
    import torch
    ctx = lsx.Context.get(0)
    for call, (n, nrhs) in enumerate(sizes):
        print(f"noalloc residual norm call={call} n={n} nrhs={nrhs}", flush=True)
        rng = np.random.default_rng(100 + call)
        A = rng.standard_normal((n, n))
        B = rng.standard_normal((n, nrhs))
        X = rng.standard_normal((n, nrhs))
        dA, dB, dX = _dev_colmajor(A), _dev_colmajor(B), _dev_colmajor(X)
        torch.cuda.synchronize()
        rc, got = _noalloc(lsx, ctx, n, nrhs, dA, dB, dX)
        assert rc == 0, lsx._abi.lib().lsx_last_error(ctx.handle)
        # the existing entry point, through a handle of the same order
        h = lsx._LUHandle(ctx, n, lsx.LSX_F64)
        want = h.residual_norm(A, B, X)
        assert got == want, (got, want)
        np.testing.assert_allclose(got[0], np.linalg.norm(B - A @ X), rtol=1e-12)
        np.testing.assert_allclose(got[1], np.linalg.norm(B), rtol=1e-13)
        # the inputs are left as they were (the no-allocation form writes no residual)
        np.testing.assert_array_equal(dB.cpu().numpy().T, B)


def test_noalloc_refusals(lsx):
    import torch
    ctx = lsx.Context.get(0)
    n = 64
    A = np.eye(n)
    B = np.ones((n, 1))
    dA, dB = _dev_colmajor(A), _dev_colmajor(B)
    from linearsolve_jl_amd import _abi
    out = (ctypes.c_double * 2)()
    L = _abi.lib()
    # host arrays would need a staging copy: refused
    Ah = np.asfortranarray(A)
    Bh = np.asfortranarray(B)
    rc = L.lsx_residual_norm_noalloc(ctx.handle, n, Ah.ctypes.data, n, 1, Bh.ctypes.data, n, Bh.ctypes.data, n, out)
    assert rc == LSX_EINVAL
    # an order beyond the scratch bound is refused before anything is read
    rc = L.lsx_residual_norm_noalloc(ctx.handle, NOALLOC_MAX_N + 1, dA.data_ptr(), NOALLOC_MAX_N + 1, 1,
                                     dB.data_ptr(), NOALLOC_MAX_N + 1, dB.data_ptr(), NOALLOC_MAX_N + 1, out)
    assert rc == LSX_EINVAL
    assert b"scratch bound" in L.lsx_last_error(ctx.handle)
    # a valid call still works after the refusals: B - I B = 0
    rc = L.lsx_residual_norm_noalloc(ctx.handle, n, dA.data_ptr(), n, 1, dB.data_ptr(), n, dB.data_ptr(), n, out)
    assert rc == 0 and out[0] == 0.0 and out[1] == np.sqrt(n)
    torch.cuda.synchronize()
