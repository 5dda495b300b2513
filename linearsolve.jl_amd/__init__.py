"""Host-side mirror of LinearSolve.jl's LinearProblem / init / solve! / LinearCache interface for
the gfx950 hot paths.

Reference interface (paths in LinearSolve.jl v5.12.0):
  LinearProblem + init + solve!          src/common.jl:700-712, 758-931, 988-1017
  LinearCache fields and setproperty!    src/common.jl:281-306, 313-360 (A set -> isfresh; b set)
  LUFactorization.solve!                 src/factorization.jl:643-736
  SimpleGMRES.solve!                     src/simplegmres.jl:375-616
  BlockDiagonalFactorization             ext/LinearSolveBlockDiagonalsExt.jl:48-185
  ReturnCode / LinearSolution            SciMLBase (build_linear_solution, src/common.jl:748-753)

`solve_(cache)` is `solve!(cache)` (Python has no `!`).  All arithmetic runs in liblsx.so on the
GPU; this module only moves arrays, keeps handles in `cache.cacheval` and maps return codes.
"""
import ctypes
import json
import enum
import math
import os
import warnings

import numpy as np

from . import _abi
from ._abi import LSX_BF16, LSX_F32, LSX_F64, LSXError, Context

__all__ = [
    "ReturnCode", "LinearProblem", "LinearCache", "LinearSolution", "init", "solve", "solve_",
    "MI355XLUFactorization", "MI355XMixedLUFactorization", "MI355XBatchedLUFactorization",
    "MI355XSimpleGMRES", "MI355XKrylovJL_GMRES", "BlockDiagonal", "CSRMatrix", "LSXError", "Context", "Server",
    "default_tol", "default_algorithm", "OperatorAssumptions", "resize_",
]


class ReturnCode(enum.IntEnum):
    """Subset of SciMLBase.ReturnCode used by the hot paths (values 0-4 match lsx_retcode; the
    post-solve residual check of src/factorization.jl:127-156 adds APosterioriSafetyFailure)."""
    Default = 0
    Success = 1
    MaxIters = 2
    Infeasible = 3
    Failure = 4
    APosterioriSafetyFailure = 5


def successful_retcode(rc):
    return rc == ReturnCode.Success


def default_tol(dtype=np.float64):
    """src/common.jl:411: sqrt(eps(T))."""
    return math.sqrt(np.finfo(dtype).eps)


# -------------------------------------------------------------------------------------------------
# array plumbing (numpy host arrays or torch tensors on the context's GPU)

def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ready(t):
    """Waits for the work torch has queued on the tensor's device.  The library reads and writes
    device arrays on its own non-blocking streams, which do not wait for torch's stream: an input
    still being written there (a transposed copy just made, a fresh randn) would otherwise be read
    half-written, and an output still being read there overwritten early."""
    import torch
    if t.is_cuda:
        torch.cuda.current_stream(t.device).synchronize()
    return t


def _colmajor(x):
    """Returns (keepalive, pointer) of a column-major float64 copy/view of a vector or matrix."""
    if _is_torch(x):
        import torch
        t = x.to(torch.float64)
        if t.dim() == 2:
            t = t.t().contiguous()  # row-major storage of A^T == column-major storage of A
        else:
            t = t.contiguous()
        return _ready(t), t.data_ptr()
    a = np.asfortranarray(np.asarray(x, dtype=np.float64))
    return a, a.ctypes.data


def _empty_like_rhs(b):
    if _is_torch(b):
        import torch
        if b.dim() == 2:
            return torch.empty((b.shape[1], b.shape[0]), dtype=torch.float64, device=b.device).t()
        return torch.empty(b.shape, dtype=torch.float64, device=b.device)
    return np.empty(b.shape, dtype=np.float64, order="F")


def _ptr_out(u):
    if _is_torch(u):
        return _ready(u).data_ptr()
    return u.ctypes.data


def _ncols(b):
    return 1 if len(b.shape) == 1 else int(b.shape[1])


def _copy(x):
    if _is_torch(x):
        return x.clone()
    return np.array(x, copy=True)


# -------------------------------------------------------------------------------------------------
# problem, cache, solution

class LinearProblem:
    """A x = b (SciMLBase.LinearProblem)."""

    def __init__(self, A, b, u0=None):
        self.A = A
        self.b = b
        self.u0 = u0


class LinearSolution:
    """Fields of SciMLBase.LinearSolution used by the reference (src/common.jl:748-753)."""

    def __init__(self, u, resid=None, retcode=ReturnCode.Success, iters=0, cache=None, stats=None):
        self.u = u
        self.resid = resid
        self.retcode = retcode
        self.iters = iters
        self.cache = cache
        self.stats = stats

    @property
    def successful(self):
        return self.retcode == ReturnCode.Success


class LinearCache:
    """LinearCache (src/common.jl:281-306).  Assigning `A` marks the cache fresh (refactorize on the
    next solve_); assigning `b` keeps the device-resident factors / Krylov workspace
    (src/common.jl:313-360)."""

    def __init__(self, A, b, u, alg, Pl, Pr, abstol, reltol, maxiters, verbose, ctx):
        object.__setattr__(self, "_A", A)
        object.__setattr__(self, "isfresh", True)
        self.b = b
        self.u = u
        self.alg = alg
        self.Pl = Pl
        self.Pr = Pr
        self.abstol = abstol
        self.reltol = reltol
        self.maxiters = maxiters
        self.verbose = verbose
        self.ctx = ctx
        self.cacheval = None

    @property
    def A(self):
        return self._A

    @A.setter
    def A(self, value):
        object.__setattr__(self, "_A", value)
        object.__setattr__(self, "isfresh", True)


CONDITIONS = ("WellConditioned", "IllConditioned", "VeryIllConditioned", "SuperIllConditioned")


class OperatorAssumptions:
    """OperatorAssumptions (src/common.jl:184-198): issq (None = from A's shape) and the
    conditioning the caller vouches for, IllConditioned by default."""

    def __init__(self, issq=None, condition="IllConditioned"):
        if condition not in CONDITIONS:
            raise ValueError(f"condition must be one of {CONDITIONS}")
        self.issq = issq
        self.condition = condition


def default_algorithm(A, b, assumptions=None):
    """The algorithm solve(prob) picks without one: a device slot of defaultalg mapped onto this
    build's algorithms, with the conditioning thresholds of the reference's dense route
    (src/default.jl:411-477: LU for Well- and IllConditioned, QR for VeryIllConditioned, SVD for
    SuperIllConditioned) instead of its GPU route's QR for the default IllConditioned
    (:277-306), and KrylovJL_GMRES for square operators without a direct solver (:320-343):
      dense square (numpy or torch)  -> MI355XLUFactorization
      CSRMatrix                      -> MI355XKrylovJL_GMRES (the device GMRES)
      BlockDiagonal                  -> MI355XBatchedLUFactorization (blockwise LU)
    QR, SVD and the non-square Krylov methods are not built here: those choices raise."""
    assumptions = assumptions or OperatorAssumptions()
    if isinstance(A, BlockDiagonal):
        return MI355XBatchedLUFactorization()
    if isinstance(A, CSRMatrix):
        if A.shape[0] != A.shape[1]:
            raise NotImplementedError("non-square sparse systems (LSMR / CRAIGMR) are not built")
        return MI355XKrylovJL_GMRES()
    shape = tuple(A.shape)
    issq = (shape[0] == shape[1]) if assumptions.issq is None else assumptions.issq
    if not issq or len(shape) != 2 or shape[0] != shape[1]:
        raise NotImplementedError("non-square dense systems take QR in the reference; QR is not built")
    if assumptions.condition in ("VeryIllConditioned", "SuperIllConditioned"):
        raise NotImplementedError(f"{assumptions.condition} takes QR / SVD in the reference; they are not built")
    return MI355XLUFactorization()


def init(prob, alg=None, *, alias_A=False, alias_b=False, abstol=None, reltol=None, maxiters=None,
         Pl=None, Pr=None, verbose=False, device=0, assumptions=None):
    """SciMLBase.init -> __init (src/common.jl:758-931): copies A and b unless aliased (dense and
    Krylov algorithms default to no aliasing, :525-544; simplegmres.jl:185-186), zero u, default
    tolerances sqrt(eps) (:411) and maxiters = length(b) (:764).  Without alg, the default
    algorithm for A (default_algorithm)."""
    if alg is None:
        alg = default_algorithm(prob.A, prob.b, assumptions)
    A = prob.A if alias_A else (_copy(prob.A) if not isinstance(prob.A, (CSRMatrix, BlockDiagonal)) else prob.A)
    b = prob.b if alias_b else _copy(prob.b)
    if isinstance(alg, MI355XSimpleGMRES) and len(b.shape) == 2:
        # src/common.jl:614-625: Krylov methods take a vector right-hand side only
        raise ValueError("SimpleGMRES does not support a matrix right-hand side b")
    if isinstance(alg, MI355XKrylovJL_GMRES) and len(b.shape) == 2:
        # the reference hands a matrix b to Krylov.jl's block GMRES (src/iterative_wrappers.jl:705-706)
        raise NotImplementedError("block GMRES for a matrix right-hand side is not built")
    u = _empty_like_rhs(b)
    if _is_torch(u):
        u.zero_()
    else:
        u[...] = 0.0
    tol = default_tol()
    n = b.shape[0]
    cache = LinearCache(A, b, u, alg, Pl, Pr, tol if abstol is None else abstol,
                        tol if reltol is None else reltol, n if maxiters is None else maxiters,
                        verbose, Context.get(device))
    cache.cacheval = alg.init_cacheval(cache)
    return cache


def solve_(cache, adjoint=False):
    """solve!(cache) (src/common.jl:1012-1017) -> solve!(cache, cache.alg).  adjoint=True solves
    adjoint(A) x = b (src/common.jl:1012-1027): a fresh cache is factorized first, then the
    factorization is reused where the algorithm holds one (MI355XLUFactorization: getrs with
    trans = 'T' on the device, src/adjoint_factorization.jl:162-176); otherwise the problem
    with adjoint(A) is solved (Krylov methods with their preconditioners swapped and adjointed,
    src/adjoint.jl:78-93; the others refactorized, src/adjoint_factorization.jl:245-264).
    Writes cache.u; retcode Success."""
    if not adjoint:
        return cache.alg.solve_(cache)
    alg = cache.alg
    if cache.isfresh:
        alg.solve_(cache)
    if isinstance(alg, MI355XLUFactorization) and not cache.isfresh:
        cache.cacheval.solve_adjoint(cache.b, cache.u)
    else:
        At = _adjoint_of(cache.A)
        if isinstance(alg, (MI355XSimpleGMRES, MI355XKrylovJL_GMRES)):
            # the GMRES methods have a right slot: Pl' = adjoint(Pr), Pr' = adjoint(Pl) (diagonal: themselves)
            sol = solve(LinearProblem(At, _copy(cache.b)), alg, abstol=cache.abstol, reltol=cache.reltol,
                        maxiters=cache.maxiters, Pl=cache.Pr, Pr=cache.Pl, device=cache.ctx.device)
        else:
            sol = solve(LinearProblem(At, _copy(cache.b)), alg, abstol=cache.abstol, reltol=cache.reltol,
                        device=cache.ctx.device)
        if _is_torch(cache.u):
            cache.u.copy_(sol.u if _is_torch(sol.u) else __import__("torch").from_numpy(np.asarray(sol.u)))
        else:
            cache.u[...] = _np(sol.u)
    return LinearSolution(cache.u, None, ReturnCode.Success, 0, cache)


def resize_(cache, i):
    """resize!(cache, i) (src/common.jl:362-366): the cache is marked fresh; the device handle is
    kept and replaced on the next solve only if the order changed (a handle of the same order
    and type is reused: refactorization allocates nothing on the device)."""
    object.__setattr__(cache, "isfresh", True)
    return cache


def _adjoint_of(A):
    """adjoint(A) for the matrix types the algorithms take (real data: the transpose)."""
    if isinstance(A, CSRMatrix):
        import scipy.sparse as sp
        S = sp.csr_matrix((A.vals, A.colind, A.rowptr), shape=A.shape).T.tocsr()
        S.sort_indices()
        return CSRMatrix(S.indptr, S.indices, S.data, S.shape)
    if isinstance(A, BlockDiagonal):
        return BlockDiagonal([B.T for B in A.blocks])
    if _is_torch(A):
        return A.t().contiguous()
    return np.ascontiguousarray(np.asarray(A).T)


def solve(prob, alg=None, **kwargs):
    """solve(prob, alg) = solve!(init(prob, alg)) (src/common.jl:988-993); solve(prob) picks
    default_algorithm."""
    return solve_(init(prob, alg, **kwargs))


# -------------------------------------------------------------------------------------------------
# dense LU

class _LUHandle:
    def __init__(self, ctx, n, dtype):
        self.ctx = ctx
        self.n = n
        self.dtype = dtype
        h = ctypes.c_void_p()
        ctx.check(_abi.lib().lsx_lu_create(ctx.handle, n, dtype, ctypes.byref(h)), "lsx_lu_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _abi._lib is not None:
            _abi.lib().lsx_lu_destroy(h)
            self.h = None

    def factor(self, A):
        keep, p = _colmajor(A)
        return self.ctx.check(_abi.lib().lsx_lu_factor(self.h, p, self.n), "lsx_lu_factor")

    def _out(self, u):
        """(target, pointer, ld, copy_back) for writing a column-major result into u."""
        if _is_torch(u):
            if u.dim() == 2 and not (u.stride(0) == 1 and u.stride(1) == self.n):
                import torch
                t = torch.empty((u.shape[1], u.shape[0]), dtype=torch.float64, device=u.device).t()
                return t, _ptr_out(t), self.n, True
            return u, _ptr_out(u), self.n, False
        if u.ndim == 2 and not u.flags.f_contiguous:
            t = np.empty(u.shape, dtype=np.float64, order="F")
            return t, t.ctypes.data, self.n, True
        return u, _ptr_out(u), self.n, False

    def solve(self, b, u):
        keep, p = _colmajor(b)
        nr = _ncols(b)
        t, up, ldu, back = self._out(u)
        self.ctx.check(_abi.lib().lsx_lu_solve(self.h, nr, p, self.n, up, ldu), "lsx_lu_solve")
        if back:
            u[...] = t
        return u

    def factors(self):
        F = np.empty((self.n, self.n), order="F")
        ipiv = np.empty(self.n, dtype=np.int64)
        self.ctx.check(_abi.lib().lsx_lu_get_factors(self.h, F.ctypes.data, self.n, ipiv.ctypes.data),
                       "lsx_lu_get_factors")
        return F, ipiv

    def solve_adjoint(self, b, u):
        keep, p = _colmajor(b)
        nr = _ncols(b)
        t, up, ldu, back = self._out(u)
        self.ctx.check(_abi.lib().lsx_lu_solve_adjoint(self.h, nr, p, self.n, up, ldu), "lsx_lu_solve_adjoint")
        if back:
            u[...] = t
        return u

    def refine(self, b, u, tol, max_iters, mode):
        """(info, stats): info > 0 when a refactorization in FP32 met an exactly-zero pivot."""
        keep, p = _colmajor(b)
        nr = _ncols(b)
        st = _abi.RefineStats()
        up = _ptr_out(u)
        info = self.ctx.check(_abi.lib().lsx_lu_refine(self.h, nr, p, self.n, up, self.n, tol, max_iters, mode,
                                                       ctypes.byref(st)), "lsx_lu_refine")
        return info, st

    def residual_norm(self, A, b, u):
        """(||b - A u||_2, ||b||_2) in FP64 on the device (norms over all columns)."""
        ka, pa = _colmajor(A)
        kb, pb = _colmajor(b)
        ku, pu = _colmajor(u)
        out = (ctypes.c_double * 2)()
        self.ctx.check(_abi.lib().lsx_lu_residual_norm(self.h, pa, self.n, _ncols(b), pb, self.n, pu, self.n, out),
                       "lsx_lu_residual_norm")
        return out[0], out[1]


class _DenseAlg:
    needs_concrete_A = True
    needs_square_A = True
    dtype = LSX_F64

    def init_cacheval(self, cache):
        # src/factorization.jl:840-849: a placeholder; the factorization happens in solve!
        return None

    def _fresh_handle(self, cache, n):
        cv = cache.cacheval
        if cv is None or cv.n != n or cv.dtype != self.dtype:
            cv = _LUHandle(cache.ctx, n, self.dtype)
            cache.cacheval = cv
        return cv


class MI355XLUFactorization(_DenseAlg):
    """Drop-in for LUFactorization (src/factorization.jl:306-314): FP64 partial-pivot getrf on
    gfx950, factors and pivots resident on the device until `cache.A` changes."""

    def __init__(self, pivot="RowMaximum", residualsafety=False):
        if pivot != "RowMaximum":
            raise ValueError("only RowMaximum pivoting is implemented")
        self.pivot = pivot
        self.residualsafety = residualsafety

    def solve_(self, cache):
        A = cache.A
        n = A.shape[0]
        if A.shape[0] != A.shape[1]:
            raise ValueError("MI355XLUFactorization needs a square A")
        # src/factorization.jl:646: the residual check runs on solves that (re)factorize
        check_safety = self.residualsafety and cache.isfresh
        if cache.isfresh:
            h = self._fresh_handle(cache, n)
            info = h.factor(A)
            if info > 0:
                # src/factorization.jl:716-722: Failure, cache stays fresh
                return LinearSolution(cache.u, None, ReturnCode.Failure, 0, cache)
            object.__setattr__(cache, "isfresh", False)
        cache.cacheval.solve(cache.b, cache.u)
        if check_safety:
            # src/factorization.jl:127-156: ||A y - b|| > abstol + reltol ||b|| (2-norms, the A
            # of the cache, which the device factorization leaves untouched), on the device
            rn, bn = cache.cacheval.residual_norm(cache.A, cache.b, cache.u)
            if rn > cache.abstol + cache.reltol * bn:
                return LinearSolution(cache.u, None, ReturnCode.APosterioriSafetyFailure, 0, cache)
        return LinearSolution(cache.u, None, ReturnCode.Success, 0, cache)


class MI355XMixedLUFactorization(_DenseAlg):
    """Mixed-precision LU: factor with FP32 arithmetic ("fp32": trailing products of three BF16
    pieces per operand, FP32 accuracy) or BF16 products ("bf16": two pieces, relative error about
    2^-17), FP32 accumulation, on the GPU; recover FP64 accuracy by iterative refinement or
    GMRES-IR with FP64 residuals.  API shape of OpenBLAS32MixedLUFactorization
    (src/openblas.jl:461-543), which stops at FP32 accuracy."""

    def __init__(self, precision="fp32", refinement="auto", tol=1e-13, max_iters=50):
        if precision not in ("fp32", "bf16"):
            raise ValueError("precision must be 'fp32' or 'bf16'")
        self.precision = precision
        self.dtype = LSX_F32 if precision == "fp32" else LSX_BF16
        self.mode = {"ir": 0, "gmres-ir": 1, "auto": 2}[refinement]
        self.tol = tol
        self.max_iters = max_iters
        self.last_stats = None

    def solve_(self, cache):
        A = cache.A
        n = A.shape[0]
        if cache.isfresh:
            h = self._fresh_handle(cache, n)
            info = h.factor(A)
            if info > 0:
                return LinearSolution(cache.u, None, ReturnCode.Failure, 0, cache)
            object.__setattr__(cache, "isfresh", False)
        info, st = cache.cacheval.refine(cache.b, cache.u, self.tol, self.max_iters, self.mode)
        self.last_stats = st
        if info > 0:
            # the FP32 refactorization of an escalated BF16 solve is singular: as a zero pivot
            return LinearSolution(cache.u, None, ReturnCode.Failure, 0, cache)
        rc = ReturnCode.Success if st.converged else ReturnCode.MaxIters
        return LinearSolution(cache.u, st.backward_error, rc, st.iterations, cache,
                              stats=dict(mode_used=st.mode_used, iterations=st.iterations, outer=st.outer,
                                         backward_error=st.backward_error, escalated=st.escalated))


def _np(x):
    if _is_torch(x):
        return x.detach().cpu().numpy()
    return np.asarray(x)


# -------------------------------------------------------------------------------------------------
# batched (block-diagonal) LU

class BlockDiagonal:
    """Square blocks on the diagonal (the BlockDiagonals.jl matrix type the reference's blockwise
    factorization takes)."""

    def __init__(self, blocks):
        self.blocks = [np.asarray(B, dtype=np.float64) for B in blocks]
        for B in self.blocks:
            if B.ndim != 2 or B.shape[0] != B.shape[1]:
                raise ValueError("blocks must be square")
        self.sizes = np.array([B.shape[0] for B in self.blocks], dtype=np.int64)
        n = int(self.sizes.sum())
        self.shape = (n, n)

    @classmethod
    def from_packed(cls, sizes, packed):
        obj = cls.__new__(cls)
        obj.sizes = np.asarray(sizes, dtype=np.int64)
        obj._packed = np.ascontiguousarray(packed, dtype=np.float64)
        obj.blocks = None
        n = int(obj.sizes.sum())
        obj.shape = (n, n)
        return obj

    def packed(self):
        """Blocks column-major, back to back (the lsx_batch layout)."""
        if getattr(self, "_packed", None) is None:
            self._packed = np.concatenate([np.asfortranarray(B).ravel(order="F") for B in self.blocks])
        return self._packed

    def to_dense(self):
        n = self.shape[0]
        D = np.zeros((n, n))
        off = 0
        P = self.packed()
        po = 0
        for s in self.sizes:
            s = int(s)
            D[off:off + s, off:off + s] = P[po:po + s * s].reshape((s, s), order="F")
            off += s
            po += s * s
        return D


TILE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tile_table.json")


def load_tile_table(path=TILE_TABLE, dtype=LSX_F64):
    """Size-binned panel widths written by autotune.py for the factor type dtype:
    {"bins": [[upper, bw], ...] (FP64 factors), "bins_fp32": [...] (FP32 factors)} (None if absent;
    FP32 batches fall back to the FP64 table when the file has none of their own)."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    key = "bins_fp32" if dtype == LSX_F32 and "bins_fp32" in d else "bins"
    return [(int(u), int(w)) for u, w in d[key]]


class _BatchHandle:
    def __init__(self, ctx, sizes, dtype, tiles="auto"):
        self.ctx = ctx
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        h = ctypes.c_void_p()
        ctx.check(_abi.lib().lsx_batch_create(ctx.handle, len(self.sizes), self.sizes.ctypes.data, dtype,
                                              ctypes.byref(h)), "lsx_batch_create")
        self.h = h
        table = load_tile_table(dtype=dtype) if tiles == "auto" else tiles
        if table:
            self.set_tiles(table)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _abi._lib is not None:
            _abi.lib().lsx_batch_destroy(h)
            self.h = None

    def set_tiles(self, table):
        up = np.ascontiguousarray([u for u, _ in table], dtype=np.int32)
        bw = np.ascontiguousarray([w for _, w in table], dtype=np.int32)
        self.ctx.check(_abi.lib().lsx_batch_set_tiles(self.h, len(up), up.ctypes.data, bw.ctypes.data),
                       "lsx_batch_set_tiles")

    def set_refinement(self, tol, max_iters):
        self.ctx.check(_abi.lib().lsx_batch_set_refinement(self.h, float(tol), int(max_iters)),
                       "lsx_batch_set_refinement")

    def stats(self):
        it = np.zeros(len(self.sizes), dtype=np.int32)
        be = np.zeros(len(self.sizes), dtype=np.float64)
        self.ctx.check(_abi.lib().lsx_batch_stats(self.h, it.ctypes.data, be.ctypes.data), "lsx_batch_stats")
        return it, be

    def factor_solve(self, Apacked, b, x, want_pivots=False):
        info = np.zeros(len(self.sizes), dtype=np.int64)
        ipiv = np.zeros(int(self.sizes.sum()), dtype=np.int64) if want_pivots else None
        ka, pa = _colmajor(Apacked)
        kb, pb = _colmajor(b)
        nfail = self.ctx.check(_abi.lib().lsx_batch_factor_solve(
            self.h, pa, pb, _ptr_out(x), info.ctypes.data, ipiv.ctypes.data if want_pivots else None),
            "lsx_batch_factor_solve")
        return nfail, info, ipiv


class MI355XBatchedLUFactorization(_DenseAlg):
    """Drop-in for the blockwise LU of BlockDiagonalFactorization
    (ext/LinearSolveBlockDiagonalsExt.jl:48-185): every block factored and solved in one
    persistent launch.  success = all blocks nonsingular (:101-114).  precision="fp32" factors in
    FP32 and refines every block's solution in FP64 to backward error <= tol; success then also
    needs every block to reach tol."""

    def __init__(self, precision="fp64", tol=1e-14, max_iters=10, tiles="auto"):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.dtype = LSX_F64 if precision == "fp64" else LSX_F32
        self.tol, self.max_iters, self.tiles = tol, max_iters, tiles
        self.last_info = None
        self.last_ipiv = None
        self.last_iters = None
        self.last_backward_error = None

    def solve_(self, cache):
        A = cache.A
        if not isinstance(A, BlockDiagonal):
            raise ValueError("MI355XBatchedLUFactorization takes a BlockDiagonal A")
        if cache.isfresh or cache.cacheval is None:
            cache.cacheval = _BatchHandle(cache.ctx, A.sizes, self.dtype, self.tiles)
            if self.dtype == LSX_F32:
                cache.cacheval.set_refinement(self.tol, self.max_iters)
            object.__setattr__(cache, "isfresh", False)
        nfail, info, ipiv = cache.cacheval.factor_solve(A.packed(), cache.b, cache.u, want_pivots=True)
        self.last_info, self.last_ipiv = info, ipiv
        rc = ReturnCode.Success if nfail == 0 else ReturnCode.Failure
        if self.dtype == LSX_F32:
            self.last_iters, self.last_backward_error = cache.cacheval.stats()
            if nfail == 0 and not np.all(self.last_backward_error <= self.tol):
                rc = ReturnCode.Failure
        return LinearSolution(cache.u, None, rc, 0, cache)


# -------------------------------------------------------------------------------------------------
# SimpleGMRES

class CSRMatrix:
    """Sparse matrix in CSR form (int32 indices, FP64 values)."""

    def __init__(self, rowptr, colind, vals, shape):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.colind = np.ascontiguousarray(colind, dtype=np.int32)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.shape = tuple(shape)

    @classmethod
    def from_scipy(cls, S):
        S = S.tocsr()
        S.sort_indices()
        return cls(S.indptr, S.indices, S.data, S.shape)

    @property
    def nnz(self):
        return int(self.vals.shape[0])

    def diagonal(self):
        d = np.zeros(self.shape[0])
        for i in range(self.shape[0]):
            lo, hi = self.rowptr[i], self.rowptr[i + 1]
            m = self.colind[lo:hi] == i
            if m.any():
                d[i] = self.vals[lo:hi][m][0]
        return d


class _CSRHandle:
    def __init__(self, ctx, A):
        self.ctx = ctx
        self.n = A.shape[0]
        h = ctypes.c_void_p()
        ctx.check(_abi.lib().lsx_csr_create(ctx.handle, A.shape[0], A.nnz, A.rowptr.ctypes.data,
                                            A.colind.ctypes.data, A.vals.ctypes.data, ctypes.byref(h)),
                  "lsx_csr_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _abi._lib is not None:
            _abi.lib().lsx_csr_destroy(h)
            self.h = None

    def spmv(self, x):
        kx, px = _colmajor(x)
        y = _empty_like_rhs(x)
        self.ctx.check(_abi.lib().lsx_csr_spmv(self.h, px, _ptr_out(y)), "lsx_csr_spmv")
        return y


def _diag_of(P, n):
    """Diagonal preconditioner as a vector (Pl = Diagonal(d)); None for the identity."""
    if P is None:
        return None
    if isinstance(P, (int, float)) and P == 1:
        return None
    d = np.ascontiguousarray(P, dtype=np.float64).ravel()
    if d.shape[0] != n:
        raise ValueError("diagonal preconditioner length mismatch")
    return d


class MI355XSimpleGMRES:
    """Drop-in for SimpleGMRES (src/simplegmres.jl:43-66) on a CSR matrix: restarted GMRES with
    MGS Arnoldi and symmetric Givens rotations; Krylov workspace resident on the device."""

    needs_concrete_A = True
    needs_square_A = True

    def __init__(self, restart=True, memory=20, blocksize=0, warm_start=False):
        if blocksize < 0:
            raise ValueError("blocksize must be >= 0")
        if warm_start:
            raise ValueError("warm_start is not supported (it does not seed from u0 in the reference)")
        self.restart = restart
        self.memory = memory
        self.blocksize = blocksize  # > 0: block-diagonal A, one GMRES per block (the batched layout)

    def init_cacheval(self, cache):
        return None

    def solve_(self, cache):
        A = cache.A
        if not isinstance(A, CSRMatrix):
            A = CSRMatrix.from_scipy(A)
            object.__setattr__(cache, "_A", A)
        n = A.shape[0]
        if cache.isfresh or cache.cacheval is None:
            cache.cacheval = _CSRHandle(cache.ctx, A)
            object.__setattr__(cache, "isfresh", False)
        pl = _diag_of(cache.Pl, n)
        pr = _diag_of(cache.Pr, n)
        kb, pb = _colmajor(cache.b)
        st = _abi.GmresStats()
        if self.blocksize > 0:
            if n % self.blocksize:
                raise ValueError("the blocksize must divide the size of the matrix")  # simplegmres.jl:422
            cache.ctx.check(_abi.lib().lsx_gmres_solve_batched(
                cache.cacheval.h, self.blocksize, pb, _ptr_out(cache.u), self.memory, int(bool(self.restart)),
                cache.abstol, cache.reltol, cache.maxiters, pl.ctypes.data if pl is not None else None,
                pr.ctypes.data if pr is not None else None, ctypes.byref(st)), "lsx_gmres_solve_batched")
        else:
            memory, restart, kst = _device_krylov_dims(cache, self.memory, bool(self.restart))
            cache.ctx.check(_abi.lib().lsx_gmres_solve(
                cache.cacheval.h, pb, _ptr_out(cache.u), memory, int(restart), cache.abstol,
                cache.reltol, cache.maxiters, pl.ctypes.data if pl is not None else None,
                pr.ctypes.data if pr is not None else None, ctypes.byref(st)), "lsx_gmres_solve")
            return LinearSolution(cache.u, st.resid, ReturnCode(st.retcode), int(st.iters), cache, kst)
        return LinearSolution(cache.u, st.resid, ReturnCode(st.retcode), int(st.iters), cache)


def _device_krylov_dims(cache, memory, restart):
    """(memory, restart, stats) for the device solve.  Without restart the reference grows the
    Krylov basis up to maxiters vectors (src/simplegmres.jl:222-235; Krylov.jl's gmres likewise);
    the device holds at most lsx_gmres_max_memory of them (2048: the basis in HBM, its
    Hessenberg/Givens bookkeeping in LDS or, beyond what the LDS holds, in HBM).  A solve whose
    maxiters exceeds that restarts every that many iterations instead: it says so with a
    RuntimeWarning and in the solution's stats ({"krylov_restart": that length})."""
    if restart:
        return memory, True, None
    lim = ctypes.c_int32()
    cache.ctx.check(_abi.lib().lsx_gmres_max_memory(cache.cacheval.h, ctypes.byref(lim)), "lsx_gmres_max_memory")
    if max(memory, cache.maxiters) <= lim.value:
        return memory, False, None
    warnings.warn(f"GMRES without restart: the device holds at most {lim.value} Krylov vectors, so this solve "
                  f"(maxiters = {cache.maxiters}) restarts every {lim.value} iterations", RuntimeWarning, stacklevel=3)
    return int(lim.value), True, {"krylov_restart": int(lim.value)}


class MI355XKrylovJL_GMRES:
    """Drop-in for KrylovJL_GMRES (src/iterative_wrappers.jl:200-228, the KrylovJL struct
    :142-162, its solve! :625-768) on a CSR matrix, on the device GMRES.

    KrylovJL_GMRES calls Krylov.gmres! from the Krylov.jl package, which the reference does not
    vendor.  Its published algorithm is the one SimpleGMRES restates (src/simplegmres.jl:186-192:
    "the Arnoldi loop, the Givens bookkeeping, the stopping criteria and the restart logic"):
    MGS Arnoldi on M A N, Givens QR of the Hessenberg matrix, stop when the residual estimate
    reaches atol + rtol ||M r0|| (or breakdown / itmax), which the device kernel computes.
    Options, as the reference maps them (:211-221, :680-693):
      gmres_restart = 0  -> memory = min(20, n), no restart (the basis grows to maxiters)
      gmres_restart > 0  -> memory = gmres_restart, restart every gmres_restart iterations
      memory             -> overrides the size derived from gmres_restart
      precs(A, p)        -> (Pl, Pr), diagonal preconditioners (vectors, or None for identity),
                            applied with ldiv (Pl \\ x) or, with ldiv=False, by multiplication
    The solution: retcode Success when solved, MaxIters when the iterations ran out, Failure
    otherwise (an inconsistent system); iters = niter; resid = the last residual norm (:730-766).
    warm_start=Auto starts from zero, as in standalone LinearSolve (:669-673)."""

    needs_concrete_A = True
    needs_square_A = True

    def __init__(self, gmres_restart=0, memory=None, precs=None, ldiv=True, warm_start="Auto", window=0):
        if gmres_restart < 0:
            raise ValueError("gmres_restart must be >= 0")
        if warm_start not in ("Auto", None):
            raise ValueError("only WarmStart.Auto (a cold start) is supported")
        self.gmres_restart = int(gmres_restart)
        self.memory = memory
        self.precs = precs
        self.ldiv = bool(ldiv)
        self.window = window  # accepted and unused for GMRES (:222-223)

    def init_cacheval(self, cache):
        return None

    def _preconditioners(self, cache, n):
        if self.precs is None:
            return _diag_of(cache.Pl, n), _diag_of(cache.Pr, n)
        Pl, Pr = self.precs(cache.A, None)
        pl, pr = _diag_of(Pl, n), _diag_of(Pr, n)
        if not self.ldiv:  # M * x with M = Diagonal(m) is Diagonal(1 ./ m) \ x
            pl = None if pl is None else np.ascontiguousarray(1.0 / pl)
            pr = None if pr is None else np.ascontiguousarray(1.0 / pr)
        return pl, pr

    def solve_(self, cache):
        A = cache.A
        if not isinstance(A, CSRMatrix):
            A = CSRMatrix.from_scipy(A)
            object.__setattr__(cache, "_A", A)
        n = A.shape[0]
        if cache.isfresh or cache.cacheval is None:
            cache.cacheval = _CSRHandle(cache.ctx, A)
            object.__setattr__(cache, "isfresh", False)
        pl, pr = self._preconditioners(cache, n)
        kb, pb = _colmajor(cache.b)
        if self.memory is not None:
            memory = int(self.memory)
        else:
            memory = self.gmres_restart if self.gmres_restart > 0 else min(20, n)
        memory, restart, kst = _device_krylov_dims(cache, memory, self.gmres_restart > 0)
        st = _abi.GmresStats()
        cache.ctx.check(_abi.lib().lsx_gmres_solve(
            cache.cacheval.h, pb, _ptr_out(cache.u), memory, int(restart), cache.abstol, cache.reltol,
            cache.maxiters, pl.ctypes.data if pl is not None else None, pr.ctypes.data if pr is not None else None,
            ctypes.byref(st)), "lsx_gmres_solve")
        rc = ReturnCode(st.retcode)
        if rc == ReturnCode.Infeasible:
            rc = ReturnCode.Failure
        return LinearSolution(cache.u, st.resid, rc, int(st.iters), cache, kst)


# -------------------------------------------------------------------------------------------------
# serving

class Server:
    """Continuous-batching request loop (lsx_serve_*)."""

    def __init__(self, ctx=None, hbm_budget=64 << 30, nstreams=4):
        self.ctx = ctx or Context.get(0)
        h = ctypes.c_void_p()
        self.ctx.check(_abi.lib().lsx_serve_create(self.ctx.handle, hbm_budget, nstreams, ctypes.byref(h)),
                       "lsx_serve_create")
        self.h = h
        self._keep = {}

    def close(self):
        if self.h is not None and self.h.value:
            _abi.lib().lsx_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_dense(self, A, b):
        Af = np.asfortranarray(A, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(bb)
        req = _abi.Request()
        req.kind = 0
        req.dtype = LSX_F64
        req.n = Af.shape[0]
        req.A = Af.ctypes.data
        req.b = bb.ctypes.data
        t = ctypes.c_uint64()
        self.ctx.check(_abi.lib().lsx_serve_submit(self.h, ctypes.byref(req), x.ctypes.data, ctypes.byref(t)),
                       "lsx_serve_submit")
        self._keep[t.value] = (Af, bb, x)
        return t.value

    def submit_gmres(self, A, b, pl=None, memory=30, restart=True, abstol=None, reltol=1e-8, maxiters=None):
        A = A if isinstance(A, CSRMatrix) else CSRMatrix.from_scipy(A)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(bb)
        pld = None if pl is None else np.ascontiguousarray(pl, dtype=np.float64)
        req = _abi.Request()
        req.kind = 1
        req.dtype = LSX_F64
        req.n = A.shape[0]
        req.b = bb.ctypes.data
        req.nnz = A.nnz
        req.rowptr = A.rowptr.ctypes.data
        req.colind = A.colind.ctypes.data
        req.vals = A.vals.ctypes.data
        req.pl_diag = pld.ctypes.data if pld is not None else None
        req.memory = memory
        req.restart = int(bool(restart))
        req.abstol = default_tol() if abstol is None else abstol
        req.reltol = reltol
        req.maxiters = A.shape[0] if maxiters is None else maxiters
        t = ctypes.c_uint64()
        self.ctx.check(_abi.lib().lsx_serve_submit(self.h, ctypes.byref(req), x.ctypes.data, ctypes.byref(t)),
                       "lsx_serve_submit")
        self._keep[t.value] = (A, bb, x, pld)
        return t.value

    def submit_dense_synth(self, n, seed):
        """Dense request whose A = I - g J and b are generated on the device from seed."""
        x = np.empty(int(n))
        req = _abi.Request()
        req.kind = 0
        req.dtype = LSX_F64
        req.n = int(n)
        req.synth = 1
        req.seed = int(seed)
        t = ctypes.c_uint64()
        self.ctx.check(_abi.lib().lsx_serve_submit(self.h, ctypes.byref(req), x.ctypes.data, ctypes.byref(t)),
                       "lsx_serve_submit")
        self._keep[t.value] = (None, None, x)
        return t.value

    def submit_gmres_synth(self, m, pe, seed, memory=30, restart=True, abstol=None, reltol=1e-8, maxiters=None):
        """SimpleGMRES request on the m x m convection-diffusion grid generated on the device, with
        the Jacobi preconditioner."""
        N = int(m) * int(m)
        x = np.empty(N)
        req = _abi.Request()
        req.kind = 1
        req.dtype = LSX_F64
        req.n = N
        req.synth = 1
        req.grid_m = int(m)
        req.pe = float(pe)
        req.seed = int(seed)
        req.memory = memory
        req.restart = int(bool(restart))
        req.abstol = default_tol() if abstol is None else abstol
        req.reltol = reltol
        req.maxiters = N if maxiters is None else int(maxiters)
        t = ctypes.c_uint64()
        self.ctx.check(_abi.lib().lsx_serve_submit(self.h, ctypes.byref(req), x.ctypes.data, ctypes.byref(t)),
                       "lsx_serve_submit")
        self._keep[t.value] = (None, None, x)
        return t.value

    def stats(self):
        out = np.zeros(6, dtype=np.int64)
        self.ctx.check(_abi.lib().lsx_serve_stats(self.h, out.ctypes.data), "lsx_serve_stats")
        keys = ("submitted", "completed", "batched_launches", "batched_requests", "peak_bytes", "waiting")
        return dict(zip(keys, (int(v) for v in out)))

    def poll(self, ticket):
        r = _abi.Result()
        rc = _abi.lib().lsx_serve_poll(self.h, ticket, ctypes.byref(r))
        if rc == _abi.LSX_EAGAIN:
            return None
        self.ctx.check(rc, "lsx_serve_poll")
        keep = self._keep.pop(ticket)
        x = keep[2]
        # a request that ended in an error (status < 0: a device-side wait gave up, a malformed
        # CSR, ...) has no valid result: its retcode is Failure whatever the solver reported
        retcode = ReturnCode(r.retcode) if r.status == _abi.LSX_OK else ReturnCode.Failure
        return dict(x=x, status=r.status, retcode=retcode, info=r.info, iters=r.iters,
                    resid=r.resid, t_submit=r.t_submit, t_start=r.t_start, t_done=r.t_done)

    def drain(self):
        self.ctx.check(_abi.lib().lsx_serve_drain(self.h), "lsx_serve_drain")

    def clock(self):
        return _abi.lib().lsx_serve_clock(self.h)
