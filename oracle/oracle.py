"""ctypes wrapper of the CPU oracle (lsx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the CPU-baseline leg of
bench.py, always as the checker.  The product (the HIP library and its host layer) never imports
this module.

Citations for each restated function are in lsx_oracle.c's header.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liblsx_oracle.so")
_lib = None

RET_DEFAULT, RET_SUCCESS, RET_MAXITERS, RET_INFEASIBLE = 0, 1, 2, 3


def build():
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        i64, p = ctypes.c_int64, ctypes.c_void_p
        L.ora_dgetrf.restype = i64
        L.ora_dgetrf.argtypes = [i64, i64, p, i64, p, i64]
        L.ora_dgetrs.restype = None
        L.ora_dgetrs.argtypes = [i64, i64, p, i64, p, p, i64]
        L.ora_gmres.restype = ctypes.c_int
        L.ora_gmres.argtypes = [i64, p, p, p, p, p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, i64, p, p, p, p, p]
        L.ora_sym_givens.restype = None
        L.ora_sym_givens.argtypes = [ctypes.c_double, ctypes.c_double, p, p, p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def dgetrf(A, nb=0):
    """LU with partial pivoting of a copy of A.  Returns (factors, ipiv 1-based int64, info)."""
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = F.shape
    ipiv = np.zeros(min(m, n), dtype=np.int64)
    info = lib().ora_dgetrf(m, n, _ptr(F), max(m, 1), _ptr(ipiv), nb)
    return F, ipiv, int(info)


def dgetrs(F, ipiv, B):
    """Solve with the factors of dgetrf; B is a vector or an n x nrhs matrix."""
    vec = B.ndim == 1
    X = np.array(B.reshape(B.shape[0], -1), dtype=np.float64, order="F", copy=True)
    n = F.shape[0]
    Ff = np.asfortranarray(F)
    lib().ora_dgetrs(n, X.shape[1], _ptr(Ff), n, _ptr(np.ascontiguousarray(ipiv, dtype=np.int64)),
                     _ptr(X), n)
    return X[:, 0] if vec else X


def sym_givens(a, b):
    c, s, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().ora_sym_givens(a, b, ctypes.byref(c), ctypes.byref(s), ctypes.byref(r))
    return c.value, s.value, r.value


def gmres(rowptr, colind, vals, b, memory=20, restart=True, abstol=None, reltol=None,
          maxiters=None, pl=None, pr=None, trace=False):
    """SimpleGMRES on a CSR matrix.  Returns dict(u, retcode, iters, resid[, trace])."""
    N = b.shape[0]
    eps = np.sqrt(np.finfo(np.float64).eps)
    abstol = eps if abstol is None else abstol
    reltol = eps if reltol is None else reltol
    maxiters = N if maxiters is None else maxiters
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(colind, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(N)
    it = ctypes.c_int64()
    rn = ctypes.c_double()
    tr = np.zeros(max(maxiters, 1)) if trace else None
    plc = None if pl is None else np.ascontiguousarray(pl, dtype=np.float64)
    prc = None if pr is None else np.ascontiguousarray(pr, dtype=np.float64)
    ret = lib().ora_gmres(N, _ptr(rp), _ptr(ci), _ptr(v), _ptr(bb), _ptr(x), int(memory),
                          int(bool(restart)), abstol, reltol, int(maxiters), _ptr(plc), _ptr(prc),
                          ctypes.byref(it), ctypes.byref(rn), _ptr(tr))
    out = dict(u=x, retcode=int(ret), iters=int(it.value), resid=float(rn.value))
    if trace:
        out["trace"] = tr[: it.value]
    return out
