"""Writes the committed LAPACK fixtures (tests/golden/lapack_lu.npz).

The reference's LUFactorization delegates getrf/getrs to LAPACK (OpenBLAS_jll, compat 0.3;
src/factorization.jl:633,697 and src/openblas.jl:144-151,267-275).  The same routines are run
here through scipy's bundled OpenBLAS (library run, not read): for each case the fixture holds
A, the factors, the 1-based pivots, info, and X = A \\ B for a seeded B.

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
import scipy.linalg.lapack as lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    g = np.random.default_rng(np.random.SeedSequence(20261202))
    out = {}
    for n in (1, 5, 17, 64, 129):
        out[f"gauss{n}"] = g.standard_normal((n, n))
    out["uniform100"] = g.random((100, 100))
    # decisive margins: rows scaled by distinct powers of two, shuffled
    n = 64
    A = np.diag(g.permutation(2.0 ** np.arange(1, n + 1))) @ (np.eye(n) + 0.01 * g.standard_normal((n, n)))
    out["margins64"] = A[g.permutation(n), :]
    # singular: zero column 4
    A = g.standard_normal((50, 50))
    A[:, 3] = 0.0
    out["singular50"] = A
    return out


def main():
    data = {"scipy_version": np.array(scipy.__version__)}
    g = np.random.default_rng(7)
    for name, A in cases().items():
        n = A.shape[0]
        lu, piv, info = lapack.dgetrf(A)
        B = g.standard_normal((n, 3))
        data[f"{name}_A"] = A
        data[f"{name}_lu"] = lu
        data[f"{name}_ipiv"] = (piv + 1).astype(np.int64)
        data[f"{name}_info"] = np.array(info)
        data[f"{name}_B"] = B
        if info == 0:
            data[f"{name}_X"] = lapack.dgetrs(lu, piv, B)[0]
    np.savez_compressed(os.path.join(HERE, "lapack_lu.npz"), **data)
    print("wrote", os.path.join(HERE, "lapack_lu.npz"))


if __name__ == "__main__":
    main()
