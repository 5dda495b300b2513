"""CPU checks of the C-ABI boundary: the library loads without a GPU and exports every function
include/lsx.h declares, and the Python binding declares a signature for each of them."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, load_package

HEADER = os.path.join(ROOT, "include", "lsx.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|double|const char\*)\s+(lsx_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_api():
    names = declared()
    for must in ("lsx_lu_factor", "lsx_lu_solve", "lsx_lu_refine", "lsx_batch_factor_solve", "lsx_gmres_solve",
                 "lsx_serve_submit", "lsx_serve_poll", "lsx_csr_create"):
        assert must in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lsx = load_package()
    from linearsolve_jl_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("liblsx.so not built")
    L = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _abi.SIGNATURES}
    assert set(declared()) <= bound, set(declared()) - bound


def test_loader_refuses_missing_library(tmp_path, monkeypatch):
    load_package()
    from linearsolve_jl_amd import _abi
    monkeypatch.setattr(_abi, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_abi, "_lib", None)
    with pytest.raises(ImportError):
        _abi.lib()


def test_host_api_semantics_without_gpu():
    lsx = load_package()
    import numpy as np
    # cache setters: A marks fresh, b does not (src/common.jl:313-360)
    c = lsx.LinearCache(np.eye(2), np.ones(2), np.zeros(2), None, None, None, 1e-8, 1e-8, 2, False, None)
    c.isfresh = False
    c.b = np.zeros(2)
    assert not c.isfresh
    c.A = np.eye(2)
    assert c.isfresh
    assert lsx.default_tol() == np.sqrt(np.finfo(float).eps)
    # BlockDiagonal packing round trip
    rng = np.random.default_rng(0)
    blocks = [rng.standard_normal((s, s)) for s in (2, 3, 1)]
    D = lsx.BlockDiagonal(blocks).to_dense()
    assert D.shape == (6, 6) and np.allclose(D[2:5, 2:5], blocks[1])
    # CSR from scipy keeps sorted int32 indices
    import scipy.sparse as sp
    S = sp.random(20, 20, density=0.2, format="csr", random_state=1) + sp.identity(20)
    A = lsx.CSRMatrix.from_scipy(S)
    assert A.rowptr.dtype == np.int32 and A.colind.dtype == np.int32
    np.testing.assert_allclose(A.diagonal(), S.diagonal())


def test_convection_diffusion_generator():
    load_package()
    from linearsolve_jl_amd import problems as P
    import numpy as np
    import scipy.sparse as sp
    m = 7
    rp, ci, v = P.convection_diffusion(m, 0.0)
    S = sp.csr_matrix((v, ci, rp), shape=(m * m, m * m))
    # Pe = 0: the scaled 5-point Laplacian (symmetric, diagonal 4, row sums 0 inside)
    assert abs(S - S.T).max() == 0
    assert np.all(S.diagonal() == 4.0)
    assert S.nnz == 5 * m * m - 4 * m
    rp, ci, v = P.convection_diffusion(m, 8.0)
    S = sp.csr_matrix((v, ci, rp), shape=(m * m, m * m)).toarray()
    h = 1.0 / (m + 1)
    assert np.isclose(S[10, 11], -1 + 8.0 * h / 2) and np.isclose(S[11, 10], -1 - 8.0 * h / 2)


def test_default_algorithm_routes():
    # solve(prob) without an algorithm: dense square -> LU, CSR -> KrylovJL_GMRES (the reference's
    # choice for square operators, src/default.jl:320-343), BlockDiagonal -> batched LU; the QR / SVD choices of the reference (non-square, VeryIll / SuperIllConditioned)
    # are not built and raise.  Host logic only (no device calls).
    import numpy as np
    import pytest
    from conftest import load_package
    lsx = load_package()
    A = np.eye(4)
    b = np.ones(4)
    assert isinstance(lsx.default_algorithm(A, b), lsx.MI355XLUFactorization)
    assert isinstance(lsx.default_algorithm(A, b, lsx.OperatorAssumptions(condition="WellConditioned")),
                      lsx.MI355XLUFactorization)
    S = lsx.CSRMatrix(np.arange(5, dtype=np.int32), np.arange(4, dtype=np.int32), np.ones(4), (4, 4))
    assert isinstance(lsx.default_algorithm(S, b), lsx.MI355XKrylovJL_GMRES)
    assert isinstance(lsx.default_algorithm(lsx.BlockDiagonal([np.eye(2), np.eye(2)]), b),
                      lsx.MI355XBatchedLUFactorization)
    for cond in ("VeryIllConditioned", "SuperIllConditioned"):
        with pytest.raises(NotImplementedError):
            lsx.default_algorithm(A, b, lsx.OperatorAssumptions(condition=cond))
    with pytest.raises(NotImplementedError):
        lsx.default_algorithm(np.ones((4, 3)), b)
    with pytest.raises(ValueError):
        lsx.OperatorAssumptions(condition="Unknown")


def test_tile_tables_per_factor_type(tmp_path):
    # tile_table.json holds one size-binned panel-width table per factor type; a file without an
    # FP32 table serves FP32 batches from the FP64 one, and no file means the library's default
    import json
    lsx = load_package()
    p = tmp_path / "tiles.json"
    p.write_text(json.dumps({"bins": [[64, 8], [512, 32]], "bins_fp32": [[128, 16], [512, 32]]}))
    assert lsx.load_tile_table(str(p), dtype=lsx.LSX_F64) == [(64, 8), (512, 32)]
    assert lsx.load_tile_table(str(p), dtype=lsx.LSX_F32) == [(128, 16), (512, 32)]
    p.write_text(json.dumps({"bins": [[512, 16]]}))
    assert lsx.load_tile_table(str(p), dtype=lsx.LSX_F32) == [(512, 16)]
    assert lsx.load_tile_table(str(tmp_path / "none.json")) is None
    # the committed table covers every size a batch accepts (n <= 512) with widths the library takes
    for dt in (lsx.LSX_F64, lsx.LSX_F32):
        t = lsx.load_tile_table(dtype=dt)
        if t is not None:
            assert max(u for u, _ in t) >= 512 and all(w in (8, 16, 32) for _, w in t)
