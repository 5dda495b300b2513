import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_package():
    """Imports linearsolve.jl_amd/ (a directory name that is not a Python identifier)."""
    name = "linearsolve_jl_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(ROOT, "linearsolve.jl_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def lsx():
    return load_package()


def backward_error(A, x, b):
    """max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    A = np.asarray(A)
    x = np.asarray(x).reshape(A.shape[0], -1)
    b = np.asarray(b).reshape(A.shape[0], -1)
    r = b - A @ x
    nA = np.abs(A).sum(axis=1).max()
    return max(np.abs(r[:, j]).max() / (nA * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(x.shape[1]))
