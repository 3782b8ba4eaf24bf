"""CPU checks of the C ABI library: it loads and exports every function that
include/cyclops_amd.h declares (no compute calls without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cyclops_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(ccd_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["ccd_create", "ccd_grad_hess", "ccd_update_xbeta", "ccd_fit", "ccd_fit_batch",
              "ccd_cross_validate", "ccd_server_start", "ccd_last_error"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from cyclops_amd import engine
    L = C.CDLL(engine.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_wrapper_binds():
    from cyclops_amd import engine
    L = engine.lib()
    assert L.ccd_version().decode().startswith("cyclops_amd")


def test_create_without_device_fails_cleanly():
    """With no visible device, ccd_create reports an error instead of crashing."""
    from cyclops_amd import engine
    try:
        n = engine.device_count()
    except engine.CyclopsError:
        n = 0
    if n > 0:
        pytest.skip("a device is present")
    import _cases
    d = _cases.logistic()
    with pytest.raises(engine.CyclopsError):
        engine.Engine(d)


def test_flat_prior_refuses_cross_validation():
    import pytest
    from cyclops_amd.control import create_prior
    with pytest.raises(ValueError, match="flat prior"):
        create_prior("none", use_cross_validation=True)


def test_from_matrices_assembles_like_the_reference():
    """Offset, dense, sparse (dgCMatrix slots), indicator: the columns of one dataset built
    from matrices equal the same columns built from a dense array."""
    import numpy as np
    import scipy.sparse as sp
    from cyclops_amd import data as D
    rng = np.random.default_rng(3)
    K = 50
    Xd = rng.normal(size=(K, 2))
    Xs = np.where(rng.random((K, 3)) < 0.3, rng.integers(1, 5, (K, 3)), 0).astype(float)
    Xi = (rng.random((K, 2)) < 0.3).astype(float)
    y = rng.poisson(1.0, K).astype(float)
    off = np.log(rng.uniform(1, 5, K))
    s = sp.csc_matrix(Xs)
    a = D.from_matrices(y, "pr", dense=Xd, sparse=(s.indices, s.indptr, s.data), indicator=sp.csc_matrix(Xi),
                        offset=off)
    b = D.from_dense(np.column_stack([Xd, Xs, Xi]), y, "pr",
                     formats=["dense"] * 2 + ["sparse"] * 3 + ["indicator"] * 2, offset=off)
    assert np.array_equal(a.formats, b.formats) and np.array_equal(a.col_ptr, b.col_ptr)
    assert np.array_equal(a.rows, b.rows) and np.allclose(a.values, b.values)
    assert np.array_equal(a.dense_values, b.dense_values) and a.has_offset and b.has_offset
    assert np.array_equal(a.to_dense(), b.to_dense())
