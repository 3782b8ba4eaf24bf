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


def test_precision_takes_the_reference_enum_values():
    """PrecisionType (src/cyclops/Types.h:155-160): FP64 = 0 and FP32 = 1 pass through
    unchanged, the bit counts 64 / 32 are aliases of them, and FP16 (2) is refused.
    ccd_estimate_footprint reads the precision without touching a device."""
    import ctypes as Cc
    import _cases
    from cyclops_amd import _abi, engine
    assert (_abi.FP64, _abi.FP32) == (0, 1)
    d = _cases.logistic()
    L = engine.lib()

    def footprint(p):
        desc = d.desc(p)
        out = Cc.c_double()
        rc = L.ccd_estimate_footprint(Cc.byref(desc), Cc.byref(out))
        return rc, out.value

    assert footprint(0)[0] == 0 and footprint(1)[0] == 0
    assert footprint(0) == footprint(64)
    assert footprint(1) == footprint(32)
    assert footprint(0)[1] > footprint(1)[1]   # FP64 state takes more room
    rc, _ = footprint(2)
    assert rc == _abi.ERR_INVALID
    assert "precision" in L.ccd_last_error().decode()
    assert _abi.precision_of(64) == _abi.FP64 and _abi.precision_of(_abi.FP32) == _abi.FP32
    with pytest.raises(ValueError):
        _abi.precision_of(2)


def test_stepped_batch_refuses_bad_handles_without_a_device():
    """The stepped-batch entry points check their handles before touching a device."""
    import ctypes as Cc
    from cyclops_amd import _abi, engine
    L = engine.lib()
    running = Cc.c_int32()
    assert L.ccd_batch_step(None, 1, None, Cc.byref(running)) == _abi.ERR_INVALID
    assert L.ccd_batch_end(None, None, None) == _abi.ERR_INVALID
