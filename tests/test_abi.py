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
