"""ccd_grad_hess_noalloc (reduction buffers in the library's static device scratch)
against ccd_grad_hess (the engine's own buffers), through the C ABI.

Both run the same kernels with the same launch shape, so the results are equal to the
last bit (the bar for this pair is exact equality, in FP64 and FP32). Each case makes
two calls in a row on the no-allocation entry point, the second with a smaller input
(a column with fewer entries, so fewer per-block partials, or a smaller dataset): the
partials a larger call left in the shared array must not reach a later result.
"""
import ctypes as C

import numpy as np
import pytest

import _cases
from cyclops_amd import _abi
from cyclops_amd import data as D

pytestmark = pytest.mark.gpu

MODELS = ["lr", "pr", "clr", "cox"]


@pytest.fixture(scope="module")
def E():
    from cyclops_amd import engine
    assert engine.device_count() > 0, "no HIP device"
    return engine


def _prepare(E, d, precision, weighted, seed):
    eng = E.Engine(d, precision)
    rng = np.random.default_rng(seed)
    w = (rng.random(d.K) < 0.8).astype(float) if weighted else None
    if weighted and d.model_type == _abi.CONDITIONAL_LOGISTIC:
        keep = rng.random(int(d.pid.max()) + 1) < 0.8
        w = keep[d.pid].astype(float)
    beta = rng.normal(0, 0.2, d.J)
    if d.has_offset:
        beta[0] = 1.0
    eng.set_weights(w)
    eng.fixed_terms(weighted)
    eng.compute_xbeta(beta)
    eng.remaining_stats(weighted)
    return eng


def _both(E, eng, j, weighted):
    L = E.lib()
    g0, h0 = C.c_double(), C.c_double()
    g1, h1 = C.c_double(), C.c_double()
    eng.numerator(j, weighted)
    assert L.ccd_grad_hess(eng.h, j, C.byref(g0), C.byref(h0), int(weighted)) == 0, L.ccd_last_error()
    rc = L.ccd_grad_hess_noalloc(eng.h, j, C.byref(g1), C.byref(h1), int(weighted))
    assert rc == 0, L.ccd_last_error()
    return (g0.value, h0.value), (g1.value, h1.value)


def _columns_by_size(d):
    n = np.diff(d.col_ptr)
    sparse = [j for j in range(d.J) if int(d.formats[j]) in (_abi.INDICATOR, _abi.SPARSE) and n[j] > 0]
    sparse.sort(key=lambda j: -n[j])
    dense = [j for j in range(d.J) if int(d.formats[j]) in (_abi.DENSE, _abi.INTERCEPT)]
    return dense, sparse


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("model", MODELS)
def test_noalloc_equals_grad_hess(E, model, weighted, precision):
    d = _cases.ALL[model](K=3000, seed=3) if model in ("lr", "pr") else _cases.ALL[model]()
    print(f"[case] noalloc {model} weighted={weighted} precision={precision} K={d.K}", flush=True)
    eng = _prepare(E, d, precision, weighted, 11)
    try:
        dense, sparse = _columns_by_size(d)
        # a full-length column first, then the largest and the smallest sparse column
        order = dense[-1:] + sparse[:1] + sparse[-1:]
        for j in order:
            ref, got = _both(E, eng, j, weighted)
            assert got == ref, (model, j, ref, got)
    finally:
        eng.close()


def test_noalloc_large_then_small_dataset(E):
    """A wide column pass (many blocks) on a large dataset, then a small dataset: the
    second call reuses the array with fewer partials."""
    big, _ = D.make_config(1, scale=0.1)       # 100,000 rows; column 1 holds every row
    small = _cases.poisson(K=500, seed=4)
    print(f"[case] noalloc large then small: K={big.K} then K={small.K}", flush=True)
    for d in (big, small, big):
        eng = _prepare(E, d, _abi.FP64, False, 5)
        try:
            dense, sparse = _columns_by_size(d)
            for j in sparse[:1] + sparse[-1:]:
                ref, got = _both(E, eng, j, False)
                assert got == ref, (d.K, j, ref, got)
        finally:
            eng.close()


def test_noalloc_refuses_bad_arguments(E):
    d = _cases.logistic()
    eng = _prepare(E, d, _abi.FP64, False, 1)
    L = E.lib()
    g, h = C.c_double(), C.c_double()
    try:
        assert L.ccd_grad_hess_noalloc(eng.h, d.J, C.byref(g), C.byref(h), 0) == _abi.ERR_INVALID
        assert L.ccd_grad_hess_noalloc(eng.h, 0, None, C.byref(h), 0) == _abi.ERR_INVALID
    finally:
        eng.close()
