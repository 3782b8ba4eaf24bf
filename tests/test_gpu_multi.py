"""GPU tests of fits of several datasets in one persistent launch (ccd_fit_multi):
every fit equals the same engine's own resident fit exactly, and the oracle."""
import numpy as np
import pytest

import _cases
import oracle
from cyclops_amd import data as D
from cyclops_amd.control import create_control, create_prior

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from cyclops_amd import engine
    assert engine.device_count() > 0, "no HIP device"
    return engine


def _datasets(model):
    if model == "lr":
        return [_cases.logistic(K=300, seed=21), _cases.logistic(K=4000, seed=22), D.make_config(0, scale=0.5)[0]]
    if model == "pr":
        return [_cases.poisson(K=500, seed=23), _cases.poisson(K=3000, seed=24), D.make_config(1, scale=0.02)[0]]
    if model == "clr":
        return [_cases.clr(n_strata=60, seed=25), _cases.clr(n_strata=400, seed=26), D.make_config(3, scale=0.01)[0]]
    return [_cases.cox(K=250, seed=27), _cases.cox(K=1500, seed=28), D.make_config(2, scale=0.005)[0]]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("model", ["lr", "pr", "clr", "cox"])
def test_fit_multi_equals_each_resident_fit(E, model, precision):
    ds = _datasets(model)
    print(f"[case] fit_multi {model} fp{precision} K={[d.K for d in ds]}", flush=True)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    priors = [create_prior("laplace", 0.5), create_prior("normal", 1.0), create_prior("laplace", 2.0)]
    rng = np.random.default_rng(5)
    weights = [None, (rng.random(ds[1].K) < 0.8).astype(float), None]
    engs = [E.Engine(d, precision) for d in ds]
    try:
        alone = [e.fit(p, ctl, weights=w, mode="resident") for e, p, w in zip(engs, priors, weights)]
        together = E.fit_multi(engs, priors, [ctl] * 3, weights=weights)
    finally:
        for e in engs:
            e.close()
    for i, ((b1, r1), (b2, r2)) in enumerate(zip(alone, together)):
        assert np.array_equal(b1, b2), i
        assert r1["iterations"] == r2["iterations"] and r1["log_likelihood"] == r2["log_likelihood"], i
        assert r1["return_flag"] == r2["return_flag"] == 0
    if precision == 64:
        for d, p, w, (b, r) in zip(ds[:2], priors[:2], weights[:2], together[:2]):
            bo, ro = oracle.fit(d, p, ctl, weights=w)
            assert r["iterations"] == ro["iterations"]
            assert np.max(np.abs(b - bo)) <= 1e-6 * max(np.max(np.abs(bo)), 1e-12)


def test_fit_multi_refuses_mixed_models(E):
    a = E.Engine(_cases.logistic(K=200, seed=1), 64)
    b = E.Engine(_cases.poisson(K=200, seed=2), 64)
    try:
        with pytest.raises(E.CyclopsError, match="one model"):
            E.fit_multi([a, b], [create_prior("none")] * 2, [create_control()] * 2)
        with pytest.raises(E.CyclopsError, match="distinct"):
            E.fit_multi([a, a], [create_prior("none")] * 2, [create_control()] * 2)
    finally:
        a.close()
        b.close()
