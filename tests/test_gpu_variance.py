"""GPU parity of the variance and profile-likelihood consumers against the oracle:
the Fisher information kernel (ccd_fisher_information), the batched profile searches
(ccd_profile) and per-fit fixed columns of batched fits."""
import json
import os

import numpy as np
import pytest

import _cases
import oracle
from cyclops_amd import data as D
from cyclops_amd.control import create_control, create_prior

pytestmark = pytest.mark.gpu

MODELS = ["lr", "pr", "clr", "cox"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CHI95 = 3.841458820694124 / 2.0


@pytest.fixture(scope="module")
def E():
    from cyclops_amd import engine
    assert engine.device_count() > 0, "no HIP device"
    return engine


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _beta(d, seed, scale=0.3):
    b = np.random.default_rng(seed).normal(0, scale, d.J)
    if d.has_offset:
        b[0] = 1.0
    return b


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("model", MODELS)
def test_fisher_information_matches_oracle(E, model, precision):
    d = _cases.ALL[model]()
    print(f"[case] fisher {model} fp{precision}", flush=True)
    b = _beta(d, 11)
    idx = list(range(1 if d.has_offset else 0, d.J))
    Fo = oracle.fisher_information(d, b, idx)
    eng = E.Engine(d, precision)
    try:
        Fg = eng.fisher_information(b, idx)
    finally:
        eng.close()
    assert np.array_equal(Fg, Fg.T)
    assert _rel(Fg, Fo) <= (1e-11 if precision == 64 else 1e-4), (Fg, Fo)


@pytest.mark.parametrize("cfg", [0, 3])
def test_fisher_information_row_chunks_at_scale(E, cfg):
    """Many columns (the dense block is filled a row chunk at a time, groups straddling
    chunks): the leading block equals a one-chunk call, which equals the oracle."""
    d, _ = D.make_config(cfg, scale=1.0 if cfg == 0 else 0.02)
    print(f"[case] fisher chunks cfg{cfg} K={d.K} J={d.J}", flush=True)
    rng = np.random.default_rng(3)
    b = _beta(d, 12, 0.05)
    small = sorted(rng.choice(d.J, 12, replace=False).tolist())
    rest = [j for j in range(d.J) if j not in small]
    big = small + rng.choice(rest, min(1188, len(rest)), replace=False).tolist()   # row chunks of ~30k
    eng = E.Engine(d, 64)
    try:
        Fs = eng.fisher_information(b, small)
        Fb = eng.fisher_information(b, big)
    finally:
        eng.close()
    assert _rel(Fb[:12, :12], Fs) <= 1e-11
    Fo = oracle.fisher_information(d, b, small)
    assert _rel(Fs, Fo) <= 1e-11


@pytest.mark.parametrize("model", MODELS)
def test_batch_fixed_index_matches_oracle_fixed_fit(E, model):
    """A column held at its starting value in some fits of a batch, free in others."""
    d = _cases.ALL[model]()
    print(f"[case] batch fixed_index {model}", flush=True)
    ptype = "normal"
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    col = d.J - 1
    starts = [None, 0.4, -0.7, None, 1.1]
    n = len(starts)
    betas0 = np.zeros((n, d.J))
    if d.has_offset:
        betas0[:, 0] = 1.0
    fix = np.full(n, -1, dtype=np.int32)
    for f, v in enumerate(starts):
        if v is not None:
            betas0[f, col] = v
            fix[f] = col
    variances = [0.5] * n
    eng = E.Engine(d, 64)
    try:
        betas, res = eng.fit_batch(variances, ptype, ctl, betas0=betas0, fixed_index=fix)
    finally:
        eng.close()
    for f, v in enumerate(starts):
        mask = None
        if v is not None:
            mask = np.zeros(d.J, dtype=np.uint8)
            mask[col] = 1
        bo, ro = oracle.fit(d, create_prior(ptype, 0.5), ctl, fixed=mask, beta0=betas0[f])
        assert res[f]["iterations"] == ro["iterations"], (f, res[f], ro)
        assert _rel(betas[f], bo) <= 1e-6, (f, betas[f], bo)
        if v is not None:
            assert betas[f, col] == v


def _dobson():
    with open(os.path.join(GOLDEN, "dobson_glm_published.json")) as fh:
        g = json.load(fh)
    counts = np.array(g["counts"], dtype=float)
    outcome = np.tile([0, 1, 2], 3)
    treat = np.repeat([0, 1, 2], 3)
    X = np.zeros((9, 4))
    X[outcome == 1, 0] = 1
    X[outcome == 2, 1] = 1
    X[treat == 1, 2] = 1
    X[treat == 2, 3] = 1
    return D.from_dense(X, counts, "pr", add_intercept=True), g


def test_confint_and_ses_dobson(E):
    d, g = _dobson()
    print("[case] profile dobson", flush=True)
    prior = create_prior("none")
    ctl = create_control(tolerance=1e-8)
    eng = E.Engine(d, 64)
    try:
        b, _ = eng.fit(prior, ctl, mode="resident")
        se = E.get_ses(eng, b)
        ci = E.confint(eng, prior, ctl, b)
        lo, up, ev = eng.profile(prior, ctl, b, range(5), CHI95)
    finally:
        eng.close()
    assert np.allclose(se, g["std_error"], atol=1e-4)
    pub = np.array(g["confint_95"])
    assert np.allclose(ci[:, 1:], pub, atol=1e-3), ci
    lo_o, up_o, ev_o = oracle.profile(d, prior, ctl, b, range(5), threshold=CHI95)
    assert np.allclose(lo, lo_o, atol=1e-7) and np.allclose(up, up_o, atol=1e-7), (np.c_[lo, up], np.c_[lo_o, up_o])
    assert np.array_equal(ev, ev_o)


@pytest.mark.parametrize("model", MODELS)
def test_profile_matches_oracle(E, model):
    d = _cases.ALL[model]()
    print(f"[case] profile {model}", flush=True)
    prior = create_prior("none")
    ctl = create_control(convergence_type="mittal", tolerance=1e-10, max_iterations=5000)
    cols = [d.J - 1, d.J - 2]
    bo, _ = oracle.fit(d, prior, ctl)
    eng = E.Engine(d, 64)
    try:
        lo, up, ev = eng.profile(prior, ctl, bo, cols, CHI95)
    finally:
        eng.close()
    lo_o, up_o, ev_o = oracle.profile(d, prior, ctl, bo, cols, threshold=CHI95)
    assert np.all(np.isfinite(lo)) and np.all(np.isfinite(up))
    assert np.all(lo < bo[cols]) and np.all(bo[cols] < up)
    assert np.allclose(lo, lo_o, atol=1e-6) and np.allclose(up, up_o, atol=1e-6), (np.c_[lo, up], np.c_[lo_o, up_o])
    assert np.array_equal(ev, ev_o)


def test_profile_refuses_regularized_column(E):
    d = _cases.logistic()
    prior = create_prior("laplace", 0.5)
    ctl = create_control(tolerance=1e-8)
    eng = E.Engine(d, 64)
    try:
        b, _ = eng.fit(prior, ctl, mode="resident")
        with pytest.raises(E.CyclopsError, match="regularized"):
            eng.profile(prior, ctl, b, [2], CHI95)
        lo, up, _ = eng.profile(prior, ctl, b, [2], CHI95, override_no_regularization=True)
        assert lo[0] < b[2] < up[0] or (lo[0] <= b[2] <= up[0])
        # the intercept carries no prior and needs no override
        lo, up, _ = eng.profile(prior, ctl, b, [0], CHI95)
        assert lo[0] < b[0] < up[0]
    finally:
        eng.close()


@pytest.mark.parametrize("model", MODELS)
def test_profile_likelihood_grid_matches_oracle(E, model):
    d = _cases.ALL[model]()
    print(f"[case] profile likelihood grid {model}", flush=True)
    prior = create_prior("normal", 2.0)
    ctl = create_control(convergence_type="mittal", tolerance=1e-10, max_iterations=5000)
    col = d.J - 1
    bo, _ = oracle.fit(d, prior, ctl)
    pts = bo[col] + np.linspace(-1.0, 1.0, 9)
    eng = E.Engine(d, 64)
    try:
        vg = eng.profile_likelihood(prior, ctl, bo, col, pts)
    finally:
        eng.close()
    vo = oracle.profile_likelihood(d, prior, ctl, bo, col, pts)
    assert np.allclose(vg, vo, rtol=1e-9, atol=1e-9 * np.max(np.abs(vo))), (vg, vo)
    assert np.argmax(vg) == 4   # the mode sits at the centre of the grid


def test_aconfint_and_predict_dobson(E):
    d, g = _dobson()
    prior = create_prior("none")
    ctl = create_control(tolerance=1e-8)
    eng = E.Engine(d, 64)
    try:
        b, _ = eng.fit(prior, ctl, mode="resident")
        ci = E.aconfint(eng, b, level=0.95)
        mu = E.predict(eng, b)
    finally:
        eng.close()
    se = np.array(g["std_error"])
    assert np.allclose(ci[:, 1], b - 1.959963984540054 * se, atol=2e-4)
    assert np.allclose(ci[:, 2], b + 1.959963984540054 * se, atol=2e-4)
    # fitted counts of the additive log-linear model: row total x column total / N
    outcome = np.tile([0, 1, 2], 3)
    assert np.allclose(mu, 50.0 * np.array([63.0, 40.0, 47.0])[outcome] / 150.0, rtol=1e-6)


def test_variance_entry_points_refuse_bad_arguments(E):
    d = _cases.poisson()   # column 0 is the log-time offset
    prior = create_prior("none")
    ctl = create_control(tolerance=1e-8)
    eng = E.Engine(d, 64)
    try:
        b, _ = eng.fit(prior, ctl, mode="resident")
        with pytest.raises(E.CyclopsError, match="out of range"):
            eng.fisher_information(b, [0, d.J])
        assert eng.fisher_information(b, []).shape == (0, 0)
        with pytest.raises(E.CyclopsError, match="fixed"):
            eng.profile(prior, ctl, b, [0], CHI95)          # the offset is fixed at 1
        with pytest.raises(E.CyclopsError, match="not found"):
            eng.profile(prior, ctl, b, [d.J], CHI95)
        with pytest.raises(E.CyclopsError, match="threshold"):
            eng.profile(prior, ctl, b, [1], 0.0)
        with pytest.raises(E.CyclopsError, match="fixed"):
            eng.profile_likelihood(prior, ctl, b, 0, [0.5])
        # a failed call leaves the engine usable
        lo, up, _ = eng.profile(prior, ctl, b, [1], CHI95)
        assert lo[0] < b[1] < up[0]
    finally:
        eng.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_warm_started_batch_over_many_columns_matches_oracle(E, precision):
    """Starting values on 100 sparse columns: xb of every fit is formed from the
    row-major copy of the columns in one pass."""
    d, _ = D.make_config(0, scale=0.1)
    print(f"[case] warm batch K={d.K} J={d.J} fp{precision}", flush=True)
    rng = np.random.default_rng(9)
    b0 = rng.normal(0, 0.2, (3, d.J))
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    eng = E.Engine(d, precision)
    try:
        bg, rg = eng.fit_batch([0.5, 1.0, 2.0], "normal", ctl, betas0=b0)
    finally:
        eng.close()
    for i, v in enumerate([0.5, 1.0, 2.0]):
        bo, ro = oracle.fit(d, create_prior("normal", v), ctl, beta0=b0[i])
        if precision == 64:
            assert rg[i]["iterations"] == ro["iterations"]
            assert _rel(bg[i], bo) <= 1e-6, i
        else:
            assert _rel(bg[i], bo) <= 1e-4, i
