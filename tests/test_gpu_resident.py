"""GPU parity of the device-resident fit (one persistent launch per findMode), the
batched lockstep fits and the cross-validation driver against the oracle."""
import numpy as np
import pytest

import _cases
import oracle
from cyclops_amd import _abi
from cyclops_amd.control import create_control, create_prior

pytestmark = pytest.mark.gpu

MODELS = ["lr", "pr", "clr", "cox"]


@pytest.fixture(scope="module")
def E():
    from cyclops_amd import engine
    assert engine.device_count() > 0, "no HIP device"
    return engine


def _prior(model):
    return {"lr": create_prior("laplace", 0.5), "pr": create_prior("normal", 1.0),
            "clr": create_prior("laplace", 1.0), "cox": create_prior("laplace", 0.5)}[model]


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-12)


@pytest.mark.parametrize("conv", ["mittal", "gradient", "lange"])
@pytest.mark.parametrize("model", MODELS)
def test_resident_fp64_matches_oracle(E, model, conv):
    d = _cases.ALL[model]()
    print(f"[case] resident fp64 {model} {conv}", flush=True)
    prior = _prior(model)
    ctl = create_control(convergence_type=conv, tolerance=1e-8)
    bo, ro = oracle.fit(d, prior, ctl)
    bg, rg = E.fit_cyclops_model(d, prior, ctl, precision=64, mode="resident")
    assert rg["return_flag"] == ro["return_flag"]
    assert rg["iterations"] == ro["iterations"]
    assert _rel(bg, bo) <= 1e-6, (bg, bo)
    assert abs(rg["log_likelihood"] - ro["log_likelihood"]) <= 1e-9 * abs(ro["log_likelihood"])


@pytest.mark.parametrize("model", MODELS)
def test_resident_fp32_within_tolerance(E, model):
    d = _cases.ALL[model]()
    print(f"[case] resident fp32 {model}", flush=True)
    prior = _prior(model)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    bo, ro = oracle.fit(d, prior, ctl)
    bg, rg = E.fit_cyclops_model(d, prior, ctl, precision=32, mode="resident")
    err = _rel(bg, bo)
    lle = abs(rg["log_likelihood"] - ro["log_likelihood"]) / abs(ro["log_likelihood"])
    print(f"resident fp32 {model}: coef rel err {err:.2e}, loglik rel err {lle:.2e}")
    assert err <= 1e-4
    assert lle <= 1e-7


@pytest.mark.parametrize("model", MODELS)
def test_resident_weighted_fixed(E, model):
    d = _cases.ALL[model]()
    print(f"[case] resident weighted {model}", flush=True)
    rng = np.random.default_rng(5)
    if model == "clr":
        keep = rng.random(d.pid.max() + 1) < 0.8
        w = keep[d.pid].astype(float)
    else:
        w = (rng.random(d.K) < 0.8).astype(float)
    fixed = np.zeros(d.J, dtype=np.uint8)
    fixed[-1] = 1
    prior = _prior(model)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    bo, ro = oracle.fit(d, prior, ctl, weights=w, fixed=fixed)
    bg, rg = E.fit_cyclops_model(d, prior, ctl, weights=w, fixed=fixed, precision=64, mode="resident")
    assert rg["iterations"] == ro["iterations"]
    assert _rel(bg, bo) <= 1e-6
    assert abs(rg["log_likelihood"] - ro["log_likelihood"]) <= 1e-9 * abs(ro["log_likelihood"])


def _fold_setup(d, model, n_folds=3):
    selector = _abi.SELECT_BY_PID if model == "clr" else _abi.SELECT_BY_ROW
    pid = d.pid if model == "clr" else None
    return oracle.assign_folds(d.K, pid, n_folds, selector, 7)


@pytest.mark.parametrize("n_var", [1, 2, 7, 25, 90])
@pytest.mark.parametrize("model", MODELS)
def test_batched_fits_match_independent_oracle_fits(E, model, n_var):
    """Fits of every (variance, held-out fold) pair in one launch equal the fits run alone."""
    d = _cases.ALL[model]()
    folds = 3
    rf = _fold_setup(d, model, folds)
    variances = np.exp(np.linspace(np.log(0.05), np.log(2.0), n_var))
    var_list, held = [], []
    for v in variances:
        for f in range(folds):
            var_list.append(v)
            held.append(f)
    if n_var == 1:   # a single unweighted fit through the batch entry point
        var_list, held = [variances[0]], [-1]
    n = len(var_list)
    if n > 256:
        var_list, held = var_list[:256], held[:256]
        n = 256
    print(f"[case] batched {model} n_fits={n}", flush=True)
    ptype = "normal" if model == "pr" else "laplace"
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    eng = E.Engine(d, 64)
    betas, res = eng.fit_batch(var_list, ptype, ctl, held_out_fold=held, row_fold=rf)
    eng.close()
    for i in range(0, n, max(1, n // 6)):
        w = None if held[i] < 0 else (rf != held[i]).astype(float)
        bo, ro = oracle.fit(d, create_prior(ptype, var_list[i]), ctl, weights=w)
        assert res[i]["iterations"] == ro["iterations"], (i, res[i], ro)
        assert _rel(betas[i], bo) <= 1e-6, (i, betas[i], bo)
        assert abs(res[i]["log_likelihood"] - ro["log_likelihood"]) <= 1e-9 * abs(ro["log_likelihood"])


@pytest.mark.parametrize("layout", ["1,1", "4,1", "8,4", "32,16"])
@pytest.mark.parametrize("model", MODELS)
def test_batch_layouts_agree(E, model, layout, monkeypatch):
    """Every team shape (fits per team, workgroups per team) gives the oracle's fits:
    one-workgroup teams of 1 or 4 fits, three co-resident 4-workgroup teams (the last
    one partly filled) and one 16-workgroup lockstep team."""
    d = _cases.ALL[model]()
    folds = 3
    rf = _fold_setup(d, model, folds)
    var_list, held = [], []
    for v in np.exp(np.linspace(np.log(0.05), np.log(2.0), 7)):
        for f in range(folds):
            var_list.append(v)
            held.append(f)
    print(f"[case] layout {layout} {model} n_fits={len(var_list)}", flush=True)
    ptype = "normal" if model == "pr" else "laplace"
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    monkeypatch.setenv("CCD_BATCH_LAYOUT", layout)
    eng = E.Engine(d, 64)
    betas, res = eng.fit_batch(var_list, ptype, ctl, held_out_fold=held, row_fold=rf)
    eng.close()
    for i in [0, 5, 11, 20]:
        w = (rf != held[i]).astype(float)
        bo, ro = oracle.fit(d, create_prior(ptype, var_list[i]), ctl, weights=w)
        assert res[i]["iterations"] == ro["iterations"], (i, res[i], ro)
        assert _rel(betas[i], bo) <= 1e-6, (i, betas[i], bo)
        assert abs(res[i]["log_likelihood"] - ro["log_likelihood"]) <= 1e-9 * abs(ro["log_likelihood"])


def test_co_resident_teams_are_exact_at_scale(E, monkeypatch):
    """Several multi-workgroup teams sharing the chip exchange partial sums while
    their workgroups drift apart in time; their fits must equal the fits of
    one-workgroup teams (no exchange at all) to the last bit (FP32 bulk, FP64 sums of
    FP32 terms are exact in any order here). Poisson workload at 0.2 scale: 200,000
    rows, 4,001 columns, 64 fits, 3 cycles."""
    from cyclops_amd import data as D
    d, _ = D.make_config(1, scale=0.2)
    rf = E.assign_folds(d.K, None, 4, _abi.SELECT_BY_ROW, 1001)
    var_list = [float(v) for v in np.exp(np.linspace(np.log(0.1), np.log(10.0), 16)) for _ in range(4)]
    held = [f for _ in range(16) for f in range(4)]
    ctl = create_control(convergence_type="mittal", tolerance=1e-6, max_iterations=3)
    eng = E.Engine(d, _abi.FP32)
    out = {}
    for layout in ["1,1", "16,64", "4,32", "1,16", "64,128"]:
        print(f"[case] co-resident teams layout {layout}", flush=True)
        monkeypatch.setenv("CCD_BATCH_LAYOUT", layout)
        out[layout], _ = eng.fit_batch(var_list, "normal", ctl, held_out_fold=held, row_fold=rf)
    eng.close()
    for layout, b in out.items():
        assert np.max(np.abs(b - out["1,1"])) <= 1e-12, layout


@pytest.mark.parametrize("workgroups", [1, 3, 8])
@pytest.mark.parametrize("model", ["lr", "pr"])
def test_single_fit_workgroup_count_matches_oracle(E, model, workgroups, monkeypatch):
    """A single fit of independent rows on one workgroup (no exchange, the default up to
    100k rows) and forced onto several (a grid exchange per coordinate): FP64
    coefficients within 1e-6 of the oracle and the same cycles either way."""
    d = _cases.ALL[model](K=3000, seed=31)
    print(f"[case] single fit {model} on {workgroups} workgroup(s)", flush=True)
    prior = _prior(model)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    monkeypatch.setenv("CCD_RESIDENT_WG", str(workgroups))
    eng = E.Engine(d, 64)
    try:
        b, r = eng.fit(prior, ctl, mode="resident")
    finally:
        eng.close()
    bo, ro = oracle.fit(d, prior, ctl)
    assert r["iterations"] == ro["iterations"], (r, ro)
    assert _rel(b, bo) <= 1e-6
    assert abs(r["log_likelihood"] - ro["log_likelihood"]) <= 1e-9 * abs(ro["log_likelihood"])


@pytest.mark.parametrize("model", [2, 3])
def test_grouped_multi_workgroup_matches_oracle(E, model):
    """Cox (config 2 generator) and CLR (config 3 generator) at sizes where the one-fit
    resident kernel spans several workgroups, so risk-set and stratum carries cross
    the grid exchange; FP64 coefficients within 1e-6 of the oracle, same cycles."""
    from cyclops_amd import data as D
    d, _ = D.make_config(model, scale=0.02 if model == 2 else 0.01)
    print(f"[case] grouped multi-workgroup model={d.model_type} K={d.K} J={d.J}", flush=True)
    prior = create_prior("laplace", 0.05 if model == 2 else 0.5)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8)
    eng = E.Engine(d, 64)
    b, r = eng.fit(prior, ctl, mode="resident")
    eng.close()
    bo, ro = oracle.fit(d, prior, ctl)
    assert r["iterations"] == ro["iterations"], (r, ro)
    assert _rel(b, bo) <= 1e-6
    assert abs(r["log_likelihood"] - ro["log_likelihood"]) <= 1e-9 * abs(ro["log_likelihood"])


@pytest.mark.parametrize("model,reset", [("lr", True), ("pr", True), ("clr", True), ("cox", True),
                                         ("lr", False), ("clr", False)])
def test_cross_validation_matches_oracle(E, model, reset):
    """Cold start (all fits batched, every held-out predictive log-likelihood formed on
    the device from the run's xb) for every model, and R's default warm-start chain
    (host driver over resident fits); each fold's held-out value, the grid means, the
    chosen variance and the final refit."""
    d = _cases.ALL[model]()
    print(f"[case] cross-validation {model} reset_coefficients={reset}", flush=True)
    prior = create_prior("normal" if model == "pr" else "laplace", 1.0)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8, fold=4, grid_size=5,
                         lower_limit=0.01, upper_limit=5.0, seed=123, reset_coefficients=reset)
    selector = _abi.SELECT_BY_PID if model == "clr" else _abi.SELECT_BY_ROW
    ro = oracle.cross_validate(d, prior, ctl, selector)
    eng = E.Engine(d, 64)
    rg = eng.cross_validate(prior, ctl, selector)
    eng.close()
    assert np.allclose(rg["pred"], ro["pred"], rtol=1e-8, atol=1e-10), (rg["pred"], ro["pred"])
    assert np.allclose(rg["grid"], ro["grid"], rtol=1e-8), (rg["grid"], ro["grid"])
    assert rg["best_variance"] == pytest.approx(ro["best_variance"], rel=1e-12)
    assert _rel(rg["beta"], ro["beta"]) <= 1e-6
    assert rg["n_fits"] == 20


@pytest.mark.parametrize("precision", [64, 32])
def test_cross_validation_on_the_clr_workload(E, precision):
    """The grid-search driver on the conditional-logistic workload generator (1% scale:
    5,000 strata, about 32,000 rows, 1,000 columns), 5 folds by stratum x 4 Laplace
    variances, cold start, against the oracle's driver: per-fold held-out values, grid
    means and the chosen variance (FP64 to 1e-8; FP32 bulk to 1e-6 relative, same
    choice), final refit to 1e-6 in FP64."""
    from cyclops_amd import data as D
    d, _ = D.make_config(3, scale=0.01)
    print(f"[case] cross-validation on the CLR workload K={d.K} J={d.J} fp{precision}", flush=True)
    prior = create_prior("laplace", 0.1)
    ctl = create_control(convergence_type="mittal", tolerance=1e-7, fold=5, grid_size=4,
                         lower_limit=1e-3, upper_limit=1.0, seed=77, reset_coefficients=True)
    ro = oracle.cross_validate(d, prior, ctl, _abi.SELECT_BY_PID)
    eng = E.Engine(d, precision)
    try:
        rg = eng._cross_validate(prior, ctl, _abi.SELECT_BY_PID)
    finally:
        eng.close()
    tol = 1e-8 if precision == 64 else 1e-6
    assert np.allclose(rg["pred"], ro["pred"], rtol=tol), (rg["pred"], ro["pred"])
    assert np.allclose(rg["grid"], ro["grid"], rtol=tol), (rg["grid"], ro["grid"])
    assert rg["best_variance"] == pytest.approx(ro["best_variance"], rel=1e-12)
    if precision == 64:
        assert _rel(rg["beta"], ro["beta"]) <= 1e-6


def test_fold_assignment_matches_oracle(E):
    for sel, pid in [(_abi.SELECT_BY_ROW, None), (_abi.SELECT_BY_PID, np.repeat(np.arange(50), 3))]:
        K = 150
        a = E.assign_folds(K, pid, 10, sel, 42)
        b = oracle.assign_folds(K, pid, 10, sel, 42)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("conv", ["gradient", "mittal"])
@pytest.mark.parametrize("mode", ["plain", "resident"])
def test_separable_logistic_follows_the_oracle(E, mode, conv):
    """A covariate that is non-zero only on successes (tests/testthat/test-largeBernoulli.R:24-46):
    without a prior its coefficient runs off along a flat direction of the likelihood. There
    the Hessian p(1 - p) of rows with |x beta| >> 1 cancels, so a one-ulp difference between
    the device's and the host's exp moves the steps; the stopping cycle then depends on the
    implementation (as it does between the reference's own builds). What must agree is the
    outcome: the return flag, the log-likelihood and the run-off coefficient."""
    from cyclops_amd import data as D
    rng = np.random.default_rng(666)
    K = 500
    X = (rng.random((K, 5)) < 0.3).astype(float)
    p = 1 / (1 + np.exp(-(X @ np.array([0.5, -0.3, 0.8, 0.2, -0.6]) - 0.5)))
    y = (rng.random(K) < p).astype(float)
    d = D.from_dense(np.column_stack([X, X[:, 2] * y]), y, "lr", add_intercept=True)
    print(f"[case] separable logistic {mode} {conv}", flush=True)
    ctl = create_control(convergence_type=conv)
    bo, ro = oracle.fit(d, create_prior("none"), ctl)
    bg, rg = E.fit_cyclops_model(d, create_prior("none"), ctl, precision=64, mode=mode)
    assert abs(bo[-1]) > 20 and abs(bg[-1]) > 20   # the separable coefficient has run off
    assert rg["return_flag"] == ro["return_flag"]
    assert abs(rg["log_likelihood"] - ro["log_likelihood"]) <= 1e-6 * abs(ro["log_likelihood"])
    if conv == "mittal":   # stops early enough that the steps still agree
        assert rg["iterations"] == ro["iterations"] and _rel(bg, bo) <= 1e-6, (bg, bo)


@pytest.mark.parametrize("model", MODELS)
def test_degenerate_inputs_match_oracle(E, model):
    """An empty column, a constant dense column and a tiny dataset (8 rows), through the
    plain path, the resident kernel and a batch."""
    from cyclops_amd import data as D
    rng = np.random.default_rng(77)
    for K in (8, 300):
        X = (rng.random((K, 4)) < 0.4).astype(float)
        X[:, 1] = 0.0                      # an empty indicator column
        Xd = np.column_stack([X, np.full(K, 0.5)])
        fm = ["indicator"] * 4 + ["dense"]
        if model == "lr":
            d = D.from_dense(Xd, (rng.random(K) < 0.4).astype(float), "lr", formats=fm, add_intercept=True)
        elif model == "pr":
            d = D.from_dense(Xd, rng.poisson(2.0, K).astype(float), "pr", formats=fm, add_intercept=True)
        elif model == "clr":
            pid = np.repeat(np.arange(K // 4), 4).astype(np.int32)
            y = np.zeros(K)
            y[::4] = 1.0
            d = D.from_dense(Xd, y, "clr", formats=fm, pid=pid)
        else:
            t = np.sort(rng.integers(1, 6, K))[::-1].astype(float)
            y = (rng.random(K) < 0.5).astype(float)
            order = np.lexsort((np.arange(K), y, -t))
            d = D.from_dense(Xd[order], y[order], "cox", formats=fm, time=t[order], pid=np.zeros(K, dtype=np.int32))
        print(f"[case] degenerate {model} K={K}", flush=True)
        prior = create_prior("normal", 1.0)
        ctl = create_control(convergence_type="mittal", tolerance=1e-8)
        bo, ro = oracle.fit(d, prior, ctl)
        for mode in ("plain", "resident"):
            bg, rg = E.fit_cyclops_model(d, prior, ctl, precision=64, mode=mode)
            assert rg["return_flag"] == ro["return_flag"] and rg["iterations"] == ro["iterations"], (mode, rg, ro)
            assert _rel(bg, bo) <= 1e-6, (mode, bg, bo)
        eng = E.Engine(d, 64)
        try:
            bb, rb = eng.fit_batch([1.0, 1.0, 1.0], "normal", ctl)
        finally:
            eng.close()
        for i in range(3):
            assert _rel(bb[i], bo) <= 1e-6 and rb[i]["iterations"] == ro["iterations"]


def test_fit_cyclops_model_with_cross_validation(E):
    """prior.use_cross_validation routes fitCyclopsModel through the CV driver."""
    d = _cases.logistic(K=400, seed=31)
    ctl = create_control(convergence_type="mittal", tolerance=1e-8, fold=4, grid_size=4, lower_limit=0.01,
                         upper_limit=2.0, seed=7, reset_coefficients=True)
    prior = create_prior("laplace", 1.0, use_cross_validation=True)
    b, r = E.fit_cyclops_model(d, prior, ctl)
    ro = oracle.cross_validate(d, prior, ctl, _abi.SELECT_BY_ROW)
    assert r["variance"] == pytest.approx(ro["best_variance"], rel=1e-12)
    assert _rel(b, ro["beta"]) <= 1e-6
    with pytest.raises(E.CyclopsError, match="Insufficient"):
        E.fit_cyclops_model(_cases.logistic(K=60, seed=3), prior, ctl)
