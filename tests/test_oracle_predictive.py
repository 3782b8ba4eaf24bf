"""Independent pins of the oracle's held-out predictive log-likelihood, gradient
objective and predictive estimates (CPU).

The oracle restates ModelSpecifics::getPredictiveLogLikelihood (ModelSpecifics.hpp:
652-690), getGradientObjective (:103-120) and getPredictiveEstimates (:692-707), with
the per-model terms of ModelSpecifics.h (CLR :867-873, LR :1002-1013, Breslow Cox
:1243-1256, Poisson :1583-1594). Here each is recomputed from a dense X and beta with
numpy, from the definitions:

  LR   sum_k w_k y_k (x_k b - log(1 + exp(x_k b)))          (events only)
  PR   sum_k w_k (y_k x_k b - exp(x_k b))
  CLR  sum_k w_k y_k (x_k b - log sum_{k' in stratum(k)} exp(x_k' b))
  Cox  sum_k w_k y_k (x_k b - log sum_{k' at risk, w_k' != 0} w_k' exp(x_k' b))
       (Breslow: the risk set of time t is every row of the stratum with time >= t,
       counted over the held-out rows only, O(K^2) brute force)

with the call sequence of the cross-validation driver (weights of the training rows,
xb at beta, remaining statistics, then the held-out weights). Beside the brute force:
at unit weights the predictive value equals the log-likelihood (less the Poisson
log y! terms for PR), and for models of independent rows it is additive over folds.
A mistaken variant (the full Bernoulli term for LR, risk sets over every row for Cox,
weighted CLR denominators, a missing exp in PR) fails these checks; the last test
shows it on variants planted in the numpy restatement.
"""
import numpy as np
import pytest

import _cases
import oracle


def _beta(d, seed):
    rng = np.random.default_rng(seed)
    b = rng.normal(0.0, 0.4, d.J)
    if d.has_offset:
        b[0] = 1.0
    return b


def _folds(d, seed, model):
    rng = np.random.default_rng(seed)
    if model == "clr":   # whole strata, as the BY_PID selector holds them out
        held = rng.random(int(d.pid.max()) + 1) < 0.35
        return (~held[d.pid]).astype(float)
    return (rng.random(d.K) >= 0.35).astype(float)


def _oracle_pred(d, beta, train_w, pred_w):
    e = oracle.OracleEngine(d)
    e.set_weights(train_w)
    e.compute_xbeta(beta)
    e.remaining_stats(True)
    return e.pred_loglik(pred_w), e


def _brute_pred(d, beta, w, model):
    X = d.to_dense()
    xb = X @ beta
    y = d.y
    if model == "lr":
        return float(np.sum(w * y * (xb - np.log1p(np.exp(xb)))))
    if model == "pr":
        return float(np.sum(w * (y * xb - np.exp(xb))))
    if model == "clr":
        out = 0.0
        for s in np.unique(d.pid):
            m = d.pid == s
            den = np.sum(np.exp(xb[m]))   # every row of the stratum
            out += float(np.sum(w[m] * y[m] * (xb[m] - np.log(den))))
        return out
    # Breslow Cox over the held-out rows only
    out = 0.0
    for k in range(d.K):
        if w[k] == 0.0 or y[k] == 0.0:
            continue
        risk = (d.pid == d.pid[k]) & (d.time >= d.time[k]) & (w != 0.0)
        out += w[k] * y[k] * (xb[k] - np.log(np.sum(w[risk] * np.exp(xb[risk]))))
    return out


MODELS = ["lr", "pr", "clr", "cox"]


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("model", MODELS)
def test_predictive_loglik_matches_brute_force(model, seed):
    d = _cases.ALL[model](seed=seed + 10)
    beta = _beta(d, seed)
    train = _folds(d, seed, model)
    held = 1.0 - train
    v, _ = _oracle_pred(d, beta, train, held)
    ref = _brute_pred(d, beta, held, model)
    assert v == pytest.approx(ref, rel=1e-12, abs=1e-10), (model, v, ref)


@pytest.mark.parametrize("model", MODELS)
def test_predictive_loglik_at_unit_weights_is_the_loglik(model):
    """With every row held out, the held-out value is the model's log-likelihood at
    beta (Poisson: less the fixed sum of -log y!)."""
    d = _cases.ALL[model](seed=21)
    beta = _beta(d, 5)
    ones = np.ones(d.K)
    v, e = _oracle_pred(d, beta, None, ones)
    e.set_weights(None)
    e.fixed_terms(False)   # the Poisson -log y! terms (computeFixedTermsInLogLikelihood)
    e.compute_xbeta(beta)
    e.remaining_stats(False)
    ll = e.loglik(False)
    if model == "pr":
        ll += float(sum(np.sum(np.log(np.arange(2, int(yk) + 1))) for yk in d.y))
    if model == "lr":   # events-only form: sum_y=1 log p = ll + sum_y=0 log(1 + exp(xb))
        xb = d.to_dense() @ beta
        ll += float(np.sum((d.y == 0) * np.log1p(np.exp(xb))))
    assert v == pytest.approx(ll, rel=1e-12), (model, v, ll)


@pytest.mark.parametrize("model", ["lr", "pr", "clr"])
def test_predictive_loglik_is_additive_over_folds(model):
    """Independent rows (and whole strata for CLR): the held-out values of the folds
    add up to the value with every row held out."""
    d = _cases.ALL[model](seed=33)
    beta = _beta(d, 6)
    a = _folds(d, 9, model)
    va, _ = _oracle_pred(d, beta, a, 1.0 - a)
    vb, _ = _oracle_pred(d, beta, 1.0 - a, a)
    vall, _ = _oracle_pred(d, beta, None, np.ones(d.K))
    assert va + vb == pytest.approx(vall, rel=1e-12)


def test_cox_held_out_risk_sets_are_not_additive():
    """Cox risk sets are rebuilt over the held-out rows: the fold values do not add up
    to the all-rows value (they would if the risk sets stayed over every row)."""
    d = _cases.cox(seed=34)
    beta = _beta(d, 7)
    a = _folds(d, 10, "cox")
    va, _ = _oracle_pred(d, beta, a, 1.0 - a)
    vb, _ = _oracle_pred(d, beta, 1.0 - a, a)
    vall, _ = _oracle_pred(d, beta, None, np.ones(d.K))
    assert abs(va + vb - vall) > 1e-3 * abs(vall)


@pytest.mark.parametrize("model", MODELS)
def test_gradient_objective_matches_numpy(model):
    """getGradientObjective: sum_k w_k y_k x_k b (unweighted without CV weights)."""
    d = _cases.ALL[model](seed=41)
    beta = _beta(d, 8)
    w = _folds(d, 11, model)
    xb = d.to_dense() @ beta
    e = oracle.OracleEngine(d)
    e.set_weights(w)
    e.compute_xbeta(beta)
    assert e.gradient_objective(True) == pytest.approx(float(np.sum(w * d.y * xb)), rel=1e-12)
    assert e.gradient_objective(False) == pytest.approx(float(np.sum(d.y * xb)), rel=1e-12)


@pytest.mark.parametrize("model", MODELS)
def test_predictive_estimates_match_numpy(model):
    """getPredictiveEstimates: LR the probability, PR the expected count, CLR / Cox the
    linear predictor; with weights only rows of non-zero weight are written."""
    d = _cases.ALL[model](seed=51)
    beta = _beta(d, 9)
    xb = d.to_dense() @ beta
    want = {"lr": 1.0 / (1.0 + np.exp(-xb)), "pr": np.exp(xb), "clr": xb, "cox": xb}[model]
    e = oracle.OracleEngine(d)
    e.compute_xbeta(beta)
    e.remaining_stats(False)
    assert np.allclose(e.pred_est(), want, rtol=1e-13, atol=1e-15)
    w = _folds(d, 12, model)
    got = np.full(d.K, -7.0)
    lib_out = np.ascontiguousarray(got)
    import ctypes as C
    oracle.lib().orc_pred_est(e.h, lib_out.ctypes.data_as(C.POINTER(C.c_double)),
                              np.ascontiguousarray(w).ctypes.data_as(C.POINTER(C.c_double)))
    assert np.allclose(lib_out[w != 0], want[w != 0], rtol=1e-13, atol=1e-15)
    assert np.all(lib_out[w == 0] == -7.0)


def test_planted_mistakes_fail_the_brute_force():
    """The checks above would catch the plausible mistakes: each variant below, planted
    in a copy of the numpy restatement, moves the value far beyond the tolerance."""
    d_lr, d_cox, d_clr = _cases.logistic(seed=61), _cases.cox(seed=62), _cases.clr(seed=63)
    # LR: the full Bernoulli term instead of the events-only form
    b = _beta(d_lr, 1)
    w = _folds(d_lr, 2, "lr")
    v, _ = _oracle_pred(d_lr, b, 1.0 - w, w)
    xb = d_lr.to_dense() @ b
    p = 1.0 / (1.0 + np.exp(-xb))
    full = float(np.sum(w * (d_lr.y * np.log(p) + (1 - d_lr.y) * np.log1p(-p))))
    assert abs(v - full) > 1e-3 * abs(v)
    # Cox: risk sets over every row instead of the held-out ones
    b = _beta(d_cox, 3)
    w = _folds(d_cox, 4, "cox")
    v, _ = _oracle_pred(d_cox, b, 1.0 - w, w)
    xb = d_cox.to_dense() @ b
    wrong = 0.0
    for k in range(d_cox.K):
        if w[k] and d_cox.y[k]:
            risk = (d_cox.pid == d_cox.pid[k]) & (d_cox.time >= d_cox.time[k])
            wrong += xb[k] - np.log(np.sum(np.exp(xb[risk])))
    assert abs(v - wrong) > 1e-3 * abs(v)
    # CLR: denominators over the held-out rows of a stratum only (strata split by row)
    b = _beta(d_clr, 5)
    rng = np.random.default_rng(6)
    w = (rng.random(d_clr.K) < 0.6).astype(float)
    v, _ = _oracle_pred(d_clr, b, 1.0 - w, w)
    xb = d_clr.to_dense() @ b
    wrong = 0.0
    for s in np.unique(d_clr.pid):
        m = d_clr.pid == s
        den = np.sum(w[m] * np.exp(xb[m]))
        if den > 0:
            wrong += float(np.sum(w[m] * d_clr.y[m] * (xb[m] - np.log(den))))
    assert abs(v - wrong) > 1e-3 * abs(v)
