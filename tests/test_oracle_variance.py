"""Pins of the oracle's Fisher information and profile-likelihood limits (CPU only).

Fisher information (ModelSpecifics::computeFisherInformationImpl) is checked against
finite-difference Hessians of log-likelihoods written here from the model
definitions: logistic and Poisson (it is the observed information there) and
conditional logistic with one case per stratum (where the reference's formula is the
observed information). Breslow Cox is checked through the identity the reference's
formula implies: per-pid denominators make it the conditional-logistic information
with the tie groups as strata. Profile limits (CcdInterface::profileModel) are checked
against roots found by scipy's brentq on a profile likelihood maximised here by
scipy, and against R's published profile intervals and standard errors of the
Dobson Poisson example.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.optimize import brentq, minimize

import _cases
import oracle
from cyclops_amd import data as D
from cyclops_amd.control import create_control, create_prior

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TIGHT = create_control(convergence_type="mittal", tolerance=1e-13, max_iterations=100000)
CHI95 = 3.841458820694124 / 2.0   # qchisq(0.95, 1) / 2


def ll_logistic(X, y, b):
    eta = X @ b
    return float(np.sum(y * eta - np.logaddexp(0.0, eta)))


def ll_poisson(X, y, b):
    eta = X @ b
    return float(np.sum(y * eta - np.exp(eta)))


def ll_clr(X, y, strata, b):
    eta = X @ b
    ll = 0.0
    for s in np.unique(strata):
        m = strata == s
        ll += np.sum(y[m] * eta[m]) - y[m].sum() * np.log(np.sum(np.exp(eta[m])))
    return float(ll)


def fd_hessian(f, b, idx, h=1e-4):
    """Central second differences of f over coordinates idx."""
    m = len(idx)
    H = np.zeros((m, m))
    for a in range(m):
        for c in range(a, m):
            def at(da, dc):
                x = b.copy()
                x[idx[a]] += da
                x[idx[c]] += dc
                return f(x)
            v = (at(h, h) - at(h, -h) - at(-h, h) + at(-h, -h)) / (4 * h * h)
            H[a, c] = H[c, a] = v
    return H


def _beta(d, seed, scale=0.3):
    b = np.random.default_rng(seed).normal(0, scale, d.J)
    if d.has_offset:
        b[0] = 1.0
    return b


@pytest.mark.parametrize("model", ["lr", "pr"])
def test_fisher_is_negative_hessian_independent_rows(model):
    d = _cases.ALL[model]()
    X = d.to_dense()
    b = _beta(d, 5)
    idx = list(range(1 if d.has_offset else 0, d.J))
    f = (lambda x: ll_logistic(X, d.y, x)) if model == "lr" else (lambda x: ll_poisson(X, d.y, x))
    F = oracle.fisher_information(d, b, idx)
    H = fd_hessian(f, b, idx)
    assert np.allclose(F, F.T)
    assert np.max(np.abs(F + H)) <= 1e-5 * np.max(np.abs(F)), (F, -H)


def _clr_one_case(n_strata=60, seed=8):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(2, 7, n_strata)
    K = int(sizes.sum())
    pid = np.repeat(np.arange(n_strata), sizes).astype(np.int32)
    X, fm = _cases._mixed_X(rng, K, density=0.3)
    y = np.zeros(K)
    start = 0
    for s in sizes:
        y[start + rng.integers(0, s)] = 1.0
        start += s
    return D.from_dense(X, y, "clr", formats=fm, pid=pid)


def test_fisher_is_negative_hessian_clr_one_case_per_stratum():
    d = _clr_one_case()
    X = d.to_dense()
    b = _beta(d, 6, 0.5)
    idx = list(range(d.J))
    F = oracle.fisher_information(d, b, idx)
    H = fd_hessian(lambda x: ll_clr(X, d.y, d.pid, x), b, idx)
    assert np.max(np.abs(F + H)) <= 1e-5 * np.max(np.abs(F)), (F, -H)
    # a formula without the strata cross terms is not the information
    w = np.exp(X @ b)
    den = np.bincount(d.pid, weights=w)
    no_cross = (X * (w / den[d.pid])[:, None]).T @ X
    assert np.max(np.abs(no_cross - F)) > 1e-2 * np.max(np.abs(F))


def _breslow_pids(d):
    """Tie groups of setPidForAccumulation with no weights: consecutive events of one
    stratum at the same time share a pid, every other row starts a new one."""
    pid = np.zeros(d.K, dtype=np.int32)
    cur = 0
    for k in range(1, d.K):
        same = (d.pid[k] == d.pid[k - 1] and d.y[k - 1] == 1.0 and d.time[k] == d.time[k - 1]
                and d.y[k] == d.y[k - 1])
        if not same:
            cur += 1
        pid[k] = cur
    return pid


def test_fisher_cox_is_clr_over_tie_groups():
    d = _cases.cox()
    b = _beta(d, 7, 0.4)
    idx = list(range(d.J))
    F = oracle.fisher_information(d, b, idx)
    pid = _breslow_pids(d)
    assert pid.max() + 1 < d.K   # the data has ties
    X = d.to_dense()
    c = D.from_dense(X, d.y, "clr", formats=None, pid=pid)
    # same columns in the same order (formats may differ; values do not)
    assert np.array_equal(c.to_dense(), X)
    G = oracle.fisher_information(c, b, idx)
    assert np.allclose(F, G, rtol=1e-12, atol=1e-12 * np.max(np.abs(F)))


def _bid_logistic():
    bid = np.array([1, 5, 10, 20, 30, 40, 50, 75, 100, 150, 200.0])
    n = np.array([31, 29, 27, 25, 23, 21, 19, 17, 15, 15, 15])
    yy = np.array([0, 3, 6, 7, 9, 13, 17, 12, 11, 14, 13])
    x = np.log(np.concatenate([np.repeat(bid, n - yy), np.repeat(bid, yy)]))
    y = np.concatenate([np.zeros((n - yy).sum()), np.ones(yy.sum())])
    return D.from_dense(x[:, None], y, "lr", formats=["dense"], add_intercept=True)


def test_profile_limits_are_roots_of_an_independent_profile():
    d = _bid_logistic()
    X = d.to_dense()
    prior = create_prior("none")
    b, r = oracle.fit(d, prior, TIGHT)
    mode = ll_logistic(X, d.y, b)
    assert r["log_likelihood"] == pytest.approx(mode, rel=1e-12)
    lo, up, ev = oracle.profile(d, prior, TIGHT, b, [0, 1], threshold=CHI95)

    def prof(col, x):
        other = 1 - col

        def neg(z):
            v = np.empty(2)
            v[col], v[other] = x, z[0]
            p = 1 / (1 + np.exp(-(X @ v)))
            return -ll_logistic(X, d.y, v), -np.array([X[:, other] @ (d.y - p)])
        z = minimize(neg, [b[other]], jac=True, method="BFGS", options={"gtol": 1e-10}).x
        return -neg(z)[0] - (mode - CHI95)

    for col in (0, 1):
        for lim in (lo[col], up[col]):
            root = brentq(lambda x: prof(col, x), lim - 0.05, lim + 0.05, xtol=1e-12)
            assert abs(lim - root) <= 1e-3, (col, lim, root)   # the zeroin tolerance
        assert lo[col] < b[col] < up[col]
        assert ev[col] >= 4


def test_dobson_poisson_matches_published_glm():
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
    d = D.from_dense(X, counts, "pr", add_intercept=True)
    prior = create_prior("none")
    ctl = create_control(tolerance=1e-8)
    b, _ = oracle.fit(d, prior, ctl)
    se = np.sqrt(np.diag(np.linalg.inv(oracle.fisher_information(d, b, range(5)))))
    assert np.allclose(se, g["std_error"], atol=1e-4)
    lo, up, _ = oracle.profile(d, prior, ctl, b, range(5), threshold=CHI95)
    ci = np.array(g["confint_95"])
    assert np.allclose(lo, ci[:, 0], atol=1e-3) and np.allclose(up, ci[:, 1], atol=1e-3), np.c_[lo, up]
    # without a prior the penalty changes nothing (tests/testthat/test-smallCLR.R:29-30)
    lo2, up2, _ = oracle.profile(d, prior, ctl, b, range(5), threshold=CHI95, include_penalty=False)
    assert np.array_equal(lo, lo2) and np.array_equal(up, up2)


def test_profile_likelihood_grid_matches_an_independent_profile():
    d = _bid_logistic()
    X = d.to_dense()
    prior = create_prior("none")
    b, r = oracle.fit(d, prior, TIGHT)
    pts = np.array([b[1] - 0.5, b[1] - 0.1, b[1], b[1] + 0.2, 0.0])
    vals = oracle.profile_likelihood(d, prior, TIGHT, b, 1, pts)

    def prof(x):   # intercept maximised with the slope held at x
        def neg(z):
            v = np.array([z[0], x])
            p = 1 / (1 + np.exp(-(X @ v)))
            return -ll_logistic(X, d.y, v), -np.array([X[:, 0] @ (d.y - p)])
        z = minimize(neg, [b[0]], jac=True, method="BFGS", options={"gtol": 1e-10}).x
        return -neg(z)[0]

    expect = np.array([prof(x) for x in pts])
    assert np.allclose(vals, expect, rtol=0, atol=1e-7), (vals, expect)
    assert vals[2] == pytest.approx(r["log_likelihood"], abs=1e-9)   # at the mode: the maximum
    assert np.all(vals <= r["log_likelihood"] + 1e-9)


def test_warm_start_cv_chain_reaches_the_cold_start_optima():
    """The chain only changes where each fit starts: at a tight tolerance every fold x
    grid fit reaches the same penalised optimum as from zero."""
    from cyclops_amd import _abi
    d = _cases.logistic()
    prior = create_prior("laplace", 1.0)
    out = {}
    for reset in (True, False):
        ctl = create_control(convergence_type="mittal", tolerance=1e-12, max_iterations=100000, fold=4,
                             grid_size=5, lower_limit=0.01, upper_limit=5.0, seed=123, reset_coefficients=reset)
        out[reset] = oracle.cross_validate(d, prior, ctl, _abi.SELECT_BY_ROW)
    assert np.allclose(out[True]["pred"], out[False]["pred"], rtol=0, atol=1e-7)
    assert out[True]["best_variance"] == out[False]["best_variance"]
    # and at the default tolerance the chain is not the cold start (it starts elsewhere)
    loose = {}
    for reset in (True, False):
        ctl = create_control(convergence_type="mittal", tolerance=1e-8, fold=4, grid_size=5, lower_limit=0.01,
                             upper_limit=5.0, seed=123, reset_coefficients=reset)
        loose[reset] = oracle.cross_validate(d, prior, ctl, _abi.SELECT_BY_ROW)
    assert not np.array_equal(loose[True]["pred"], loose[False]["pred"])
