"""Pins of the oracle (CPU only): each check comes from outside the oracle's own
formulas -- closed forms, textbook solvers written independently here, brute-force
Breslow likelihoods, finite differences, optimality conditions, and the known
answer stored with the reference's 17-variable logistic fixture.

Fixtures under tests/golden/ are data files held by the reference
(inst/extdata/*.txt, standalone/example/cox-tiny-test.txt).
"""
import math
import os

import numpy as np
import pytest
from scipy.optimize import minimize

import _cases
import oracle
from cyclops_amd import _abi
from cyclops_amd import data as D
from cyclops_amd.control import create_control, create_prior

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TIGHT = create_control(convergence_type="mittal", tolerance=1e-13, max_iterations=100000)


# ---------------------------------------------------------------------------
# independent log-likelihoods (written from the model definitions)
# ---------------------------------------------------------------------------
def ll_logistic(X, y, b, w=None):
    w = np.ones(len(y)) if w is None else w
    eta = X @ b
    return float(np.sum(w * (y * eta - np.logaddexp(0.0, eta))))


def ll_poisson(X, y, b, offset=None, w=None):
    w = np.ones(len(y)) if w is None else w
    eta = X @ b + (0.0 if offset is None else offset)
    return float(np.sum(w * (y * eta - np.exp(eta) - np.array([math.lgamma(v + 1.0) for v in y]))))


def ll_clr(X, y, strata, b):
    eta = X @ b
    ll = 0.0
    for s in np.unique(strata):
        m = strata == s
        ll += np.sum(y[m] * eta[m]) - y[m].sum() * np.log(np.sum(np.exp(eta[m])))
    return float(ll)


def ll_breslow(X, time, y, strata, b):
    """Breslow partial likelihood from the risk-set definition (O(K^2))."""
    eta = X @ b
    ll = 0.0
    for s in np.unique(strata):
        m = strata == s
        t, d, e = time[m], y[m], eta[m]
        for tt in np.unique(t[d == 1]):
            events = (t == tt) & (d == 1)
            risk = t >= tt
            ll += e[events].sum() - events.sum() * np.log(np.exp(e[risk]).sum())
    return float(ll)


# ---------------------------------------------------------------------------
# closed form: Dobson Poisson (tests/testthat/test-smallPoisson.R:8-12)
# ---------------------------------------------------------------------------
def test_dobson_poisson_closed_form():
    counts = np.array([18, 17, 15, 20, 10, 20, 25, 13, 12.0])
    outcome = np.tile([0, 1, 2], 3)
    treat = np.repeat([0, 1, 2], 3)
    X = np.zeros((9, 4))
    X[outcome == 1, 0] = 1
    X[outcome == 2, 1] = 1
    X[treat == 1, 2] = 1
    X[treat == 2, 3] = 1
    for fm in (["dense"] * 4, None):
        d = D.from_dense(X, counts, "pr", formats=fm, add_intercept=True)
        b, r = oracle.fit(d, create_prior("none"), TIGHT)
        # additive log-linear model on a 3x3 table: fitted = row total * col total / N
        expect = np.array([math.log(21.0), math.log(40.0 / 3.0 / 21.0), math.log(47.0 / 3.0 / 21.0), 0.0, 0.0])
        assert np.allclose(b, expect, atol=1e-5), b
        fitted = 50.0 * np.array([63.0, 40.0, 47.0])[outcome] / 150.0
        ll = float(np.sum(counts * np.log(fitted) - fitted - [math.lgamma(c + 1) for c in counts]))
        assert r["log_likelihood"] == pytest.approx(ll, rel=1e-9)


# ---------------------------------------------------------------------------
# textbook IRLS: bid logistic (tests/testthat/test-smallBernoulli.R:10-17)
# ---------------------------------------------------------------------------
def _bid_data():
    bid = np.array([1, 5, 10, 20, 30, 40, 50, 75, 100, 150, 200.0])
    n = np.array([31, 29, 27, 25, 23, 21, 19, 17, 15, 15, 15])
    yy = np.array([0, 3, 6, 7, 9, 13, 17, 12, 11, 14, 13])
    log_bid = np.log(np.concatenate([np.repeat(bid, n - yy), np.repeat(bid, yy)]))
    y = np.concatenate([np.zeros((n - yy).sum()), np.ones(yy.sum())])
    return log_bid, y


def irls_logistic(X, y, iters=100):
    b = np.zeros(X.shape[1])
    for _ in range(iters):
        p = 1 / (1 + np.exp(-(X @ b)))
        W = p * (1 - p)
        b = b + np.linalg.solve(X.T @ (W[:, None] * X), X.T @ (y - p))
    return b


@pytest.mark.parametrize("fmt", ["dense", "sparse"])
def test_bid_logistic_matches_irls(fmt):
    log_bid, y = _bid_data()
    d = D.from_dense(log_bid[:, None], y, "lr", formats=[fmt], add_intercept=True)
    b, r = oracle.fit(d, create_prior("none"), TIGHT)
    X = np.column_stack([np.ones(len(y)), log_bid])
    bi = irls_logistic(X, y)
    # CCD stops on the relative log-likelihood change, so coefficients agree to
    # the curvature-limited precision; the log-likelihood is exact at its beta
    assert np.allclose(b, bi, rtol=2e-5), (b, bi)
    assert r["log_likelihood"] == pytest.approx(ll_logistic(X, y, b), rel=1e-12)
    assert r["log_likelihood"] == pytest.approx(ll_logistic(X, y, bi), rel=1e-9)


# ---------------------------------------------------------------------------
# reference-held known answer: SAS -2 log-likelihood (test-largeBernoulli.R:20)
# ---------------------------------------------------------------------------
def test_17var_logistic_sas_known_answer():
    d = D.read_generic_text(os.path.join(GOLDEN, "CCD_LOGISTIC_TEST_17var.txt"), "lr", first_field="row")
    assert d.K == 22296 and d.J == 18   # 17 covariates + intercept (test-largeBernoulli.R:14-16)
    b, r = oracle.fit(d, create_prior("none"), create_control(convergence_type="mittal", tolerance=1e-10))
    assert r["return_flag"] == _abi.SUCCESS
    assert r["log_likelihood"] * 2 == pytest.approx(-1578.046, rel=1e-4)   # the reference's own tolerance


# ---------------------------------------------------------------------------
# CLR: independent Newton on the reference's infert fixture
# ---------------------------------------------------------------------------
def newton_clr(X, y, strata, iters=60):
    b = np.zeros(X.shape[1])
    us = np.unique(strata)
    for _ in range(iters):
        g = np.zeros_like(b)
        H = np.zeros((len(b), len(b)))
        eta = X @ b
        for s in us:
            m = strata == s
            e = np.exp(eta[m])
            p = e / e.sum()
            nE = y[m].sum()
            xm = X[m]
            mu = p @ xm
            g += y[m] @ xm - nE * mu
            H -= nE * ((xm * p[:, None]).T @ xm - np.outer(mu, mu))
        b = b - np.linalg.solve(H, g)
    return b


def test_infert_clr_matches_newton():
    d = D.read_generic_text(os.path.join(GOLDEN, "infert_ccd.txt"), "clr", first_field="stratum")
    X = d.to_dense()
    b, r = oracle.fit(d, create_prior("none"), TIGHT)
    bn = newton_clr(X, d.y, d.pid)
    assert np.allclose(b, bn, rtol=2e-5), (b, bn)
    assert r["log_likelihood"] == pytest.approx(ll_clr(X, d.y, d.pid, b), rel=1e-12)
    # one case per matched set, so this is also the exact conditional likelihood of
    # survival::clogit (spontaneous, induced) used by tests/testthat/test-smallCLR.R
    assert b == pytest.approx([1.98587, 1.40901], rel=1e-4)


def test_breslow_clr_with_multiple_cases_matches_newton():
    d = D.read_csv_clr(os.path.join(GOLDEN, "test1-clr.txt"))
    X = d.to_dense()
    b, r = oracle.fit(d, create_prior("none"), TIGHT)
    bn = newton_clr(X, d.y, d.pid)
    assert np.allclose(b, bn, rtol=2e-5, atol=1e-6), (b, bn)


# ---------------------------------------------------------------------------
# Cox: brute-force Breslow on the reference's inline tables (test-smallCox.R)
# ---------------------------------------------------------------------------
COX_TABLES = {
    "no_ties": (":6-16", [(4, 1, 0, 0), (3.5, 1, 2, 0), (3, 0, 0, 1), (2.5, 1, 0, 1), (2, 1, 1, 1), (1.5, 0, 1, 0),
                          (1, 1, 1, 0)]),
    "time_ties": (":48-58", [(4, 1, 0, 0), (3, 1, 2, 0), (3, 0, 0, 1), (2, 1, 0, 1), (2, 0, 1, 1), (1, 0, 1, 0),
                             (1, 1, 1, 0)]),
    "failure_ties": (":72-80", [(4, 1, 0, 0), (3, 1, 2, 0), (3, 0, 0, 1), (2, 1, 0, 1), (2, 1, 1, 1), (2, 1, 0, 0)]),
    "failure_ties_after": (":94-103", [(4, 1, 0, 0), (3, 1, 2, 0), (3, 0, 0, 1), (2, 1, 0, 1), (2, 1, 1, 1),
                                       (1, 0, 1, 0), (1, 1, 1, 0)]),
}


def _cox_from_rows(rows, strata=None, fmt=("dense", "dense"), cols=(2, 3)):
    a = np.array(rows, dtype=float)
    time, y = a[:, 0], a[:, 1]
    X = a[:, list(cols)]
    s = np.zeros(len(a), dtype=int) if strata is None else np.asarray(strata)
    order = D.sort_cox(time, y, X, s)
    d = D.from_dense(X[order], y[order], "cox", formats=list(fmt)[:X.shape[1]], time=time[order],
                     pid=s[order].astype(np.int32))
    return d, X[order], time[order], y[order], s[order]


def _cox_mle(X, time, y, s):
    f = lambda b: -ll_breslow(X, time, y, s, b)
    res = minimize(f, np.zeros(X.shape[1]), method="BFGS", options={"gtol": 1e-11})
    return res.x


@pytest.mark.parametrize("name", list(COX_TABLES))
def test_cox_breslow_matches_brute_force(name):
    _, rows = COX_TABLES[name]
    d, X, time, y, s = _cox_from_rows(rows)
    b, r = oracle.fit(d, create_prior("none"), TIGHT)
    bm = _cox_mle(X, time, y, s)
    assert np.allclose(b, bm, atol=2e-5), (b, bm)
    assert r["log_likelihood"] == pytest.approx(ll_breslow(X, time, y, s, b), rel=1e-10)


def test_cox_strata_matches_brute_force():
    rows = COX_TABLES["failure_ties_after"][1]
    strata = [int(r[3]) for r in rows]   # strata(x2), test-smallCox.R:117-139
    d, X, time, y, s = _cox_from_rows(rows, strata=strata, fmt=("sparse",), cols=(2,))
    b, r = oracle.fit(d, create_prior("none"), TIGHT)
    bm = _cox_mle(X, time, y, s)
    assert np.allclose(b, bm, atol=2e-5), (b, bm)


def test_cox_tiny_fixture_file():
    d = D.read_generic_text(os.path.join(GOLDEN, "cox-tiny-test.txt"), "cox", first_field="time", dense=True)
    order = D.sort_cox(d.time, d.y, None)
    assert np.all(order == np.arange(d.K))   # the fixture is already in (time desc, y asc) order
    X = d.to_dense()
    b, _ = oracle.fit(d, create_prior("none"), TIGHT)
    bm = _cox_mle(X, d.time, d.y, np.zeros(d.K, dtype=int))
    assert np.allclose(b, bm, atol=2e-5)


def test_cox_weighted_duplicate_rows_equal_unweighted():
    """test-smallCox.R:142-186: duplicated rows with 0/1 weights give the unduplicated fit."""
    rows = COX_TABLES["failure_ties_after"][1]
    strata = [int(r[3]) for r in rows]
    d1, *_ = _cox_from_rows(rows, strata=strata, fmt=("dense",), cols=(2,))
    b1, _ = oracle.fit(d1, create_prior("none"), TIGHT)
    rows2 = [r for r in rows for _ in range(2)]
    strata2 = [s for s in strata for _ in range(2)]
    a = np.array(rows2, dtype=float)
    x = a[:, 2].copy()
    x[0], x[4] = 5.0, 6.0   # the weight-0 copies carry different covariates
    w = np.tile([0.0, 1.0], len(rows))
    s = np.array(strata2)
    order = np.lexsort((np.arange(len(a)), a[:, 1], -a[:, 0], s))
    d2 = D.from_dense(x[order][:, None], a[order, 1], "cox", formats=["dense"], time=a[order, 0],
                      pid=s[order].astype(np.int32))
    b2, _ = oracle.fit(d2, create_prior("none"), TIGHT, weights=w[order])
    assert np.allclose(b1, b2, atol=1e-7)


# ---------------------------------------------------------------------------
# finite differences: oracle gradient/Hessian vs an independent log-likelihood
# ---------------------------------------------------------------------------
def _independent_ll(d, beta, w=None):
    X = d.to_dense()
    if d.model_type == _abi.LOGISTIC:
        return ll_logistic(X, d.y, beta, w)
    if d.model_type == _abi.POISSON:
        return ll_poisson(X, d.y, beta, w=w)
    if d.model_type == _abi.CONDITIONAL_LOGISTIC:
        return ll_clr(X, d.y, d.pid, beta)
    keep = np.ones(d.K, bool) if w is None else w > 0
    return ll_breslow(X[keep], d.time[keep], d.y[keep], d.pid[keep], beta)


@pytest.mark.parametrize("model", ["lr", "pr", "clr", "cox"])
def test_gradient_hessian_finite_differences(model):
    d = _cases.ALL[model]()
    rng = np.random.default_rng(3)
    beta = rng.normal(0, 0.2, d.J)
    if d.has_offset:
        beta[0] = 1.0
    o = oracle.OracleEngine(d)
    o.set_weights(None)
    o.fixed_terms(False)
    o.compute_xbeta(beta)
    o.remaining_stats(False)
    ll0 = _independent_ll(d, beta)
    # the oracle's log-likelihood equals the independent one (Poisson adds the
    # log y! fixed term; LR/PR/CLR/Cox match exactly)
    assert o.loglik(False) == pytest.approx(ll0, rel=1e-10)
    eps = 1e-5
    for j in range(d.J):
        if d.has_offset and j == 0:
            continue
        o.numerator(j)
        g, h = o.grad_hess(j)
        e = np.zeros(d.J)
        e[j] = eps
        lp, lm = _independent_ll(d, beta + e), _independent_ll(d, beta - e)
        g_fd = -(lp - lm) / (2 * eps)
        h_fd = -(lp - 2 * ll0 + lm) / eps ** 2
        assert g == pytest.approx(g_fd, rel=1e-5, abs=1e-5), (model, j, g, g_fd)
        assert h == pytest.approx(h_fd, rel=2e-3, abs=1e-3), (model, j, h, h_fd)


@pytest.mark.parametrize("model", ["lr", "pr", "cox"])
def test_weighted_loglik_matches_independent(model):
    d = _cases.ALL[model]()
    rng = np.random.default_rng(8)
    w = (rng.random(d.K) < 0.7).astype(float)
    beta = rng.normal(0, 0.2, d.J)
    if d.has_offset:
        beta[0] = 1.0
    o = oracle.OracleEngine(d)
    o.set_weights(w)
    o.fixed_terms(True)
    o.compute_xbeta(beta)
    o.remaining_stats(True)
    assert o.loglik(True) == pytest.approx(_independent_ll(d, beta, w), rel=1e-10)


# ---------------------------------------------------------------------------
# optimality: Laplace KKT and Normal stationarity at the returned beta
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("model", ["lr", "pr", "clr", "cox"])
def test_laplace_kkt(model):
    d = _cases.ALL[model]()
    var = 0.3
    lam = math.sqrt(2.0 / var)
    b, r = oracle.fit(d, create_prior("laplace", var), TIGHT)
    o = oracle.OracleEngine(d)
    o.set_weights(None)
    o.fixed_terms(False)
    o.compute_xbeta(b)
    o.remaining_stats(False)
    from cyclops_amd.control import intercept_index
    ii = intercept_index(d)
    for j in range(d.J):
        if d.has_offset and j == 0:
            continue
        o.numerator(j)
        g, _ = o.grad_hess(j)
        if j == ii:
            assert abs(g) < 1e-5
        elif b[j] == 0.0:
            assert abs(g) <= lam * (1 + 1e-6)
        else:
            assert g == pytest.approx(-lam * np.sign(b[j]), rel=1e-5, abs=1e-5)


@pytest.mark.parametrize("model", ["lr", "pr", "clr", "cox"])
def test_normal_prior_stationarity(model):
    d = _cases.ALL[model]()
    var = 0.5
    b, r = oracle.fit(d, create_prior("normal", var), TIGHT)
    o = oracle.OracleEngine(d)
    o.set_weights(None)
    o.fixed_terms(False)
    o.compute_xbeta(b)
    o.remaining_stats(False)
    from cyclops_amd.control import intercept_index
    ii = intercept_index(d)
    for j in range(d.J):
        if d.has_offset and j == 0:
            continue
        o.numerator(j)
        g, _ = o.grad_hess(j)
        target = 0.0 if j == ii else -b[j] / var
        assert g == pytest.approx(target, abs=1e-5)


# ---------------------------------------------------------------------------
# CV fold assignment (CrossValidationSelector.cpp:19-189)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("selector", [_abi.SELECT_BY_ROW, _abi.SELECT_BY_PID])
def test_fold_sizes_and_partition(selector):
    K = 103
    pid = np.repeat(np.arange(41), 3)[:K] if selector == _abi.SELECT_BY_PID else None
    f = oracle.assign_folds(K, pid, 10, selector, 1234)
    N = K if pid is None else len(np.unique(pid))
    ids = np.arange(K) if pid is None else pid
    fold_of_id = {}
    for k in range(K):
        fold_of_id.setdefault(ids[k], f[k])
        assert fold_of_id[ids[k]] == f[k]      # a stratum never straddles folds
    sizes = np.bincount(list(fold_of_id.values()), minlength=10)
    expect = [N // 10 + (1 if i < N % 10 else 0) for i in range(10)]
    assert list(sizes) == expect              # interval sizes of :30-42


def test_fold_assignment_product_matches_oracle():
    from cyclops_amd import engine
    for sel, pid in [(_abi.SELECT_BY_ROW, None), (_abi.SELECT_BY_PID, np.repeat(np.arange(50), 3))]:
        a = engine.assign_folds(150, pid, 10, sel, 42)
        b = oracle.assign_folds(150, pid, 10, sel, 42)
        assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
# Prior step: the minimiser of the penalised quadratic model of the coordinate,
# g d + h d^2 / 2 + penalty(beta + d), found by brute force. The Laplace step may
# not carry beta across zero in one step (it stops at zero), so for beta != 0 the
# minimum is taken over the half-line that keeps the sign of beta.
# ---------------------------------------------------------------------------
def _argmin(f, lo, hi, n=200001):
    x = np.linspace(lo, hi, n)
    return x[np.argmin(f(x))], (hi - lo) / (n - 1)


@pytest.mark.parametrize("seed", range(6))
def test_prior_step_minimises_penalised_model(seed):
    rng = np.random.default_rng(seed)
    for _ in range(40):
        g = rng.normal(0, 3)
        h = rng.uniform(0.2, 4)
        var = float(np.exp(rng.uniform(np.log(0.05), np.log(5))))
        lam = np.sqrt(2 / var)
        beta = 0.0 if rng.random() < 0.4 else rng.normal(0, 1.5)
        q = lambda d: g * d + 0.5 * h * d * d
        R = 2 * (abs(g) + lam + abs(beta) / var) / h + 2 * abs(beta) + 1   # holds every minimiser
        # none
        d0, step = _argmin(q, -R, R)
        assert abs(oracle.prior_delta("none", var, g, h, beta) - d0) <= 2 * step
        # normal
        dn, step = _argmin(lambda d: q(d) + (beta + d) ** 2 / (2 * var), -R, R)
        assert abs(oracle.prior_delta("normal", var, g, h, beta) - dn) <= 2 * step
        # laplace
        if beta == 0:
            dl, step = _argmin(lambda d: q(d) + lam * np.abs(d), -R, R)
        elif beta > 0:
            dl, step = _argmin(lambda d: q(d) + lam * (beta + d), -beta, R)
        else:
            dl, step = _argmin(lambda d: q(d) - lam * (beta + d), -R, -beta)
        assert abs(oracle.prior_delta("laplace", var, g, h, beta) - dl) <= 2 * step, (g, h, var, beta)
