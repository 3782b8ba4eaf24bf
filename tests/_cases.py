"""Small seeded datasets covering every model and column format (shared by tests)."""
import numpy as np

from cyclops_amd import _abi
from cyclops_amd import data as D


def _mixed_X(rng, K, n_ind=3, n_sparse=1, n_dense=1, density=0.15):
    cols, fm = [], []
    for _ in range(n_ind):
        cols.append((rng.random(K) < density).astype(float)); fm.append("indicator")
    for _ in range(n_sparse):
        m = rng.random(K) < density * 1.5
        cols.append(np.where(m, rng.integers(1, 5, size=K).astype(float), 0.0)); fm.append("sparse")
    for _ in range(n_dense):
        cols.append(rng.standard_normal(K) * 0.5); fm.append("dense")
    return np.column_stack(cols), fm


def logistic(K=400, seed=1, intercept=True):
    rng = np.random.default_rng(seed)
    X, fm = _mixed_X(rng, K)
    beta = rng.normal(0, 0.7, X.shape[1])
    p = 1 / (1 + np.exp(-(X @ beta - 0.5)))
    y = (rng.random(K) < p).astype(float)
    return D.from_dense(X, y, "lr", formats=fm, add_intercept=intercept)


def poisson(K=400, seed=2, offset=True):
    rng = np.random.default_rng(seed)
    X, fm = _mixed_X(rng, K)
    t = rng.uniform(1, 30, K)
    beta = rng.normal(0, 0.3, X.shape[1])
    y = rng.poisson(t * np.exp(X @ beta - 1.0)).astype(float)
    return D.from_dense(X, y, "pr", formats=fm, add_intercept=not offset,
                        offset=np.log(t) if offset else None)


def clr(n_strata=80, seed=3):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(2, 7, n_strata)
    K = int(sizes.sum())
    pid = np.repeat(np.arange(n_strata), sizes).astype(np.int32)
    X, fm = _mixed_X(rng, K, density=0.25)
    beta = rng.normal(0, 0.6, X.shape[1])
    xb = X @ beta
    y = np.zeros(K)
    start = 0
    for s in sizes:
        w = np.exp(xb[start:start + s]); w /= w.sum()
        y[start + rng.choice(s, p=w)] = 1.0
        if s > 4 and rng.random() < 0.3:   # a second case in some strata (Breslow ties)
            y[start + rng.integers(0, s)] = 1.0
        start += s
    return D.from_dense(X, y, "clr", formats=fm, pid=pid)


def cox(K=300, seed=4, n_strata=2):
    rng = np.random.default_rng(seed)
    X, fm = _mixed_X(rng, K, density=0.25)
    beta = rng.normal(0, 0.5, X.shape[1])
    T = rng.exponential(1.0, K) / np.exp(X @ beta) * 10
    C = rng.uniform(0, 20, K)
    y = (T <= C).astype(float)
    time = np.round(np.minimum(T, C))          # integer days -> many ties
    strata = rng.integers(0, n_strata, K)
    order = np.lexsort((np.arange(K), y, -time, strata))
    return D.from_dense(X[order], y[order], "cox", formats=fm, time=time[order],
                        pid=strata[order].astype(np.int32))


ALL = {"lr": logistic, "pr": poisson, "clr": clr, "cox": cox}
