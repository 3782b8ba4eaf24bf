"""Column-store datasets in the reference's layout, fixture readers and the
seeded synthetic generators for the benchmark workloads.

Layout (CompressedDataColumn, src/cyclops/CompressedDataMatrix.h:40-241): each
column is DENSE, SPARSE, INDICATOR or INTERCEPT; SPARSE/INDICATOR columns hold
sorted int32 row ids (CSC), SPARSE columns also hold values. Row order follows
the reference's sort contracts (R/NewDataConversion.R:125-166): LR/PR by row,
CLR by (stratum, row), Cox by (stratum, time descending, y ascending, row).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi

_MODEL_FROM_NAME = _abi.MODEL_NAMES


@dataclass
class CyclopsData:
    model_type: int
    K: int
    formats: np.ndarray            # int32 [J]
    col_ptr: np.ndarray            # int64 [J+1]
    rows: np.ndarray               # int32 [nnz]
    values: Optional[np.ndarray]   # float64 [nnz] or None
    dense_ptr: Optional[np.ndarray]   # int64 [J+1] or None
    dense_values: Optional[np.ndarray]
    y: np.ndarray                  # float64 [K]
    time: Optional[np.ndarray] = None
    pid: Optional[np.ndarray] = None
    has_offset: bool = False
    names: Optional[List[str]] = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def J(self) -> int:
        return int(self.formats.shape[0])

    @property
    def nnz(self) -> int:
        return int(self.col_ptr[-1])

    def column(self, j):
        """(rows, x) of column j; DENSE/INTERCEPT return all K rows."""
        f = int(self.formats[j])
        if f in (_abi.INDICATOR, _abi.SPARSE):
            a, b = int(self.col_ptr[j]), int(self.col_ptr[j + 1])
            r = self.rows[a:b]
            x = self.values[a:b] if (f == _abi.SPARSE) else np.ones(b - a)
            return r, x
        if f == _abi.DENSE:
            a = int(self.dense_ptr[j])
            return np.arange(self.K, dtype=np.int32), self.dense_values[a:a + self.K]
        return np.arange(self.K, dtype=np.int32), np.ones(self.K)

    def to_dense(self) -> np.ndarray:
        X = np.zeros((self.K, self.J))
        for j in range(self.J):
            r, x = self.column(j)
            X[r, j] = x
        return X

    def desc(self, precision: int = _abi.FP64) -> _abi.DataDesc:
        """ctypes view; the arrays stay owned by this object."""
        d = _abi.DataDesc()
        d.model_type = self.model_type
        d.precision = precision
        d.K = self.K
        d.J = self.J
        d.has_offset = 1 if self.has_offset else 0
        d.format = _abi.ptr(self.formats, C.c_int32)
        d.col_ptr = _abi.ptr(self.col_ptr, C.c_int64)
        d.rows = _abi.ptr(self.rows, C.c_int32)
        d.values = _abi.ptr(self.values, C.c_double)
        d.dense_ptr = _abi.ptr(self.dense_ptr, C.c_int64)
        d.dense_values = _abi.ptr(self.dense_values, C.c_double)
        d.y = _abi.ptr(self.y, C.c_double)
        d.time = _abi.ptr(self.time, C.c_double)
        d.pid = _abi.ptr(self.pid, C.c_int32)
        return d

    def footprint_bytes(self) -> float:
        b = self.rows.nbytes + self.col_ptr.nbytes + self.y.nbytes
        for a in (self.values, self.dense_values, self.time, self.pid):
            if a is not None:
                b += a.nbytes
        return float(b)


def _finish(model, K, formats, cols_rows, cols_vals, dense_cols, y, time=None, pid=None,
            has_offset=False, names=None) -> CyclopsData:
    """Assemble a CyclopsData from per-column lists.

    cols_rows[j]: int array (INDICATOR/SPARSE) or None; cols_vals[j]: values for SPARSE;
    dense_cols[j]: K values for DENSE."""
    J = len(formats)
    counts = np.array([0 if cols_rows[j] is None else len(cols_rows[j]) for j in range(J)], dtype=np.int64)
    col_ptr = np.zeros(J + 1, dtype=np.int64)
    np.cumsum(counts, out=col_ptr[1:])
    nnz = int(col_ptr[-1])
    rows = np.empty(nnz, dtype=np.int32)
    any_sparse = any(int(f) == _abi.SPARSE for f in formats)
    values = np.ones(nnz, dtype=np.float64) if any_sparse else None
    for j in range(J):
        if cols_rows[j] is not None and counts[j]:
            rows[col_ptr[j]:col_ptr[j + 1]] = cols_rows[j]
            if any_sparse and int(formats[j]) == _abi.SPARSE:
                values[col_ptr[j]:col_ptr[j + 1]] = cols_vals[j]
    dense_ptr = None
    dense_values = None
    n_dense = sum(1 for f in formats if int(f) == _abi.DENSE)
    if n_dense:
        dense_ptr = np.zeros(J + 1, dtype=np.int64)
        off = 0
        for j in range(J):
            dense_ptr[j] = off
            if int(formats[j]) == _abi.DENSE:
                off += K
        dense_ptr[J] = off
        dense_values = np.empty(off, dtype=np.float64)
        for j in range(J):
            if int(formats[j]) == _abi.DENSE:
                dense_values[dense_ptr[j]:dense_ptr[j] + K] = dense_cols[j]
    return CyclopsData(
        model_type=model, K=int(K), formats=np.asarray(formats, dtype=np.int32), col_ptr=col_ptr,
        rows=rows, values=values, dense_ptr=dense_ptr, dense_values=dense_values,
        y=np.ascontiguousarray(y, dtype=np.float64),
        time=None if time is None else np.ascontiguousarray(time, dtype=np.float64),
        pid=None if pid is None else np.ascontiguousarray(pid, dtype=np.int32),
        has_offset=has_offset, names=names)


def from_dense(X, y, model, formats=None, time=None, pid=None, add_intercept=False,
               offset=None, names=None) -> CyclopsData:
    """Build a dataset from a dense K x P matrix.

    formats: per column 'dense' | 'sparse' | 'indicator' (default: inferred like
    ModelData.cpp:245-370 -- all non-zeros equal 1 -> INDICATOR, else SPARSE).
    add_intercept: prepend an INTERCEPT column (R/NewDataConversion.R:198).
    offset: K values, stored as a DENSE column 0 with beta fixed at 1 (already on
    the log scale; ModelData.cpp:640-651)."""
    model = _MODEL_FROM_NAME.get(model, model)
    X = np.asarray(X, dtype=np.float64)
    K, P = X.shape
    fm = {"dense": _abi.DENSE, "sparse": _abi.SPARSE, "indicator": _abi.INDICATOR}
    formats_out, rows_l, vals_l, dense_l = [], [], [], []
    nm = []
    if offset is not None:
        formats_out.append(_abi.DENSE); rows_l.append(None); vals_l.append(None)
        dense_l.append(np.asarray(offset, dtype=np.float64)); nm.append("(offset)")
    if add_intercept:
        formats_out.append(_abi.INTERCEPT); rows_l.append(None); vals_l.append(None); dense_l.append(None)
        nm.append("(Intercept)")
    for p in range(P):
        col = X[:, p]
        if formats is None:
            nzv = col[col != 0]
            f = _abi.INDICATOR if np.all(nzv == 1.0) else _abi.SPARSE
        else:
            f = fm[formats[p]] if isinstance(formats[p], str) else int(formats[p])
        formats_out.append(f)
        if f == _abi.DENSE:
            rows_l.append(None); vals_l.append(None); dense_l.append(col.copy())
        else:
            r = np.nonzero(col)[0].astype(np.int32)
            rows_l.append(r); vals_l.append(col[r]); dense_l.append(None)
        nm.append(names[p] if names is not None else f"x{p + 1}")
    return _finish(model, K, formats_out, rows_l, vals_l, dense_l, y, time=time, pid=pid,
                   has_offset=offset is not None, names=nm)


def _csc_slots(m, K, with_values):
    """(i, p[, x]) of a column-compressed matrix given as a scipy sparse matrix or as the
    dgCMatrix slots themselves (0-based row ids i, column pointers p)."""
    if hasattr(m, "tocsc"):
        c = m.tocsc()
        c.sort_indices()
        if c.shape[0] != K:
            raise ValueError("sparse matrix rows must equal len(y)")
        return c.indices.astype(np.int32), c.indptr.astype(np.int64), (c.data.astype(np.float64) if with_values else None)
    if with_values:
        i, pp, x = m
        return np.asarray(i, dtype=np.int32), np.asarray(pp, dtype=np.int64), np.asarray(x, dtype=np.float64)
    i, pp = m[0], m[1]
    return np.asarray(i, dtype=np.int32), np.asarray(pp, dtype=np.int64), None


def from_matrices(y, model, dense=None, sparse=None, indicator=None, time=None, pid=None, offset=None,
                  names=None) -> CyclopsData:
    """cyclopsModelData's column assembly (RcppModelData.cpp:671-737, 751-900): the offset
    column (when given, beta fixed at 1), then the dense matrix's columns (K x Pd, row
    major here), then the SPARSE columns of a column-compressed matrix (dgCMatrix slots
    i, p, x or a scipy sparse matrix), then the INDICATOR columns of another (slots i, p;
    every stored entry is 1)."""
    model = _MODEL_FROM_NAME.get(model, model)
    y = np.asarray(y, dtype=np.float64)
    K = len(y)
    formats, rows_l, vals_l, dense_l = [], [], [], []
    if offset is not None:
        formats.append(_abi.DENSE); rows_l.append(None); vals_l.append(None)
        dense_l.append(np.asarray(offset, dtype=np.float64))
    if dense is not None:
        Xd = np.asarray(dense, dtype=np.float64).reshape(K, -1)
        for q in range(Xd.shape[1]):
            formats.append(_abi.DENSE); rows_l.append(None); vals_l.append(None); dense_l.append(Xd[:, q])
    for m, fmt in ((sparse, _abi.SPARSE), (indicator, _abi.INDICATOR)):
        if m is None:
            continue
        i, pp, x = _csc_slots(m, K, fmt == _abi.SPARSE)
        for q in range(len(pp) - 1):
            a, b = int(pp[q]), int(pp[q + 1])
            formats.append(fmt)
            rows_l.append(i[a:b])
            vals_l.append(x[a:b] if x is not None else None)
            dense_l.append(None)
    if not formats:
        raise ValueError("no covariates")
    return _finish(model, K, formats, rows_l, vals_l, dense_l, y, time=time, pid=pid,
                   has_offset=offset is not None, names=names)


def read_generic_text(path, model, add_intercept=None, first_field="row", sep=None,
                      dense=False) -> CyclopsData:
    """Reader for the reference's generic text fixtures (io/NewGenericInputReader.h:16-217):
    one row per line, `<id> <y> [col:value ...]`; a leading '#' line may name options
    (add_intercept, stratum_label). first_field: 'row' | 'stratum' | 'time'.
    Columns are ordered by covariate id; a column whose values are all 1 is INDICATOR,
    otherwise SPARSE (ModelData.cpp:245-370); `dense` makes every column DENSE."""
    model = _MODEL_FROM_NAME.get(model, model)
    ids, ys, entries = [], [], []
    header_opts = ""
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line:
                continue
            if line.startswith("#") or line.startswith("header"):
                header_opts += line
                continue
            parts = line.split(sep)
            ids.append(float(parts[0]))
            ys.append(float(parts[1]))
            ent = []
            for tok in parts[2:]:
                c, v = tok.split(":")
                ent.append((int(c), float(v)))
            entries.append(ent)
    if add_intercept is None:
        add_intercept = "add_intercept" in header_opts
    K = len(ys)
    cov_ids = sorted({c for ent in entries for c, _ in ent})
    col_index = {c: i for i, c in enumerate(cov_ids)}
    X = np.zeros((K, len(cov_ids)))
    for k, ent in enumerate(entries):
        for c, v in ent:
            X[k, col_index[c]] = v
    y = np.array(ys)
    time = pid = None
    if first_field == "stratum":
        pid = np.array(ids, dtype=np.int32)
    elif first_field == "time":
        time = np.array(ids)
        pid = np.zeros(K, dtype=np.int32)
    formats = ["dense"] * len(cov_ids) if dense else None
    return from_dense(X, y, model, formats=formats, time=time, pid=pid, add_intercept=add_intercept,
                      names=[str(c) for c in cov_ids])


def read_csv_clr(path) -> CyclopsData:
    """inst/extdata/test1-clr.txt: `stratum,y,x1..x10` comma-separated, 0/1 covariates."""
    a = np.loadtxt(path, delimiter=",")
    pid = a[:, 0].astype(np.int32)
    y = a[:, 1]
    X = a[:, 2:]
    order = np.argsort(pid, kind="stable")
    return from_dense(X[order], y[order], "clr", pid=pid[order])


def sort_cox(time, y, X, strata=None):
    """Row order for Cox: (stratum, time descending, y ascending, row)."""
    K = len(time)
    s = np.zeros(K) if strata is None else np.asarray(strata)
    order = np.lexsort((np.arange(K), y, -np.asarray(time), s))
    return order


# ---------------------------------------------------------------------------
# Synthetic workloads (seeded). Bernoulli(p) row selection per column is drawn
# with geometric gaps, which gives a Binomial(K, p) count of uniformly chosen,
# sorted rows -- the same law as sampling the count and then rows without
# replacement.
# ---------------------------------------------------------------------------

def _bernoulli_rows(rng, K, p):
    if p >= 1.0:
        return np.arange(K, dtype=np.int32)
    if p <= 0.0:
        return np.zeros(0, dtype=np.int32)
    n_exp = K * p
    m = int(n_exp + 6.0 * math.sqrt(n_exp + 1.0) + 16)
    while True:
        gaps = rng.geometric(p, size=m)
        pos = np.cumsum(gaps, dtype=np.int64) - 1
        if pos[-1] >= K:
            return pos[pos < K].astype(np.int32)
        m *= 2


def zipf_probabilities(P, s=1.1, mean_p=0.02):
    """p_r = min(1, c r^-s) with c solved so that mean(p) = mean_p."""
    r = np.arange(1, P + 1, dtype=np.float64)
    base = r ** (-s)
    lo, hi = 0.0, 1e6
    for _ in range(200):
        c = 0.5 * (lo + hi)
        m = np.minimum(1.0, c * base).mean()
        if m < mean_p:
            lo = c
        else:
            hi = c
    return np.minimum(1.0, 0.5 * (lo + hi) * base)


def _indicator_columns(rng, K, probs):
    cols = [_bernoulli_rows(rng, K, float(p)) for p in probs]
    return cols


def _xbeta(K, cols, beta_idx, beta_val):
    xb = np.zeros(K)
    for j, b in zip(beta_idx, beta_val):
        xb[cols[j]] += b
    return xb


def make_logistic(K=100_000, P=1_000, density=0.01, n_true=50, sd=0.5, intercept=-2.0, seed=1000,
                  probs=None):
    """LR workload: P indicator columns, INTERCEPT column first (unpenalised)."""
    rng = np.random.default_rng(seed)
    probs = np.full(P, density) if probs is None else probs
    cols = _indicator_columns(rng, K, probs)
    bidx = rng.choice(P, size=min(n_true, P), replace=False)
    bval = rng.normal(0.0, sd, size=len(bidx))
    xb = intercept + _xbeta(K, cols, bidx, bval)
    y = (rng.random(K) < 1.0 / (1.0 + np.exp(-xb))).astype(np.float64)
    formats = [_abi.INTERCEPT] + [_abi.INDICATOR] * P
    d = _finish(_abi.LOGISTIC, K, formats, [None] + cols, [None] * (P + 1), [None] * (P + 1), y)
    true_beta = np.zeros(P + 1)
    true_beta[0] = intercept
    true_beta[1 + bidx] = bval
    return d, true_beta


def make_poisson(K=1_000_000, P=20_000, mean_density=0.02, zipf_s=1.1, n_true=200, sd=0.3,
                 seed=1001, probs=None):
    """PR workload with a log(time-at-risk) offset stored as DENSE column 0 (beta fixed 1)."""
    rng = np.random.default_rng(seed)
    probs = zipf_probabilities(P, zipf_s, mean_density) if probs is None else probs
    cols = _indicator_columns(rng, K, probs)
    t = rng.uniform(1.0, 365.0, size=K)
    off = np.log(t)
    bidx = rng.choice(P, size=min(n_true, P), replace=False)
    bval = rng.normal(0.0, sd, size=len(bidx))
    xb = off + _xbeta(K, cols, bidx, bval)
    y = rng.poisson(np.exp(xb)).astype(np.float64)
    formats = [_abi.DENSE] + [_abi.INDICATOR] * P
    d = _finish(_abi.POISSON, K, formats, [None] + cols, [None] * (P + 1), [off] + [None] * P, y,
                has_offset=True)
    true_beta = np.zeros(P + 1)
    true_beta[0] = 1.0
    true_beta[1 + bidx] = bval
    return d, true_beta


def make_cox(K=2_000_000, P=50_000, density=0.005, n_dense=10, n_true=100, sd=0.2,
             censor_frac=0.7, seed=1002, probs=None):
    """Breslow Cox workload: P indicator columns + n_dense DENSE N(0,1) columns, one stratum,
    integer-day times, rows sorted (time desc, y asc, row)."""
    rng = np.random.default_rng(seed)
    probs = np.full(P, density) if probs is None else probs
    cols = _indicator_columns(rng, K, probs)
    dense = [rng.standard_normal(K).astype(np.float32).astype(np.float64) for _ in range(n_dense)]
    n_all = P + n_dense
    bidx = rng.choice(n_all, size=min(n_true, n_all), replace=False)
    bval = rng.normal(0.0, sd, size=len(bidx))
    xb = np.zeros(K)
    for j, b in zip(bidx, bval):
        if j < P:
            xb[cols[j]] += b
        else:
            xb += b * dense[j - P]
    T = rng.exponential(1.0, size=K) / (np.exp(xb) / 1000.0)
    day = np.round(T)
    # censoring C ~ U(0, cmax) with cmax chosen for the requested censored fraction
    u = rng.random(K)
    lo, hi = 1e-3, 1e9
    for _ in range(100):
        cmax = math.sqrt(lo * hi)
        frac = np.mean(u * cmax < T)
        if frac > censor_frac:
            lo = cmax
        else:
            hi = cmax
    cmax = math.sqrt(lo * hi)
    C_ = np.round(u * cmax)
    event = (T <= u * cmax).astype(np.float64)
    time = np.where(event == 1.0, day, C_)
    order = np.lexsort((np.arange(K), event, -time))
    inv = np.empty(K, dtype=np.int64)
    inv[order] = np.arange(K)
    cols_s = [np.sort(inv[c]).astype(np.int32) for c in cols]
    dense_s = [dv[order] for dv in dense]
    formats = [_abi.INDICATOR] * P + [_abi.DENSE] * n_dense
    d = _finish(_abi.COX, K, formats, cols_s + [None] * n_dense, [None] * n_all,
                [None] * P + dense_s, event[order], time=time[order], pid=np.zeros(K, dtype=np.int32))
    true_beta = np.zeros(n_all)
    true_beta[bidx] = bval
    return d, true_beta


def make_clr(n_strata=500_000, P=100_000, density=0.003, n_true=300, sd=0.3, max_controls=10,
             seed=1003, probs=None):
    """CLR workload: strata of 1 case + U{1..max_controls} controls, rows sorted (stratum,row)."""
    rng = np.random.default_rng(seed)
    sizes = 1 + rng.integers(1, max_controls + 1, size=n_strata)
    K = int(sizes.sum())
    pid = np.repeat(np.arange(n_strata, dtype=np.int32), sizes)
    probs = np.full(P, density) if probs is None else probs
    cols = _indicator_columns(rng, K, probs)
    bidx = rng.choice(P, size=min(n_true, P), replace=False)
    bval = rng.normal(0.0, sd, size=len(bidx))
    xb = _xbeta(K, cols, bidx, bval)
    # one case per stratum drawn with probability proportional to exp(xb)
    w = np.exp(xb - xb.max())
    starts = np.zeros(n_strata + 1, dtype=np.int64)
    np.cumsum(sizes, out=starts[1:])
    cw = np.cumsum(w)
    tot_before = np.concatenate([[0.0], cw])[starts[:-1]]
    tot_after = cw[starts[1:] - 1]
    target = tot_before + rng.random(n_strata) * (tot_after - tot_before)
    case = np.searchsorted(cw, target, side="right")
    case = np.clip(case, starts[:-1], starts[1:] - 1)
    y = np.zeros(K)
    y[case] = 1.0
    formats = [_abi.INDICATOR] * P
    d = _finish(_abi.CONDITIONAL_LOGISTIC, K, formats, cols, [None] * P, [None] * P, y, pid=pid)
    true_beta = np.zeros(P)
    true_beta[bidx] = bval
    return d, true_beta


def make_config(index, scale=1.0, seed=None):
    """The five benchmark workloads (indices 0..3; 4 is the serving mix) at full size when
    scale == 1. `scale` shrinks rows and columns for tests (density kept)."""
    s = seed if seed is not None else 1000 + index
    if index == 0:
        return make_logistic(K=max(200, int(100_000 * scale)), P=max(10, int(1_000 * scale)),
                             n_true=max(2, int(50 * scale)), seed=s)
    if index == 1:
        return make_poisson(K=max(200, int(1_000_000 * scale)), P=max(10, int(20_000 * scale)),
                            n_true=max(2, int(200 * scale)), seed=s)
    if index == 2:
        return make_cox(K=max(200, int(2_000_000 * scale)), P=max(10, int(50_000 * scale)),
                        n_true=max(2, int(100 * scale)), seed=s)
    if index == 3:
        return make_clr(n_strata=max(30, int(500_000 * scale)), P=max(10, int(100_000 * scale)),
                        n_true=max(2, int(300 * scale)), seed=s)
    raise ValueError(index)


def make_request(r, max_rows=2_000_000, max_cols=200_000, seed_base=5_000_000):
    """One request of the serving mix: model LR/PR/Cox with probabilities .5/.3/.2,
    N ~ logU[1e4, max_rows], P ~ logU[1e3, max_cols], density ~ U[0.2%, 2%] with a
    Zipf(1.1) column profile, Laplace or Normal prior with variance ~ logU[1e-3, 1]."""
    rng = np.random.default_rng(seed_base + r)
    u = rng.random()
    model = "lr" if u < 0.5 else ("pr" if u < 0.8 else "cox")
    K = int(round(math.exp(rng.uniform(math.log(1e4), math.log(max(1e4, max_rows))))))
    P = int(round(math.exp(rng.uniform(math.log(1e3), math.log(max(1e3, max_cols))))))
    dens = rng.uniform(0.002, 0.02)
    prior = "laplace" if rng.random() < 0.5 else "normal"
    var = math.exp(rng.uniform(math.log(1e-3), math.log(1.0)))
    probs = zipf_probabilities(P, 1.1, dens)
    seed = seed_base + r
    if model == "lr":
        d, _ = make_logistic(K=K, P=P, n_true=min(50, P), seed=seed, probs=probs)
    elif model == "pr":
        d, _ = make_poisson(K=K, P=P, n_true=min(200, P), seed=seed, probs=probs)
    else:
        d, _ = make_cox(K=K, P=P, n_dense=0, n_true=min(100, P), seed=seed, probs=probs)
    return {"model": model, "data": d, "prior": prior, "variance": var}
