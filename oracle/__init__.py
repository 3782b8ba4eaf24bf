"""THE ORACLE (test infrastructure only).

ctypes wrapper of oracle/liboracle.so, the FP64 CPU restatement of the
reference's CCD hot path (oracle/ccd_oracle.cpp lists the reference file:line
each part follows). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package; it checks the engine and is never
the thing measured or shipped.

Every function here is pinned by a test under tests/ (closed forms, textbook
solvers, brute force, finite differences, KKT, the reference's stored SAS
log-likelihood); none is "parity unpinned".
"""
import ctypes as C
import os

import numpy as np

from cyclops_amd import _abi
from cyclops_amd.control import FitArgsHolder, create_control, create_prior, intercept_index

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build as _b
            _b.build()
        L = C.CDLL(path)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_abi.DataDesc)]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        L.orc_num_pids.restype = C.c_int64
        L.orc_num_pids.argtypes = [vp]
        L.orc_set_weights.argtypes = [vp, dp, C.c_int]
        L.orc_fixed_terms.argtypes = [vp, C.c_int]
        L.orc_compute_xbeta.argtypes = [vp, dp]
        L.orc_remaining_stats.argtypes = [vp, C.c_int]
        L.orc_numerator.argtypes = [vp, C.c_int, C.c_int]
        L.orc_grad_hess.argtypes = [vp, C.c_int, dp, dp, C.c_int]
        L.orc_update_xbeta.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        L.orc_loglik.restype = C.c_double
        L.orc_loglik.argtypes = [vp, C.c_int]
        L.orc_pred_loglik.restype = C.c_double
        L.orc_pred_loglik.argtypes = [vp, dp]
        L.orc_gradient_objective.restype = C.c_double
        L.orc_gradient_objective.argtypes = [vp, C.c_int]
        L.orc_pred_est.argtypes = [vp, dp, dp]
        L.orc_get_xbeta.argtypes = [vp, dp]
        L.orc_get_denom.argtypes = [vp, dp]
        L.orc_fit.restype = C.c_int
        L.orc_fit.argtypes = [vp, C.POINTER(_abi.FitArgs), dp, C.POINTER(_abi.FitResult)]
        L.orc_time_cycles.restype = C.c_double
        L.orc_time_cycles.argtypes = [vp, C.POINTER(_abi.FitArgs), dp, C.c_int]
        L.orc_cv_assign_folds.argtypes = [C.c_int64, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                          C.c_int64, C.POINTER(C.c_int32)]
        L.orc_cross_validate.restype = C.c_int
        L.orc_cross_validate.argtypes = [vp, C.POINTER(_abi.FitArgs), C.POINTER(_abi.CvArgs), dp, dp, dp,
                                         C.POINTER(_abi.FitResult), dp]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleEngine:
    """The oracle's ModelSpecifics<Model,double> (one instance per dataset)."""

    def __init__(self, data):
        self.data = data
        self._desc = data.desc(_abi.FP64)
        self.h = lib().orc_create(C.byref(self._desc))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def N(self):
        return int(lib().orc_num_pids(self.h))

    def set_weights(self, w=None):
        self._w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        lib().orc_set_weights(self.h, None if self._w is None else _d(self._w), 0 if w is None else 1)

    def fixed_terms(self, use_w=False):
        lib().orc_fixed_terms(self.h, int(use_w))

    def compute_xbeta(self, beta):
        b = np.ascontiguousarray(beta, dtype=np.float64)
        lib().orc_compute_xbeta(self.h, _d(b))

    def remaining_stats(self, use_w=False):
        lib().orc_remaining_stats(self.h, int(use_w))

    def numerator(self, j, use_w=False):
        lib().orc_numerator(self.h, j, int(use_w))

    def grad_hess(self, j, use_w=False):
        g, h = C.c_double(), C.c_double()
        lib().orc_grad_hess(self.h, j, C.byref(g), C.byref(h), int(use_w))
        return g.value, h.value

    def update_xbeta(self, j, delta, use_w=False):
        lib().orc_update_xbeta(self.h, j, float(delta), int(use_w))

    def loglik(self, use_w=False):
        return lib().orc_loglik(self.h, int(use_w))

    def pred_loglik(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        return lib().orc_pred_loglik(self.h, _d(w))

    def gradient_objective(self, use_w=False):
        return lib().orc_gradient_objective(self.h, int(use_w))

    def pred_est(self, w=None):
        out = np.zeros(self.data.K)
        ww = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        lib().orc_pred_est(self.h, _d(out), None if ww is None else _d(ww))
        return out

    def xbeta(self):
        out = np.zeros(self.data.K)
        lib().orc_get_xbeta(self.h, _d(out))
        return out

    def denominators(self):
        out = np.zeros(self.N)
        lib().orc_get_denom(self.h, _d(out))
        return out


def _args(data, prior, control, weights=None, fixed=None):
    return FitArgsHolder(data.J, prior, control, weights=weights, fixed=fixed,
                         intercept_index=intercept_index(data))


def fit(data, prior=None, control=None, weights=None, fixed=None, beta0=None):
    """FP64 CPU fit (CyclicCoordinateDescent::update). Returns (beta, result dict)."""
    prior = prior or create_prior("none")
    control = control or create_control()
    hold = _args(data, prior, control, weights, fixed)
    eng = OracleEngine(data)
    beta = np.zeros(data.J) if beta0 is None else np.array(beta0, dtype=np.float64)
    res = _abi.FitResult()
    rc = lib().orc_fit(eng.h, C.byref(hold.args), _d(beta), C.byref(res))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return beta, res.as_dict()


def fisher_information(data, beta, indices):
    """Fisher information of columns `indices` at beta, unit weights: [m][m] (FP64 CPU)."""
    L = lib()
    L.orc_fisher_information.restype = C.c_int
    L.orc_fisher_information.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int,
                                         C.POINTER(C.c_double)]
    eng = OracleEngine(data)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    m = len(idx)
    out = np.zeros((m, m))
    rc = L.orc_fisher_information(eng.h, _d(b), idx.ctypes.data_as(C.POINTER(C.c_int)), m, _d(out))
    if rc != 0:
        raise RuntimeError(L.orc_last_error().decode())
    return out


def profile(data, prior, control, beta, indices, threshold=1.920729, include_penalty=True):
    """Profile-likelihood limits (CcdInterface::profileModel, one CCD per bound started at
    the mode): returns (lower, upper, evaluations) per column of `indices`."""
    L = lib()
    L.orc_profile.restype = C.c_int
    L.orc_profile.argtypes = [C.c_void_p, C.POINTER(_abi.FitArgs), C.POINTER(C.c_double), C.POINTER(C.c_int),
                              C.c_int, C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_int)]
    hold = _args(data, prior, control)
    eng = OracleEngine(data)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    idx = np.ascontiguousarray(indices, dtype=np.int32)
    m = len(idx)
    lo, up = np.zeros(m), np.zeros(m)
    ev = np.zeros(m, dtype=np.int32)
    rc = L.orc_profile(eng.h, C.byref(hold.args), _d(b), idx.ctypes.data_as(C.POINTER(C.c_int)), m,
                       float(threshold), int(bool(include_penalty)), _d(lo), _d(up),
                       ev.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        raise RuntimeError(L.orc_last_error().decode())
    return lo, up, ev


def profile_likelihood(data, prior, control, beta, index, points, include_penalty=True):
    """Profile log-likelihood of column `index` at each of `points` (evaluateProfileModel,
    one clone per point started from the fit with the column held at 0)."""
    L = lib()
    L.orc_profile_likelihood.restype = C.c_int
    L.orc_profile_likelihood.argtypes = [C.c_void_p, C.POINTER(_abi.FitArgs), C.POINTER(C.c_double), C.c_int,
                                         C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double)]
    hold = _args(data, prior, control)
    eng = OracleEngine(data)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    x = np.ascontiguousarray(points, dtype=np.float64)
    out = np.zeros(len(x))
    rc = L.orc_profile_likelihood(eng.h, C.byref(hold.args), _d(b), int(index), _d(x), len(x),
                                  int(bool(include_penalty)), _d(out))
    if rc != 0:
        raise RuntimeError(L.orc_last_error().decode())
    return out


def prior_delta(prior_type, variance, g, h, beta):
    """The prior's coordinate step (CovariatePrior.h getDelta) for gradient g, Hessian h."""
    L = lib()
    L.orc_prior_delta.restype = C.c_double
    L.orc_prior_delta.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    return L.orc_prior_delta(_abi.PRIOR_NAMES[prior_type], variance, g, h, beta)


def time_coordinates(data, prior, control, n, weights=None):
    """Seconds for the first n coordinates of the first cycle from a cold start, one core."""
    hold = _args(data, prior, control, weights=weights)
    eng = OracleEngine(data)
    L = lib()
    L.orc_time_coordinates.restype = C.c_double
    L.orc_time_coordinates.argtypes = [C.c_void_p, C.POINTER(_abi.FitArgs), C.c_int]
    return L.orc_time_coordinates(eng.h, C.byref(hold.args), int(n))


def time_cycles(data, prior, control, cycles, beta0=None, weights=None):
    """Seconds for exactly `cycles` CCD cycles on one CPU core."""
    hold = _args(data, prior, control, weights=weights)
    eng = OracleEngine(data)
    beta = np.zeros(data.J) if beta0 is None else np.array(beta0, dtype=np.float64)
    s = lib().orc_time_cycles(eng.h, C.byref(hold.args), _d(beta), int(cycles))
    return s, beta


def assign_folds(K, pid, fold, selector, seed):
    out = np.zeros(K, dtype=np.int32)
    p = np.ascontiguousarray(pid if pid is not None else np.arange(K), dtype=np.int32)
    lib().orc_cv_assign_folds(K, p.ctypes.data_as(C.POINTER(C.c_int32)), fold, selector, seed,
                              out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def cross_validate(data, prior, control, selector):
    hold = _args(data, prior, control)
    eng = OracleEngine(data)
    cv = _abi.CvArgs()
    cv.fold = control.fold
    cv.grid_size = control.grid_size
    cv.lower_limit = control.lower_limit
    cv.upper_limit = control.upper_limit
    cv.selector = selector
    cv.warm_start = 0 if control.reset_coefficients else 1
    cv.seed = control.seed
    grid = np.zeros(control.grid_size)
    pred = np.zeros(control.grid_size * control.fold)
    beta = np.zeros(data.J)
    res = _abi.FitResult()
    best = C.c_double()
    rc = lib().orc_cross_validate(eng.h, C.byref(hold.args), C.byref(cv), _d(grid), _d(pred), _d(beta),
                                  C.byref(res), C.byref(best))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return {"grid": grid, "pred": pred.reshape(control.grid_size, control.fold), "beta": beta,
            "result": res.as_dict(), "best_variance": best.value}
