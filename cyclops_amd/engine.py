"""Python mirror of the reference's fitting interface over the C ABI.

    Engine          <-> bsccs::AbstractModelSpecifics (per-coordinate calls)
    fit_cyclops_model <-> fitCyclopsModel (R/ModelFit.R:65-293) for one fit
    fit_batch       <-> the fold x grid fits of a cross-validation, batched
    cross_validate  <-> CcdInterface::runCrossValidation (CcdInterface.cpp:731-793)

Every call goes through libcyclops_amd.so; there is no CPU fallback: a missing
library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _abi
from .control import FitArgsHolder, create_control, create_prior, intercept_index

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcyclops_amd.so")
_LIB = None


class CyclopsError(RuntimeError):
    pass


def lib():
    """Load libcyclops_amd.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CyclopsError(f"{LIB_PATH} is missing: build it with `python -m cyclops_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        i32 = C.c_int32
        L.ccd_last_error.restype = C.c_char_p
        L.ccd_version.restype = C.c_char_p
        L.ccd_device_count.argtypes = [C.POINTER(C.c_int)]
        L.ccd_create.argtypes = [C.POINTER(_abi.DataDesc), C.c_int, C.POINTER(vp)]
        L.ccd_clone.argtypes = [vp, C.POINTER(vp)]
        L.ccd_destroy.argtypes = [vp]
        L.ccd_get_dims.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(i32), C.POINTER(C.c_int64)]
        L.ccd_set_weights.argtypes = [vp, dp, C.c_int]
        L.ccd_fixed_terms.argtypes = [vp, C.c_int]
        L.ccd_compute_xbeta.argtypes = [vp, dp]
        L.ccd_zero_xbeta.argtypes = [vp]
        L.ccd_axpy_xbeta.argtypes = [vp, C.c_double, C.c_int]
        L.ccd_remaining_stats.argtypes = [vp, C.c_int]
        L.ccd_numerator.argtypes = [vp, C.c_int, C.c_int]
        L.ccd_grad_hess.argtypes = [vp, C.c_int, dp, dp, C.c_int]
        L.ccd_grad_hess_noalloc.argtypes = [vp, C.c_int, dp, dp, C.c_int]
        L.ccd_update_xbeta.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        L.ccd_loglik.argtypes = [vp, C.c_int, dp]
        L.ccd_pred_loglik.argtypes = [vp, dp, dp]
        L.ccd_pred_est.argtypes = [vp, dp, dp]
        L.ccd_gradient_objective.argtypes = [vp, C.c_int, dp]
        L.ccd_get_xbeta.argtypes = [vp, dp]
        L.ccd_save_xbeta.argtypes = [vp]
        L.ccd_get_xbeta_save.argtypes = [vp, dp]
        L.ccd_fit.argtypes = [vp, C.POINTER(_abi.FitArgs), C.c_int, dp, C.POINTER(_abi.FitResult)]
        L.ccd_fit_batch.argtypes = [vp, C.POINTER(_abi.FitArgs), C.POINTER(_abi.BatchArgs), dp,
                                    C.POINTER(_abi.FitResult)]
        L.ccd_cross_validate.argtypes = [vp, C.POINTER(_abi.FitArgs), C.POINTER(_abi.CvArgs), dp, dp, dp,
                                         C.POINTER(_abi.FitResult), C.POINTER(_abi.CvResult)]
        L.ccd_cv_assign_folds.argtypes = [C.c_int64, C.POINTER(i32), i32, i32, C.c_int64, C.POINTER(i32)]
        L.ccd_fisher_information.argtypes = [vp, dp, C.POINTER(i32), i32, dp]
        L.ccd_fit_multi.argtypes = [C.POINTER(vp), i32, C.POINTER(_abi.FitArgs), C.POINTER(dp),
                                    C.POINTER(_abi.FitResult)]
        L.ccd_profile.argtypes = [vp, C.POINTER(_abi.FitArgs), dp, C.POINTER(i32), i32, C.c_double, C.c_int,
                                  C.c_int, dp, dp, C.POINTER(i32)]
        L.ccd_profile_likelihood.argtypes = [vp, C.POINTER(_abi.FitArgs), dp, i32, dp, i32, C.c_int, dp]
        L.ccd_get_kernel_stats.argtypes = [vp, C.POINTER(_abi.KernelStats)]
        L.ccd_reset_kernel_stats.argtypes = [vp]
        L.ccd_estimate_footprint.argtypes = [C.POINTER(_abi.DataDesc), dp]
        L.ccd_copy_bandwidth.argtypes = [C.c_int, C.c_double, i32, dp]
        L.ccd_batch_begin.argtypes = [vp, C.POINTER(_abi.FitArgs), C.POINTER(_abi.BatchArgs), dp, C.POINTER(vp)]
        L.ccd_batch_step.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(i32)]
        L.ccd_batch_end.argtypes = [vp, dp, C.POINTER(_abi.FitResult)]
        L.ccd_batch_layout.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
        if hasattr(L, "ccd_server_start"):
            L.ccd_server_start.argtypes = [C.POINTER(_abi.ServerConfig), C.POINTER(vp)]
            L.ccd_server_submit.argtypes = [vp, C.POINTER(_abi.DataDesc), C.POINTER(_abi.FitArgs), dp,
                                            C.POINTER(C.c_int64)]
            L.ccd_server_poll.argtypes = [vp, C.c_int64, C.c_int, dp, C.POINTER(_abi.FitResult),
                                          C.POINTER(_abi.RequestStats)]
            L.ccd_server_release.argtypes = [vp, C.c_int64]
            L.ccd_server_stop.argtypes = [vp]
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise CyclopsError(f"ccd error {rc}: {lib().ccd_last_error().decode()}")


def device_count() -> int:
    n = C.c_int(0)
    check(lib().ccd_device_count(C.byref(n)))
    return n.value


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def copy_bandwidth(device=0, nbytes=1 << 30, reps=10):
    """Achievable HBM bandwidth of the device (GB/s, read + write) from a copy kernel."""
    v = C.c_double()
    check(lib().ccd_copy_bandwidth(int(device), float(nbytes), int(reps), C.byref(v)))
    return v.value


class Engine:
    """Device engine for one dataset: the AbstractModelSpecifics interface."""

    def __init__(self, data, precision=_abi.FP64, device=0, _handle=None):
        self.data = data
        self.precision = _abi.precision_of(precision)
        if _handle is not None:
            self.h = _handle
            return
        self._desc = data.desc(self.precision)
        h = C.c_void_p()
        check(lib().ccd_create(C.byref(self._desc), device, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().ccd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clone(self):
        h = C.c_void_p()
        check(lib().ccd_clone(self.h, C.byref(h)))
        return Engine(self.data, self.precision, _handle=h)

    def dims(self):
        K, J, N = C.c_int64(), C.c_int32(), C.c_int64()
        check(lib().ccd_get_dims(self.h, C.byref(K), C.byref(J), C.byref(N)))
        return K.value, J.value, N.value

    # --- per-coordinate interface -------------------------------------------
    def set_weights(self, w=None):
        self._w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        check(lib().ccd_set_weights(self.h, None if self._w is None else _d(self._w), 0 if w is None else 1))

    def fixed_terms(self, use_w=False):
        check(lib().ccd_fixed_terms(self.h, int(use_w)))

    def compute_xbeta(self, beta):
        b = np.ascontiguousarray(beta, dtype=np.float64)
        check(lib().ccd_compute_xbeta(self.h, _d(b)))

    def zero_xbeta(self):
        check(lib().ccd_zero_xbeta(self.h))

    def axpy_xbeta(self, beta, j):
        check(lib().ccd_axpy_xbeta(self.h, float(beta), int(j)))

    def remaining_stats(self, use_w=False):
        check(lib().ccd_remaining_stats(self.h, int(use_w)))

    def numerator(self, j, use_w=False):
        check(lib().ccd_numerator(self.h, int(j), int(use_w)))

    def grad_hess(self, j, use_w=False):
        g, h = C.c_double(), C.c_double()
        check(lib().ccd_grad_hess(self.h, int(j), C.byref(g), C.byref(h), int(use_w)))
        return g.value, h.value

    def update_xbeta(self, j, delta, use_w=False):
        check(lib().ccd_update_xbeta(self.h, int(j), float(delta), int(use_w)))

    def loglik(self, use_w=False):
        v = C.c_double()
        check(lib().ccd_loglik(self.h, int(use_w), C.byref(v)))
        return v.value

    def pred_loglik(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        v = C.c_double()
        check(lib().ccd_pred_loglik(self.h, _d(w), C.byref(v)))
        return v.value

    def pred_est(self, w=None):
        out = np.zeros(self.data.K)
        ww = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        check(lib().ccd_pred_est(self.h, _d(out), None if ww is None else _d(ww)))
        return out

    def gradient_objective(self, use_w=False):
        v = C.c_double()
        check(lib().ccd_gradient_objective(self.h, int(use_w), C.byref(v)))
        return v.value

    def xbeta(self):
        out = np.zeros(self.data.K)
        check(lib().ccd_get_xbeta(self.h, _d(out)))
        return out

    def save_xbeta(self):
        check(lib().ccd_save_xbeta(self.h))

    def xbeta_save(self):
        out = np.zeros(self.data.K)
        check(lib().ccd_get_xbeta_save(self.h, _d(out)))
        return out

    # --- whole fits ------------------------------------------------------------
    def fit(self, prior=None, control=None, weights=None, fixed=None, beta0=None, mode="plain"):
        prior = prior or create_prior("none")
        control = control or create_control()
        hold = FitArgsHolder(self.data.J, prior, control, weights=weights, fixed=fixed,
                             intercept_index=intercept_index(self.data))
        beta = np.zeros(self.data.J) if beta0 is None else np.array(beta0, dtype=np.float64)
        res = _abi.FitResult()
        m = _abi.MODE_PLAIN if mode == "plain" else _abi.MODE_RESIDENT
        check(lib().ccd_fit(self.h, C.byref(hold.args), m, _d(beta), C.byref(res)))
        return beta, res.as_dict()

    def _batch_args(self, variances, prior_type, control, held_out_fold, row_fold, betas0, exclude, fixed_index):
        control = control or create_control()
        n = len(variances)
        prior = create_prior(prior_type, float(variances[0]), exclude=exclude)
        keep = {"hold": FitArgsHolder(self.data.J, prior, control, intercept_index=intercept_index(self.data)),
                "var": np.ascontiguousarray(variances, dtype=np.float64),
                "hof": None if held_out_fold is None else np.ascontiguousarray(held_out_fold, dtype=np.int32),
                "rf": None if row_fold is None else np.ascontiguousarray(row_fold, dtype=np.int32),
                "fi": None if fixed_index is None else np.ascontiguousarray(fixed_index, dtype=np.int32)}
        b = _abi.BatchArgs()
        b.n_fits = n
        b.variances = _d(keep["var"])
        b.held_out_fold = _abi.ptr(keep["hof"], C.c_int32)
        b.row_fold = _abi.ptr(keep["rf"], C.c_int32)
        b.fixed_index = _abi.ptr(keep["fi"], C.c_int32)
        betas = np.zeros((n, self.data.J)) if betas0 is None else np.array(betas0, dtype=np.float64)
        return b, np.ascontiguousarray(betas), keep

    def fit_batch(self, variances, prior_type="laplace", control=None, held_out_fold=None, row_fold=None,
                  betas0=None, exclude=(), fixed_index=None):
        n = len(variances)
        b, betas, keep = self._batch_args(variances, prior_type, control, held_out_fold, row_fold, betas0, exclude,
                                          fixed_index)
        results = (_abi.FitResult * n)()
        check(lib().ccd_fit_batch(self.h, C.byref(keep["hold"].args), C.byref(b), _d(betas), results))
        return betas, [r.as_dict() for r in results]

    def begin_batch(self, variances, prior_type="laplace", control=None, held_out_fold=None, row_fold=None,
                    betas0=None, exclude=(), fixed_index=None):
        """The fits of fit_batch, advanced by SteppedBatch.step(cycles) (ccd_batch_begin)."""
        b, betas, keep = self._batch_args(variances, prior_type, control, held_out_fold, row_fold, betas0, exclude,
                                          fixed_index)
        h = C.c_void_p()
        check(lib().ccd_batch_begin(self.h, C.byref(keep["hold"].args), C.byref(b), _d(betas), C.byref(h)))
        return SteppedBatch(self, h, len(variances))

    def cross_validate(self, prior, control, selector=None, refit_fp64=True):
        """Grid-search CV (fold x grid fits batched on this engine) and the final fit at
        the best variance. With FP32 bulk and refit_fp64, the final fit runs on an FP64
        engine over the same data."""
        out = self._cross_validate(prior, control, selector)
        if refit_fp64 and self.precision == _abi.FP32:
            e64 = Engine(self.data, _abi.FP64)
            try:
                p = create_prior(prior.prior_type, out["best_variance"], exclude=prior.exclude)
                out["beta"], out["result"] = e64.fit(p, control, mode="resident")
                out["refit_precision"] = _abi.FP64
            finally:
                e64.close()
        return out

    def _cross_validate(self, prior, control, selector=None):
        if selector is None:
            # R/ModelFit.R:174-188: "auto" -> byPid for conditional models, byRow otherwise
            selector = _abi.SELECT_BY_PID if self.data.model_type in (_abi.CONDITIONAL_LOGISTIC,) else \
                _abi.SELECT_BY_ROW
        hold = FitArgsHolder(self.data.J, prior, control, intercept_index=intercept_index(self.data))
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
        beta = np.zeros(self.data.J)
        res = _abi.FitResult()
        cvr = _abi.CvResult()
        check(lib().ccd_cross_validate(self.h, C.byref(hold.args), C.byref(cv), _d(grid), _d(pred), _d(beta),
                                       C.byref(res), C.byref(cvr)))
        return {"grid": grid, "pred": pred.reshape(control.grid_size, control.fold), "beta": beta,
                "result": res.as_dict(), "best_variance": cvr.best_variance, "best_value": cvr.best_value,
                "seconds": cvr.seconds, "n_fits": cvr.n_fits}

    # --- variance and profile likelihood ------------------------------------------
    def fisher_information(self, beta, indices):
        """Fisher information of columns `indices` at beta ([m][m]; unit row weights):
        cyclopsGetFisherInformation (RcppCyclopsInterface.cpp:176-200)."""
        b = np.ascontiguousarray(beta, dtype=np.float64)
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        m = len(idx)
        out = np.zeros((m, m))
        check(lib().ccd_fisher_information(self.h, _d(b), idx.ctypes.data_as(C.POINTER(C.c_int32)), m, _d(out)))
        return out

    def profile(self, prior, control, beta, indices, threshold, include_penalty=True,
                override_no_regularization=False):
        """Profile-likelihood limits of columns `indices` around the mode beta:
        cyclopsProfileModel (CcdInterface::profileModel). Returns (lower, upper, evaluations)."""
        hold = FitArgsHolder(self.data.J, prior, control, intercept_index=intercept_index(self.data))
        b = np.ascontiguousarray(beta, dtype=np.float64)
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        m = len(idx)
        lo, up = np.zeros(m), np.zeros(m)
        ev = np.zeros(m, dtype=np.int32)
        check(lib().ccd_profile(self.h, C.byref(hold.args), _d(b), idx.ctypes.data_as(C.POINTER(C.c_int32)), m,
                                float(threshold), int(bool(include_penalty)), int(bool(override_no_regularization)),
                                _d(lo), _d(up), ev.ctypes.data_as(C.POINTER(C.c_int32))))
        return lo, up, ev

    def profile_likelihood(self, prior, control, beta, index, points, include_penalty=True):
        """Profile log-likelihood of column `index` at `points`: cyclopsGetProfileLikelihood
        (CcdInterface::evaluateProfileModel), all points as one batched fit."""
        hold = FitArgsHolder(self.data.J, prior, control, intercept_index=intercept_index(self.data))
        b = np.ascontiguousarray(beta, dtype=np.float64)
        x = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros(len(x))
        check(lib().ccd_profile_likelihood(self.h, C.byref(hold.args), _d(b), int(index), _d(x), len(x),
                                           int(bool(include_penalty)), _d(out)))
        return out

    def kernel_stats(self):
        st = _abi.KernelStats()
        check(lib().ccd_get_kernel_stats(self.h, C.byref(st)))
        return {"launches": st.launches, "device_ms": st.device_ms, "algorithmic_bytes": st.algorithmic_bytes,
                "coordinates": st.coordinates}

    def reset_kernel_stats(self):
        check(lib().ccd_reset_kernel_stats(self.h))


class SteppedBatch:
    """A batch of fits advanced a few CCD cycles at a time (ccd_batch_step); end()
    returns what fit_batch would have returned for the same fits."""

    def __init__(self, engine, handle, n):
        self.engine = engine
        self.h = handle
        self.n = n

    def step(self, cycles=1):
        """Runs up to `cycles` more cycles; returns (fits still running, done flags [n])."""
        done = np.zeros(self.n, dtype=np.int32)
        running = C.c_int32()
        check(lib().ccd_batch_step(self.h, int(cycles), done.ctypes.data_as(C.POINTER(C.c_int32)),
                                   C.byref(running)))
        return running.value, done.astype(bool)

    def layout(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().ccd_batch_layout(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"fits_per_team": a.value, "workgroups_per_team": b.value, "teams": c.value}

    def end(self):
        betas = np.zeros((self.n, self.engine.data.J))
        results = (_abi.FitResult * self.n)()
        h, self.h = self.h, None
        check(lib().ccd_batch_end(h, _d(betas), results))
        return betas, [r.as_dict() for r in results]

    def abandon(self):
        if self.h:
            h, self.h = self.h, None
            check(lib().ccd_batch_end(h, None, None))

    def __del__(self):
        try:
            self.abandon()
        except Exception:
            pass


def fit_cyclops_model(data, prior=None, control=None, weights=None, fixed=None, beta0=None,
                      precision=_abi.FP64, mode="plain", device=0):
    """One fit on the device: fitCyclopsModel (R/ModelFit.R:65-293). With
    prior.use_cross_validation the prior variance is chosen by cross-validation first
    (:262-274, .cyclopsRunCrossValidation) and the result carries it as "variance"."""
    prior = prior or create_prior("none")
    control = control or create_control()
    eng = Engine(data, precision, device)
    try:
        if prior.use_cross_validation:
            selector = control.selector
            if selector == "auto":   # R/ModelFit.R:174-188
                selector = "byPid" if data.model_type == _abi.CONDITIONAL_LOGISTIC else "byRow"
            count = data.K if (selector == "byRow" or data.pid is None) else len(np.unique(data.pid))
            if control.min_cv_data > count:
                raise CyclopsError("Insufficient data count for cross validation")
            if weights is not None or fixed is not None or beta0 is not None:
                raise CyclopsError("cross-validation here takes no weights, fixed or starting coefficients")
            sel = _abi.SELECT_BY_PID if selector == "byPid" else _abi.SELECT_BY_ROW
            out = eng.cross_validate(prior, control, sel)
            res = dict(out["result"])
            res["variance"] = out["best_variance"]
            return out["beta"], res
        return eng.fit(prior, control, weights=weights, fixed=fixed, beta0=beta0, mode=mode)
    finally:
        eng.close()


def fit_multi(engines, priors, controls, weights=None, betas0=None):
    """One fit per engine (each its own dataset), all as shared persistent launches
    (ccd_fit_multi). Returns [(beta, result dict)] in engine order; each equals the
    engine's own resident fit."""
    n = len(engines)
    weights = weights or [None] * n
    holds = [FitArgsHolder(e.data.J, p, c, weights=w, intercept_index=intercept_index(e.data))
             for e, p, c, w in zip(engines, priors, controls, weights)]
    args = (_abi.FitArgs * n)(*[h.args for h in holds])
    betas = [np.zeros(e.data.J) if betas0 is None or betas0[i] is None else np.array(betas0[i], dtype=np.float64)
             for i, e in enumerate(engines)]
    ptrs = (C.POINTER(C.c_double) * n)(*[_d(b) for b in betas])
    hs = (C.c_void_p * n)(*[e.h for e in engines])
    res = (_abi.FitResult * n)()
    check(lib().ccd_fit_multi(hs, n, args, ptrs, res))
    return [(b, r.as_dict()) for b, r in zip(betas, res)]


def _covariates(data, covariates):
    """Column indices of the covariates R names by number or label (.checkCovariates);
    None means every covariate (the offset column is not one)."""
    first = 1 if data.has_offset else 0
    if covariates is None:
        return list(range(first, data.J))
    out = []
    for c in covariates:
        if isinstance(c, str):
            if not data.names or c not in data.names:
                raise CyclopsError(f"Variable {c} not found.")
            out.append(data.names.index(c))
        else:
            out.append(int(c))
    return out


def get_fisher_information(engine, beta, covariates=None):
    """.cyclopsGetFisherInformation (R/RcppExports.R:56-58) at the fitted beta."""
    return engine.fisher_information(beta, _covariates(engine.data, covariates))


def vcov(engine, beta):
    """vcov.cyclopsFit (R/ModelFit.R:884-894): the inverse Fisher information of every covariate."""
    return np.linalg.inv(get_fisher_information(engine, beta))


def get_ses(engine, beta, covariates=None):
    """getSEs (R/ModelFit.R:726-736): sqrt of the diagonal of the inverse Fisher information."""
    return np.sqrt(np.diag(np.linalg.inv(get_fisher_information(engine, beta, covariates))))


def confint(engine, prior, control, beta, parm=None, level=0.95, override_no_regularization=False,
            include_penalty=True):
    """confint.cyclopsFit (R/ModelFit.R:762-799): profile-likelihood limits at `level`
    (threshold qchisq(level, 1) / 2). Returns an array [m][3]: column index, lower, upper;
    NaN where a limit is not bracketed (R reports NA)."""
    from scipy.stats import chi2
    if level < 0.01 or level > 0.99:
        raise CyclopsError("level must be between 0 and 1")
    cols = _covariates(engine.data, parm)
    lo, up, _ = engine.profile(prior, control, beta, cols, chi2.ppf(level, 1) / 2.0, include_penalty,
                               override_no_regularization)
    return np.column_stack([np.array(cols, dtype=np.float64), lo, up])


def aconfint(engine, beta, parm=None, level=0.95):
    """aconfint (R/ModelFit.R:848-869): asymptotic limits beta +- qnorm((1 - level) / 2) * SE,
    the SEs from the inverse Fisher information of every covariate. Returns [m][3]:
    column index, lower, upper."""
    from scipy.stats import norm
    if level < 0.01 or level > 0.99:
        raise CyclopsError("level must be between 0 and 1")
    allc = _covariates(engine.data, None)
    ses = np.sqrt(np.diag(np.linalg.inv(engine.fisher_information(beta, allc))))
    cols = _covariates(engine.data, parm)
    se = np.array([ses[allc.index(c)] for c in cols])
    a = (1.0 - level) / 2.0
    b = np.asarray(beta, dtype=np.float64)[cols]
    return np.column_stack([np.array(cols, dtype=np.float64), b + se * norm.ppf(a), b + se * norm.ppf(1.0 - a)])


def predict(engine, beta):
    """predict.cyclopsFit on the data of the fit (R/Predict.R:38-46, CcdInterface::predictModel):
    getPredictiveEstimates at beta (LR: probability, PR: expected count per row)."""
    engine.compute_xbeta(beta)
    engine.remaining_stats()
    return engine.pred_est()


def get_cyclops_profile_log_likelihood(engine, prior, control, beta, parm, x, include_penalty=True):
    """getCyclopsProfileLogLikelihood (R/ModelFit.R:816-826): the profile log-likelihood
    of one covariate at each value of x."""
    col = _covariates(engine.data, [parm])[0]
    return engine.profile_likelihood(prior, control, beta, col, x, include_penalty)


def assign_folds(K, pid, fold, selector, seed):
    out = np.zeros(K, dtype=np.int32)
    p = None if pid is None else np.ascontiguousarray(pid, dtype=np.int32)
    check(lib().ccd_cv_assign_folds(K, _abi.ptr(p, C.c_int32), fold, selector, seed,
                                    out.ctypes.data_as(C.POINTER(C.c_int32))))
    return out


class Server:
    """Fit-serving loop (ccd_server_*): requests of any model and size are admitted
    by HBM budget and run concurrently on the device; poll() returns each result
    with its submit / start / end times."""

    def __init__(self, slots=4, hbm_budget_bytes=200e9, mode="resident", device=0, pack=0):
        cfg = _abi.ServerConfig()
        cfg.device = device
        cfg.slots = slots
        cfg.hbm_budget_bytes = float(hbm_budget_bytes)
        cfg.mode = {"plain": _abi.MODE_PLAIN, "resident": _abi.MODE_RESIDENT}[mode]
        cfg.pack = int(pack)
        h = C.c_void_p()
        check(lib().ccd_server_start(C.byref(cfg), C.byref(h)))
        self.h = h
        self._J = {}

    def submit(self, data, prior=None, control=None, precision=_abi.FP64, weights=None, fixed=None, beta0=None):
        prior = prior or create_prior("laplace", 1.0)
        control = control or create_control()
        desc = data.desc(precision)
        hold = FitArgsHolder(data.J, prior, control, weights=weights, fixed=fixed,
                             intercept_index=intercept_index(data))
        b0 = None if beta0 is None else np.ascontiguousarray(beta0, dtype=np.float64)
        rid = C.c_int64()
        check(lib().ccd_server_submit(self.h, C.byref(desc), C.byref(hold.args),
                                      None if b0 is None else _d(b0), C.byref(rid)))
        self._J[rid.value] = data.J
        return rid.value

    def poll(self, request_id, wait=True, release=True):
        """(beta, result, stats), or None while the request runs and wait is False.
        With release (the default) the server forgets the request once it is read."""
        J = self._J.get(request_id, 0)
        beta = np.zeros(J)
        res = _abi.FitResult()
        st = _abi.RequestStats()
        rc = lib().ccd_server_poll(self.h, request_id, int(wait), _d(beta) if J else None, C.byref(res),
                                   C.byref(st))
        if rc == _abi.ERR_STATE and not wait:
            return None
        if release and rc in (0, _abi.ERR_HIP, _abi.ERR_INVALID, _abi.ERR_NOMEM, _abi.ERR_UNSUPPORTED):
            err = lib().ccd_last_error().decode() if rc != 0 else ""
            lib().ccd_server_release(self.h, request_id)
            self._J.pop(request_id, None)
            if rc != 0:
                raise CyclopsError(f"ccd error {rc}: {err}")
        check(rc)
        return beta, res.as_dict(), {"submit": st.submit_time, "start": st.start_time, "end": st.end_time,
                                     "footprint_bytes": st.footprint_bytes}

    def release(self, request_id):
        """Forgets a finished request (ccd_server_release)."""
        check(lib().ccd_server_release(self.h, request_id))
        self._J.pop(request_id, None)

    def stop(self):
        if getattr(self, "h", None):
            check(lib().ccd_server_stop(self.h))
            self.h = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
