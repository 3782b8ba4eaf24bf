"""Prior and control objects mirroring the reference's R front end:
createPrior (R/ModelFit.R:574-618) and createControl (R/ModelFit.R:489-544).
They turn into the C ABI's ccd_fit_args."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi


@dataclass
class Prior:
    prior_type: str = "none"          # none | laplace | normal
    variance: float = 1.0
    exclude: Sequence[int] = ()       # column indices left unpenalised (NoPrior)
    use_cross_validation: bool = False


@dataclass
class Control:
    tolerance: float = 1e-6
    convergence_type: str = "gradient"   # gradient | lange | mittal
    max_iterations: int = 1000
    initial_bound: float = 2.0
    max_bound_count: int = 5
    fold: int = 10
    grid_size: int = 10
    lower_limit: float = 0.01
    upper_limit: float = 20.0
    seed: int = 1
    selector: str = "auto"               # auto | byRow | byPid
    reset_coefficients: bool = False     # R's default: cross-validation fits warm-start in a chain
    min_cv_data: int = 100               # fewer rows (byRow) or strata (byPid) refuse cross-validation


def create_prior(prior_type="none", variance=1.0, exclude=(), use_cross_validation=False) -> Prior:
    if prior_type not in _abi.PRIOR_NAMES:
        raise ValueError(f"unknown prior type {prior_type!r}")
    if prior_type == "none" and use_cross_validation:   # R/ModelFit.R:596-598
        raise ValueError("Cannot perform cross validation with a flat prior")
    return Prior(prior_type, float(variance), tuple(exclude), use_cross_validation)


def create_control(**kw) -> Control:
    c = Control(**kw)
    if c.convergence_type not in _abi.CONV_NAMES:
        raise ValueError(f"unknown convergence type {c.convergence_type!r}")
    return c


class FitArgsHolder:
    """Owns the numpy buffers a ccd_fit_args points to."""

    def __init__(self, J: int, prior: Prior, control: Control, weights=None, fixed=None,
                 intercept_index: Optional[int] = None):
        self.exclude = np.zeros(J, dtype=np.uint8)
        for j in prior.exclude:
            self.exclude[j] = 1
        # the intercept is never penalised (R/ModelFit.R:120-139)
        if intercept_index is not None and prior.prior_type != "none":
            self.exclude[intercept_index] = 1
        self.fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        self.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        a = _abi.FitArgs()
        a.prior_type = _abi.PRIOR_NAMES[prior.prior_type]
        a.variance = prior.variance
        a.exclude = _abi.ptr(self.exclude, C.c_uint8)
        a.fixed = _abi.ptr(self.fixed, C.c_uint8)
        a.convergence_type = _abi.CONV_NAMES[control.convergence_type]
        a.tolerance = control.tolerance
        a.max_iterations = control.max_iterations
        a.initial_bound = control.initial_bound
        a.max_bound_count = control.max_bound_count
        a.weights = _abi.ptr(self.weights, C.c_double)
        self.args = a


def intercept_index(data) -> Optional[int]:
    for j, f in enumerate(data.formats):
        if int(f) == _abi.INTERCEPT:
            return j
    return None
