"""ctypes mirror of the structs in include/cyclops_amd.h.

The numeric constants repeat the reference's enumerations (src/cyclops/Types.h,
src/cyclops/CompressedDataMatrix.h:36-38) so values can be passed through unchanged.
"""
import ctypes as C

CCD_OK = 0
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_UNSUPPORTED = 1, 2, 3, 4, 5

# ModelType (src/cyclops/Types.h:134-147)
POISSON = 2
LOGISTIC = 3
CONDITIONAL_LOGISTIC = 4
COX = 8
MODEL_NAMES = {"pr": POISSON, "lr": LOGISTIC, "clr": CONDITIONAL_LOGISTIC, "cox": COX}

# FormatType
DENSE, SPARSE, INDICATOR, INTERCEPT = 0, 1, 2, 3

# PriorType
PRIOR_NONE, PRIOR_LAPLACE, PRIOR_NORMAL = 0, 1, 2
PRIOR_NAMES = {"none": PRIOR_NONE, "laplace": PRIOR_LAPLACE, "normal": PRIOR_NORMAL}

# ConvergenceType
CONV_GRADIENT, CONV_LANGE, CONV_MITTAL = 0, 1, 2
CONV_NAMES = {"gradient": CONV_GRADIENT, "lange": CONV_LANGE, "mittal": CONV_MITTAL}

# UpdateReturnFlags
SUCCESS, FAIL, MAX_ITERATIONS, ILLCONDITIONED = 0, 1, 2, 3

# PrecisionType (src/cyclops/Types.h:155-160): FP64 = 0, FP32 = 1. The library also
# accepts the bit counts 64 and 32.
FP64, FP32 = 0, 1
FP64_BITS, FP32_BITS = 64, 32



def precision_of(p):
    """FP64 / FP32 for a PrecisionType value or a bit count (64 / 32)."""
    if p in (FP64, FP64_BITS):
        return FP64
    if p in (FP32, FP32_BITS):
        return FP32
    raise ValueError(f"precision must be FP64 ({FP64}) or FP32 ({FP32}), or 64 / 32; got {p!r}")


MODE_PLAIN, MODE_RESIDENT = 0, 1
SELECT_BY_ROW, SELECT_BY_PID = 0, 1


class DataDesc(C.Structure):
    _fields_ = [
        ("model_type", C.c_int32),
        ("precision", C.c_int32),
        ("K", C.c_int64),
        ("J", C.c_int32),
        ("has_offset", C.c_int32),
        ("format", C.POINTER(C.c_int32)),
        ("col_ptr", C.POINTER(C.c_int64)),
        ("rows", C.POINTER(C.c_int32)),
        ("values", C.POINTER(C.c_double)),
        ("dense_ptr", C.POINTER(C.c_int64)),
        ("dense_values", C.POINTER(C.c_double)),
        ("y", C.POINTER(C.c_double)),
        ("time", C.POINTER(C.c_double)),
        ("pid", C.POINTER(C.c_int32)),
    ]


class FitArgs(C.Structure):
    _fields_ = [
        ("prior_type", C.c_int32),
        ("variance", C.c_double),
        ("exclude", C.POINTER(C.c_uint8)),
        ("fixed", C.POINTER(C.c_uint8)),
        ("convergence_type", C.c_int32),
        ("tolerance", C.c_double),
        ("max_iterations", C.c_int32),
        ("initial_bound", C.c_double),
        ("max_bound_count", C.c_int32),
        ("weights", C.POINTER(C.c_double)),
    ]


class FitResult(C.Structure):
    _fields_ = [
        ("return_flag", C.c_int32),
        ("iterations", C.c_int32),
        ("find_mode_calls", C.c_int32),
        ("pad_", C.c_int32),
        ("log_likelihood", C.c_double),
        ("log_prior", C.c_double),
        ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {
            "return_flag": self.return_flag,
            "iterations": self.iterations,
            "find_mode_calls": self.find_mode_calls,
            "log_likelihood": self.log_likelihood,
            "log_prior": self.log_prior,
            "seconds": self.seconds,
        }


class BatchArgs(C.Structure):
    _fields_ = [
        ("n_fits", C.c_int32),
        ("variances", C.POINTER(C.c_double)),
        ("held_out_fold", C.POINTER(C.c_int32)),
        ("row_fold", C.POINTER(C.c_int32)),
        ("fixed_index", C.POINTER(C.c_int32)),
    ]


class CvArgs(C.Structure):
    _fields_ = [
        ("fold", C.c_int32),
        ("grid_size", C.c_int32),
        ("lower_limit", C.c_double),
        ("upper_limit", C.c_double),
        ("selector", C.c_int32),
        ("warm_start", C.c_int32),
        ("seed", C.c_int64),
    ]


class CvResult(C.Structure):
    _fields_ = [
        ("best_variance", C.c_double),
        ("best_value", C.c_double),
        ("seconds", C.c_double),
        ("n_fits", C.c_int32),
        ("pad_", C.c_int32),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("slots", C.c_int32),
        ("hbm_budget_bytes", C.c_double),
        ("mode", C.c_int32),
        ("pack", C.c_int32),
    ]


class RequestStats(C.Structure):
    _fields_ = [
        ("submit_time", C.c_double),
        ("start_time", C.c_double),
        ("end_time", C.c_double),
        ("footprint_bytes", C.c_double),
    ]


class KernelStats(C.Structure):
    _fields_ = [
        ("launches", C.c_double),
        ("device_ms", C.c_double),
        ("algorithmic_bytes", C.c_double),
        ("coordinates", C.c_double),
    ]


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if a is None:
        return C.POINTER(ctype)()
    return a.ctypes.data_as(C.POINTER(ctype))
