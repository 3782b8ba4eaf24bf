/*
 * cyclops_amd.h -- C ABI of the MI355X cyclic-coordinate-descent (CCD) engine.
 *
 * The engine replaces the per-coordinate hot path that the reference reaches
 * through the virtual interface bsccs::AbstractModelSpecifics
 * (src/cyclops/engine/AbstractModelSpecifics.h:36-122) and the driver loop
 * CyclicCoordinateDescent::findMode (src/cyclops/CyclicCoordinateDescent.cpp:909-1216).
 *
 * Conventions
 *   - Plain C: opaque handles, int status codes (CCD_OK == 0), no exceptions
 *     cross this boundary; ccd_last_error() returns the message of the last
 *     failing call on the calling thread.
 *   - Enumerations use the reference's numeric values (ModelType, FormatType,
 *     ConvergenceType, UpdateReturnFlags in src/cyclops/Types.h and
 *     src/cyclops/CompressedDataMatrix.h:36-38) so a caller can pass its own
 *     enum values through unchanged.
 *   - Caller buffers are borrowed only for the duration of a call; data is
 *     copied into HBM by ccd_create().
 *   - One handle owns one HIP stream; distinct handles may be used from
 *     distinct host threads, one handle must not be used concurrently.
 */
#ifndef CYCLOPS_AMD_H_
#define CYCLOPS_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CCD_OK                0
#define CCD_ERR_INVALID       1   /* bad argument / unsupported combination   */
#define CCD_ERR_HIP           2   /* HIP runtime error                         */
#define CCD_ERR_NOMEM         3   /* device allocation failed / over budget   */
#define CCD_ERR_STATE         4   /* call order violated                       */
#define CCD_ERR_UNSUPPORTED   5   /* operation the reference also rejects     */

/* ---- reference enumerations (src/cyclops/Types.h:134-147) -------------- */
#define CCD_MODEL_POISSON               2
#define CCD_MODEL_LOGISTIC              3
#define CCD_MODEL_CONDITIONAL_LOGISTIC  4
#define CCD_MODEL_COX                   8   /* Breslow-tied Cox (reference COX) */

/* FormatType (src/cyclops/CompressedDataMatrix.h:36-38) */
#define CCD_FORMAT_DENSE      0
#define CCD_FORMAT_SPARSE     1
#define CCD_FORMAT_INDICATOR  2
#define CCD_FORMAT_INTERCEPT  3

/* PriorType (src/cyclops/Types.h:100-105) */
#define CCD_PRIOR_NONE     0
#define CCD_PRIOR_LAPLACE  1
#define CCD_PRIOR_NORMAL   2

/* ConvergenceType (src/cyclops/Types.h:109-115) */
#define CCD_CONV_GRADIENT  0
#define CCD_CONV_LANGE     1
#define CCD_CONV_MITTAL    2

/* UpdateReturnFlags (src/cyclops/Types.h:124-130) */
#define CCD_SUCCESS         0
#define CCD_FAIL            1
#define CCD_MAX_ITERATIONS  2
#define CCD_ILLCONDITIONED  3

/* Precision of the bulk arithmetic: PrecisionType (src/cyclops/Types.h:155-160,
 * FP64 = 0, FP32 = 1). The bit counts 64 and 32 are accepted as aliases; FP16 (2)
 * is refused with CCD_ERR_INVALID. */
#define CCD_FP64       0
#define CCD_FP32       1
#define CCD_FP64_BITS  64
#define CCD_FP32_BITS  32

/* ---- data description ---------------------------------------------------
 * Column store in the reference's layout (CompressedDataColumn,
 * src/cyclops/CompressedDataMatrix.h:40-241): every column is DENSE, SPARSE,
 * INDICATOR or INTERCEPT. SPARSE and INDICATOR columns hold sorted int32 row
 * ids in rows[col_ptr[j] .. col_ptr[j+1]); SPARSE columns also hold values at
 * the same positions of `values`. DENSE columns hold K values at
 * dense_values[dense_ptr[j] .. dense_ptr[j]+K). INTERCEPT columns hold
 * nothing. y, time and pid are row vectors of length K.
 *   LR / PR : pid may be NULL (every row is its own patient, N = K).
 *   CLR     : pid = stratum id, rows sorted by (stratum, row)
 *             (R/NewDataConversion.R:138-149).
 *   COX     : pid = stratum id, time = survival time, rows sorted by
 *             (stratum, time descending, y ascending, row)
 *             (R/NewDataConversion.R:151-166).
 * has_offset = 1 means column 0 is the offset covariate with beta fixed at 1
 * (CyclicCoordinateDescent.cpp:218-221; ModelData.cpp:640-651).
 */
typedef struct ccd_data_desc {
    int32_t        model_type;
    int32_t        precision;       /* CCD_FP64 / CCD_FP32 (or 64 / 32) */
    int64_t        K;               /* rows */
    int32_t        J;               /* columns */
    int32_t        has_offset;
    const int32_t* format;          /* [J] */
    const int64_t* col_ptr;         /* [J+1] into rows/values */
    const int32_t* rows;            /* [col_ptr[J]] */
    const double*  values;          /* [col_ptr[J]] or NULL when no SPARSE column */
    const int64_t* dense_ptr;       /* [J+1] into dense_values, or NULL */
    const double*  dense_values;    /* or NULL */
    const double*  y;               /* [K] */
    const double*  time;            /* [K] (COX) or NULL */
    const int32_t* pid;             /* [K] or NULL */
} ccd_data_desc;

/* Arguments of one fit: createControl + createPrior
 * (R/ModelFit.R:489-618; ModeFindingArguments, CcdInterface.h:70-94). */
typedef struct ccd_fit_args {
    int32_t        prior_type;        /* CCD_PRIOR_* */
    double         variance;          /* prior variance (Laplace: lambda = sqrt(2/var)) */
    const uint8_t* exclude;           /* [J] 1 = column unpenalised (NoPrior), or NULL */
    const uint8_t* fixed;             /* [J] 1 = beta held fixed, or NULL */
    int32_t        convergence_type;  /* CCD_CONV_* */
    double         tolerance;         /* default 1e-6 */
    int32_t        max_iterations;    /* default 1000 */
    double         initial_bound;     /* default 2.0 */
    int32_t        max_bound_count;   /* default 5 */
    const double*  weights;           /* [K] row weights, or NULL (no weights) */
} ccd_fit_args;

typedef struct ccd_fit_result {
    int32_t return_flag;      /* CCD_SUCCESS / CCD_MAX_ITERATIONS / CCD_ILLCONDITIONED */
    int32_t iterations;       /* cycles of the last findMode */
    int32_t find_mode_calls;  /* findMode invocations inside update() (bound retries) */
    int32_t pad_;
    double  log_likelihood;
    double  log_prior;
    double  seconds;          /* wall time of the fit, host clock */
} ccd_fit_result;

typedef struct ccd_engine* ccd_handle;

/* ---- library ------------------------------------------------------------ */
const char* ccd_last_error(void);
const char* ccd_version(void);
int ccd_device_count(int* count);

/* ---- engine lifetime: AbstractModelSpecifics::factory / clone / dtor ----
 * factory: AbstractModelSpecifics.cpp:128-154 (returns nullptr for an
 * unsupported combination -> CCD_ERR_UNSUPPORTED here).
 * clone:   ModelSpecifics.hpp:98-101 (independent engine over the same data). */
int ccd_create(const ccd_data_desc* data, int device, ccd_handle* out);
int ccd_clone(ccd_handle h, ccd_handle* out);
int ccd_destroy(ccd_handle h);
int ccd_get_dims(ccd_handle h, int64_t* K, int32_t* J, int64_t* N);

/* ---- per-coordinate interface: one entry per AbstractModelSpecifics virtual.
 * use_weights mirrors the reference's useCrossValidation / useWeights flag. */

/* setWeights(double* w, double* cenWeights, bool useCV)  AbstractModelSpecifics.h:65 */
int ccd_set_weights(ccd_handle h, const double* weights, int use_weights);
/* computeFixedTermsInLogLikelihood + computeFixedTermsInGradientAndHessian
 * AbstractModelSpecifics.h:83-85 */
int ccd_fixed_terms(ccd_handle h, int use_weights);
/* CyclicCoordinateDescent::computeXBeta = zeroXBeta + axpyXBeta per column
 * (CyclicCoordinateDescent.cpp:1452-1470; ModelSpecifics.hpp:296-299,393-423) */
int ccd_compute_xbeta(ccd_handle h, const double* beta);
int ccd_zero_xbeta(ccd_handle h);
int ccd_axpy_xbeta(ccd_handle h, double beta, int j);
/* computeRemainingStatistics(bool)  AbstractModelSpecifics.h:81 */
int ccd_remaining_stats(ccd_handle h, int use_weights);
/* computeNumeratorForGradient(int, bool)  AbstractModelSpecifics.h:72 */
int ccd_numerator(ccd_handle h, int j, int use_weights);
/* computeGradientAndHessian(int, double*, double*, bool)  AbstractModelSpecifics.h:67 */
int ccd_grad_hess(ccd_handle h, int j, double* gradient, double* hessian, int use_weights);
/* ccd_grad_hess without allocation: the same gradient and Hessian, from the same
 * kernels and launch shape, with the per-block partial sums, the completion counter
 * and the result in a static device array of the library (CCD_NOALLOC_SCRATCH_BYTES
 * per device, held until the process ends) instead of the engine's buffers, copied
 * back through the engine's pinned result buffer. The call allocates no device or
 * host memory. No counterpart in the reference: for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid
 * allocation). Every call on a device shares that device's array: calls that
 * overlap on one device race, so run them on one host thread or serialise them. A
 * column pass that would need more scratch than the array holds is refused with
 * CCD_ERR_INVALID. */
#define CCD_NOALLOC_SCRATCH_BYTES 65536
int ccd_grad_hess_noalloc(ccd_handle h, int j, double* gradient, double* hessian, int use_weights);
/* updateXBeta(double, int, bool)  AbstractModelSpecifics.h:77 */
int ccd_update_xbeta(ccd_handle h, int j, double delta, int use_weights);
/* getLogLikelihood(bool)  AbstractModelSpecifics.h:87 */
int ccd_loglik(ccd_handle h, int use_weights, double* out);
/* getPredictiveLogLikelihood(double*)  AbstractModelSpecifics.h:89 */
int ccd_pred_loglik(ccd_handle h, const double* weights, double* out);
/* getPredictiveEstimates(double* y, double* w)  AbstractModelSpecifics.h:91 */
int ccd_pred_est(ccd_handle h, double* y_out, const double* weights);
/* getGradientObjective(bool)  AbstractModelSpecifics.h:93 */
int ccd_gradient_objective(ccd_handle h, int use_weights, double* out);
/* getXBeta / saveXBeta / getXBetaSave  AbstractModelSpecifics.h:114-118 */
int ccd_get_xbeta(ccd_handle h, double* out);
int ccd_save_xbeta(ccd_handle h);
int ccd_get_xbeta_save(ccd_handle h, double* out);

/* ---- whole fits ----------------------------------------------------------
 * CyclicCoordinateDescent::update -> findMode (CyclicCoordinateDescent.cpp:586-615,
 * 909-1216). beta is read as the starting point (the offset column is forced
 * to 1) and overwritten with the estimate.
 *   CCD_MODE_PLAIN    : host loop, one kernel launch per engine call, exactly
 *                       the reference's call pattern through the interface
 *                       above (the "plain per-fit GPU path").
 *   CCD_MODE_RESIDENT : the whole update() runs inside one persistent kernel;
 *                       g/h reductions, the prior step, bounds, convergence
 *                       test and Cox risk-set rescans never leave the device.
 */
#define CCD_MODE_PLAIN     0
#define CCD_MODE_RESIDENT  1
int ccd_fit(ccd_handle h, const ccd_fit_args* args, int mode,
            double* beta, ccd_fit_result* result);

/* ---- batched fits ----------------------------------------------------------
 * n_fits independent fits over the same data, advanced in lockstep by one
 * persistent launch (replaces the per-thread clones of
 * AbstractCrossValidationDriver::drive, AbstractCrossValidationDriver.cpp:44-129).
 * Fit f uses prior variance variances[f] and row weights
 *   w_f[k] = (row_fold[k] == held_out_fold[f]) ? 0 : 1     (held_out_fold[f] >= 0)
 *   w_f[k] = 1                                              (held_out_fold[f] <  0)
 * Everything else comes from `common` (its `weights` field must be NULL).
 * fixed_index[f] >= 0 holds that column of fit f at its starting value, as
 * CyclicCoordinateDescent::setFixedBeta (CyclicCoordinateDescent.cpp:384-386)
 * does for one fit (the profile-likelihood evaluations of CcdInterface.cpp:237-245).
 * betas: [n_fits][J], in (starting values) and out. */
typedef struct ccd_batch_args {
    int32_t        n_fits;
    const double*  variances;      /* [n_fits] */
    const int32_t* held_out_fold;  /* [n_fits] or NULL (= all -1) */
    const int32_t* row_fold;       /* [K] fold of each row, or NULL */
    const int32_t* fixed_index;    /* [n_fits] column held fixed in fit f, -1 for none; or NULL */
} ccd_batch_args;
int ccd_fit_batch(ccd_handle h, const ccd_fit_args* common, const ccd_batch_args* batch,
                  double* betas, ccd_fit_result* results);

/* ---- stepped batched fits --------------------------------------------------
 * The fits of ccd_fit_batch (same arguments, same results) advanced by the caller
 * a given number of CCD cycles per call, so that a caller can watch fits finish,
 * interleave its own work between cycles, or time single cycles. The cycle loop is
 * findMode's (CyclicCoordinateDescent.cpp:1008-1021, 1202-1207); no reference entry
 * point steps it from outside. All fits share one persistent run whose state stays
 * in HBM between steps; where the steps break changes no result.
 *   ccd_batch_begin : set-up (copies every argument; betas0 [n_fits][J] are the
 *                     starting values). While the batch is open, the engine runs no
 *                     other whole fit (CCD_ERR_STATE) and cannot be destroyed.
 *   ccd_batch_step  : up to `cycles` more cycles of every unfinished fit. done[f]
 *                     (optional, [n_fits]) becomes 1 once fit f has converged or hit
 *                     max_iterations; *running receives the fits still running.
 *   ccd_batch_end   : runs what remains, applies update()'s bound retries
 *                     (CyclicCoordinateDescent.cpp:599-614) and returns betas
 *                     [n_fits][J] and results, then frees the batch. betas and
 *                     results may both be NULL to abandon the batch.
 *   ccd_batch_layout: the team shape the batch runs with (fits per team, workgroups
 *                     per team, teams). */
typedef struct ccd_batch_run* ccd_batch_handle;
int ccd_batch_begin(ccd_handle h, const ccd_fit_args* common, const ccd_batch_args* batch,
                    const double* betas0, ccd_batch_handle* out);
int ccd_batch_step(ccd_batch_handle b, int32_t cycles, int32_t* done, int32_t* running);
int ccd_batch_end(ccd_batch_handle b, double* betas, ccd_fit_result* results);
int ccd_batch_layout(ccd_batch_handle b, int32_t* fits_per_team, int32_t* workgroups_per_team, int32_t* teams);

/* ---- fits of several datasets in one launch ---------------------------------
 * One fit per handle (its own data, args[i] and starting betas[i] [J_i], overwritten
 * with the estimate), as CyclicCoordinateDescent::update (CyclicCoordinateDescent.cpp:
 * 586-615) on each dataset, run as shared persistent launches: every fit keeps the
 * team shape of its own resident fit (so the results equal ccd_fit(..., RESIDENT)),
 * and the fits are packed into grids whose workgroups are all resident together.
 * The handles must be distinct engines of one model type and precision on one device;
 * ccd_fit_args.weights may differ per fit. This is the batching of uneven datasets
 * into one fused launch; a caller would use it in place of fitting the datasets one
 * after another. */
int ccd_fit_multi(const ccd_handle* handles, int32_t n, const ccd_fit_args* args, double* const* betas,
                  ccd_fit_result* results);

/* ---- variance and profile likelihood ---------------------------------------
 * Fisher information of the columns indices[0..m) at beta:
 * CyclicCoordinateDescent::computeFisherInformation (CyclicCoordinateDescent.cpp:1288-1300)
 * over ModelSpecifics::computeFisherInformation (ModelSpecifics.hpp:1204-1352), with
 * unit row weights (the reference refuses cross-validation weights, :1206-1207; the
 * engine's weights are reset to none). R's vcov / getSEs invert it
 * (R/ModelFit.R:732-733, 887-888). out: [m][m], row-major, symmetric. */
int ccd_fisher_information(ccd_handle h, const double* beta, const int32_t* indices, int32_t m, double* out);

/* Profile-likelihood confidence limits: CcdInterface::profileModel
 * (CcdInterface.cpp:262-434, its NEW_PROFILE branch) with OptimizationProfile (:223-260) and RZeroIn
 * (utils/RZeroIn.h:39-74, 183-288: bracketSignChange, then Brent's zeroin with tolerance 1e-3).
 * beta is the mode (the fit of `args`); for column indices[i], lower[i] / upper[i]
 * are where the log-likelihood maximised with that column held fixed (plus the
 * log prior when include_penalty) falls `threshold` below its value at the mode
 * (R's confint: threshold = qchisq(level, 1) / 2, R/ModelFit.R:772). NaN where no
 * sign change is bracketed in 21 steps. evaluations[i]: profile fits for column i.
 * Every bound search starts from the mode, as the reference's per-thread clones do
 * (CcdInterface.cpp:319-325), and warm-starts each fit from its previous one; all
 * searches advance together, each round of evaluations being one batched fit.
 * A profiled column that carries a prior is refused unless override_no_regularization
 * (:300-304). args->weights must be NULL. */
int ccd_profile(ccd_handle h, const ccd_fit_args* args, const double* beta, const int32_t* indices, int32_t m,
                double threshold, int include_penalty, int override_no_regularization,
                double* lower, double* upper, int32_t* evaluations);

/* Profile log-likelihood on a grid: CcdInterface::evaluateProfileModel
 * (CcdInterface.cpp:527-607), R's getCyclopsProfileLogLikelihood (R/ModelFit.R:816-826).
 * values[i] = log-likelihood (plus the log prior when include_penalty) of the fit
 * with column `index` held at points[i]. As the reference, the column is first held
 * at 0 starting from beta (:570); every point's fit then starts from that fit, as
 * the per-thread clones made after it do (:572-577). All points run as one batched
 * fit. args->weights must be NULL. */
int ccd_profile_likelihood(ccd_handle h, const ccd_fit_args* args, const double* beta, int32_t index,
                           const double* points, int32_t n, int include_penalty, double* values);

/* ---- cross-validation (grid search) --------------------------------------
 * CrossValidationSelector (drivers/CrossValidationSelector.cpp:19-189) with
 * GridSearchCrossValidationDriver (drivers/GridSearchCrossValidationDriver.cpp:38-44,
 * 96-138,234-244) and AbstractCrossValidationDriver::doCrossValidationStep /
 * resetForOptimal (AbstractCrossValidationDriver.cpp:31-42,131-248).
 * warm_start = 0 (resetCoefficients = TRUE): every fit starts from zero, so all
 *   fold x grid-point fits are independent and run as one batched launch.
 * warm_start = 1 (resetCoefficients = FALSE, R's default, one thread): the fits run
 *   in the reference's task order (grid point, then fold), each starting from the
 *   previous fit's coefficients, and a fit that fails restarts the chain from zero
 *   (:201-203, 224-226); one persistent launch per fit.
 * The final refit at the optimum starts from zero. */
#define CCD_SELECT_BY_ROW  0
#define CCD_SELECT_BY_PID  1
typedef struct ccd_cv_args {
    int32_t fold;             /* folds (default 10) */
    int32_t grid_size;        /* grid points (default 10) */
    double  lower_limit;      /* smallest variance */
    double  upper_limit;      /* largest variance */
    int32_t selector;         /* CCD_SELECT_BY_ROW / CCD_SELECT_BY_PID */
    int32_t warm_start;       /* 1 = resetCoefficients FALSE (chain), 0 = cold start (batched) */
    int64_t seed;             /* mt19937 seed of the fold permutation */
} ccd_cv_args;
typedef struct ccd_cv_result {
    double  best_variance;
    double  best_value;       /* max over grid of mean predictive log-likelihood */
    double  seconds;
    int32_t n_fits;
    int32_t pad_;
} ccd_cv_result;
/* grid_values: [grid_size] mean predictive log-likelihood per grid point
 * pred_loglik: [grid_size][fold] or NULL; beta: [J] final refit. */
int ccd_cross_validate(ccd_handle h, const ccd_fit_args* common, const ccd_cv_args* cv,
                       double* grid_values, double* pred_loglik, double* beta,
                       ccd_fit_result* final_result, ccd_cv_result* cv_result);
/* Fold assignment used by ccd_cross_validate: row_fold[k] in [0, fold). */
int ccd_cv_assign_folds(int64_t K, const int32_t* pid, int32_t fold, int32_t selector,
                        int64_t seed, int32_t* row_fold);

/* ---- serving ---------------------------------------------------------------
 * A request loop that keeps design matrices and iterate state resident in HBM
 * and admits fits of mixed model type and size while others run: a request is
 * admitted when a slot is free and its HBM footprint fits the budget; admitted
 * fits run concurrently on the device. */
typedef struct ccd_server* ccd_server_handle;
typedef struct ccd_server_config {
    int32_t device;
    int32_t slots;            /* concurrent fits on the device */
    double  hbm_budget_bytes; /* admission budget */
    int32_t mode;             /* CCD_MODE_RESIDENT (default) or CCD_MODE_PLAIN */
    int32_t pack;             /* 0 or 1: one request per launch. n > 1 (resident mode): up to n
                                 queued requests of one model and precision whose resident fit is
                                 one workgroup share a persistent launch (ccd_fit_multi); each
                                 result equals its own launch's, and the members finish together */
} ccd_server_config;
typedef struct ccd_request_stats {
    double submit_time;       /* seconds since server start */
    double start_time;        /* admitted to the device */
    double end_time;          /* result ready */
    double footprint_bytes;   /* HBM charged against the budget */
} ccd_request_stats;
int ccd_server_start(const ccd_server_config* cfg, ccd_server_handle* out);
/* data and args are copied before return; beta0 may be NULL (cold start). */
int ccd_server_submit(ccd_server_handle s, const ccd_data_desc* data, const ccd_fit_args* args,
                      const double* beta0, int64_t* request_id);
/* wait != 0 blocks until the request is done. Returns CCD_ERR_STATE while the
 * request is still running and wait == 0. beta_out: [J] or NULL. */
int ccd_server_poll(ccd_server_handle s, int64_t request_id, int wait, double* beta_out,
                    ccd_fit_result* result, ccd_request_stats* stats);
/* Forgets a finished request (its coefficients, result and timestamps); its id is
 * unknown afterwards. CCD_ERR_STATE while it still runs. A request's inputs are
 * freed when it finishes; a long-running server releases each result once read. */
int ccd_server_release(ccd_server_handle s, int64_t request_id);
int ccd_server_stop(ccd_server_handle s);

/* Estimated HBM bytes ccd_create() allocates for `data` (admission control). */
int ccd_estimate_footprint(const ccd_data_desc* data, double* bytes);

/* ---- timing hooks for measurement -----------------------------------------
 * Average device duration (HIP events on the handle's stream) of the last fit's
 * dominant kernel, and the algorithmic bytes it moved per launch. */
typedef struct ccd_kernel_stats {
    double  launches;         /* kernel launches (or persistent-kernel coordinate steps) */
    double  device_ms;        /* summed device time, HIP events */
    double  algorithmic_bytes;/* summed algorithmic bytes */
    double  coordinates;      /* coordinate updates performed */
} ccd_kernel_stats;
int ccd_get_kernel_stats(ccd_handle h, ccd_kernel_stats* stats);
int ccd_reset_kernel_stats(ccd_handle h);
/* Achievable HBM bandwidth of `device`: a copy kernel over two buffers of `bytes`
 * each (16-byte loads and stores), timed with HIP events over `reps` launches after
 * one untimed launch; *gbs = 2 * bytes * reps / seconds / 1e9 (read + write). No
 * reference counterpart: a measurement hook that puts kernel figures beside the
 * bandwidth the device delivers. */
int ccd_copy_bandwidth(int device, double bytes, int32_t reps, double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* CYCLOPS_AMD_H_ */
