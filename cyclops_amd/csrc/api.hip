// api.hip -- C ABI (include/cyclops_amd.h), the host CCD driver of the plain
// per-fit path, fold assignment and the cross-validation driver.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <vector>

struct ccd_engine {
    ccd::EngineBase* e = nullptr;
    ~ccd_engine() { delete e; }
};

struct ccd_batch_run {
    ccd_engine* h = nullptr;
    std::unique_ptr<ccd::SteppedBatch> b;
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return CCD_OK;
    } catch (const ccd::HipError& ex) {
        g_err = ex.what();
        return CCD_ERR_HIP;
    } catch (const std::invalid_argument& ex) {
        g_err = ex.what();
        return CCD_ERR_INVALID;
    } catch (const std::bad_alloc& ex) {
        g_err = "out of host memory";
        return CCD_ERR_NOMEM;
    } catch (const std::logic_error& ex) {
        g_err = ex.what();
        return CCD_ERR_UNSUPPORTED;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return CCD_ERR_STATE;
    }
}

ccd::EngineBase& E(ccd_handle h) {
    if (!h || !h->e) throw std::invalid_argument("null engine handle");
    return *h->e;
}
}  // namespace

namespace ccd {

void setLastError(const std::string& msg) { g_err = msg; }

// ---------------------------------------------------------------------------
// Priors (CovariatePrior.h:186-188 NoPrior, 240-273 LaplacePrior, 410-415 NormalPrior)
// ---------------------------------------------------------------------------
double priorDelta(int type, double variance, double g, double h, double beta) {
    if (type == CCD_PRIOR_NONE) return -(g / h);
    if (type == CCD_PRIOR_NORMAL) return -(g + beta / variance) / (h + 1.0 / variance);
    const double lambda = std::sqrt(2.0 / variance);
    const double negU = -(g - lambda) / h;
    const double posU = -(g + lambda) / h;
    const int sb = beta == 0 ? 0 : (beta < 0 ? -1 : 1);
    double delta;
    if (sb == 0) {
        if (negU < 0) delta = negU;
        else if (posU > 0) delta = posU;
        else delta = 0;
    } else {
        delta = (sb < 0) ? negU : posU;
        const double nb = beta + delta;
        const int sn = nb == 0 ? 0 : (nb < 0 ? -1 : 1);
        if (sn != sb) delta = -beta;
    }
    return delta;
}

double priorLogDensity(int type, double variance, double x) {
    if (type == CCD_PRIOR_NONE) return 0.0;
    if (type == CCD_PRIOR_LAPLACE) {
        const double l = std::sqrt(2.0 / variance);
        return std::log(0.5 * l) - l * std::abs(x);
    }
    return -0.5 * std::log(2.0 * M_PI * variance) - 0.5 * x * x / variance;
}

static void checkArgs(const EngineBase& eng, const ccd_fit_args& a) {
    if (a.prior_type < CCD_PRIOR_NONE || a.prior_type > CCD_PRIOR_NORMAL)
        throw std::invalid_argument("Unknown prior type");
    if (a.prior_type != CCD_PRIOR_NONE && !(a.variance > 0)) throw std::invalid_argument("prior variance must be > 0");
    if (a.convergence_type < CCD_CONV_GRADIENT || a.convergence_type > CCD_CONV_MITTAL)
        throw std::invalid_argument("Unknown convergence criterion");
    if (a.max_iterations < 1) throw std::invalid_argument("max_iterations must be >= 1");
    (void)eng;
}

// ---------------------------------------------------------------------------
// Host CCD driver over the engine: the plain per-fit path.
// CyclicCoordinateDescent::update (:586-615) / findMode (:909-1216) /
// ccdUpdateBeta (:1394-1413) / applyBounds (:1503-1520) /
// performCheckConvergence (:839-906).
// ---------------------------------------------------------------------------
void fitPlain(EngineBase& eng, const ccd_fit_args& a, double* betaIO, ccd_fit_result& r) {
    checkArgs(eng, a);
    const auto t0 = std::chrono::steady_clock::now();
    const int J = eng.J();
    std::vector<double> beta(betaIO, betaIO + J), hDelta(J);
    std::vector<char> fix(J, 0);
    std::vector<int> ptype(J, a.prior_type);
    for (int j = 0; j < J; ++j) {
        if (a.exclude && a.exclude[j]) ptype[j] = CCD_PRIOR_NONE;
        if (a.fixed && a.fixed[j]) fix[j] = 1;
    }
    if (eng.hasOffset()) { beta[0] = 1.0; fix[0] = 1; }
    const bool useW = a.weights != nullptr;
    auto logPrior = [&]() {
        double s = 0.0;
        for (int j = 0; j < J; ++j) s += priorLogDensity(ptype[j], a.variance, beta[j]);
        return s;
    };
    auto objective = [&](double* ll) -> double {
        if (a.convergence_type == CCD_CONV_GRADIENT) return eng.gradientObjective(useW);
        const double l = eng.logLik(useW);
        if (ll) *ll = l;
        if (a.convergence_type == CCD_CONV_MITTAL) return l;
        return l + logPrior();
    };
    // lazy set-up (findMode :923-941)
    eng.setWeights(a.weights, useW);
    eng.fixedTerms(useW);
    eng.computeXBeta(beta.data());
    eng.remainingStats(useW);

    double initialBound = a.initial_bound;
    int count = 0, flag = CCD_SUCCESS, iteration = 0;
    double lastLL = 0.0;
    bool llKnown = false;
    while (true) {
        std::fill(hDelta.begin(), hDelta.end(), initialBound);
        iteration = 0;
        double lastObj = objective(nullptr);
        bool done = false;
        while (!done) {
            for (int j = 0; j < J; ++j) {
                if (fix[j]) continue;
                eng.numerator(j, useW);
                double g, h;
                eng.gradHess(j, &g, &h, useW);
                if (h < 0.0) { g = 0.0; h = 0.0; }
                double delta = priorDelta(ptype[j], a.variance, g, h, beta[j]);
                if (delta < -hDelta[j]) delta = -hDelta[j];
                else if (delta > hDelta[j]) delta = hDelta[j];
                hDelta[j] = std::max(std::max(std::abs(delta) * 2, hDelta[j] / 2), 1E-3);
                if (delta != 0.0) {
                    beta[j] += delta;
                    eng.updateXBeta(delta, j, useW);
                    eng.counters.coordinates += 1;
                }
            }
            ++iteration;
            double ll = 0.0;
            const double thisObj = objective(&ll);
            llKnown = a.convergence_type != CCD_CONV_GRADIENT;
            lastLL = ll;
            double conv;
            bool ill = false;
            if (thisObj != thisObj) { conv = 0.0; ill = true; }
            else conv = std::abs(thisObj - lastObj) / (std::abs(thisObj) + 1.0);
            lastObj = thisObj;
            if (a.tolerance > 0 && conv < a.tolerance) { flag = ill ? CCD_ILLCONDITIONED : CCD_SUCCESS; done = true; }
            else if (iteration == a.max_iterations) { flag = CCD_MAX_ITERATIONS; done = true; }
        }
        ++count;
        if (flag == CCD_ILLCONDITIONED && count < a.max_bound_count) {
            initialBound /= 10.0;   // resetBeta (:256-264)
            for (int j = eng.hasOffset() ? 1 : 0; j < J; ++j) beta[j] = 0.0;
            eng.computeXBeta(beta.data());
            eng.remainingStats(useW);
        } else {
            break;
        }
    }
    std::copy(beta.begin(), beta.end(), betaIO);
    r.return_flag = flag;
    r.iterations = iteration;
    r.find_mode_calls = count;
    r.log_likelihood = llKnown ? lastLL : eng.logLik(useW);
    r.log_prior = logPrior();
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------------------
// CrossValidationSelector fold assignment: permutation of ids by std::shuffle with
// std::mt19937(seed) (CrossValidationSelector.cpp:45-52, 138-145; AbstractSelector.cpp),
// contiguous intervals of floor(n/fold) (+1 for the first n mod fold) (:30-42).
// ---------------------------------------------------------------------------
void assignFolds(int64_t K, const int32_t* pidIn, int fold, int selector, int64_t seed, int32_t* rowFold) {
    if (fold < 1) throw std::invalid_argument("fold must be >= 1");
    std::vector<int> ids(K);
    int64_t N;
    if (selector == CCD_SELECT_BY_PID) {
        if (!pidIn) throw std::invalid_argument("BY_PID selection needs pid");
        int cur = 0;
        int32_t last = pidIn[0];
        for (int64_t k = 0; k < K; ++k) {
            if (pidIn[k] != last) { ++cur; last = pidIn[k]; }
            ids[k] = cur;
        }
        N = cur + 1;
    } else {
        for (int64_t k = 0; k < K; ++k) ids[k] = (int)k;
        N = K;
    }
    std::vector<int64_t> start;
    int64_t index = 0;
    const int64_t fraction = N / fold, extra = N - fraction * fold;
    for (int i = 0; i < fold; ++i) {
        start.push_back(index);
        index += fraction;
        if (i < extra) ++index;
    }
    start.push_back(N);
    std::vector<int> perm(N);
    for (int64_t i = 0; i < N; ++i) perm[i] = (int)i;
    std::mt19937 prng;
    prng.seed((std::mt19937::result_type)seed);
    std::shuffle(perm.begin(), perm.end(), prng);
    std::vector<int32_t> foldOfId(N);
    for (int f = 0; f < fold; ++f)
        for (int64_t i = start[f]; i < start[f + 1]; ++i) foldOfId[perm[i]] = f;
    for (int64_t k = 0; k < K; ++k) rowFold[k] = foldOfId[ids[k]];
}

// ---------------------------------------------------------------------------
// Grid-search cross-validation: GridSearchCrossValidationDriver::doCrossValidationLoop
// (:96-138) with computeGridPoint (:38-44), AbstractCrossValidationDriver::
// doCrossValidationStep (:131-248), computePointEstimate (:250-261), findMax
// (:234-244) and resetForOptimal (:31-42) followed by the final fit
// (CcdInterface.cpp:777-789). Cold start (resetCoefficients = TRUE): the fold x
// grid-point fits are independent and run as one batched launch. Warm start (R's
// default): one chain in the reference's task order, each fit a persistent launch
// starting from the previous fit's coefficients.
// ---------------------------------------------------------------------------
static void crossValidate(EngineBase& eng, const ccd_fit_args& a, const ccd_cv_args& cv, double* gridValues,
                          double* predLL, double* betaOut, ccd_fit_result& fin, ccd_cv_result& out) {
    checkArgs(eng, a);
    if (cv.fold < 1 || cv.grid_size < 1) throw std::invalid_argument("fold and grid_size must be >= 1");
    if (cv.selector == CCD_SELECT_BY_PID && !(eng.model() == CCD_MODEL_CONDITIONAL_LOGISTIC || eng.model() == CCD_MODEL_COX))
        throw std::invalid_argument("BY_PID selection needs strata");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t K = eng.K();
    const int J = eng.J();
    std::vector<int32_t> rowFold(K);
    const std::vector<int32_t>* strata = nullptr;
    std::vector<int32_t> pidHost;
    if (cv.selector == CCD_SELECT_BY_PID) {
        pidHost.resize(K);
        eng.getStrata(pidHost.data());
        strata = &pidHost;
    }
    assignFolds(K, strata ? strata->data() : nullptr, cv.fold, cv.selector, cv.seed, rowFold.data());
    std::vector<double> points(cv.grid_size);
    for (int step = 0; step < cv.grid_size; ++step) {
        if (cv.grid_size == 1) points[step] = cv.upper_limit;
        else {
            const double stepSize = (std::log(cv.upper_limit) - std::log(cv.lower_limit)) / (cv.grid_size - 1);
            points[step] = std::exp(std::log(cv.lower_limit) + step * stepSize);
        }
    }
    const int nFits = cv.grid_size * cv.fold;
    if (cv.warm_start) {
        std::vector<double> beta(J, 0.0), w(K), comp(K);
        if (eng.hasOffset()) beta[0] = 1.0;
        double bestValue = -std::numeric_limits<double>::infinity(), bestPoint = points[0];
        for (int step = 0; step < cv.grid_size; ++step) {
            double total = 0.0;
            int count = 0;
            for (int f = 0; f < cv.fold; ++f) {
                for (int64_t k = 0; k < K; ++k) {
                    w[k] = rowFold[k] == f ? 0.0 : 1.0;
                    comp[k] = 1.0 - w[k];
                }
                ccd_fit_args fa = a;
                fa.weights = w.data();
                fa.variance = points[step];
                ccd_fit_result r{};
                eng.fitResident(fa, beta.data(), r);   // starts from the previous fit's beta
                double v = std::numeric_limits<double>::quiet_NaN();
                if (r.return_flag == CCD_SUCCESS) {
                    eng.setWeights(w.data(), true);
                    eng.computeXBeta(beta.data());
                    eng.remainingStats(true);
                    v = eng.predLogLik(comp.data());
                } else {   // cold start for stability (:224-226)
                    std::fill(beta.begin(), beta.end(), 0.0);
                    if (eng.hasOffset()) beta[0] = 1.0;
                }
                if (predLL) predLL[step * cv.fold + f] = v;
                if (v == v) { total += v; ++count; }
            }
            const double value = total / (double)count;
            gridValues[step] = value;
            if (step == 0 || value > bestValue) { bestValue = value; bestPoint = points[step]; }
        }
        std::fill(beta.begin(), beta.end(), 0.0);   // resetForOptimal: cold start
        if (eng.hasOffset()) beta[0] = 1.0;
        ccd_fit_args fa = a;
        fa.weights = nullptr;
        fa.variance = bestPoint;
        eng.fitResident(fa, beta.data(), fin);
        std::copy(beta.begin(), beta.end(), betaOut);
        out.best_variance = bestPoint;
        out.best_value = bestValue;
        out.n_fits = nFits;
        out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return;
    }
    std::vector<double> var(nFits), betas((size_t)nFits * J, 0.0);
    std::vector<int32_t> held(nFits);
    for (int step = 0; step < cv.grid_size; ++step)
        for (int f = 0; f < cv.fold; ++f) {
            var[step * cv.fold + f] = points[step];
            held[step * cv.fold + f] = f;
        }
    if (eng.hasOffset())
        for (int i = 0; i < nFits; ++i) betas[(size_t)i * J] = 1.0;
    std::vector<ccd_fit_result> res(nFits);
    ccd_batch_args b{nFits, var.data(), held.data(), rowFold.data(), nullptr};
    ccd_fit_args common = a;
    common.weights = nullptr;
    // every fit's predictive log-likelihood on its held-out rows is formed on the device
    // right after its run, from the xb the run leaves in HBM
    // (CyclicCoordinateDescent::getNewPredictiveLogLikelihood, :319-355)
    std::vector<double> pred(nFits, 0.0);
    eng.fitBatch(common, b, betas.data(), res.data(), true, pred.data());
    double bestValue = -std::numeric_limits<double>::infinity(), bestPoint = points[0];
    for (int step = 0; step < cv.grid_size; ++step) {
        double total = 0.0;
        int count = 0;
        for (int f = 0; f < cv.fold; ++f) {
            const int i = step * cv.fold + f;
            double v = std::numeric_limits<double>::quiet_NaN();
            if (res[i].return_flag == CCD_SUCCESS) v = pred[i];
            if (predLL) predLL[i] = v;
            if (v == v) { total += v; ++count; }
        }
        const double value = total / (double)count;
        gridValues[step] = value;
        if (step == 0 || value > bestValue) { bestValue = value; bestPoint = points[step]; }
    }
    // resetForOptimal + final fit on all rows
    std::vector<double> beta(J, 0.0);
    if (eng.hasOffset()) beta[0] = 1.0;
    ccd_fit_args fa = a;
    fa.weights = nullptr;
    fa.variance = bestPoint;
    eng.fitResident(fa, beta.data(), fin);
    std::copy(beta.begin(), beta.end(), betaOut);
    out.best_variance = bestPoint;
    out.best_value = bestValue;
    out.n_fits = nFits;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace ccd

// Plain copy: 16-byte loads and stores, grid-stride (bandwidth measurement).
__global__ __launch_bounds__(256) void kCopy16(const float4* __restrict__ a, float4* __restrict__ b, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* ccd_last_error(void) { return g_err.c_str(); }
const char* ccd_version(void) { return "cyclops_amd 0.1.0 (gfx950)"; }

int ccd_device_count(int* count) {
    return guard([&] { CCD_HIP(hipGetDeviceCount(count)); });
}

int ccd_create(const ccd_data_desc* data, int device, ccd_handle* out) {
    return guard([&] {
        if (!data || !out) throw std::invalid_argument("null argument");
        int n = 0;
        CCD_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw std::invalid_argument("no such device");
        auto* h = new ccd_engine();
        try {
            h->e = ccd::createEngine(*data, device);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int ccd_clone(ccd_handle h, ccd_handle* out) {
    return guard([&] {
        auto* c = new ccd_engine();
        c->e = E(h).clone();
        *out = c;
    });
}

int ccd_destroy(ccd_handle h) {
    return guard([&] {
        if (h && h->e && h->e->steppedOpen())
            throw std::runtime_error("a stepped batch is open on this engine: end it first");
        delete h;
    });
}

int ccd_get_dims(ccd_handle h, int64_t* K, int32_t* J, int64_t* N) {
    return guard([&] {
        auto& e = E(h);
        if (K) *K = e.K();
        if (J) *J = e.J();
        if (N) *N = e.N();
    });
}

int ccd_set_weights(ccd_handle h, const double* w, int use) {
    return guard([&] { E(h).setWeights(w, use != 0 && w != nullptr); });
}
int ccd_fixed_terms(ccd_handle h, int use) {
    return guard([&] { E(h).fixedTerms(use != 0); });
}
int ccd_compute_xbeta(ccd_handle h, const double* beta) {
    return guard([&] { E(h).computeXBeta(beta); });
}
int ccd_zero_xbeta(ccd_handle h) {
    return guard([&] { E(h).zeroXBeta(); });
}
int ccd_axpy_xbeta(ccd_handle h, double beta, int j) {
    return guard([&] {
        if (j < 0 || j >= E(h).J()) throw std::invalid_argument("column out of range");
        E(h).axpyXBeta(beta, j);
    });
}
int ccd_remaining_stats(ccd_handle h, int use) {
    return guard([&] { E(h).remainingStats(use != 0); });
}
int ccd_numerator(ccd_handle h, int j, int use) {
    return guard([&] {
        if (j < 0 || j >= E(h).J()) throw std::invalid_argument("column out of range");
        E(h).numerator(j, use != 0);
    });
}
int ccd_grad_hess(ccd_handle h, int j, double* g, double* hh, int use) {
    return guard([&] {
        if (j < 0 || j >= E(h).J()) throw std::invalid_argument("column out of range");
        E(h).gradHess(j, g, hh, use != 0);
    });
}
int ccd_grad_hess_noalloc(ccd_handle h, int j, double* g, double* hh, int use) {
    return guard([&] {
        if (!g || !hh) throw std::invalid_argument("null argument");
        if (j < 0 || j >= E(h).J()) throw std::invalid_argument("column out of range");
        E(h).gradHessNoAlloc(j, g, hh, use != 0);
    });
}
int ccd_update_xbeta(ccd_handle h, int j, double delta, int use) {
    return guard([&] {
        if (j < 0 || j >= E(h).J()) throw std::invalid_argument("column out of range");
        E(h).updateXBeta(delta, j, use != 0);
    });
}
int ccd_loglik(ccd_handle h, int use, double* out) {
    return guard([&] { *out = E(h).logLik(use != 0); });
}
int ccd_pred_loglik(ccd_handle h, const double* w, double* out) {
    return guard([&] {
        if (!w) throw std::invalid_argument("weights required");
        *out = E(h).predLogLik(w);
    });
}
int ccd_pred_est(ccd_handle h, double* y, const double* w) {
    return guard([&] { E(h).predEstimates(y, w); });
}
int ccd_gradient_objective(ccd_handle h, int use, double* out) {
    return guard([&] { *out = E(h).gradientObjective(use != 0); });
}
int ccd_get_xbeta(ccd_handle h, double* out) {
    return guard([&] { E(h).getXBeta(out); });
}
int ccd_save_xbeta(ccd_handle h) {
    return guard([&] { E(h).saveXBeta(); });
}
int ccd_get_xbeta_save(ccd_handle h, double* out) {
    return guard([&] { E(h).getXBetaSave(out); });
}

int ccd_fit(ccd_handle h, const ccd_fit_args* args, int mode, double* beta, ccd_fit_result* result) {
    return guard([&] {
        if (!args || !beta || !result) throw std::invalid_argument("null argument");
        auto& e = E(h);
        std::memset(result, 0, sizeof(*result));
        if (mode == CCD_MODE_PLAIN) ccd::fitPlain(e, *args, beta, *result);
        else if (mode == CCD_MODE_RESIDENT) {
            ccd::checkArgs(e, *args);
            e.fitResident(*args, beta, *result);
        } else throw std::invalid_argument("unknown mode");
    });
}

int ccd_fit_batch(ccd_handle h, const ccd_fit_args* common, const ccd_batch_args* batch, double* betas,
                  ccd_fit_result* results) {
    return guard([&] {
        if (!common || !batch || !betas || !results) throw std::invalid_argument("null argument");
        if (common->weights) throw std::invalid_argument("batched fits take weights from row_fold");
        if (batch->n_fits < 1) throw std::invalid_argument("n_fits must be >= 1");
        ccd::checkArgs(E(h), *common);
        E(h).fitBatch(*common, *batch, betas, results);
    });
}

int ccd_batch_begin(ccd_handle h, const ccd_fit_args* common, const ccd_batch_args* batch, const double* betas0,
                    ccd_batch_handle* out) {
    return guard([&] {
        if (!common || !batch || !betas0 || !out) throw std::invalid_argument("null argument");
        if (common->weights) throw std::invalid_argument("batched fits take weights from row_fold");
        if (batch->n_fits < 1) throw std::invalid_argument("n_fits must be >= 1");
        ccd::checkArgs(E(h), *common);
        std::unique_ptr<ccd_batch_run> r(new ccd_batch_run());
        r->h = h;
        r->b.reset(E(h).beginStepped(*common, *batch, betas0));
        *out = r.release();
    });
}

int ccd_batch_step(ccd_batch_handle b, int32_t cycles, int32_t* done, int32_t* running) {
    return guard([&] {
        if (!b || !b->b) throw std::invalid_argument("null batch handle");
        const int n = b->b->step(cycles, done);
        if (running) *running = n;
    });
}

int ccd_batch_end(ccd_batch_handle b, double* betas, ccd_fit_result* results) {
    return guard([&] {
        if (!b) throw std::invalid_argument("null batch handle");
        std::unique_ptr<ccd_batch_run> own(b);
        if (!betas && !results) return;   // abandoned: the destructor releases the engine
        if (!betas || !results) throw std::invalid_argument("betas and results go together");
        std::memset(results, 0, sizeof(ccd_fit_result) * (size_t)own->b->fits());
        own->b->finish(betas, results);
    });
}

int ccd_batch_layout(ccd_batch_handle b, int32_t* fits_per_team, int32_t* workgroups_per_team, int32_t* teams) {
    return guard([&] {
        if (!b || !b->b || !fits_per_team || !workgroups_per_team || !teams) throw std::invalid_argument("null argument");
        b->b->layout(fits_per_team, workgroups_per_team, teams);
    });
}

int ccd_fit_multi(const ccd_handle* handles, int32_t n, const ccd_fit_args* args, double* const* betas,
                  ccd_fit_result* results) {
    return guard([&] {
        if (!handles || !args || !betas || !results) throw std::invalid_argument("null argument");
        if (n < 1) throw std::invalid_argument("n must be >= 1");
        std::vector<ccd::EngineBase*> e(n);
        for (int i = 0; i < n; ++i) {
            if (!betas[i]) throw std::invalid_argument("null argument");
            e[i] = &E(handles[i]);
            ccd::checkArgs(*e[i], args[i]);
            std::memset(&results[i], 0, sizeof(results[i]));
        }
        ccd::fitMulti(e.data(), n, args, betas, results);
    });
}

int ccd_fisher_information(ccd_handle h, const double* beta, const int32_t* indices, int32_t m, double* out) {
    return guard([&] {
        if (!beta || (m > 0 && (!indices || !out))) throw std::invalid_argument("null argument");
        if (m < 0) throw std::invalid_argument("m must be >= 0");
        E(h).fisherInformation(beta, indices, m, out);
    });
}

int ccd_profile(ccd_handle h, const ccd_fit_args* args, const double* beta, const int32_t* indices, int32_t m,
                double threshold, int include_penalty, int override_no_regularization, double* lower,
                double* upper, int32_t* evaluations) {
    return guard([&] {
        if (!args || !beta || (m > 0 && (!indices || !lower || !upper || !evaluations)))
            throw std::invalid_argument("null argument");
        if (m < 0) throw std::invalid_argument("m must be >= 0");
        if (!(threshold > 0.0)) throw std::invalid_argument("threshold must be > 0");
        ccd::checkArgs(E(h), *args);
        ccd::profileModel(E(h), *args, beta, indices, m, threshold, include_penalty != 0,
                          override_no_regularization != 0, lower, upper, evaluations);
    });
}

int ccd_profile_likelihood(ccd_handle h, const ccd_fit_args* args, const double* beta, int32_t index,
                           const double* points, int32_t n, int include_penalty, double* values) {
    return guard([&] {
        if (!args || !beta || (n > 0 && (!points || !values))) throw std::invalid_argument("null argument");
        if (n < 0) throw std::invalid_argument("n must be >= 0");
        ccd::checkArgs(E(h), *args);
        ccd::profileLikelihood(E(h), *args, beta, index, points, n, include_penalty != 0, values);
    });
}

int ccd_cross_validate(ccd_handle h, const ccd_fit_args* common, const ccd_cv_args* cv, double* grid_values,
                       double* pred_loglik, double* beta, ccd_fit_result* final_result, ccd_cv_result* cv_result) {
    return guard([&] {
        if (!common || !cv || !grid_values || !beta || !final_result || !cv_result)
            throw std::invalid_argument("null argument");
        std::memset(final_result, 0, sizeof(*final_result));
        std::memset(cv_result, 0, sizeof(*cv_result));
        ccd::crossValidate(E(h), *common, *cv, grid_values, pred_loglik, beta, *final_result, *cv_result);
    });
}

int ccd_cv_assign_folds(int64_t K, const int32_t* pid, int32_t fold, int32_t selector, int64_t seed,
                        int32_t* rowFold) {
    return guard([&] {
        if (K < 1 || !rowFold) throw std::invalid_argument("bad arguments");
        ccd::assignFolds(K, pid, fold, selector, seed, rowFold);
    });
}

int ccd_get_kernel_stats(ccd_handle h, ccd_kernel_stats* st) {
    return guard([&] {
        auto& c = E(h).counters;
        st->launches = c.launches;
        st->device_ms = c.deviceMs;
        st->algorithmic_bytes = c.bytes;
        st->coordinates = c.coordinates;
    });
}
int ccd_reset_kernel_stats(ccd_handle h) {
    return guard([&] {
        E(h).counters = ccd::KernelCounters();
        E(h).timing = true;
    });
}

int ccd_copy_bandwidth(int device, double bytes, int32_t reps, double* gbs) {
    return guard([&] {
        if (!gbs || !(bytes >= 16.0) || reps < 1) throw std::invalid_argument("bad arguments");
        CCD_HIP(hipSetDevice(device));
        const int64_t n = (int64_t)(bytes / 16.0);
        float4 *a = nullptr, *b = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&]() {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
        };
        try {
            CCD_HIP(hipMalloc(&a, (size_t)n * 16));
            CCD_HIP(hipMalloc(&b, (size_t)n * 16));
            CCD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            CCD_HIP(hipEventCreate(&e0));
            CCD_HIP(hipEventCreate(&e1));
            CCD_HIP(hipMemsetAsync(a, 0, (size_t)n * 16, s));
            const int blocks = (int)std::min<int64_t>(8192, (n + 255) / 256);
            kCopy16<<<blocks, 256, 0, s>>>(a, b, n);
            CCD_HIP(hipGetLastError());
            CCD_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) kCopy16<<<blocks, 256, 0, s>>>(a, b, n);
            CCD_HIP(hipGetLastError());
            CCD_HIP(hipEventRecord(e1, s));
            CCD_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CCD_HIP(hipEventElapsedTime(&ms, e0, e1));
            *gbs = 2.0 * (double)n * 16.0 * reps / ((double)ms * 1e-3) / 1e9;
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int ccd_estimate_footprint(const ccd_data_desc* d, double* bytes) {
    return guard([&] {
        if (!d || !bytes) throw std::invalid_argument("null argument");
        const double s = ccd::precisionOf(d->precision) == CCD_FP32 ? 4.0 : 8.0;
        const double nnz = (double)d->col_ptr[d->J];
        double nDense = 0;
        bool sparse = false;
        for (int j = 0; j < d->J; ++j) {
            if (d->format[j] == CCD_FORMAT_DENSE) nDense += 1;
            if (d->format[j] == CCD_FORMAT_SPARSE) sparse = true;
        }
        double b = nnz * 4.0 + (sparse ? nnz * s : 0.0) + nDense * d->K * s;   // design
        b += d->K * (s + 4.0 + 4.0);                                           // y, strata
        b += d->K * (4.0 * s + 8.0);                                           // xb, save, exp, w, pred w
        b += (d->K + 1) * (5.0 * 8.0 + 12.0);                                  // per-pid state
        b += d->J * 24.0 + (1 << 20);
        *bytes = b;
    });
}

}  // extern "C"
