// profile.hip -- profile-likelihood confidence limits over batched fits.
//
// CcdInterface::profileModel (CcdInterface.cpp:262-434, the NEW_PROFILE branch): for
// each profiled column two bound searches (directions +1 and -1), each of which
// brackets a sign change of
//     y(x) = logLik(fit with beta_col held at x) [+ logPrior] + threshold - mode
// (OptimizationProfile::objective, :238-249) and then finds its root. The reference
// runs the searches one after another on per-thread clones; here every search is a
// state machine and all of them advance together: each round gathers the point every
// unfinished search wants evaluated and runs them as ONE batched persistent fit
// (ccd_fit_batch with a per-fit fixed column), warm-started from that search's
// previous evaluation. Every search begins at the mode, as a clone does (clones copy
// beta, CyclicCoordinateDescent.cpp:115-119; findMode resets the bounds, :943).
#include "engine.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace ccd {
namespace {

// One bound search: RZeroIn::bracketSignChange (utils/RZeroIn.h:50-74: displacement
// max(|x0|, 0.01), multiplier 0.1 doubling per step, NaN after 21 steps) followed by
// RZeroIn::getRoot -> R_zeroin2 (:39-48, 183-288), Brent's root finder (R. P. Brent,
// "Algorithms for Minimization without Derivatives", 1973, ch. 4; the zeroin of
// Forsythe, Malcolm and Moler, 1977) with tolerance 1e-3 and 1000 iterations. The
// steps are the published ones, in the order the reference takes them, so each search
// asks for the same points as the reference's. `point` is the next abscissa to
// evaluate while the search is not done; take() consumes its objective value.
struct BoundSearch {
    int out = 0;        // result slot: 2 * column position + (direction < 0)
    int col = 0;
    double dir = 1.0, x0 = 0.0, obj0 = 0.0;
    enum Phase { BRACKET, ROOT, DONE } phase = BRACKET;
    // bracketing
    double disp = 0.0, mult = 0.1, x1 = 0.0, o1 = 0.0;
    int steps = 0;
    // root finding: b is the best estimate, c brackets the root with b, a is the previous b
    double a = 0.0, b = 0.0, c = 0.0, fa = 0.0, fb = 0.0, fc = 0.0;
    int left = 0;
    static constexpr double kTol = 1e-3;
    static constexpr int kMaxIt = 1000;
    double point = 0.0;
    double result = std::nan("");
    int evals = 0;
    std::vector<double> beta;   // starting values of the next evaluation

    void start() {
        disp = std::max(std::fabs(x0), 0.01);
        mult = 0.1;
        x1 = x0;
        o1 = obj0;
        steps = 0;
        phase = BRACKET;
        bracketNext();
    }
    void finish(double r) {
        result = r;
        phase = DONE;
    }
    void bracketNext() {
        if (o1 > 0) {
            point = x0 + dir * disp * mult;
            return;
        }
        rootStart();
    }
    void bracketTake(double y) {
        x1 = point;
        o1 = y;
        if (std::isnan(o1) || steps > 20) {
            finish(std::nan(""));
            return;
        }
        ++steps;
        mult *= 2.0;
        bracketNext();
    }
    void rootStart() {
        double xl = x0, xu = x1, fl = obj0, fu = o1;
        if (xu < xl) {
            std::swap(xl, xu);
            std::swap(fl, fu);
        }
        a = xl;
        b = xu;
        fa = fl;
        fb = fu;
        c = a;
        fc = fa;
        if (fa == 0.0) { finish(a); return; }
        if (fb == 0.0) { finish(b); return; }
        left = kMaxIt + 1;
        phase = ROOT;
        rootNext();
    }
    void rootNext() {
        if (left == 0) {   // iterations exhausted: the best estimate stands
            finish(b);
            return;
        }
        --left;
        const double prev = b - a;
        if (std::fabs(fc) < std::fabs(fb)) {   // b must be the better end
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        const double tolAct = 2.0 * DBL_EPSILON * std::fabs(b) + kTol / 2.0;
        double step = (c - b) / 2.0;
        if (std::fabs(step) <= tolAct || fb == 0.0) {
            finish(b);
            return;
        }
        if (std::fabs(prev) >= tolAct && std::fabs(fa) > std::fabs(fb)) {
            // inverse interpolation: secant through (a, b) when a == c, else
            // inverse quadratic through a, b, c; accepted only inside [b, c] and
            // shorter than half the step before last
            const double cb = c - b;
            double p, q;
            if (a == c) {
                const double s = fb / fa;
                p = cb * s;
                q = 1.0 - s;
            } else {
                const double qa = fa / fc, r = fb / fc, s = fb / fa;
                p = s * (cb * qa * (qa - r) - (b - a) * (r - 1.0));
                q = (qa - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            else p = -p;
            if (p < 0.75 * cb * q - std::fabs(tolAct * q) / 2.0 && p < std::fabs(prev * q / 2.0)) step = p / q;
        }
        if (std::fabs(step) < tolAct) step = step > 0.0 ? tolAct : -tolAct;
        a = b;
        fa = fb;
        point = b + step;
    }
    void rootTake(double y) {
        b = point;
        fb = y;
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
            c = a;
            fc = fa;
        }
        rootNext();
    }
    void take(double y) {
        ++evals;
        if (phase == BRACKET) bracketTake(y);
        else rootTake(y);
    }
};

}  // namespace

void profileModel(EngineBase& eng, const ccd_fit_args& a, const double* beta, const int32_t* idx, int m,
                  double threshold, bool includePenalty, bool overrideNoRegularization, double* lower,
                  double* upper, int32_t* evaluations) {
    if (a.weights) throw std::invalid_argument("profiling takes no row weights");
    const int J = eng.J();
    for (int i = 0; i < m; ++i) {
        const int j = idx[i];
        if (j < 0 || j >= J) throw std::invalid_argument("Variable " + std::to_string(j) + " not found.");
        if ((eng.hasOffset() && j == 0) || (a.fixed && a.fixed[j]))
            throw std::invalid_argument("Cannot estimate confidence interval for a fixed coefficient");
        const bool regularized = a.prior_type != CCD_PRIOR_NONE && !(a.exclude && a.exclude[j]);
        if (regularized && !overrideNoRegularization)   // CcdInterface.cpp:300-304
            throw std::invalid_argument("Variable " + std::to_string(j) + " is regularized.");
    }
    auto logPrior = [&](const double* b) {   // JointPrior::logDensity over every column
        double lp = 0.0;
        for (int j = 0; j < J; ++j)
            lp += priorLogDensity((a.exclude && a.exclude[j]) ? CCD_PRIOR_NONE : a.prior_type, a.variance, b[j]);
        return lp;
    };
    std::vector<double> mode(beta, beta + J);
    if (eng.hasOffset()) mode[0] = 1.0;
    // value at the mode (CcdInterface.cpp:268-272)
    eng.setWeights(nullptr, false);
    eng.fixedTerms(false);
    eng.computeXBeta(mode.data());
    eng.remainingStats(false);
    double maxValue = eng.logLik(false);
    if (includePenalty) maxValue += logPrior(mode.data());

    std::vector<BoundSearch> S(2 * (size_t)m);
    for (int i = 0; i < m; ++i)
        for (int d = 0; d < 2; ++d) {
            BoundSearch& q = S[2 * i + d];
            q.out = 2 * i + d;
            q.col = idx[i];
            q.dir = d == 0 ? 1.0 : -1.0;   // upper, then lower, as the reference's bound list (:306-307)
            q.x0 = mode[idx[i]];
            q.obj0 = threshold;            // OptimizationProfile::getMaximum (:250-252)
            q.beta = mode;
            q.start();
        }
    std::vector<double> betas, var;
    std::vector<int32_t> fixIdx;
    std::vector<ccd_fit_result> res;
    std::vector<int> live;
    for (;;) {
        live.clear();
        for (int t = 0; t < (int)S.size(); ++t)
            if (S[t].phase != BoundSearch::DONE) live.push_back(t);
        if (live.empty()) break;
        const int n = (int)live.size();
        betas.assign((size_t)n * J, 0.0);
        var.assign(n, a.variance);
        fixIdx.assign(n, -1);
        res.assign(n, ccd_fit_result{});
        for (int f = 0; f < n; ++f) {
            BoundSearch& q = S[live[f]];
            std::copy(q.beta.begin(), q.beta.end(), betas.begin() + (size_t)f * J);
            betas[(size_t)f * J + q.col] = q.point;   // setBeta (:240)
            fixIdx[f] = q.col;                         // setFixedBeta(index, true) (:241)
        }
        ccd_batch_args ba{};
        ba.n_fits = n;
        ba.variances = var.data();
        ba.fixed_index = fixIdx.data();
        ccd_fit_args common = a;
        common.weights = nullptr;
        eng.fitBatch(common, ba, betas.data(), res.data(), false);
        for (int f = 0; f < n; ++f) {
            BoundSearch& q = S[live[f]];
            const double* bf = betas.data() + (size_t)f * J;
            q.beta.assign(bf, bf + J);
            double y = res[f].log_likelihood + threshold - maxValue;   // :244-247
            if (includePenalty) y += res[f].log_prior;
            q.take(y);
        }
    }
    for (int i = 0; i < m; ++i) {
        upper[i] = S[2 * i].result;
        lower[i] = S[2 * i + 1].result;
        evaluations[i] = S[2 * i].evals + S[2 * i + 1].evals;
    }
}

void profileLikelihood(EngineBase& eng, const ccd_fit_args& a, const double* beta, int index,
                       const double* points, int n, bool includePenalty, double* values) {
    if (a.weights) throw std::invalid_argument("profiling takes no row weights");
    const int J = eng.J();
    if (index < 0 || index >= J) throw std::invalid_argument("Variable " + std::to_string(index) + " not found.");
    if ((eng.hasOffset() && index == 0) || (a.fixed && a.fixed[index]))
        throw std::invalid_argument("Cannot profile a fixed coefficient");
    if (n <= 0) return;
    ccd_fit_args common = a;
    common.weights = nullptr;
    // the column held at 0, from beta (CcdInterface.cpp:570)
    std::vector<double> start(beta, beta + J);
    if (eng.hasOffset()) start[0] = 1.0;
    start[index] = 0.0;
    {
        double var = a.variance;
        int32_t fix = index;
        ccd_batch_args one{};
        one.n_fits = 1;
        one.variances = &var;
        one.fixed_index = &fix;
        ccd_fit_result r{};
        eng.fitBatch(common, one, start.data(), &r, false);
    }
    // every point from that fit, all in one batch (OptimizationProfile::objective
    // with max = threshold = 0, :238-249)
    std::vector<double> betas((size_t)n * J), var(n, a.variance);
    std::vector<int32_t> fix(n, index);
    std::vector<ccd_fit_result> res(n);
    for (int i = 0; i < n; ++i) {
        std::copy(start.begin(), start.end(), betas.begin() + (size_t)i * J);
        betas[(size_t)i * J + index] = points[i];
    }
    ccd_batch_args ba{};
    ba.n_fits = n;
    ba.variances = var.data();
    ba.fixed_index = fix.data();
    eng.fitBatch(common, ba, betas.data(), res.data(), false);
    for (int i = 0; i < n; ++i) values[i] = res[i].log_likelihood + (includePenalty ? res[i].log_prior : 0.0);
}

}  // namespace ccd
