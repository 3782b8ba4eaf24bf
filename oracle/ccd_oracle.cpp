/*
 * ccd_oracle.cpp -- THE ORACLE. Test infrastructure only.
 *
 * A plain FP64 CPU restatement of the reference's cyclic-coordinate-descent hot
 * path, used to check the HIP engine. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never does.
 *
 * What it restates (reference file:line):
 *   engine       src/cyclops/engine/ModelSpecifics.hpp
 *                  setWeights              :427-484
 *                  computeXjY              :486-524
 *                  fixed log-lik terms     :570-587
 *                  getLogLikelihood        :605-650
 *                  getPredictiveLogLik     :652-690
 *                  getPredictiveEstimates  :692-707
 *                  computeGradientAndHessian(Impl) :710-766, 891-1201
 *                  computeNumeratorForGradient     :1353-1464
 *                  updateXBeta(Impl)       :1466-1586
 *                  computeRemainingStatistics      :1588-1633
 *                  computeAccumlatedDenominator    :1670-1715
 *                  setPidForAccumulation   :1809-1868
 *                  setupSparseIndices      :1870-1892
 *   model traits src/cyclops/engine/ModelSpecifics.h  (Logistic :540-568,
 *                  ConditionalLogisticRegression :799-880,
 *                  LogisticRegression :970-1056, BreslowTiedCox :1162-1257,
 *                  PoissonRegression :1479-1603)
 *   priors       src/cyclops/priors/CovariatePrior.h :160-447
 *   driver       src/cyclops/CyclicCoordinateDescent.cpp
 *                  init :200-239, resetBeta :256-264, update :586-615,
 *                  performCheckConvergence :839-906, findMode :909-1216,
 *                  ccdUpdateBeta :1394-1413, computeXBeta :1452-1470,
 *                  applyBounds :1503-1520, getNewPredictiveLogLikelihood :319-355
 *   CV           src/cyclops/drivers/CrossValidationSelector.cpp :19-189,
 *                  AbstractSelector.cpp (mt19937 seeding),
 *                  AbstractCrossValidationDriver.cpp :31-42,131-261,
 *                  GridSearchCrossValidationDriver.cpp :38-44,96-138,234-244
 *
 * Pins (tests/test_oracle_*.py): closed-form Poisson MLE (Dobson), textbook IRLS
 * logistic fit, brute-force Breslow partial likelihood and its derivatives on the
 * reference's inline Cox tables, finite differences of the log-likelihood for every
 * model and column format, Laplace KKT conditions, the SAS -2 log-likelihood stored
 * with the reference's 17-variable logistic fixture, CLR against an independent
 * Newton solver on the reference's infert fixture, and the held-out predictive
 * log-likelihood, gradient objective and predictive estimates against numpy
 * restatements from a dense X (Cox held-out risk sets by O(K^2) brute force), their
 * unit-weight identities with the log-likelihood and their additivity over folds.
 */
#include "../include/cyclops_amd.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

enum { DENSE = 0, SPARSE = 1, INDICATOR = 2, INTERCEPT = 3 };
enum { POISSON = 2, LOGISTIC = 3, CLR = 4, COX = 8 };

struct Data {
    int model = 0;
    int64_t K = 0;
    int J = 0;
    bool hasOffset = false;
    std::vector<int> fmt;
    std::vector<int64_t> colPtr;
    std::vector<int32_t> rows;
    std::vector<double> vals;
    std::vector<int64_t> densePtr;
    std::vector<double> dense;
    std::vector<double> y, time;
    std::vector<int32_t> pid;   // renumbered: a new id each time the input id changes
    int64_t nPatients = 0;
};

static Data makeData(const ccd_data_desc* d) {
    Data o;
    o.model = d->model_type;
    o.K = d->K;
    o.J = d->J;
    o.hasOffset = d->has_offset != 0;
    o.fmt.assign(d->format, d->format + o.J);
    o.colPtr.assign(d->col_ptr, d->col_ptr + o.J + 1);
    const int64_t nnz = o.colPtr[o.J];
    o.rows.assign(d->rows, d->rows + nnz);
    if (d->values) o.vals.assign(d->values, d->values + nnz);
    if (d->dense_ptr) {
        o.densePtr.assign(d->dense_ptr, d->dense_ptr + o.J + 1);
        o.dense.assign(d->dense_values, d->dense_values + o.densePtr[o.J]);
    } else {
        o.densePtr.assign(o.J + 1, 0);
    }
    o.y.assign(d->y, d->y + o.K);
    if (d->time) o.time.assign(d->time, d->time + o.K);
    else o.time.assign(o.K, 0.0);
    o.pid.resize(o.K);
    // RcppModelData.cpp:878-899: renumber pids into contiguous runs
    if (d->pid == nullptr || o.model == LOGISTIC || o.model == POISSON) {
        for (int64_t k = 0; k < o.K; ++k) o.pid[k] = (int32_t)k;
        o.nPatients = o.K;
    } else {
        int cur = 0;
        int32_t last = d->pid[0];
        for (int64_t k = 0; k < o.K; ++k) {
            if (d->pid[k] != last) { ++cur; last = d->pid[k]; }
            o.pid[k] = cur;
        }
        o.nPatients = cur + 1;
    }
    return o;
}

// Visit the entries of column j as (row, x). DENSE and INTERCEPT visit all K rows.
template <class F>
static void forEach(const Data& d, int j, F f) {
    switch (d.fmt[j]) {
    case INDICATOR:
        for (int64_t e = d.colPtr[j]; e < d.colPtr[j + 1]; ++e) f((int64_t)d.rows[e], 1.0);
        break;
    case SPARSE:
        for (int64_t e = d.colPtr[j]; e < d.colPtr[j + 1]; ++e) f((int64_t)d.rows[e], d.vals[e]);
        break;
    case DENSE: {
        const double* v = d.dense.data() + d.densePtr[j];
        for (int64_t k = 0; k < d.K; ++k) f(k, v[k]);
        break;
    }
    case INTERCEPT:
        for (int64_t k = 0; k < d.K; ++k) f(k, 1.0);
        break;
    }
}

static int64_t nonZeros(const Data& d, int j) {   // CompressedDataMatrix.cpp:93-100
    return (d.fmt[j] == DENSE || d.fmt[j] == INTERCEPT) ? d.K : d.colPtr[j + 1] - d.colPtr[j];
}
static bool indicatorLike(int fmt) { return fmt == INDICATOR || fmt == INTERCEPT; }

// ---------------------------------------------------------------------------
// Engine: ModelSpecifics<Model, double>
// ---------------------------------------------------------------------------
class Specifics {
public:
    explicit Specifics(const Data& data) : d(data) {
        K = d.K;
        N = d.nPatients;
        xb.assign(K, 0.0);
        e.assign(K, 0.0);
        kW.assign(K, 1.0);
        XjY.assign(d.J, 0.0);
        pid = d.pid;
        // the per-column pid lists are read only by the Cox gradient
        if (cumulative()) setPidForAccumulation(nullptr);
        const size_t len = (size_t)N + 1 + 16;
        denom.assign(len, 0.0);
        numer.assign(len, 0.0);
        numer2.assign(len, 0.0);
        nW.assign(len, 0.0);
    }

    bool cumulative() const { return d.model == COX; }
    bool independent() const { return d.model == LOGISTIC || d.model == POISSON; }
    int64_t group(int64_t k) const { return independent() ? k : pid[k]; }

    // ModelSpecifics.h: observationCount (LR/PR 1, CLR/Cox y)
    double observationCount(double y) const { return independent() ? 1.0 : y; }
    double denomNull() const { return d.model == LOGISTIC ? 1.0 : 0.0; }

    // ModelSpecifics.hpp:1809-1868
    void setPidForAccumulation(const double* w) {
        pid = d.pid;
        accReset.clear();
        const int ignore = -1;
        int64_t index = 0;
        while (w != nullptr && index < K && w[index] == 0.0) { pid[index] = ignore; ++index; }
        if (index == K) {   // no row with non-zero weight
            N = 0;
            accReset.push_back(0);
            for (int64_t k = 0; k < K; ++k) pid[k] = 0;
            setupSparseIndices(0);
            return;
        }
        int lastPid = pid[index];
        double lastTime = d.time[index];
        double lastEvent = d.y[index];
        int cur = 0;
        pid[index] = 0;
        for (int64_t k = index + 1; k < K; ++k) {
            if (w == nullptr || w[k] != 0.0) {
                int nextPid = pid[k];
                if (nextPid != lastPid) {
                    ++cur;
                    accReset.push_back(cur);
                    lastPid = nextPid;
                } else if (!(lastEvent == 1.0 && lastTime == d.time[k] && lastEvent == d.y[k])) {
                    ++cur;
                }
                lastTime = d.time[k];
                lastEvent = d.y[k];
                pid[k] = cur;
            } else {
                pid[k] = ignore;
            }
        }
        ++cur;
        accReset.push_back(cur);
        N = cur;
        if (w != nullptr)
            for (int64_t k = 0; k < K; ++k) if (pid[k] == ignore) pid[k] = (int)N;
        setupSparseIndices((int)N);
        const size_t len = (size_t)N + 1 + 16;
        if (denom.size() < len) { denom.resize(len); numer.resize(len); numer2.resize(len); nW.resize(len); }
    }

    // ModelSpecifics.hpp:1870-1892
    void setupSparseIndices(int max) {
        sparseIdx.assign(d.J, {});
        sparseNull.assign(d.J, 0);
        for (int j = 0; j < d.J; ++j) {
            if (d.fmt[j] == DENSE || d.fmt[j] == INTERCEPT) { sparseNull[j] = 1; continue; }
            std::set<int> unique;
            for (int64_t e2 = d.colPtr[j]; e2 < d.colPtr[j + 1]; ++e2) {
                const int k = d.rows[e2];
                const int i = (k < (int64_t)pid.size()) ? pid[k] : k;
                if (i < max) unique.insert(i);
            }
            sparseIdx[j].assign(unique.begin(), unique.end());
        }
    }

    // ModelSpecifics.hpp:427-484
    void setWeights(const double* w, bool useCV) {
        if (useCV) kW.assign(w, w + K);
        else std::fill(kW.begin(), kW.end(), 1.0);
        if (cumulative()) setPidForAccumulation(useCV ? w : nullptr);
        std::fill(nW.begin(), nW.end(), 0.0);
        for (int64_t k = 0; k < K; ++k) nW[group(k)] += observationCount(d.y[k]) * kW[k];
    }

    // ModelSpecifics.hpp:486-524 (+570-587 for the Poisson fixed terms)
    void fixedTerms(bool useCV) {
        for (int j = 0; j < d.J; ++j) {
            double s = 0.0;
            forEach(d, j, [&](int64_t k, double x) {
                const double yy = cumulative() ? observationCount(d.y[k]) : d.y[k];
                s += useCV ? x * yy * kW[k] : x * yy;
            });
            XjY[j] = s;
        }
        fixedTerm = 0.0;
        if (d.model == POISSON) {
            for (int64_t k = 0; k < K; ++k) {
                double c = 0.0;
                for (int i = 2; i <= (int)d.y[k]; ++i) c -= std::log((double)i);
                fixedTerm += useCV ? c * kW[k] : c;
            }
        }
    }

    void zeroXBeta() { std::fill(xb.begin(), xb.end(), 0.0); }
    void axpyXBeta(double beta, int j) {   // ModelSpecifics.hpp:393-423
        if (beta != 0.0) forEach(d, j, [&](int64_t k, double x) { xb[k] += beta * x; });
    }

    // ModelSpecifics.hpp:1588-1633
    void remainingStats(bool useW) {
        std::fill(denom.begin(), denom.begin() + N, denomNull());
        if (cumulative()) {
            for (int64_t k = 0; k < K; ++k) {
                e[k] = std::exp(xb[k]);
                denom[pid[k]] += useW ? kW[k] * e[k] : e[k];
            }
        } else {
            for (int64_t k = 0; k < K; ++k) {
                e[k] = std::exp(xb[k]);
                denom[group(k)] += e[k];
            }
        }
        accumulateDenominator();
    }

    // ModelSpecifics.hpp:1670-1694
    void accumulateDenominator() {
        if (!cumulative()) return;
        accDenom.assign((size_t)N + 1, 0.0);
        double total = 0.0;
        size_t r = 0;
        for (int64_t i = 0; i < N; ++i) {
            if (r < accReset.size() && accReset[r] == i) { total = 0.0; ++r; }
            total += denom[i];
            accDenom[i] = total;
        }
    }

    // ModelSpecifics.hpp:1353-1464
    void numerator(int j, bool useW) {
        if (!cumulative()) return;
        if (d.fmt[j] == INDICATOR || d.fmt[j] == SPARSE) {
            for (int i : sparseIdx[j]) { numer[i] = 0.0; numer2[i] = 0.0; }
        } else {
            std::fill(numer.begin(), numer.begin() + N, 0.0);
            std::fill(numer2.begin(), numer2.begin() + N, 0.0);
        }
        const bool ind = d.fmt[j] == INDICATOR;
        forEach(d, j, [&](int64_t k, double x) {
            const double w = useW ? kW[k] : 1.0;
            numer[pid[k]] += w * e[k] * x;
            if (!ind) numer2[pid[k]] += w * e[k] * x * x;
        });
    }

    // ModelSpecifics.hpp:710-766 and 891-1201
    void gradHess(int j, double* og, double* oh, bool useW) {
        if (nonZeros(d, j) == 0) { *og = 0.0; *oh = 0.0; return; }
        const bool ind = indicatorLike(d.fmt[j]);
        double g = 0.0, h = 0.0;
        if (cumulative()) {
            // Breslow: t = A/D, g += nE t, h += nE t (1-t) | nE (A2/D - t^2)
            auto inc = [&](double a1, double a2, double D, double nE) {
                const double t = a1 / D;
                const double gi = nE * t;
                g += gi;
                h += ind ? gi * (1.0 - t) : nE * (a2 / D - t * t);
            };
            std::vector<int> all;
            const std::vector<int>* it = nullptr;
            if (sparseNull[j]) {
                all.resize(N);
                for (int64_t i = 0; i < N; ++i) all[i] = (int)i;
                it = &all;
            } else {
                it = &sparseIdx[j];
            }
            const bool gapFill = !sparseNull[j];   // IteratorType::isSparse
            if (!it->empty()) {
                double acc1 = 0.0, acc2 = 0.0;
                size_t r = 0;
                while (r < accReset.size() && accReset[r] < (*it)[0]) ++r;
                for (size_t p = 0; p < it->size(); ++p) {
                    int64_t i = (*it)[p];
                    if (r < accReset.size() && accReset[r] <= i) { acc1 = 0.0; acc2 = 0.0; ++r; }
                    acc1 += numer[i];
                    acc2 += numer2[i];
                    inc(acc1, acc2, accDenom[i], nW[i]);
                    if (gapFill) {
                        const int64_t next = (p + 1 < it->size()) ? (*it)[p + 1] : N;
                        for (++i; i < next; ++i) {
                            if (r < accReset.size() && accReset[r] <= i) { acc1 = 0.0; acc2 = 0.0; ++r; }
                            inc(acc1, acc2, accDenom[i], nW[i]);
                        }
                    }
                }
            }
        } else if (independent()) {
            forEach(d, j, [&](int64_t k, double x) {
                const double n1 = e[k] * x;
                const double n2 = ind ? 0.0 : e[k] * x * x;
                const double w = nW[k];
                if (d.model == LOGISTIC) {   // Logistic<>::incrementGradientAndHessian
                    const double D = denom[k];
                    const double gi = n1 / D;
                    const double hi = ind ? gi * (1.0 - gi) : (n2 / D - gi * gi);
                    g += useW ? w * gi : gi;
                    h += useW ? w * hi : hi;
                } else {                      // PoissonRegression::incrementGradientAndHessian
                    g += useW ? w * n1 : n1;
                    h += useW ? w * (ind ? n1 : n2) : (ind ? n1 : n2);
                }
            });
        } else {
            // CLR: entries in row order, runs with equal stratum summed (:1127-1179)
            bool first = true;
            int64_t key = -1;
            double n1 = 0.0, n2 = 0.0;
            auto close = [&]() {
                const double D = denom[key], nE = nW[key];
                const double t = n1 / D;
                const double gi = nE * t;
                g += gi;
                h += ind ? gi * (1.0 - t) : nE * (n2 / D - t * t);
            };
            forEach(d, j, [&](int64_t k, double x) {
                const int64_t kk = pid[k];
                if (!first && kk != key) { close(); n1 = 0.0; n2 = 0.0; }
                first = false;
                key = kk;
                n1 += e[k] * x;
                if (!ind) n2 += e[k] * x * x;
            });
            if (!first) close();
        }
        g -= XjY[j];
        *og = g;
        *oh = h;
    }

    // ModelSpecifics.hpp:1524-1586
    void updateXBeta(double delta, int j, bool useW) {
        forEach(d, j, [&](int64_t k, double x) {
            xb[k] += delta * x;
            if (cumulative()) {
                const double oldE = useW ? kW[k] * e[k] : e[k];
                e[k] = std::exp(xb[k]);
                const double newE = useW ? kW[k] * e[k] : e[k];
                denom[pid[k]] += newE - oldE;
            } else {
                const double oldE = e[k];
                e[k] = std::exp(xb[k]);
                denom[group(k)] += e[k] - oldE;
            }
        });
        accumulateDenominator();
    }

    // ModelSpecifics.hpp:605-650
    double logLik(bool useCV) {
        double ll = 0.0;
        for (int64_t k = 0; k < K; ++k) {
            const int yi = (int)d.y[k];
            const double c = (d.model == COX) ? (yi != 1 ? 0.0 : xb[k]) : yi * xb[k];
            ll += useCV ? c * kW[k] : c;
        }
        for (int64_t i = 0; i < N; ++i) {
            if (d.model == POISSON) ll -= nW[i] * denom[i];
            else if (d.model == COX) ll -= nW[i] * std::log(accDenom[i]);
            else ll -= nW[i] * std::log(denom[i]);
        }
        if (d.model == POISSON) ll += fixedTerm;
        return ll;
    }

    double gradientObjective(bool useCV) {   // ModelSpecifics.hpp:103-120
        double c = 0.0;
        for (int64_t k = 0; k < K; ++k) c += useCV ? xb[k] * d.y[k] * kW[k] : xb[k] * d.y[k];
        return c;
    }

    // ModelSpecifics.hpp:652-690
    double predLogLik(const double* w) {
        std::vector<double> saved;
        if (cumulative()) {
            saved = kW;
            setPidForAccumulation(w);
            setWeights(w, true);
            remainingStats(true);
        }
        double ll = 0.0;
        for (int64_t k = 0; k < K; ++k) {
            const double yk = d.y[k], wk = w[k], x = xb[k];
            switch (d.model) {
            case LOGISTIC: ll += yk * wk * (x - std::log(denom[k])); break;
            case POISSON:  ll += (yk * x - std::exp(x)) * wk; break;
            case CLR:      ll += yk * wk * (x - std::log(denom[pid[k]])); break;
            case COX:      ll += (wk == 0.0) ? 0.0 : yk * wk * (x - std::log(accDenom[pid[k]])); break;
            }
        }
        if (cumulative()) {
            setPidForAccumulation(saved.data());
            setWeights(saved.data(), true);
            remainingStats(true);
        }
        return ll;
    }

    void predEstimates(double* out, const double* w) {   // ModelSpecifics.hpp:692-707
        for (int64_t k = 0; k < K; ++k) {
            if (w && w[k] == 0.0) continue;
            const double x = xb[k];
            if (d.model == LOGISTIC) { const double t = std::exp(x); out[k] = t / (t + 1.0); }
            else if (d.model == POISSON) out[k] = std::exp(x);
            else out[k] = x;
        }
    }

    const Data& d;
    int64_t K, N;
    // ModelSpecifics::computeFisherInformationImpl (ModelSpecifics.hpp:1269-1352) for
    // every pair CyclicCoordinateDescent::computeFisherInformation asks for
    // (CyclicCoordinateDescent.cpp:1288-1300); remainingStats with unit weights first.
    // Increments: Logistic (ModelSpecifics.h:521-535) w (g - g^2), g = e / denom;
    // PoissonRegression (:1497-1507) w e; Survival (:488-497, CLR and Cox) w e / denom of
    // the row's group, less the strata cross terms sum_n C_a[n] C_b[n] / denom[n]^2 with
    // C_a[n] = sum of x_a e over the rows of group n (getSubjectSpecificHessianIterator,
    // ModelSpecifics.hpp:1245-1267). Plain loops over dense copies of the columns.
    void fisher(const int* idx, int m, double* out) {
        std::vector<std::vector<double>> X(m, std::vector<double>(K, 0.0));
        for (int a = 0; a < m; ++a) forEach(d, idx[a], [&](int64_t k, double x) { X[a][k] = x; });
        std::vector<double> w(K);
        for (int64_t k = 0; k < K; ++k) {
            if (d.model == LOGISTIC) { const double g = e[k] / denom[k]; w[k] = g - g * g; }
            else if (d.model == POISSON) w[k] = e[k];
            else w[k] = e[k] / denom[pid[k]];
        }
        std::vector<std::vector<double>> C;
        if (!independent()) {
            C.assign(m, std::vector<double>(N, 0.0));
            for (int a = 0; a < m; ++a)
                for (int64_t k = 0; k < K; ++k) C[a][pid[k]] += X[a][k] * e[k];
        }
        for (int a = 0; a < m; ++a)
            for (int b = a; b < m; ++b) {
                double v = 0.0;
                for (int64_t k = 0; k < K; ++k) v += w[k] * X[a][k] * X[b][k];
                if (!independent()) {
                    double cross = 0.0;
                    for (int64_t n = 0; n < N; ++n) cross += C[a][n] * C[b][n] / (denom[n] * denom[n]);
                    v -= cross;
                }
                out[(size_t)a * m + b] = out[(size_t)b * m + a] = v;
            }
    }

    std::vector<double> xb, e, denom, numer, numer2, accDenom, nW, kW, XjY;
    std::vector<int> pid, accReset;
    std::vector<std::vector<int>> sparseIdx;
    std::vector<char> sparseNull;
    double fixedTerm = 0.0;
};

// ---------------------------------------------------------------------------
// Priors: CovariatePrior.h:186-188 (None), 240-273 (Laplace), 410-415 (Normal)
// ---------------------------------------------------------------------------
struct Prior {
    int type = 0;
    double variance = 1.0;
    double lambda() const { return std::sqrt(2.0 / variance); }
    static int sign(double x) { return x == 0 ? 0 : (x < 0 ? -1 : 1); }
    double delta(double g, double h, double beta) const {
        if (type == CCD_PRIOR_NONE) return -(g / h);
        if (type == CCD_PRIOR_NORMAL) return -(g + beta / variance) / (h + 1.0 / variance);
        const double lam = lambda();
        const double negU = -(g - lam) / h;
        const double posU = -(g + lam) / h;
        const int s = sign(beta);
        double dlt;
        if (s == 0) {
            if (negU < 0) dlt = negU;
            else if (posU > 0) dlt = posU;
            else dlt = 0;
        } else {
            dlt = (s < 0) ? negU : posU;
            if (sign(beta + dlt) != s) dlt = -beta;
        }
        return dlt;
    }
    double logDensity(double x) const {
        if (type == CCD_PRIOR_NONE) return 0.0;
        if (type == CCD_PRIOR_LAPLACE) { const double l = lambda(); return std::log(0.5 * l) - l * std::abs(x); }
        return -0.5 * std::log(2.0 * M_PI * variance) - 0.5 * x * x / variance;
    }
};

// ---------------------------------------------------------------------------
// Driver: CyclicCoordinateDescent (findMode / update)
// ---------------------------------------------------------------------------
class CCD {
public:
    CCD(const Data& data, Specifics& s) : d(data), ms(s) {
        J = d.J;
        beta.assign(J, 0.0);
        hDelta.assign(J, 2.0);
        fixBeta.assign(J, 0);
        priors.assign(J, Prior());
        if (d.hasOffset) { beta[0] = 1.0; fixBeta[0] = 1; xBetaKnown = false; }
        else xBetaKnown = true;
    }

    void setWeights(const double* w) {   // CyclicCoordinateDescent.cpp:496-519
        if (w == nullptr) { weights.clear(); useCV = false; }
        else { weights.assign(w, w + d.K); useCV = true; }
        validWeights = false;
        suffKnown = false;
    }
    void setBeta(const double* b) { beta.assign(b, b + J); xBetaKnown = false; suffKnown = false; }
    void resetBeta() {   // :256-264
        for (int j = d.hasOffset ? 1 : 0; j < J; ++j) beta[j] = 0.0;
        computeXBeta();
        suffKnown = false;
    }
    void computeXBeta() {   // :1452-1470
        ms.zeroXBeta();
        for (int j = 0; j < J; ++j) ms.axpyXBeta(beta[j], j);
    }
    void computeNEvents() { ms.setWeights(useCV ? weights.data() : nullptr, useCV); }
    void lazy() {   // checkAllLazyFlags :388-406
        if (!xBetaKnown) { computeXBeta(); xBetaKnown = true; suffKnown = false; }
        if (!validWeights) { computeNEvents(); ms.fixedTerms(useCV); validWeights = true; }
        if (!suffKnown) { ms.remainingStats(useCV); suffKnown = true; }
    }
    double logLik() { lazy(); return ms.logLik(useCV); }
    double logPrior() { double s = 0; for (int j = 0; j < J; ++j) s += priors[j].logDensity(beta[j]); return s; }
    double objective(int type) {   // :541-558
        if (type == CCD_CONV_GRADIENT) return ms.gradientObjective(useCV);
        if (type == CCD_CONV_MITTAL) return logLik();
        if (type == CCD_CONV_LANGE) return logLik() + logPrior();
        throw std::invalid_argument("Invalid convergence type");
    }
    double applyBounds(double delta, int j) {   // :1503-1520
        if (delta < -hDelta[j]) delta = -hDelta[j];
        else if (delta > hDelta[j]) delta = hDelta[j];
        double inter = std::max(std::abs(delta) * 2, hDelta[j] / 2);
        inter = std::max(inter, 1E-3);
        hDelta[j] = inter;
        return delta;
    }
    double ccdUpdateBeta(int j) {   // :1394-1413
        ms.numerator(j, useCV);
        double g, h;
        ms.gradHess(j, &g, &h, useCV);
        if (h < 0.0) { g = 0.0; h = 0.0; }
        return priors[j].delta(g, h, beta[j]);
    }

    void findMode(int maxIt, int convType, double eps) {   // :909-1216
        if (convType < CCD_CONV_GRADIENT || convType > CCD_CONV_MITTAL)
            throw std::invalid_argument("Unknown convergence criterion");
        if (!validWeights) { computeNEvents(); ms.fixedTerms(useCV); validWeights = true; }
        if (!xBetaKnown) { computeXBeta(); xBetaKnown = true; suffKnown = false; }
        if (!suffKnown) { ms.remainingStats(useCV); suffKnown = true; }
        std::fill(hDelta.begin(), hDelta.end(), initialBound);
        int iteration = 0;
        double lastObj = objective(convType);
        bool done = false;
        while (!done) {
            for (int j = 0; j < J; ++j) {
                if (!fixBeta[j]) {
                    double delta = ccdUpdateBeta(j);
                    delta = applyBounds(delta, j);
                    if (delta != 0.0) {
                        suffKnown = false;
                        beta[j] += delta;
                        ms.updateXBeta(delta, j, useCV);
                        suffKnown = true;
                    }
                }
            }
            ++iteration;
            // performCheckConvergence :839-906
            double conv;
            bool ill = false;
            const double thisObj = objective(convType);
            if (thisObj != thisObj) { conv = 0.0; ill = true; }
            else conv = std::abs(thisObj - lastObj) / (std::abs(thisObj) + 1.0);
            lastObj = thisObj;
            lastLogLik = logLik();
            lastLogPrior = logPrior();
            if (eps > 0 && conv < eps) { lastFlag = ill ? CCD_ILLCONDITIONED : CCD_SUCCESS; done = true; }
            else if (iteration == maxIt) { lastFlag = CCD_MAX_ITERATIONS; done = true; }
        }
        lastIterations = iteration;
    }

    // Seconds for the first n coordinates of the first cycle (the loop body of
    // findMode above), for a bounded CPU timing sample.
    double timeCoordinates(int n) {
        if (!validWeights) { computeNEvents(); ms.fixedTerms(useCV); validWeights = true; }
        if (!xBetaKnown) { computeXBeta(); xBetaKnown = true; suffKnown = false; }
        if (!suffKnown) { ms.remainingStats(useCV); suffKnown = true; }
        std::fill(hDelta.begin(), hDelta.end(), initialBound);
        const auto t0 = std::chrono::steady_clock::now();
        int done = 0;
        for (int j = 0; j < J && done < n; ++j) {
            if (fixBeta[j]) continue;
            double delta = ccdUpdateBeta(j);
            delta = applyBounds(delta, j);
            if (delta != 0.0) {
                beta[j] += delta;
                ms.updateXBeta(delta, j, useCV);
            }
            ++done;
        }
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void update(const ccd_fit_args& a) {   // :586-615
        initialBound = a.initial_bound;
        int count = 0;
        bool done = false;
        while (!done) {
            findMode(a.max_iterations, a.convergence_type, a.tolerance);
            ++count;
            if (lastFlag == CCD_ILLCONDITIONED && count < a.max_bound_count) {
                initialBound /= 10.0;
                resetBeta();
            } else {
                done = true;
            }
        }
        findModeCalls = count;
    }

    double predLogLik(const double* w) {   // :319-355
        computeXBeta();
        xBetaKnown = true;
        suffKnown = false;
        ms.remainingStats(useCV);
        suffKnown = true;
        return ms.predLogLik(w);
    }

    const Data& d;
    Specifics& ms;
    int J;
    std::vector<double> beta, hDelta, weights;
    std::vector<char> fixBeta;
    std::vector<Prior> priors;
    bool useCV = false, validWeights = false, suffKnown = false, xBetaKnown = true;
    double initialBound = 2.0;
    int lastFlag = CCD_SUCCESS, lastIterations = 0, findModeCalls = 0;
    double lastLogLik = 0.0, lastLogPrior = 0.0;
};

static void configure(CCD& ccd, const ccd_fit_args* a) {
    for (int j = 0; j < ccd.J; ++j) {
        ccd.priors[j].type = (a->exclude && a->exclude[j]) ? CCD_PRIOR_NONE : a->prior_type;
        ccd.priors[j].variance = a->variance;
        if (a->fixed && a->fixed[j]) ccd.fixBeta[j] = 1;
    }
}

// CrossValidationSelector.cpp:19-189, AbstractSelector.cpp: fold assignment.
// Returns row_fold[k] = index of the fold in which row k is held out.
static void assignFolds(int64_t K, const int32_t* pidIn, int fold, int selector, int64_t seed,
                        int32_t* rowFold) {
    std::vector<int> ids(K);
    int64_t N;
    if (selector == CCD_SELECT_BY_PID) {
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

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface used by tests and the bench's CPU baseline (via ctypes)
// ---------------------------------------------------------------------------
struct orc_handle {
    orc::Data data;
    orc::Specifics* ms = nullptr;
    ~orc_handle() { delete ms; }
};

static thread_local std::string g_orc_err;

extern "C" {

const char* orc_last_error(void) { return g_orc_err.c_str(); }

void* orc_create(const ccd_data_desc* d) {
    try {
        auto* h = new orc_handle();
        h->data = orc::makeData(d);
        h->ms = new orc::Specifics(h->data);
        return h;
    } catch (const std::exception& ex) {
        g_orc_err = ex.what();
        return nullptr;
    }
}
void orc_destroy(void* h) { delete static_cast<orc_handle*>(h); }

#define MS (static_cast<orc_handle*>(h)->ms)
int64_t orc_num_pids(void* h) { return MS->N; }
void orc_set_weights(void* h, const double* w, int useW) { MS->setWeights(w, useW != 0); }
void orc_fixed_terms(void* h, int useW) { MS->fixedTerms(useW != 0); }
void orc_compute_xbeta(void* h, const double* beta) {
    MS->zeroXBeta();
    for (int j = 0; j < MS->d.J; ++j) MS->axpyXBeta(beta[j], j);
}
void orc_remaining_stats(void* h, int useW) { MS->remainingStats(useW != 0); }
void orc_numerator(void* h, int j, int useW) { MS->numerator(j, useW != 0); }
void orc_grad_hess(void* h, int j, double* g, double* hh, int useW) { MS->gradHess(j, g, hh, useW != 0); }
void orc_update_xbeta(void* h, int j, double delta, int useW) { MS->updateXBeta(delta, j, useW != 0); }
double orc_loglik(void* h, int useW) { return MS->logLik(useW != 0); }
double orc_pred_loglik(void* h, const double* w) { return MS->predLogLik(w); }
double orc_gradient_objective(void* h, int useW) { return MS->gradientObjective(useW != 0); }
void orc_pred_est(void* h, double* out, const double* w) { MS->predEstimates(out, w); }
void orc_get_xbeta(void* h, double* out) { std::copy(MS->xb.begin(), MS->xb.end(), out); }
void orc_get_denom(void* h, double* out) {   // denomPid (or accDenomPid for Cox) over N pids
    if (MS->cumulative()) std::copy(MS->accDenom.begin(), MS->accDenom.begin() + MS->N, out);
    else std::copy(MS->denom.begin(), MS->denom.begin() + MS->N, out);
}
#undef MS

/* Full fit (update -> findMode) in FP64 on the CPU. */
int orc_fit(void* hv, const ccd_fit_args* a, double* beta, ccd_fit_result* r) {
    try {
        auto* h = static_cast<orc_handle*>(hv);
        const auto t0 = std::chrono::steady_clock::now();
        orc::CCD ccd(h->data, *h->ms);
        orc::configure(ccd, a);
        std::vector<double> b(beta, beta + h->data.J);
        if (h->data.hasOffset) b[0] = 1.0;
        ccd.setBeta(b.data());
        if (a->weights) ccd.setWeights(a->weights);
        ccd.update(*a);
        std::copy(ccd.beta.begin(), ccd.beta.end(), beta);
        r->return_flag = ccd.lastFlag;
        r->iterations = ccd.lastIterations;
        r->find_mode_calls = ccd.findModeCalls;
        r->log_likelihood = ccd.lastLogLik;
        r->log_prior = ccd.lastLogPrior;
        r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    } catch (const std::exception& ex) {
        g_orc_err = ex.what();
        return 1;
    }
}

/* Per-cycle timing: run exactly `cycles` CCD cycles (no convergence test) from
 * beta and report seconds; used to extrapolate CPU fits that take too long. */
double orc_time_coordinates(void* hv, const ccd_fit_args* a, int n) {
    auto* h = static_cast<orc_handle*>(hv);
    orc::CCD ccd(h->data, *h->ms);
    orc::configure(ccd, a);
    std::vector<double> b(h->data.J, 0.0);
    if (h->data.hasOffset) b[0] = 1.0;
    ccd.setBeta(b.data());
    if (a->weights) ccd.setWeights(a->weights);
    ccd.initialBound = a->initial_bound;
    ccd.lazy();   // set-up outside the timed region
    return ccd.timeCoordinates(n);
}

/* Fisher information of columns idx[0..m) at beta, unit weights: out[m][m]. */
int orc_fisher_information(void* hv, const double* beta, const int* idx, int m, double* out) {
    try {
        auto* h = static_cast<orc_handle*>(hv);
        auto* ms = h->ms;
        ms->setWeights(nullptr, false);
        ms->zeroXBeta();
        for (int j = 0; j < h->data.J; ++j) ms->axpyXBeta(beta[j], j);
        ms->remainingStats(false);
        ms->fisher(idx, m, out);
        return 0;
    } catch (const std::exception& ex) {
        g_orc_err = ex.what();
        return 1;
    }
}

/* CcdInterface::profileModel (CcdInterface.cpp:262-434) for one thread per bound:
 * each bound search runs on its own CCD started at the mode (the reference's clones,
 * :319-325), evaluating OptimizationProfile::objective (:238-249) one point at a time
 * inside RZeroIn::bracketSignChange (utils/RZeroIn.h:50-74) and Brent's zeroin
 * (R_zeroin2, :183-288; tolerance 1e-3 as CcdInterface.cpp:345). */
int orc_profile(void* hv, const ccd_fit_args* a, const double* beta, const int* idx, int m, double threshold,
                int includePenalty, double* lower, double* upper, int* evals) {
    try {
        auto* h = static_cast<orc_handle*>(hv);
        const int J = h->data.J;
        std::vector<double> mode(beta, beta + J);
        if (h->data.hasOffset) mode[0] = 1.0;
        double maxValue;
        {
            orc::CCD base(h->data, *h->ms);
            orc::configure(base, a);
            base.setBeta(mode.data());
            maxValue = base.logLik() + (includePenalty ? base.logPrior() : 0.0);
        }
        for (int i = 0; i < m; ++i) {
            const int col = idx[i];
            for (int side = 0; side < 2; ++side) {
                const double dir = side == 0 ? 1.0 : -1.0;
                orc::CCD ccd(h->data, *h->ms);
                orc::configure(ccd, a);
                ccd.setBeta(mode.data());
                int n = 0;
                auto objective = [&](double x) {
                    ++n;
                    ccd.beta[col] = x;
                    ccd.xBetaKnown = false;
                    ccd.suffKnown = false;
                    ccd.fixBeta[col] = 1;
                    ccd.update(*a);
                    ccd.fixBeta[col] = 0;
                    return ccd.logLik() + threshold - maxValue + (includePenalty ? ccd.logPrior() : 0.0);
                };
                const double x0 = mode[col], obj0 = threshold;
                // bracket: steps of 0.1, 0.2, 0.4, ... times max(|x0|, 0.01)
                const double disp = std::max(std::abs(x0), 0.01);
                double mult = 0.1, x1 = x0, o1 = obj0, pt;
                int steps = 0;
                bool bracketed = true;
                while (o1 > 0) {
                    x1 = x0 + dir * disp * mult;
                    o1 = objective(x1);
                    if (std::isnan(o1) || steps > 20) { bracketed = false; break; }
                    ++steps;
                    mult *= 2.0;
                }
                if (!bracketed) {
                    pt = std::nan("");
                } else {
                    // Brent's method on [x0, x1] (ends ordered), as published
                    double aa = x0, bb = x1, fa = obj0, fb = o1;
                    if (bb < aa) { std::swap(aa, bb); std::swap(fa, fb); }
                    double cc = aa, fc = fa;
                    const double tol = 1e-3;
                    if (fa == 0.0) pt = aa;
                    else if (fb == 0.0) pt = bb;
                    else {
                        pt = std::nan("");
                        for (int it = 0; it <= 1000; ++it) {
                            const double prev = bb - aa;
                            if (std::abs(fc) < std::abs(fb)) { aa = bb; bb = cc; cc = aa; fa = fb; fb = fc; fc = fa; }
                            const double tolAct = 2.0 * DBL_EPSILON * std::abs(bb) + tol / 2.0;
                            double step = (cc - bb) / 2.0;
                            if (std::abs(step) <= tolAct || fb == 0.0) { pt = bb; break; }
                            if (std::abs(prev) >= tolAct && std::abs(fa) > std::abs(fb)) {
                                const double cb = cc - bb;
                                double p, q;
                                if (aa == cc) { const double t = fb / fa; p = cb * t; q = 1.0 - t; }
                                else {
                                    const double qq = fa / fc, r = fb / fc, t = fb / fa;
                                    p = t * (cb * qq * (qq - r) - (bb - aa) * (r - 1.0));
                                    q = (qq - 1.0) * (r - 1.0) * (t - 1.0);
                                }
                                if (p > 0.0) q = -q; else p = -p;
                                if (p < 0.75 * cb * q - std::abs(tolAct * q) / 2.0 && p < std::abs(prev * q / 2.0))
                                    step = p / q;
                            }
                            if (std::abs(step) < tolAct) step = step > 0.0 ? tolAct : -tolAct;
                            aa = bb; fa = fb;
                            bb += step;
                            fb = objective(bb);
                            if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) { cc = aa; fc = fa; }
                        }
                        if (std::isnan(pt)) pt = bb;   // iterations exhausted
                    }
                }
                if (side == 0) upper[i] = pt;
                else lower[i] = pt;
                evals[i] = (side == 0 ? 0 : evals[i]) + n;
            }
        }
        return 0;
    } catch (const std::exception& ex) {
        g_orc_err = ex.what();
        return 1;
    }
}

/* CcdInterface::evaluateProfileModel (CcdInterface.cpp:527-607) with one clone per
 * point: the column is held at 0 starting from beta (:570), each point's fit starts
 * from that state (:572-577), values[i] = logLik (+ logPrior) of the fit. */
int orc_profile_likelihood(void* hv, const ccd_fit_args* a, const double* beta, int index, const double* points,
                           int n, int includePenalty, double* values) {
    try {
        auto* h = static_cast<orc_handle*>(hv);
        const int J = h->data.J;
        std::vector<double> start(beta, beta + J);
        if (h->data.hasOffset) start[0] = 1.0;
        auto evaluate = [&](std::vector<double>& b, double x) {
            orc::CCD ccd(h->data, *h->ms);
            orc::configure(ccd, a);
            b[index] = x;
            ccd.setBeta(b.data());
            ccd.fixBeta[index] = 1;
            ccd.update(*a);
            ccd.fixBeta[index] = 0;
            b.assign(ccd.beta.begin(), ccd.beta.end());
            return ccd.logLik() + (includePenalty ? ccd.logPrior() : 0.0);
        };
        evaluate(start, 0.0);
        for (int i = 0; i < n; ++i) {
            std::vector<double> b = start;
            values[i] = evaluate(b, points[i]);
        }
        return 0;
    } catch (const std::exception& ex) {
        g_orc_err = ex.what();
        return 1;
    }
}

double orc_prior_delta(int type, double variance, double g, double h, double beta) {
    orc::Prior p;
    p.type = type;
    p.variance = variance;
    return p.delta(g, h, beta);
}

double orc_time_cycles(void* hv, const ccd_fit_args* a, double* beta, int cycles) {
    auto* h = static_cast<orc_handle*>(hv);
    orc::CCD ccd(h->data, *h->ms);
    orc::configure(ccd, a);
    std::vector<double> b(beta, beta + h->data.J);
    if (h->data.hasOffset) b[0] = 1.0;
    ccd.setBeta(b.data());
    if (a->weights) ccd.setWeights(a->weights);
    ccd.initialBound = a->initial_bound;
    ccd.lazy();   // set-up outside the timed region
    const auto t0 = std::chrono::steady_clock::now();
    ccd.findMode(cycles, CCD_CONV_MITTAL, -1.0);   // eps <= 0: never converges early
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::copy(ccd.beta.begin(), ccd.beta.end(), beta);
    return s;
}

int orc_cv_assign_folds(int64_t K, const int32_t* pid, int32_t fold, int32_t selector, int64_t seed,
                        int32_t* rowFold) {
    orc::assignFolds(K, pid, fold, selector, seed, rowFold);
    return 0;
}

/* Grid-search cross-validation, single-threaded reference semantics
 * (AbstractCrossValidationDriver.cpp:131-248): one CCD runs every fit in task order;
 * cold start resets beta before each fit (:201-203), a warm start keeps the previous
 * fit's beta; a failed fit resets beta either way (:224-226). */
int orc_cross_validate(void* hv, const ccd_fit_args* a, const ccd_cv_args* cv, double* gridValues,
                       double* predLL, double* beta, ccd_fit_result* r, double* bestVariance) {
    try {
        auto* h = static_cast<orc_handle*>(hv);
        const auto& d = h->data;
        std::vector<int32_t> rowFold(d.K);
        std::vector<int32_t> pidIn(d.K);
        for (int64_t k = 0; k < d.K; ++k) pidIn[k] = d.pid[k];
        orc::assignFolds(d.K, pidIn.data(), cv->fold, cv->selector, cv->seed, rowFold.data());
        orc::CCD ccd(d, *h->ms);
        orc::configure(ccd, a);
        double bestV = 0.0, bestValue = -std::numeric_limits<double>::infinity();
        bool haveBest = false;
        for (int step = 0; step < cv->grid_size; ++step) {
            double point;
            if (cv->grid_size == 1) point = cv->upper_limit;
            else {
                const double stepSize = (std::log(cv->upper_limit) - std::log(cv->lower_limit)) / (cv->grid_size - 1);
                point = std::exp(std::log(cv->lower_limit) + step * stepSize);
            }
            for (int j = 0; j < d.J; ++j) ccd.priors[j].variance = point;
            double total = 0.0;
            int count = 0;
            for (int f = 0; f < cv->fold; ++f) {
                std::vector<double> w(d.K);
                for (int64_t k = 0; k < d.K; ++k) w[k] = (rowFold[k] == f) ? 0.0 : 1.0;
                ccd.setWeights(w.data());
                if (!cv->warm_start) ccd.resetBeta();
                ccd.update(*a);
                double v = std::numeric_limits<double>::quiet_NaN();
                if (ccd.lastFlag == CCD_SUCCESS) {
                    for (auto& x : w) x = 1.0 - x;
                    v = ccd.predLogLik(w.data());
                } else {
                    ccd.resetBeta();
                }
                if (predLL) predLL[step * cv->fold + f] = v;
                if (v == v) { total += v; ++count; }
            }
            const double value = total / count;
            gridValues[step] = value;
            if (!haveBest || value > bestValue) { bestValue = value; bestV = point; haveBest = true; }
        }
        // resetForOptimal + final fit (AbstractCrossValidationDriver.cpp:31-42; CcdInterface.cpp:777-789)
        ccd.setWeights(nullptr);
        for (int j = 0; j < d.J; ++j) ccd.priors[j].variance = bestV;
        ccd.resetBeta();
        ccd.update(*a);
        std::copy(ccd.beta.begin(), ccd.beta.end(), beta);
        r->return_flag = ccd.lastFlag;
        r->iterations = ccd.lastIterations;
        r->find_mode_calls = ccd.findModeCalls;
        r->log_likelihood = ccd.lastLogLik;
        r->log_prior = ccd.lastLogPrior;
        *bestVariance = bestV;
        return 0;
    } catch (const std::exception& ex) {
        g_orc_err = ex.what();
        return 1;
    }
}

}  // extern "C"
