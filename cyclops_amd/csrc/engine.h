// engine.h -- host-side engine of the MI355X CCD library.
//
// EngineBase is the device counterpart of bsccs::AbstractModelSpecifics
// (src/cyclops/engine/AbstractModelSpecifics.h:36-122): one method per virtual.
// EngineT<Real> holds the per-fit state in HBM; DevData<Real> holds the design
// matrix, shared by clones (ModelSpecifics.hpp:98-101 clones over the same data).
#pragma once

#include "common.h"
#include "../../include/cyclops_amd.h"

#include <memory>
#include <mutex>
#include <vector>

namespace ccd {

// Device copy of the data (column store in the reference's layout).
template <typename Real>
struct DevData {
    int model = 0;
    int device = 0;
    int64_t K = 0;
    int J = 0;
    bool hasOffset = false;
    int64_t nnz = 0;
    std::vector<int32_t> fmt;
    std::vector<int64_t> colPtr, densePtr;
    std::vector<double> y, time;          // host copies (pid construction, fixed terms)
    std::vector<int32_t> strat;           // renumbered strata per row
    std::vector<int32_t> stratStart;      // [nStrata + 1]
    int64_t nStrata = 0;
    // device
    int32_t* dRows = nullptr;
    Real* dVals = nullptr;
    Real* dDense = nullptr;
    Real* dY = nullptr;
    int32_t* dFmt = nullptr;
    int64_t* dColPtr = nullptr;
    int64_t* dDensePtr = nullptr;
    int32_t* dStrat = nullptr;
    int32_t* dStratStart = nullptr;
    double bytes = 0.0;
    // Row-major copy of the INDICATOR/SPARSE entries, built on first use: the entries
    // of row k are [dRowPtr[k], dRowPtr[k+1]), ascending in column (warm-started xb).
    std::mutex rowsMu;
    int64_t* dRowPtr = nullptr;
    int32_t* dRowCol = nullptr;
    Real* dRowVal = nullptr;
    void buildRowMajor(hipStream_t s);

    ~DevData();
};

struct KernelCounters {
    double launches = 0, deviceMs = 0, bytes = 0, coordinates = 0;
};

// A batch of fits advanced by the caller a given number of CCD cycles at a time
// (ccd_batch_begin / ccd_batch_step / ccd_batch_end; resident.inl).
class SteppedBatch {
public:
    virtual ~SteppedBatch() = default;
    // Runs up to `cycles` more cycles of every unfinished fit; doneOut[f] (optional)
    // is 1 once fit f has converged or stopped. Returns the fits still running.
    virtual int step(int cycles, int32_t* doneOut) = 0;
    // Runs the remaining cycles, applies update()'s bound retries and returns the
    // coefficients and results (as ccd_fit_batch).
    virtual void finish(double* betas, ccd_fit_result* r) = 0;
    virtual int fits() const = 0;
    virtual int cycles() const = 0;
    virtual void layout(int32_t* fitsPerTeam, int32_t* workgroupsPerTeam, int32_t* teams) const = 0;
};

class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual int model() const = 0;
    virtual int precision() const = 0;
    virtual int64_t K() const = 0;
    virtual int J() const = 0;
    virtual int64_t N() const = 0;
    virtual bool hasOffset() const = 0;
    virtual int format(int j) const = 0;
    virtual int64_t nonZeros(int j) const = 0;
    virtual int device() const = 0;
    virtual hipStream_t stream() const = 0;

    virtual void setWeights(const double* w, bool useW) = 0;
    virtual void fixedTerms(bool useW) = 0;
    virtual void zeroXBeta() = 0;
    virtual void axpyXBeta(double beta, int j) = 0;
    virtual void computeXBeta(const double* beta) = 0;
    virtual void remainingStats(bool useW) = 0;
    virtual void numerator(int j, bool useW) = 0;
    virtual void gradHess(int j, double* g, double* h, bool useW) = 0;
    // gradHess with its reduction buffers in the static scratch of the device
    virtual void gradHessNoAlloc(int j, double* g, double* h, bool useW) = 0;
    virtual void updateXBeta(double delta, int j, bool useW) = 0;
    virtual double logLik(bool useW) = 0;
    virtual double predLogLik(const double* w) = 0;
    virtual void predEstimates(double* out, const double* w) = 0;
    virtual double gradientObjective(bool useW) = 0;
    virtual void getXBeta(double* out) = 0;
    virtual void saveXBeta() = 0;
    virtual void getXBetaSave(double* out) = 0;
    virtual EngineBase* clone() const = 0;
    virtual double footprintBytes() const = 0;
    virtual void getStrata(int32_t* out) const = 0;   // renumbered strata per row

    // Resident / batched fits (resident.hip).
    virtual void fitResident(const ccd_fit_args& a, double* beta, ccd_fit_result& r) = 0;
    // autotune: time candidate team shapes on the first batch of each size (callers
    // issuing many small batches of varying size, such as profiling, pass false)
    // predHeldOut [n_fits] (optional, needs row_fold): each fit's predictive
    // log-likelihood on its held-out rows, formed on the device from the run's xb
    virtual void fitBatch(const ccd_fit_args& a, const ccd_batch_args& b, double* betas,
                          ccd_fit_result* r, bool autotune = true, double* predHeldOut = nullptr) = 0;
    // The same fits as fitBatch, stepped by the caller; the engine runs no other
    // resident fit until the returned batch has finished or is deleted.
    virtual SteppedBatch* beginStepped(const ccd_fit_args& a, const ccd_batch_args& b, const double* betas0) = 0;
    virtual bool steppedOpen() const = 0;

    // Fisher information of columns idx[0..m) at beta, unit row weights (the engine's
    // weights are reset): out[m][m]. fisher.inl.
    virtual void fisherInformation(const double* beta, const int32_t* idx, int m, double* out) = 0;

    KernelCounters counters;
    bool timing = false;
};

EngineBase* createEngine(const ccd_data_desc& d, int device);

// Device address of the no-allocation scratch of `device`, which must be current;
// looked up once per device (engine.hip).
void* noAllocScratch(int device);

// PrecisionType of a data description (FP64 = 0, FP32 = 1, or the bit counts 64 / 32)
// as CCD_FP64 / CCD_FP32; anything else is refused.
inline int precisionOf(int p) {
    if (p == CCD_FP64 || p == CCD_FP64_BITS) return CCD_FP64;
    if (p == CCD_FP32 || p == CCD_FP32_BITS) return CCD_FP32;
    throw std::invalid_argument("precision must be CCD_FP64 (0) or CCD_FP32 (1); 64 and 32 are accepted too");
}

// What a one-fit persistent launch over `d` needs on `device`: its workgroups, the
// fraction of the device's co-resident capacity they take, and the HBM of its
// workspace beyond the engine itself (admission control of the server).
struct ResidentNeeds {
    int workgroups;
    double capacityFraction;
    double workspaceBytes;
};
ResidentNeeds residentNeeds(const ccd_data_desc& d, int device, int maxWorkgroups = 0);

// Host CCD driver over an EngineBase: CyclicCoordinateDescent::update/findMode
// (CyclicCoordinateDescent.cpp:586-615, 909-1216). This is the plain per-fit path.
void fitPlain(EngineBase& eng, const ccd_fit_args& a, double* beta, ccd_fit_result& r);

// Single fits of several datasets (engines of one model and precision on one
// device) as shared persistent launches (resident.inl).
void fitMulti(EngineBase* const* engs, int n, const ccd_fit_args* args, double* const* betas, ccd_fit_result* res);

// Profile-likelihood limits (CcdInterface::profileModel, CcdInterface.cpp:262-434): all
// bracket and root searches advance together, one batched fit per round (profile.hip).
void profileModel(EngineBase& eng, const ccd_fit_args& a, const double* beta, const int32_t* idx, int m,
                  double threshold, bool includePenalty, bool overrideNoRegularization, double* lower,
                  double* upper, int32_t* evaluations);

// Profile log-likelihood at grid points (CcdInterface::evaluateProfileModel,
// CcdInterface.cpp:527-607): one batched fit over the points (profile.hip).
void profileLikelihood(EngineBase& eng, const ccd_fit_args& a, const double* beta, int index,
                       const double* points, int n, bool includePenalty, double* values);

// Fold assignment of CrossValidationSelector (drivers/CrossValidationSelector.cpp:19-189).
void assignFolds(int64_t K, const int32_t* pid, int fold, int selector, int64_t seed, int32_t* rowFold);

// Prior step (CovariatePrior.h:186-188, 240-273, 410-415), shared by host drivers.
double priorDelta(int type, double variance, double g, double h, double beta);
double priorLogDensity(int type, double variance, double beta);

// Message returned by ccd_last_error() on the calling thread.
void setLastError(const std::string& msg);

// Device memory arena of the calling thread. While one is installed, engines
// created on this thread carve their device buffers from it (frees are no-ops; the
// owner resets `used` when the engines are gone), run on its stream and use its
// pinned host buffer, so they call no hipMalloc / hipFree / hipHostMalloc, which
// synchronise the whole device and would wait for other threads' kernels.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    hipStream_t stream = nullptr;
    double* host = nullptr;   // >= 64 doubles, pinned
};
void setThreadArena(Arena* a);   // nullptr: plain hipMalloc / hipFree
Arena* threadArena();
// Most workgroups a single fit's persistent launch may take on the calling thread (0:
// no limit beyond the device). The server sets it on its workers, so that no one
// request holds the whole device's co-resident capacity.
void setThreadMaxWorkgroups(int g);
int threadMaxWorkgroups();

}  // namespace ccd
