// engine.hip -- device engine behind the C ABI (include/cyclops_amd.h).
//
// EngineT<Real> implements every AbstractModelSpecifics virtual
// (src/cyclops/engine/AbstractModelSpecifics.h:36-122) with the kernels in
// kernels.h. Per-row state (xBeta, offsExpXBeta, weights) is kept in the bulk
// precision Real; per-pid state (denominators, risk-set sums, event counts) is
// kept in FP64 so the Breslow risk-set scans and the strata sums are exact to
// double rounding even when Real = float.
#include "engine.h"
#include "kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>

namespace ccd {

// ---------------------------------------------------------------------------
// allocation helpers
// ---------------------------------------------------------------------------
static thread_local Arena* g_arena = nullptr;
void setThreadArena(Arena* a) { g_arena = a; }
Arena* threadArena() { return g_arena; }
static thread_local int g_maxWG = 0;
void setThreadMaxWorkgroups(int g) { g_maxWG = g > 0 ? g : 0; }
int threadMaxWorkgroups() { return g_maxWG; }

static bool inArena(const void* p) {
    const char* c = static_cast<const char*>(p);
    return g_arena && c >= g_arena->base && c < g_arena->base + g_arena->cap;
}

template <typename T>
static T* dalloc(size_t n, double& bytes) {
    T* p = nullptr;
    const size_t sz = std::max<size_t>(n, 1) * sizeof(T);
    if (g_arena) {
        const size_t off = (g_arena->used + 255) & ~size_t(255);
        if (off + sz <= g_arena->cap) {
            p = reinterpret_cast<T*>(g_arena->base + off);
            g_arena->used = off + sz;
            bytes += (double)sz;
            return p;
        }
    }
    CCD_HIP(hipMalloc(reinterpret_cast<void**>(&p), sz));
    bytes += (double)sz;
    return p;
}
template <typename T>
static void dfree(T*& p) {
    if (p && !inArena(p)) (void)hipFree(p);
    p = nullptr;
}
template <typename T>
static void h2d(T* d, const T* h, size_t n, hipStream_t s) {
    if (n) CCD_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename Real>
static void h2dConvert(Real* d, const double* h, size_t n, hipStream_t s) {
    if constexpr (std::is_same<Real, double>::value) {
        h2d(d, h, n, s);
        CCD_HIP(hipStreamSynchronize(s));
    } else {
        std::vector<Real> tmp(h, h + n);
        h2d(d, tmp.data(), n, s);
        CCD_HIP(hipStreamSynchronize(s));
    }
}

static int gridFor(int64_t n, int cap = 4096) {
    int64_t g = (n + BLOCK - 1) / BLOCK;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

template <typename Real>
DevData<Real>::~DevData() {
    (void)hipSetDevice(device);
    dfree(dRows);
    dfree(dVals);
    dfree(dDense);
    dfree(dY);
    dfree(dFmt);
    dfree(dColPtr);
    dfree(dDensePtr);
    dfree(dStrat);
    dfree(dStratStart);
    dfree(dRowPtr);
    dfree(dRowCol);
    dfree(dRowVal);
}

// Counting sort of the column-major entries by row, columns visited in order, so
// every row's entries stay ascending in column.
template <typename Real>
void DevData<Real>::buildRowMajor(hipStream_t s) {
    std::lock_guard<std::mutex> lock(rowsMu);
    if (dRowPtr) return;
    std::vector<int32_t> rows(nnz);
    std::vector<Real> vals;
    if (nnz) CCD_HIP(hipMemcpyAsync(rows.data(), dRows, nnz * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (dVals && nnz) {
        vals.resize(nnz);
        CCD_HIP(hipMemcpyAsync(vals.data(), dVals, nnz * sizeof(Real), hipMemcpyDeviceToHost, s));
    }
    CCD_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> ptr(K + 1, 0);
    for (int j = 0; j < J; ++j)
        if (fmt[j] == F_INDICATOR || fmt[j] == F_SPARSE)
            for (int64_t e = colPtr[j]; e < colPtr[j + 1]; ++e) ++ptr[rows[e] + 1];
    for (int64_t k = 0; k < K; ++k) ptr[k + 1] += ptr[k];
    const int64_t n = ptr[K];
    std::vector<int64_t> next(ptr.begin(), ptr.end() - 1);
    std::vector<int32_t> col(n);
    std::vector<Real> val(n);
    for (int j = 0; j < J; ++j) {
        if (!(fmt[j] == F_INDICATOR || fmt[j] == F_SPARSE)) continue;
        for (int64_t e = colPtr[j]; e < colPtr[j + 1]; ++e) {
            const int64_t q = next[rows[e]]++;
            col[q] = j;
            val[q] = fmt[j] == F_SPARSE ? vals[e] : Real(1);
        }
    }
    double b = 0.0;
    dRowPtr = dalloc<int64_t>(K + 1, b);
    dRowCol = dalloc<int32_t>(n, b);
    dRowVal = dalloc<Real>(n, b);
    h2d(dRowPtr, ptr.data(), K + 1, s);
    h2d(dRowCol, col.data(), n, s);
    h2d(dRowVal, val.data(), n, s);
    CCD_HIP(hipStreamSynchronize(s));
    bytes += b;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point (ccd_grad_hess_noalloc): one copy per
// device, held until the process ends. Layout: completion counter at 0, result
// (g, h) at 256, per-block partials from 512. In host code `scratch` names a host
// stand-in, not device memory: only the address hipGetSymbolAddress returns for the
// current device is ever passed to a kernel or a copy.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[CCD_NOALLOC_SCRATCH_BYTES];

__global__ void kZeroCounter(unsigned* c) {
    if (threadIdx.x == 0) *c = 0u;
}

void* noAllocScratch(int device) {
    static std::mutex mu;
    static std::vector<void*> ptrs;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptrs.size() <= device) ptrs.resize(device + 1, nullptr);
    if (!ptrs[device]) {   // once per device, with the device current (the caller set it)
        int cur = -1;
        CCD_HIP(hipGetDevice(&cur));
        if (cur != device) throw std::logic_error("noAllocScratch: device is not current");
        void* p = nullptr;
        CCD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch array");
        ptrs[device] = p;
    }
    return ptrs[device];
}

// Validate the description and copy it to HBM.
template <typename Real>
static std::shared_ptr<DevData<Real>> makeDevData(const ccd_data_desc& d, int device) {
    auto dd = std::make_shared<DevData<Real>>();
    dd->device = device;
    dd->model = d.model_type;
    dd->K = d.K;
    dd->J = d.J;
    dd->hasOffset = d.has_offset != 0;
    if (d.K <= 0 || d.J <= 0) throw std::invalid_argument("K and J must be positive");
    if (d.K >= (int64_t)INT32_MAX) throw std::invalid_argument("K must fit int32 row ids");
    // other model types: the factory's nullptr (AbstractModelSpecifics.cpp:128-154), which
    // RcppCyclopsInterface.cpp:890-892 reports as "Invalid model type." -> CCD_ERR_UNSUPPORTED
    if (!(d.model_type == M_POISSON || d.model_type == M_LOGISTIC || d.model_type == M_CLR ||
          d.model_type == M_COX))
        throw std::logic_error("Invalid model type.");
    if (!d.format || !d.col_ptr || !d.y) throw std::invalid_argument("format, col_ptr and y are required");
    const int64_t K = d.K;
    const int J = d.J;
    dd->fmt.assign(d.format, d.format + J);
    dd->colPtr.assign(d.col_ptr, d.col_ptr + J + 1);
    dd->densePtr.assign(J + 1, 0);
    bool anySparse = false;
    if (dd->colPtr[0] != 0) throw std::invalid_argument("col_ptr[0] must be 0");
    for (int j = 0; j < J; ++j) {
        const int f = dd->fmt[j];
        if (f < F_DENSE || f > F_INTERCEPT) throw std::invalid_argument("bad column format");
        if (dd->colPtr[j + 1] < dd->colPtr[j]) throw std::invalid_argument("col_ptr not monotone");
        if ((f == F_DENSE || f == F_INTERCEPT) && dd->colPtr[j + 1] != dd->colPtr[j])
            throw std::invalid_argument("DENSE/INTERCEPT columns hold no row ids");
        if (f == F_DENSE && (!d.dense_ptr || !d.dense_values))
            throw std::invalid_argument("DENSE column without dense_ptr/dense_values");
        if (f == F_SPARSE) anySparse = true;
    }
    if (anySparse && !d.values) throw std::invalid_argument("SPARSE column without values");
    dd->nnz = dd->colPtr[J];
    // every row id in range and sorted within its column (out-of-range ids would
    // be out-of-bounds device accesses, unsorted ids would break the run logic)
    for (int j = 0; j < J; ++j) {
        int64_t prev = -1;
        for (int64_t e = dd->colPtr[j]; e < dd->colPtr[j + 1]; ++e) {
            const int64_t r = d.rows[e];
            if (r < 0 || r >= K) throw std::invalid_argument("row id out of range in column " + std::to_string(j));
            if (r <= prev) throw std::invalid_argument("row ids not strictly increasing in column " + std::to_string(j));
            prev = r;
        }
    }
    dd->y.assign(d.y, d.y + K);
    if (d.model_type == M_COX) {
        if (!d.time) throw std::invalid_argument("COX requires time");
        dd->time.assign(d.time, d.time + K);
    }
    // strata: RcppModelData.cpp:878-899 renumbering
    dd->strat.assign(K, 0);
    if (d.model_type == M_CLR || d.model_type == M_COX) {
        if (!d.pid && d.model_type == M_CLR) throw std::invalid_argument("CLR requires pid (strata)");
        int cur = 0;
        int32_t last = d.pid ? d.pid[0] : 0;
        for (int64_t k = 0; k < K; ++k) {
            const int32_t v = d.pid ? d.pid[k] : 0;
            if (v != last) { ++cur; last = v; }
            dd->strat[k] = cur;
        }
        dd->nStrata = cur + 1;
        dd->stratStart.assign(dd->nStrata + 1, 0);
        int64_t s = -1;
        for (int64_t k = 0; k < K; ++k)
            if (dd->strat[k] != s) { s = dd->strat[k]; dd->stratStart[s] = (int32_t)k; }
        dd->stratStart[dd->nStrata] = (int32_t)K;
    }
    CCD_HIP(hipSetDevice(device));
    hipStream_t s;
    const bool ownStream = g_arena == nullptr;
    if (ownStream) CCD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    else s = g_arena->stream;
    double& B = dd->bytes;
    dd->dRows = dalloc<int32_t>(dd->nnz, B);
    h2d(dd->dRows, d.rows, dd->nnz, s);
    if (anySparse) {
        dd->dVals = dalloc<Real>(dd->nnz, B);
        h2dConvert(dd->dVals, d.values, dd->nnz, s);
    }
    int64_t denseTotal = 0;
    for (int j = 0; j < J; ++j) {
        dd->densePtr[j] = denseTotal;
        if (dd->fmt[j] == F_DENSE) denseTotal += K;
    }
    dd->densePtr[J] = denseTotal;
    if (denseTotal) {
        std::vector<double> packed(denseTotal);
        for (int j = 0; j < J; ++j)
            if (dd->fmt[j] == F_DENSE)
                std::memcpy(packed.data() + dd->densePtr[j], d.dense_values + d.dense_ptr[j], K * sizeof(double));
        dd->dDense = dalloc<Real>(denseTotal, B);
        h2dConvert(dd->dDense, packed.data(), denseTotal, s);
    }
    dd->dY = dalloc<Real>(K, B);
    h2dConvert(dd->dY, d.y, K, s);
    dd->dFmt = dalloc<int32_t>(J, B);
    h2d(dd->dFmt, dd->fmt.data(), J, s);
    dd->dColPtr = dalloc<int64_t>(J + 1, B);
    h2d(dd->dColPtr, dd->colPtr.data(), J + 1, s);
    dd->dDensePtr = dalloc<int64_t>(J + 1, B);
    h2d(dd->dDensePtr, dd->densePtr.data(), J + 1, s);
    if (dd->model == M_CLR || dd->model == M_COX) {
        dd->dStrat = dalloc<int32_t>(K, B);
        h2d(dd->dStrat, dd->strat.data(), K, s);
        dd->dStratStart = dalloc<int32_t>(dd->nStrata + 1, B);
        h2d(dd->dStratStart, dd->stratStart.data(), dd->nStrata + 1, s);
    }
    CCD_HIP(hipStreamSynchronize(s));
    if (ownStream) CCD_HIP(hipStreamDestroy(s));
    return dd;
}

// ---------------------------------------------------------------------------
// dispatch helpers
// ---------------------------------------------------------------------------
#define CCD_FMT_CASE(F, ...) \
    case F: {                 \
        constexpr int FMT = F; \
        __VA_ARGS__;           \
    } break;
#define CCD_DISPATCH_FMT(f, ...)                     \
    switch (f) {                                     \
        CCD_FMT_CASE(F_DENSE, __VA_ARGS__)           \
        CCD_FMT_CASE(F_SPARSE, __VA_ARGS__)          \
        CCD_FMT_CASE(F_INDICATOR, __VA_ARGS__)       \
        CCD_FMT_CASE(F_INTERCEPT, __VA_ARGS__)       \
        default: throw std::logic_error("bad format"); \
    }

// ---------------------------------------------------------------------------
// EngineT
// ---------------------------------------------------------------------------
template <typename Real>
class EngineT : public EngineBase {
public:
    EngineT(std::shared_ptr<DevData<Real>> data) : D(std::move(data)) {
        CCD_HIP(hipSetDevice(D->device));
        arena = g_arena;
        if (arena) s = arena->stream;
        else CCD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        CCD_HIP(hipEventCreate(&ev0));
        CCD_HIP(hipEventCreate(&ev1));
        const int64_t K = D->K;
        capN = (D->model == M_COX) ? K : (D->model == M_CLR ? D->nStrata : K);
        xb = dalloc<Real>(K, bytes);
        xbSave = dalloc<Real>(K, bytes);
        ex = dalloc<Real>(K, bytes);
        kw = dalloc<Real>(K, bytes);
        predW = dalloc<double>(K, bytes);
        if (D->model == M_CLR || D->model == M_COX) {
            denom = dalloc<double>(capN + 1, bytes);
            nW = dalloc<double>(capN + 1, bytes);
            R1 = dalloc<double>(capN + 1, bytes);
            R2 = dalloc<double>(capN + 1, bytes);
        }
        if (D->model == M_COX) {
            Dacc = dalloc<double>(capN + 1, bytes);
            pidArr = dalloc<int32_t>(K, bytes);
            pstart = dalloc<int32_t>(capN + 1, bytes);
            pseg = dalloc<int32_t>(capN + 1, bytes);
            scanParts = dalloc<Seg2>((capN + TILE - 1) / TILE + 1, bytes);
        }
        XjY = dalloc<double>(D->J, bytes);
        parts = dalloc<Partial>(std::max(MAX_RED_BLOCKS, MAX_GRAD_BLOCKS), bytes);
        counter = dalloc<unsigned>(1, bytes);
        dOut = dalloc<double>(8, bytes);
        if (arena) hOut = arena->host;
        else CCD_HIP(hipHostMalloc(reinterpret_cast<void**>(&hOut), 8 * sizeof(double), hipHostMallocDefault));
        CCD_HIP(hipMemsetAsync(counter, 0, sizeof(unsigned), s));
        CCD_HIP(hipMemsetAsync(xb, 0, K * sizeof(Real), s));
        CCD_HIP(hipMemsetAsync(xbSave, 0, K * sizeof(Real), s));
        hXjY.assign(D->J, 0.0);
        setWeights(nullptr, false);
        CCD_HIP(hipStreamSynchronize(s));
    }

    ~EngineT() override {
        (void)hipSetDevice(D->device);
        (void)hipStreamSynchronize(s);
        releaseResident();
        dfree(xb); dfree(xbSave); dfree(ex); dfree(kw); dfree(predW);
        dfree(denom); dfree(nW); dfree(R1); dfree(R2); dfree(Dacc);
        dfree(pidArr); dfree(pstart); dfree(pseg); dfree(scanParts);
        dfree(XjY); dfree(parts); dfree(counter); dfree(dOut);
        if (hOut && !arena) (void)hipHostFree(hOut);
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
        if (!arena) (void)hipStreamDestroy(s);
    }

    int model() const override { return D->model; }
    int precision() const override { return sizeof(Real) == 8 ? CCD_FP64 : CCD_FP32; }
    int64_t K() const override { return D->K; }
    int J() const override { return D->J; }
    int64_t N() const override { return Ncur; }
    bool hasOffset() const override { return D->hasOffset; }
    int format(int j) const override { return D->fmt[j]; }
    int64_t nonZeros(int j) const override {   // CompressedDataMatrix.cpp:93-100
        const int f = D->fmt[j];
        return (f == F_DENSE || f == F_INTERCEPT) ? D->K : D->colPtr[j + 1] - D->colPtr[j];
    }
    int device() const override { return D->device; }
    hipStream_t stream() const override { return s; }
    double footprintBytes() const override { return D->bytes + bytes; }

    EngineBase* clone() const override { return new EngineT<Real>(D); }
    void getStrata(int32_t* out) const override { std::copy(D->strat.begin(), D->strat.end(), out); }

    Col<Real> col(int j) const {
        const int f = D->fmt[j];
        const int64_t a = D->colPtr[j];
        switch (f) {
        case F_INDICATOR: return Col<Real>{D->dRows + a, nullptr, D->colPtr[j + 1] - a};
        case F_SPARSE: return Col<Real>{D->dRows + a, D->dVals + a, D->colPtr[j + 1] - a};
        case F_DENSE: return Col<Real>{nullptr, D->dDense + D->densePtr[j], D->K};
        default: return Col<Real>{nullptr, nullptr, D->K};
        }
    }

    // ---- setWeights (ModelSpecifics.hpp:427-484) ----------------------------
    void setWeights(const double* w, bool useW) override {
        CCD_HIP(hipSetDevice(D->device));
        const int64_t K = D->K;
        weighted = useW && w != nullptr;
        if (weighted) hKW.assign(w, w + K);
        else hKW.assign(K, 1.0);
        h2dConvert(kw, hKW.data(), K, s);
        if (D->model == M_LOGISTIC || D->model == M_POISSON) {
            Ncur = K;
        } else if (D->model == M_CLR) {
            Ncur = D->nStrata;
            std::vector<double> n(Ncur + 1, 0.0);
            for (int64_t k = 0; k < K; ++k) n[D->strat[k]] += D->y[k] * hKW[k];
            h2d(nW, n.data(), Ncur + 1, s);
            CCD_HIP(hipStreamSynchronize(s));
        } else {
            setPidForAccumulation(weighted ? hKW.data() : nullptr);
        }
    }

    // Breslow pid construction (ModelSpecifics.hpp:1809-1868): tied events in a
    // stratum share a pid; a new stratum starts a new risk-set segment. Rows with
    // zero weight carry no weight anywhere; they are attached to the pid of the
    // preceding kept row so pids stay non-decreasing along rows.
    void setPidForAccumulation(const double* w) {
        const int64_t K = D->K;
        std::vector<int32_t> pid(K, -1);
        int64_t index = 0;
        while (w != nullptr && index < K && w[index] == 0.0) ++index;
        if (index == K) throw std::invalid_argument("all weights are zero");
        std::vector<int32_t> ps;
        int lastStrat = D->strat[index];
        double lastTime = D->time[index];
        double lastEvent = D->y[index];
        int cur = 0;
        pid[index] = 0;
        ps.push_back(D->strat[index]);
        for (int64_t k = index + 1; k < K; ++k) {
            if (w == nullptr || w[k] != 0.0) {
                if (D->strat[k] != lastStrat) {
                    ++cur;
                    ps.push_back(D->strat[k]);
                    lastStrat = D->strat[k];
                } else if (!(lastEvent == 1.0 && lastTime == D->time[k] && lastEvent == D->y[k])) {
                    ++cur;
                    ps.push_back(D->strat[k]);
                }
                lastTime = D->time[k];
                lastEvent = D->y[k];
                pid[k] = cur;
            }
        }
        Ncur = cur + 1;
        int32_t last = 0;
        for (int64_t k = 0; k < K; ++k) {
            if (pid[k] < 0) pid[k] = last;
            else last = pid[k];
        }
        std::vector<int32_t> st(Ncur + 1, 0);
        int32_t prev = -1;
        for (int64_t k = 0; k < K; ++k)
            if (pid[k] != prev) { prev = pid[k]; st[prev] = (int32_t)k; }
        st[0] = 0;
        st[Ncur] = (int32_t)K;
        std::vector<double> n(Ncur + 1, 0.0);
        for (int64_t k = 0; k < K; ++k) n[pid[k]] += D->y[k] * (w ? w[k] : 1.0);
        ps.push_back(-1);
        hPid = pid;
        h2d(pidArr, pid.data(), K, s);
        h2d(pstart, st.data(), Ncur + 1, s);
        h2d(pseg, ps.data(), Ncur + 1, s);
        h2d(nW, n.data(), Ncur + 1, s);
        CCD_HIP(hipStreamSynchronize(s));
    }

    // ---- fixed terms (ModelSpecifics.hpp:486-524, 570-587) ------------------
    void fixedTerms(bool useW) override {
        CCD_HIP(hipSetDevice(D->device));
        const bool W = useW && weighted;
        const int grid = std::min(D->J, 4096);
        if (W)
            kXjY<Real, true><<<grid, BLOCK, 0, s>>>(D->J, D->K, D->dFmt, D->dColPtr, D->dDensePtr, D->dRows,
                                                    D->dVals, D->dDense, D->dY, kw, XjY);
        else
            kXjY<Real, false><<<grid, BLOCK, 0, s>>>(D->J, D->K, D->dFmt, D->dColPtr, D->dDensePtr, D->dRows,
                                                     D->dVals, D->dDense, D->dY, kw, XjY);
        CCD_HIP(hipGetLastError());
        CCD_HIP(hipMemcpyAsync(hXjY.data(), XjY, D->J * sizeof(double), hipMemcpyDeviceToHost, s));
        fixedTerm = 0.0;
        if (D->model == M_POISSON) {
            for (int64_t k = 0; k < D->K; ++k) {
                double c = 0.0;
                for (int i = 2; i <= (int)D->y[k]; ++i) c -= std::log((double)i);
                fixedTerm += W ? c * hKW[k] : c;
            }
        }
        CCD_HIP(hipStreamSynchronize(s));
    }

    // ---- xBeta --------------------------------------------------------------
    void zeroXBeta() override {
        CCD_HIP(hipSetDevice(D->device));
        CCD_HIP(hipMemsetAsync(xb, 0, D->K * sizeof(Real), s));
        CCD_HIP(hipStreamSynchronize(s));
    }
    void axpyLaunch(double beta, int j) {
        if (beta == 0.0) return;
        const Col<Real> c = col(j);
        if (c.n == 0) return;
        CCD_DISPATCH_FMT(D->fmt[j], kAxpy<Real, FMT><<<gridFor(c.n), BLOCK, 0, s>>>(c, (Real)beta, xb));
        CCD_HIP(hipGetLastError());
    }
    void axpyXBeta(double beta, int j) override {
        CCD_HIP(hipSetDevice(D->device));
        axpyLaunch(beta, j);
        CCD_HIP(hipStreamSynchronize(s));
    }
    // CyclicCoordinateDescent::computeXBeta: zero, then axpy column by column in order.
    void computeXBeta(const double* beta) override {
        CCD_HIP(hipSetDevice(D->device));
        CCD_HIP(hipMemsetAsync(xb, 0, D->K * sizeof(Real), s));
        for (int j = 0; j < D->J; ++j) axpyLaunch(beta[j], j);
        CCD_HIP(hipStreamSynchronize(s));
    }
    void getXBeta(double* out) override { copyRows(xb, out); }
    void saveXBeta() override {
        CCD_HIP(hipSetDevice(D->device));
        CCD_HIP(hipMemcpyAsync(xbSave, xb, D->K * sizeof(Real), hipMemcpyDeviceToDevice, s));
        CCD_HIP(hipStreamSynchronize(s));
    }
    void getXBetaSave(double* out) override { copyRows(xbSave, out); }
    void copyRows(const Real* src, double* out) {
        CCD_HIP(hipSetDevice(D->device));
        std::vector<Real> tmp(D->K);
        CCD_HIP(hipMemcpyAsync(tmp.data(), src, D->K * sizeof(Real), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipStreamSynchronize(s));
        for (int64_t k = 0; k < D->K; ++k) out[k] = (double)tmp[k];
    }

    // ---- computeRemainingStatistics (ModelSpecifics.hpp:1588-1633) ----------
    void remainingStatsAsync(bool useW) {
        const bool W = useW && weighted;
        if (D->model == M_LOGISTIC || D->model == M_POISSON) {
            kExp<Real><<<gridFor(D->K), BLOCK, 0, s>>>(D->K, xb, ex);
        } else if (D->model == M_CLR) {
            kDenomByPid<Real, false, true><<<gridFor(Ncur), BLOCK, 0, s>>>(Ncur, D->dStratStart, xb, ex, kw, denom,
                                                                          nW, R1, R2);
        } else {
            if (W)
                kDenomByPid<Real, true, false><<<gridFor(Ncur), BLOCK, 0, s>>>(Ncur, pstart, xb, ex, kw, denom,
                                                                              nW, R1, R2);
            else
                kDenomByPid<Real, false, false><<<gridFor(Ncur), BLOCK, 0, s>>>(Ncur, pstart, xb, ex, kw, denom,
                                                                               nW, R1, R2);
            rescanAsync();
        }
        CCD_HIP(hipGetLastError());
    }
    void remainingStats(bool useW) override {
        CCD_HIP(hipSetDevice(D->device));
        remainingStatsAsync(useW);
        CCD_HIP(hipStreamSynchronize(s));
    }

    // Breslow risk sets: D = forward segmented scan of denom; R1, R2 = backward
    // segmented scans of nE/D and nE/D^2.
    void rescanAsync() {
        const int64_t n = Ncur;
        const int nb = (int)((n + TILE - 1) / TILE);
        ScanFwd f{n, denom, pseg, Dacc};
        kScanReduce<ScanFwd><<<nb, BLOCK, 0, s>>>(f, scanParts);
        kScanPartials<<<1, BLOCK, 0, s>>>(nb, scanParts);
        kScanApply<ScanFwd><<<nb, BLOCK, 0, s>>>(f, scanParts);
        ScanBwd b{n, nW, Dacc, pseg, R1, R2};
        kScanReduce<ScanBwd><<<nb, BLOCK, 0, s>>>(b, scanParts);
        kScanPartials<<<1, BLOCK, 0, s>>>(nb, scanParts);
        kScanApply<ScanBwd><<<nb, BLOCK, 0, s>>>(b, scanParts);
    }

    void numerator(int, bool) override {
        // Folded into gradHess: the grouped gradient pass forms the per-pid
        // numerators and their risk-set prefix sums in registers.
    }

    // ---- computeGradientAndHessian (ModelSpecifics.hpp:710-766, 891-1201) ----
    // Launch of the column pass of j; per-block partials, the completion counter (zero on
    // entry, left at zero) and the result (g, h) go to the given buffers.
    void gradHessAsync(int j, bool useW, Partial* parts, unsigned* counter, double* dOut) {
        const bool W = useW && weighted;
        const Col<Real> c = col(j);
        const bool indep = D->model == M_LOGISTIC || D->model == M_POISSON;
        int nb = (int)std::min<int64_t>(indep ? MAX_GRAD_BLOCKS : MAX_RED_BLOCKS,
                                        std::max<int64_t>(1, (c.n + TILE * 2 - 1) / (TILE * 2)));
        const int64_t per = (c.n + nb - 1) / nb;
        const double xjy = hXjY[j];
        const int m = D->model;
        if (m == M_LOGISTIC || m == M_POISSON) {
            CCD_DISPATCH_FMT(D->fmt[j], {
                if (m == M_LOGISTIC) {
                    if (W) kGradIndep<Real, M_LOGISTIC, FMT, true><<<nb, BLOCK, 0, s>>>(c, ex, kw, parts, counter, xjy, dOut, per);
                    else kGradIndep<Real, M_LOGISTIC, FMT, false><<<nb, BLOCK, 0, s>>>(c, ex, kw, parts, counter, xjy, dOut, per);
                } else {
                    if (W) kGradIndep<Real, M_POISSON, FMT, true><<<nb, BLOCK, 0, s>>>(c, ex, kw, parts, counter, xjy, dOut, per);
                    else kGradIndep<Real, M_POISSON, FMT, false><<<nb, BLOCK, 0, s>>>(c, ex, kw, parts, counter, xjy, dOut, per);
                }
            });
        } else {
            const int* pid = (m == M_CLR) ? D->dStrat : pidArr;
            const bool coxW = (m == M_COX) && W;
            CCD_DISPATCH_FMT(D->fmt[j], {
                if (coxW) kGradGrouped<Real, FMT, true><<<nb, BLOCK, 0, s>>>(c, ex, kw, pid, D->dStrat, R1, R2, parts, counter, xjy, dOut, per);
                else kGradGrouped<Real, FMT, false><<<nb, BLOCK, 0, s>>>(c, ex, kw, pid, D->dStrat, R1, R2, parts, counter, xjy, dOut, per);
            });
        }
        CCD_HIP(hipGetLastError());
    }
    void gradHess(int j, double* g, double* h, bool useW) override {
        if (nonZeros(j) == 0) { *g = 0.0; *h = 0.0; return; }
        CCD_HIP(hipSetDevice(D->device));
        if (timing) CCD_HIP(hipEventRecord(ev0, s));
        gradHessAsync(j, useW, parts, counter, dOut);
        if (timing) CCD_HIP(hipEventRecord(ev1, s));
        CCD_HIP(hipMemcpyAsync(hOut, dOut, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipStreamSynchronize(s));
        if (timing) {
            float ms = 0.f;
            CCD_HIP(hipEventElapsedTime(&ms, ev0, ev1));
            counters.deviceMs += ms;
            counters.launches += 1;
            counters.bytes += gradBytes(j);
        }
        *g = hOut[0];
        *h = hOut[1];
    }

    // gradHess with the reduction buffers in the library's static scratch of this device
    // (ccd_grad_hess_noalloc): allocates nothing.
    void gradHessNoAlloc(int j, double* g, double* h, bool useW) override {
        if (nonZeros(j) == 0) { *g = 0.0; *h = 0.0; return; }
        CCD_HIP(hipSetDevice(D->device));
        char* base = static_cast<char*>(noAllocScratch(D->device));
        unsigned* cnt = reinterpret_cast<unsigned*>(base);
        double* out = reinterpret_cast<double*>(base + 256);
        Partial* pr = reinterpret_cast<Partial*>(base + 512);
        const bool indep = D->model == M_LOGISTIC || D->model == M_POISSON;
        const size_t need = 512 + sizeof(Partial) * (size_t)(indep ? MAX_GRAD_BLOCKS : MAX_RED_BLOCKS);
        if (need > (size_t)CCD_NOALLOC_SCRATCH_BYTES)
            throw std::invalid_argument("the column pass needs " + std::to_string(need) +
                                        " bytes of scratch; the no-allocation entry point has " +
                                        std::to_string(CCD_NOALLOC_SCRATCH_BYTES));
        kZeroCounter<<<1, 64, 0, s>>>(cnt);   // the last call's contents stay in the array
        CCD_HIP(hipGetLastError());
        gradHessAsync(j, useW, pr, cnt, out);
        CCD_HIP(hipMemcpyAsync(hOut, out, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipStreamSynchronize(s));
        *g = hOut[0];
        *h = hOut[1];
    }

    // Compulsory bytes of one gradient pass over column j: row ids (+values),
    // the gathered per-row state, and for grouped models pid/segment ids and
    // the per-pid R1/R2 (FP64).
    double gradBytes(int j) const {
        const double n = (double)nonZeros(j);
        const int f = D->fmt[j];
        const double s_ = sizeof(Real);
        double b = n * s_;                                      // offsExpXBeta gather
        if (f == F_INDICATOR || f == F_SPARSE) b += n * 4.0;    // row ids
        if (f == F_SPARSE || f == F_DENSE) b += n * s_;         // values
        if (weighted && (D->model != M_CLR)) b += n * s_;       // weights
        if (D->model == M_CLR || D->model == M_COX) b += n * (4.0 + 4.0 + 16.0);   // pid, seg, R1, R2
        return b;
    }

    // ---- updateXBeta (ModelSpecifics.hpp:1466-1586) ------------------------
    void updateXBetaAsync(double delta, int j, bool useW) {
        const bool W = useW && weighted;
        const Col<Real> c = col(j);
        if (c.n == 0) return;
        const Real rd = (Real)delta;
        const int m = D->model;
        if (m == M_LOGISTIC || m == M_POISSON) {
            CCD_DISPATCH_FMT(D->fmt[j], kUpdateIndep<Real, FMT><<<gridFor(c.n), BLOCK, 0, s>>>(c, rd, xb, ex));
        } else if (m == M_CLR) {
            CCD_DISPATCH_FMT(D->fmt[j], (kUpdateGrouped<Real, FMT, false, true><<<gridFor(c.n), BLOCK, 0, s>>>(c, rd, xb, ex, kw, D->dStrat, denom, nW, R1, R2)));
        } else {
            if (W) {
                CCD_DISPATCH_FMT(D->fmt[j], (kUpdateGrouped<Real, FMT, true, false><<<gridFor(c.n), BLOCK, 0, s>>>(c, rd, xb, ex, kw, pidArr, denom, nW, R1, R2)));
            } else {
                CCD_DISPATCH_FMT(D->fmt[j], (kUpdateGrouped<Real, FMT, false, false><<<gridFor(c.n), BLOCK, 0, s>>>(c, rd, xb, ex, kw, pidArr, denom, nW, R1, R2)));
            }
            rescanAsync();
        }
        CCD_HIP(hipGetLastError());
    }
    void updateXBeta(double delta, int j, bool useW) override {
        CCD_HIP(hipSetDevice(D->device));
        if (timing) CCD_HIP(hipEventRecord(ev0, s));
        updateXBetaAsync(delta, j, useW);
        if (timing) CCD_HIP(hipEventRecord(ev1, s));
        CCD_HIP(hipStreamSynchronize(s));
        if (timing && delta != 0.0 && nonZeros(j) > 0) {
            float ms = 0.f;
            CCD_HIP(hipEventElapsedTime(&ms, ev0, ev1));
            counters.deviceMs += ms;
            counters.launches += 1;
            counters.bytes += updateBytes(j);
        }
    }
    // Compulsory bytes of one update of column j: row ids (+values) and a
    // read-modify-write of xb plus the write of exp(xb) per entry (grouped models:
    // pid and the per-pid denominator, FP64).
    double updateBytes(int j) const {
        const double n = (double)nonZeros(j);
        const int f = D->fmt[j];
        const double s_ = sizeof(Real);
        double b = n * 3.0 * s_;
        if (f == F_INDICATOR || f == F_SPARSE) b += n * 4.0;
        if (f == F_SPARSE || f == F_DENSE) b += n * s_;
        if (D->model == M_CLR || D->model == M_COX) b += n * (4.0 + 16.0);
        return b;
    }

    // ---- reductions ---------------------------------------------------------
    template <int OP>
    double rowReduce(bool W, double add) {
        const int64_t len = std::max<int64_t>(D->K, Ncur);
        const int nb = gridFor(len, MAX_RED_BLOCKS);
        const double* den = (D->model == M_COX) ? Dacc : denom;
        const int* pid = (D->model == M_COX) ? pidArr : D->dStrat;
#define CCD_ROWRED(MODEL)                                                                                 \
    if (W) kRowReduce<Real, MODEL, OP, true><<<nb, BLOCK, 0, s>>>(D->K, Ncur, D->dY, xb, ex, kw, predW, pid, \
                                                                    nW, den, parts, counter, add, dOut);   \
    else kRowReduce<Real, MODEL, OP, false><<<nb, BLOCK, 0, s>>>(D->K, Ncur, D->dY, xb, ex, kw, predW, pid, \
                                                                   nW, den, parts, counter, add, dOut);
        switch (D->model) {
        case M_LOGISTIC: CCD_ROWRED(M_LOGISTIC) break;
        case M_POISSON: CCD_ROWRED(M_POISSON) break;
        case M_CLR: CCD_ROWRED(M_CLR) break;
        default: CCD_ROWRED(M_COX) break;
        }
#undef CCD_ROWRED
        CCD_HIP(hipGetLastError());
        CCD_HIP(hipMemcpyAsync(hOut, dOut, sizeof(double), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipStreamSynchronize(s));
        return hOut[0];
    }

    double logLik(bool useW) override {   // ModelSpecifics.hpp:605-650
        CCD_HIP(hipSetDevice(D->device));
        return rowReduce<OP_LOGLIK>(useW && weighted, D->model == M_POISSON ? fixedTerm : 0.0);
    }
    double gradientObjective(bool useW) override {   // ModelSpecifics.hpp:103-120
        CCD_HIP(hipSetDevice(D->device));
        return rowReduce<OP_GRADOBJ>(useW && weighted, 0.0);
    }
    double predLogLik(const double* w) override {   // ModelSpecifics.hpp:652-690
        CCD_HIP(hipSetDevice(D->device));
        const int64_t K = D->K;
        if (D->model == M_COX) {
            const std::vector<double> saved = hKW;
            setWeights(w, true);
            remainingStatsAsync(true);
            h2d(predW, w, K, s);
            const double v = rowReduce<OP_PRED>(false, 0.0);
            setWeights(saved.data(), true);
            remainingStats(true);
            return v;
        }
        h2d(predW, w, K, s);
        return rowReduce<OP_PRED>(false, 0.0);
    }
    void predEstimates(double* out, const double* w) override {   // ModelSpecifics.hpp:692-707
        CCD_HIP(hipSetDevice(D->device));
        const int64_t K = D->K;
        std::vector<double> tmp(K);
        switch (D->model) {
        case M_LOGISTIC: kPredEst<Real, M_LOGISTIC><<<gridFor(K), BLOCK, 0, s>>>(K, xb, predW); break;
        case M_POISSON: kPredEst<Real, M_POISSON><<<gridFor(K), BLOCK, 0, s>>>(K, xb, predW); break;
        default: kPredEst<Real, M_CLR><<<gridFor(K), BLOCK, 0, s>>>(K, xb, predW); break;
        }
        CCD_HIP(hipGetLastError());
        CCD_HIP(hipMemcpyAsync(tmp.data(), predW, K * sizeof(double), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipStreamSynchronize(s));
        for (int64_t k = 0; k < K; ++k)
            if (!w || w[k] != 0.0) out[k] = tmp[k];
    }

    // resident.hip
    void fitResident(const ccd_fit_args& a, double* beta, ccd_fit_result& r) override;
    void fitBatch(const ccd_fit_args& a, const ccd_batch_args& b, double* betas, ccd_fit_result* r,
                  bool autotune, double* predHeldOut) override;
    SteppedBatch* beginStepped(const ccd_fit_args& a, const ccd_batch_args& b, const double* betas0) override;
    bool steppedOpen() const override { return stepped != nullptr; }
    // fisher.inl
    void fisherInformation(const double* beta, const int32_t* idx, int m, double* out) override;
    void releaseResident();

    std::shared_ptr<DevData<Real>> D;
    hipStream_t s = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double bytes = 0.0;
    int64_t capN = 0, Ncur = 0;
    Real *xb = nullptr, *xbSave = nullptr, *ex = nullptr, *kw = nullptr;
    double* predW = nullptr;
    double *denom = nullptr, *nW = nullptr, *R1 = nullptr, *R2 = nullptr, *Dacc = nullptr;
    int32_t *pidArr = nullptr, *pstart = nullptr, *pseg = nullptr;
    Seg2* scanParts = nullptr;
    double* XjY = nullptr;
    Partial* parts = nullptr;
    unsigned* counter = nullptr;
    double* dOut = nullptr;
    double* hOut = nullptr;
    std::vector<double> hKW, hXjY;
    std::vector<int32_t> hPid;
    bool weighted = false;
    double fixedTerm = 0.0;
    // resident-fit workspace (resident.hip)
    void* res = nullptr;
    SteppedBatch* stepped = nullptr;   // open stepped batch using `res`, if any
    Arena* arena = nullptr;   // the creating thread's arena, if any (its stream and host buffer)
};

}  // namespace ccd

#include "resident.inl"
#include "fisher.inl"

namespace ccd {

ResidentNeeds residentNeeds(const ccd_data_desc& d, int device, int maxWorkgroups) {
    ResidentNeeds n{};
    int G = chooseWorkgroups(d.K, 1, device, d.model_type);
    if (maxWorkgroups > 0 && G > maxWorkgroups) G = maxWorkgroups;
    int cus = 0;
    CCD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int occ = precisionOf(d.precision) == CCD_FP32 ? occupancyOf<float>(d.model_type, 1) : occupancyOf<double>(d.model_type, 1);
    n.workgroups = G;
    n.capacityFraction = (double)G / std::max(1, occ * cus);
    const double s = precisionOf(d.precision) == CCD_FP32 ? 4.0 : 8.0;
    const bool grouped = d.model_type == CCD_MODEL_CONDITIONAL_LOGISTIC || d.model_type == CCD_MODEL_COX;
    n.workspaceBytes = d.K * (s + (s == 4.0 ? 2.0 : 0.0) + 12.0) + (grouped ? (d.K + 1) * 5.0 * 8.0 : 0.0) +
                       (double)d.J * ((G + 1) * 4.0 + G * 16.0 + 32.0 * 8.0 + 64.0) + (1 << 20);
    return n;
}

EngineBase* createEngine(const ccd_data_desc& d, int device) {
    if (precisionOf(d.precision) == CCD_FP64) return new EngineT<double>(makeDevData<double>(d, device));
    return new EngineT<float>(makeDevData<float>(d, device));
}

}  // namespace ccd
