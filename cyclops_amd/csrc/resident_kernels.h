// resident_kernels.h -- persistent CCD kernels: a whole findMode (or B fits of the
// same data in lockstep) inside one launch.
//
// Work split: workgroup w owns the rows [rowStart[w], rowStart[w+1]) (aligned to
// stratum / pid boundaries for CLR and Cox) and so owns every per-row and per-pid
// value of those rows for every fit. Per coordinate j:
//   1. each workgroup reduces the entries of column j that fall in its rows;
//   2. the per-workgroup partials meet in a grid all-reduce (epoch-tagged slots,
//      no atomics, fixed reduction order -> every workgroup sees identical g, h);
//   3. the FP64 prior step and trust-region bound give delta (identical in every
//      workgroup, or computed once by the fit's owner and broadcast);
//   4. each workgroup applies delta to its own rows; Cox then re-forms the risk-set
//      sums with two more grid scans.
// State is fit-minor: xb[row * BP + fit], pid arrays [pid * BP + fit].
#pragma once

#include "kernels.h"
#include "../../include/cyclops_amd.h"

namespace ccd {
namespace res {

// Ordered reduction state of a column pass (see kGradGrouped for the algebra):
// g, h sums; l quadratic term; m pre-reset R2 sum; t trailing segment sum; r reset.
struct Mon {
    double g, h, l, m, t, r;
};
__device__ __forceinline__ Mon monZero() { return Mon{0, 0, 0, 0, 0, 0}; }
__device__ __forceinline__ Mon monCombine(const Mon& a, const Mon& b) {
    Mon c;
    c.g = a.g + b.g;
    c.h = a.h + b.h;
    c.l = a.l + b.l + 2.0 * a.t * b.m;
    c.m = (a.r != 0.0) ? a.m : a.m + b.m;
    c.t = (b.r != 0.0) ? b.t : a.t + b.t;
    c.r = (a.r != 0.0 || b.r != 0.0) ? 1.0 : 0.0;
    return c;
}

// Record of a contiguous pid range for the Breslow sweep of one column
// (ModelSpecifics.hpp:904-962 accumulates, pid by pid, D = sum of w e^xb, N1 = sum of
// w x e^xb and N2 = sum of w x^2 e^xb from the stratum start and adds
// nW N1 / D to the gradient and nW (N2 / D - (N1 / D)^2) to the Hessian). With N1,
// N2 counted from the range start (D already carries the ranges before):
//   gl = sum a N1, hl = sum (a N2 - b N1^2) with a = nW / D, b = a / D;
//   af, bf, xf = sums of a, b, b N1 over the pids before the range's first stratum
//   start, which a numerator carry (K1, K2) from the left reaches;
//   t1, t2, f = numerator totals after the range's last start, and whether it has one.
// Joining L then R adds L's numerator totals to R's first stratum:
//   gl = L.gl + R.gl + L.t1 R.af, hl = L.hl + R.hl + L.t2 R.af - 2 L.t1 R.xf - L.t1^2 R.bf,
// so the records of all ranges joined in order give the column's g and h.
struct SweepRec {
    double gl, hl, af, bf, xf, t1, t2;
    int f, pad;
};
__device__ __forceinline__ SweepRec recZero() { return SweepRec{0, 0, 0, 0, 0, 0, 0, 0, 0}; }
__device__ __forceinline__ SweepRec recCombine(const SweepRec& L, const SweepRec& R) {
    SweepRec u;
    u.gl = L.gl + R.gl + L.t1 * R.af;
    u.hl = L.hl + R.hl + L.t2 * R.af - 2.0 * L.t1 * R.xf - L.t1 * L.t1 * R.bf;
    u.af = L.f ? L.af : L.af + R.af;
    u.bf = L.f ? L.bf : L.bf + R.bf;
    u.xf = L.f ? L.xf : L.xf + R.xf + L.t1 * R.bf;
    u.t1 = R.f ? R.t1 : L.t1 + R.t1;
    u.t2 = R.f ? R.t2 : L.t2 + R.t2;
    u.f = L.f | R.f;
    u.pad = 0;
    return u;
}

template <typename Real>
struct Params {
    // data
    int64_t K;
    int J;
    int G;            // workgroups in the launch
    int B;            // fits
    int64_t N;        // pids (CLR: strata; Cox: tie groups); K for LR/PR
    const int* fmt;
    const int64_t* colPtr;
    const int64_t* densePtr;
    const int* rows;
    const Real* vals;
    const Real* dense;
    const Real* y;
    const int* seg;       // stratum per row (CLR/Cox)
    const int* pid;       // pid per row (CLR: stratum, Cox: tie group)
    const int* pstart;    // [N+1] first row of each pid
    const int* pseg;      // [N] stratum of each pid
    const unsigned char* pflag;   // [N] 1 where a pid begins a stratum (Cox)
    const int* rowStart;  // [G+1]
    const int* pidStart;  // [G+1]
    const int* colOff;    // [J*(G+1)] entry offset of workgroup w inside column j (INDICATOR/SPARSE)
    const long long* info;   // [J][4] per column: fmt | flags << 32, nnz | next << 32, c0, 0
    const long long* rng;    // [J][G][2] absolute entry range of workgroup w in column j
    int firstJ;           // first column that is not fixed (J if none)
    // weights
    const Real* kw;       // [K] (single fit) or null
    const int* rowFold;   // [K] or null
    // [nnz] fold of the row of each INDICATOR/SPARSE entry (rowFold gathered once per
    // batch; a sequential byte per entry in the column pass instead of a random row
    // gather), or null (no folds)
    const signed char* entFold;
    const int* held;      // [B]
    const int* fixJ;      // [B] column held at its starting value in fit f (-1: none), or null
    // fit state
    Real* xb;             // [K*BP]
    unsigned short* xbLo; // [K*BP] FP32 runs: low part of xb as bfloat16 bits (xb + low carries the sum), else null
    double* denom;        // [N*BP] (CLR/Cox)
    double* Dacc;         // [N*BP] (Cox)
    double* R1;           // [N*BP]
    double* R2;           // [N*BP]
    double* nW;           // [N*BP]
    // Cox with one fit per team (the accumulated sweep): the column's per-pid
    // numerators, zero outside a coordinate, and the last stratum start of each
    // wave range of each workgroup (-1: none)
    Real* numer;          // [N]
    Real* numer2;         // [N]
    Real* nWs;            // [N] the event weights nW in the bulk precision (the sweep streams them)
    const int* waveLast;  // [G*WAVES]
    double* beta;         // [B*J]
    double* hDelta;       // [B*J]
    const double* var;    // [B]
    const double* initBound;   // [B]
    const unsigned char* exclude;   // [J]
    const unsigned char* fixed;     // [J]
    const double* xjyTot;   // [J]
    const double* xjyFold;  // [J*F] or null
    int F;
    const double* xjyW;     // [J] (weighted single fit) or null
    const double* fixedTermFit;   // [B] Poisson fixed term per fit
    int priorType;
    int convType;
    double tol;
    int maxIt;
    // results
    int* flag;            // [B]
    int* done;            // [B] owner-mode convergence state
    int* ctlIter;         // [B]
    int* ctlFlag;         // [B]
    double* ctlLast;      // [B]
    int resume;           // continue a run started by an earlier launch
    int cycleBudget;      // cycles this launch may run
    int* allDoneOut;      // cleared by any team with a fit still running
    int w;                // workgroup index inside its team (set in the kernel)
    int* iters;           // [B]
    double* ll;           // [B]
    // sync / scratch
    unsigned* epoch;      // [G]
    double* slots;        // [2][G][BP][SLOTW]
    double* bcast;        // [2][BP][4]
    int* err;
    unsigned long long deadlineTicks;
    int allRead;
    int timing;           // phase timers of workgroup 0 (stats[1..4]) are recorded
    int tagged;           // per-coordinate partials go through tagged slot words (writeTagged)
    // counters
    unsigned long long* steps;
    double* stats;        // [0] compulsory bytes moved; [1..5] phase ticks of workgroup 0
};

constexpr int SLOTW = 8;   // doubles per slot: a Mon (6) or a sweep record (8)

__device__ __forceinline__ unsigned long long now() { return wall_clock64(); }

// Cross-workgroup words (epochs, slots, broadcasts, beta and bound state) are read
// and written only with relaxed agent-scope atomics, which are coherent across the
// XCDs per word without flushing or invalidating whole L2 caches. Before a
// workgroup publishes an epoch, every wave waits for its stores to be acknowledged
// (drainStores), so a reader that sees the epoch sees the data; readers issue their
// data loads only after the epoch load has returned.
// Per-row and per-pid state is touched only by the workgroup that owns it.
__device__ __forceinline__ void ast(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ald(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void publish(unsigned* epoch, unsigned seq, int w) {
    drainStores();   // this workgroup's slot/broadcast/beta stores land before the epoch
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(&epoch[w], seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until every workgroup has published `seq`. Lane w watches epoch[w] and
// stops loading it once it has arrived, so a round of polling costs one load per
// workgroup still missing rather than one per team member. Returns false on
// timeout or a raised error flag (every waiter then leaves the kernel).
__device__ __forceinline__ bool waitAll(const unsigned* epoch, unsigned seq, int G, int* err,
                                        unsigned long long ticks) {
    const unsigned long long deadline = now() + ticks;
    int bad = 0;
    for (int w = threadIdx.x; w < G && !bad; w += BLOCK) {
        unsigned spins = 0;
        while (__hip_atomic_load(&epoch[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < seq) {
            if ((++spins & 63u) == 0) {   // another workgroup's timeout ends this wait too
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { bad = 1; break; }
                if (now() > deadline) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bad = 1;
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (__syncthreads_or(bad)) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

// change d of the denominator of pid p: counted into its wave range's total when p
// lies after the range's last stratum start
template <typename Real>
__device__ __forceinline__ void sweepDelta(const Params<Real>& P, int64_t p, double d, double* wsum) {
    const int64_t q0 = P.pidStart[P.w], q1 = P.pidStart[P.w + 1];
    const int64_t cw = (q1 - q0 + WAVES - 1) / WAVES;
    if (cw == 0) return;
    const int wv = (int)((p - q0) / cw);
    if (p >= P.waveLast[P.w * WAVES + wv]) {
#pragma unroll
        for (int i = 0; i < WAVES; ++i)
            if (i == wv) wsum[i] += d;
    }
}

template <typename Real>
__device__ __forceinline__ Real weightOf(const Params<Real>& P, int64_t k, int f) {
    if (P.rowFold) return (P.rowFold[k] == P.held[f]) ? Real(0) : Real(1);
    if (P.kw) return P.kw[k];
    return Real(1);
}

}  // namespace res
}  // namespace ccd

namespace ccd {
namespace res {

constexpr int MAXG = 512;

template <int BP>
struct Geo {
    static constexpr int SUB = BP < 64 ? BP : 64;   // lanes per entry
    static constexpr int NSG = BLOCK / SUB;         // sub-groups per workgroup
    static constexpr int FPL = BP / SUB;            // fits per lane
    // fit i of lane sl of a sub-group: a lane holds FPL adjacent fits
    __host__ __device__ static constexpr int fit(int sl, int i) { return sl * FPL + i; }
};

template <typename Real, int FMT>
__device__ __forceinline__ int64_t rowAt(const Params<Real>& P, int64_t e) {
    if constexpr (FMT == F_INDICATOR || FMT == F_SPARSE) return P.rows[e];
    else return e;
}
template <typename Real, int FMT>
__device__ __forceinline__ Real xAt(const Params<Real>& P, int j, int64_t e) {
    if constexpr (FMT == F_SPARSE) return P.vals[e];
    else if constexpr (FMT == F_DENSE) return P.dense[P.densePtr[j] + e];
    else return Real(1);
}

// Entry range of workgroup w in column j.
template <typename Real>
__device__ __forceinline__ void colRange(const Params<Real>& P, int j, int w, int64_t& a, int64_t& b, int64_t& c0) {
    const int f = P.fmt[j];
    if (f == F_INDICATOR || f == F_SPARSE) {
        c0 = P.colPtr[j];
        const int* off = P.colOff + (size_t)j * (P.G + 1);
        a = c0 + off[w];
        b = c0 + off[w + 1];
    } else {
        c0 = 0;
        a = P.rowStart[w];
        b = P.rowStart[w + 1];
    }
}

// Sub-group chunk of [a, b); for grouped models the boundaries move forward to the
// next change of pid so that a run of equal pid is never split.
template <typename Real, int FMT, bool GROUPED>
__device__ __forceinline__ void subChunk(const Params<Real>& P, int64_t a, int64_t b, int sg, int nsg,
                                         int64_t& n0, int64_t& n1) {
    const int64_t c = (b - a + nsg - 1) / nsg;
    n0 = a + (int64_t)sg * c;
    n1 = n0 + c;
    if (n0 > b) n0 = b;
    if (n1 > b) n1 = b;
    if constexpr (GROUPED) {
        while (n0 > a && n0 < b && P.pid[rowAt<Real, FMT>(P, n0)] == P.pid[rowAt<Real, FMT>(P, n0 - 1)]) ++n0;
        while (n1 > a && n1 < b && P.pid[rowAt<Real, FMT>(P, n1)] == P.pid[rowAt<Real, FMT>(P, n1 - 1)]) ++n1;
    }
}

// Entries are fetched U at a time (row ids, values, weights, per-fit state and
// per-pid ratios all in flight together) and then combined strictly in entry
// order, so the arithmetic is the same as a one-entry-at-a-time walk.
template <int BP>
struct Unroll {
    static constexpr int U = Geo<BP>::FPL >= 4 ? 4 : (BP >= 16 ? 16 : 8);
};

// Lane geometry per model. Wide teams of independent-row models give each lane 4
// fits of an entry (one 16-byte xb gather per lane), so 4x more entries of a
// column are in flight per workgroup than with one fit per lane (for 4 fits per
// team that costs more registers than it gains); grouped models keep Geo<BP>,
// whose ordered run sums need one fit per lane for BP < 256.
template <int BP, int MODEL>
struct GeoM {
    static constexpr bool WIDE = (MODEL == M_LOGISTIC || MODEL == M_POISSON) && BP >= 16;
    static constexpr int SUB = WIDE ? (BP / 4 < 64 ? BP / 4 : 64) : Geo<BP>::SUB;
    static constexpr int NSG = BLOCK / SUB;
    static constexpr int FPL = BP / SUB;
    // a lane holds FPL adjacent fits, so one entry's state is one vector load per lane
    __host__ __device__ static constexpr int fit(int sl, int i) { return sl * FPL + i; }
};

// FPL adjacent per-fit values of one row (fit-minor state): one vector access.
template <typename T, int N>
struct alignas(sizeof(T) * N) VecN {
    T v[N];
};
template <int BP, int MODEL>
struct UnrollM {
    static constexpr int U = GeoM<BP, MODEL>::FPL >= 4 ? 4 : (BP >= 16 ? 16 : 8);
};

// U consecutive entries of one sub-group's chunk with everything the column pass
// and the update need; held in registers across the grid exchange when a chunk
// has at most U entries (the update then only stores).
template <typename Real, int MODEL, int BP, int U = UnrollM<BP, MODEL>::U>
struct Batch {
    using GE = GeoM<BP, MODEL>;
    int kk[U];
    Real xx[U];
    Real xv[U][GE::FPL];
    Real xl[U][GE::FPL];   // low parts of xv (compensated FP32 runs, fast path)
    Real ww[U][GE::FPL];
    int sgm[U], pp[U];
    double r1[U][GE::FPL], r2[U][GE::FPL];
    bool val[U];
};

// Row ids and values of entries [e, e + U) of the chunk ending at n1.
template <typename Real, int MODEL, int BP, int FMT, int U>
__device__ __forceinline__ void loadRows(const Params<Real>& P, int j, int64_t e, int64_t n1,
                                         Batch<Real, MODEL, BP, U>& bt) {
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        bt.val[u] = e + u < n1;
        bt.kk[u] = bt.val[u] ? (int)rowAt<Real, FMT>(P, e + u) : 0;
        bt.xx[u] = bt.val[u] ? xAt<Real, FMT>(P, j, e + u) : Real(0);
        if constexpr (GROUPED) {
            bt.sgm[u] = bt.val[u] ? P.seg[bt.kk[u]] : 0;
            bt.pp[u] = bt.val[u] ? P.pid[bt.kk[u]] : 0;
        }
    }
}

// Per-fit state of the rows of a batch (xb, weights, risk-set ratios); LO also
// fetches the low parts of xb for an update from registers (they are requested with
// the gathers, so they cost no latency of their own).
// ef: the entry folds of the batch's entries (entryFolds) or null (gathered by row).
template <typename Real, int MODEL, int BP, int U, bool LO = false>
__device__ __forceinline__ void loadState(const Params<Real>& P, const bool (&on)[GeoM<BP, MODEL>::FPL],
                                          Batch<Real, MODEL, BP, U>& bt, const signed char* ef = nullptr) {
    using GE = GeoM<BP, MODEL>;
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
    const int sl = threadIdx.x % GE::SUB;
    constexpr int FPL = GE::FPL;
    if constexpr (FPL > 1) {
        // the lane's adjacent fits in one access each; values of fits that are not
        // active are kept as loaded (an update from registers writes them back as they are)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t q = (int64_t)bt.kk[u] * BP + GE::fit(sl, 0);
            VecN<Real, FPL> v{};
            if (bt.val[u]) v = *reinterpret_cast<const VecN<Real, FPL>*>(P.xb + q);
#pragma unroll
            for (int i = 0; i < FPL; ++i) bt.xv[u][i] = v.v[i];
            if constexpr (LO) {
                VecN<unsigned short, FPL> lo{};
                if (bt.val[u]) lo = *reinterpret_cast<const VecN<unsigned short, FPL>*>(P.xbLo + q);
#pragma unroll
                for (int i = 0; i < FPL; ++i) bt.xl[u][i] = (Real)bf16Value(lo.v[i]);
            }
        }
    }
    // row weights (weightOf): the held-out fold of each of this lane's fits once per
    // call and one fold (or weight) per row, instead of both per row and fit
    int heldF[FPL];
    int rowF[U];
    Real rowW[U];
    if constexpr (MODEL != M_CLR) {
#pragma unroll
        for (int i = 0; i < FPL; ++i) {
            const int f = GE::fit(sl, i);
            heldF[i] = (P.rowFold && f < P.B) ? P.held[f] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rowF[u] = (P.rowFold && bt.val[u]) ? (ef ? (int)ef[u] : P.rowFold[bt.kk[u]]) : -1;
            rowW[u] = (!P.rowFold && P.kw && bt.val[u]) ? P.kw[bt.kk[u]] : Real(1);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int i = 0; i < GE::FPL; ++i) {
            const int f = GE::fit(sl, i);
            const bool go = bt.val[u] && on[i];
            if constexpr (FPL == 1) {
                bt.xv[u][i] = go ? P.xb[(int64_t)bt.kk[u] * BP + f] : Real(0);
                if constexpr (LO) bt.xl[u][i] = go ? (Real)bf16Value(P.xbLo[(int64_t)bt.kk[u] * BP + f]) : Real(0);
            }
            if constexpr (MODEL != M_CLR) {
                const Real wt = P.rowFold ? ((rowF[u] == heldF[i]) ? Real(0) : Real(1)) : rowW[u];
                bt.ww[u][i] = go ? wt : Real(0);
            }
            if constexpr (GROUPED && !(MODEL == M_COX && BP == 1)) {   // (the Cox sweep reads no ratios)
                const size_t q = (size_t)bt.pp[u] * BP + f;
                bt.r1[u][i] = go ? P.R1[q] : 0.0;
                bt.r2[u][i] = go ? P.R2[q] : 0.0;
            }
        }
}

// Entry folds from entry e of a column of format FMT on (DENSE and INTERCEPT entries
// are rows in order, so their row folds are read in order already).
template <typename Real, int FMT>
__device__ __forceinline__ const signed char* entryFolds(const Params<Real>& P, int64_t e) {
    if constexpr (FMT == F_INDICATOR || FMT == F_SPARSE) return P.entFold ? P.entFold + e : nullptr;
    else return nullptr;
}

template <typename Real, int MODEL, int BP, int FMT, int U, bool LO = false>
__device__ __forceinline__ void loadBatch(const Params<Real>& P, int j, int64_t e, int64_t n1,
                                          const bool (&on)[GeoM<BP, MODEL>::FPL], Batch<Real, MODEL, BP, U>& bt) {
    loadRows<Real, MODEL, BP, FMT, U>(P, j, e, n1, bt);
    loadState<Real, MODEL, BP, U, LO>(P, on, bt, entryFolds<Real, FMT>(P, e));
}

// xb + delta * x. FP32 runs keep xb as an FP32 value plus a low part (the rounding
// remainder, stored as a bfloat16), add in FP64 and split again, so no update is
// lost to rounding: an FP32 xb alone loses every step under half an ulp of xb and
// drifts from X beta over hundreds of cycles, which moved FP32 coefficients by up to
// 1e-3 relative against FP64 on the 1e6-row Poisson workload; with the low part they
// stay within 3e-5. The column pass reads the FP32 value only. FP64 runs add in
// place, as the reference does.
template <typename Real>
__device__ __forceinline__ void addXb(Real& hi, Real& lo, double delta, Real x) {
    if constexpr (sizeof(Real) == 4) {
        const double sum = (double)hi + (double)lo + delta * (double)x;
        hi = (Real)sum;
        lo = (Real)(sum - (double)hi);
    } else {
        hi = hi + (Real)delta * x;
    }
}

// Gradient terms of a batch, strictly in entry order.
template <typename Real, int MODEL, int BP, int U>
__device__ __forceinline__ void accumBatch(const Batch<Real, MODEL, BP, U>& bt, const bool (&on)[GeoM<BP, MODEL>::FPL],
                                           Mon (&acc)[GeoM<BP, MODEL>::FPL], int& prevSeg, bool& hasPrev) {
    using GE = GeoM<BP, MODEL>;
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!bt.val[u]) break;
        const Real x = bt.xx[u];
        bool reset = false;
        if constexpr (GROUPED) {
            reset = hasPrev && bt.sgm[u] != prevSeg;
            prevSeg = bt.sgm[u];
            hasPrev = true;
        }
#pragma unroll
        for (int i = 0; i < GE::FPL; ++i) {
            if (!on[i]) continue;
            const Real ee = dexp(bt.xv[u][i]);
            if constexpr (MODEL == M_LOGISTIC) {
                const Real wt = bt.ww[u][i];
                const Real pr = ee / (Real(1) + ee);
                acc[i].g += (double)(wt * pr * x);
                acc[i].h += (double)(wt * x * x * pr * (Real(1) - pr));
            } else if constexpr (MODEL == M_POISSON) {
                const Real wt = bt.ww[u][i];
                acc[i].g += (double)(wt * ee * x);
                acc[i].h += (double)(wt * ee * x * x);
            } else {
                Real uu = ee * x;
                if constexpr (MODEL == M_COX) uu *= bt.ww[u][i];
                const double uv = (double)uu;
                if (reset) {
                    acc[i].t = 0.0;
                    acc[i].r = 1.0;
                }
                const double sv = acc[i].t + uv;
                acc[i].l += bt.r2[u][i] * uv * (2.0 * sv - uv);
                if (acc[i].r == 0.0) acc[i].m += bt.r2[u][i] * uv;
                acc[i].t = sv;
                acc[i].g += uv * bt.r1[u][i];
                acc[i].h += uv * (double)x * bt.r1[u][i];
            }
        }
    }
}

// Column pass over the sub-group chunk [n0, n1) (run-aligned for grouped models).
template <typename Real, int MODEL, int BP, int FMT>
__device__ __forceinline__ void gradPass(const Params<Real>& P, int j, int64_t n0, int64_t n1, int64_t c0,
                                         Mon (&acc)[GeoM<BP, MODEL>::FPL], const bool (&on)[GeoM<BP, MODEL>::FPL]) {
    constexpr int U = UnrollM<BP, MODEL>::U;
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
    int prevSeg = 0;
    bool hasPrev = false;
    if constexpr (GROUPED) {
        if (n0 < n1 && n0 > c0) {
            prevSeg = P.seg[rowAt<Real, FMT>(P, n0 - 1)];
            hasPrev = true;
        }
    }
    if constexpr (!GROUPED) {
        // the next batch's row ids and values are requested before this batch's
        // gathers, so their latencies overlap
        Batch<Real, MODEL, BP, U> bt;
        if (n0 < n1) loadRows<Real, MODEL, BP, FMT, U>(P, j, n0, n1, bt);
        for (int64_t e = n0; e < n1; e += U) {
            Batch<Real, MODEL, BP, U> nx;
            if (e + U < n1) loadRows<Real, MODEL, BP, FMT, U>(P, j, e + U, n1, nx);
            loadState<Real, MODEL, BP, U>(P, on, bt, entryFolds<Real, FMT>(P, e));
            accumBatch<Real, MODEL, BP, U>(bt, on, acc, prevSeg, hasPrev);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                bt.kk[u] = nx.kk[u];
                bt.xx[u] = nx.xx[u];
                bt.val[u] = nx.val[u];
            }
        }
    } else {
        for (int64_t e = n0; e < n1; e += U) {
            Batch<Real, MODEL, BP, U> bt;
            loadBatch<Real, MODEL, BP, FMT, U>(P, j, e, n1, on, bt);
            accumBatch<Real, MODEL, BP, U>(bt, on, acc, prevSeg, hasPrev);
        }
    }
}

// Update of a chunk that fitted in one batch: xb from the registers of the column
// pass (fits that moved were active there), per-run denominators as updatePass.
// LOW: the low parts of xb (FP32 runs) came with the registers; otherwise they are
// read here. A lane holding several fits writes them back in one store per entry,
// the fits that did not move with the values they had.
template <typename Real, int MODEL, int BP, int U, bool LOW>
__device__ __forceinline__ void applyBatch(const Params<Real>& P, const Batch<Real, MODEL, BP, U>& bt,
                                           const double (&dl)[GeoM<BP, MODEL>::FPL], double* wsum = nullptr) {
    using GE = GeoM<BP, MODEL>;
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
    constexpr bool COMP = sizeof(Real) == 4;   // compensated xb
    constexpr int FPL = GE::FPL;
    const int sl = threadIdx.x % GE::SUB;
    double acc[FPL];
    bool anyLane = false;
#pragma unroll
    for (int i = 0; i < FPL; ++i) {
        acc[i] = 0.0;
        anyLane = anyLane || dl[i] != 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!bt.val[u]) break;
        const int64_t q0 = (int64_t)bt.kk[u] * BP + GE::fit(sl, 0);
        VecN<Real, FPL> nvv;
        VecN<unsigned short, FPL> nlv{};
        if constexpr (COMP) {
            if constexpr (LOW) {
#pragma unroll
                for (int i = 0; i < FPL; ++i) nlv.v[i] = bf16Bits((float)bt.xl[u][i]);
            } else if constexpr (FPL > 1) {
                if (anyLane) nlv = *reinterpret_cast<const VecN<unsigned short, FPL>*>(P.xbLo + q0);
            }
        }
#pragma unroll
        for (int i = 0; i < FPL; ++i) {
            nvv.v[i] = bt.xv[u][i];
            if (dl[i] == 0.0) continue;
            const Real old = bt.xv[u][i];
            Real nv = old, nl = Real(0);
            if constexpr (COMP) {
                if constexpr (!LOW && FPL == 1) nlv.v[i] = P.xbLo[q0 + i];
                nl = (Real)bf16Value(nlv.v[i]);
            }
            addXb<Real>(nv, nl, dl[i], bt.xx[u]);
            nvv.v[i] = nv;
            if constexpr (COMP) nlv.v[i] = bf16Bits((float)nl);
            if constexpr (FPL == 1) {
                P.xb[q0 + i] = nv;
                if constexpr (COMP) P.xbLo[q0 + i] = nlv.v[i];
            }
            if constexpr (GROUPED) {
                Real diff = dexp(nv) - dexp(old);
                if constexpr (MODEL == M_COX) diff *= bt.ww[u][i];
                acc[i] += (double)diff;
            }
        }
        if constexpr (FPL > 1) {
            if (anyLane) {
                *reinterpret_cast<VecN<Real, FPL>*>(P.xb + q0) = nvv;
                if constexpr (COMP) *reinterpret_cast<VecN<unsigned short, FPL>*>(P.xbLo + q0) = nlv;
            }
        }
        if constexpr (GROUPED) {
            const int pn = (u + 1 < U && bt.val[u + 1]) ? bt.pp[u + 1] : -1;
            if (pn != bt.pp[u]) {   // last entry of a run of equal pid
                const int p = bt.pp[u];
#pragma unroll
                for (int i = 0; i < FPL; ++i) {
                    if (dl[i] == 0.0) continue;
                    const int f = GE::fit(sl, i);
                    const size_t q = (size_t)p * BP + f;
                    const double dn = P.denom[q] + acc[i];
                    P.denom[q] = dn;
                    if constexpr (MODEL == M_COX && BP == 1)
                        if (wsum) sweepDelta(P, p, acc[i], wsum);
                    if constexpr (MODEL == M_CLR) {
                        const double nw = P.nW[q];
                        const double r = nw == 0.0 ? 0.0 : nw / dn;
                        P.R1[q] = r;
                        P.R2[q] = nw == 0.0 ? 0.0 : r / dn;
                    }
                    acc[i] = 0.0;
                }
            }
        }
    }
}

template <typename Real, int MODEL, int BP, int FMT>
__device__ __forceinline__ void updatePass(const Params<Real>& P, int j, int64_t n0, int64_t n1,
                                           const double (&dl)[GeoM<BP, MODEL>::FPL], double* wsum = nullptr) {
    using GE = GeoM<BP, MODEL>;
    constexpr int U = UnrollM<BP, MODEL>::U;
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
    constexpr int FPL = GE::FPL;
    const int sl = threadIdx.x % GE::SUB;
    double acc[GE::FPL];
    bool anyLane = false;   // some fit of this lane moved
    int heldF[FPL];         // Cox row weights (weightOf): held-out folds of the lane's fits
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) {
        acc[i] = 0.0;
        anyLane = anyLane || dl[i] != 0.0;
        if constexpr (MODEL == M_COX) {
            const int f = GE::fit(sl, i);
            heldF[i] = (P.rowFold && f < P.B) ? P.held[f] : -1;
        }
    }
    // independent rows: the next step's row ids and values are requested before this
    // step's gathers, so the two latencies overlap (one round trip per step)
    int64_t nk[U];
    Real nx[U];
    bool nv[U];
    if constexpr (!GROUPED) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            nv[u] = n0 + u < n1;
            nk[u] = nv[u] ? rowAt<Real, FMT>(P, n0 + u) : 0;
            nx[u] = nv[u] ? xAt<Real, FMT>(P, j, n0 + u) : Real(0);
        }
    }
    for (int64_t e = n0; e < n1; e += U) {
        int64_t kk[U];
        Real xx[U], xv[U][GE::FPL], xl[U][GE::FPL], ww[U][GE::FPL];
        int pp[U], pn[U];
        bool val[U];
        if constexpr (!GROUPED) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                kk[u] = nk[u];
                xx[u] = nx[u];
                val[u] = nv[u];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t q = e + U + u;
                nv[u] = q < n1;
                nk[u] = nv[u] ? rowAt<Real, FMT>(P, q) : 0;
                nx[u] = nv[u] ? xAt<Real, FMT>(P, j, q) : Real(0);
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                val[u] = e + u < n1;
                kk[u] = val[u] ? rowAt<Real, FMT>(P, e + u) : 0;
                xx[u] = val[u] ? xAt<Real, FMT>(P, j, e + u) : Real(0);
                pp[u] = val[u] ? P.pid[kk[u]] : -1;
                pn[u] = (val[u] && e + u + 1 < n1) ? P.pid[rowAt<Real, FMT>(P, e + u + 1)] : -1;
            }
        }
        if constexpr (FPL > 1) {
            // the lane's adjacent fits in one access each, all kept as loaded
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t q = kk[u] * BP + GE::fit(sl, 0);
                VecN<Real, FPL> v{};
                if (val[u] && anyLane) v = *reinterpret_cast<const VecN<Real, FPL>*>(P.xb + q);
                VecN<unsigned short, FPL> lo{};
                if constexpr (sizeof(Real) == 4)
                    if (val[u] && anyLane) lo = *reinterpret_cast<const VecN<unsigned short, FPL>*>(P.xbLo + q);
#pragma unroll
                for (int i = 0; i < FPL; ++i) {
                    xv[u][i] = v.v[i];
                    xl[u][i] = (Real)bf16Value(lo.v[i]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < GE::FPL; ++i) {
                const int f = GE::fit(sl, i);
                const bool go = val[u] && dl[i] != 0.0;
                if constexpr (FPL == 1) {
                    xv[u][i] = go ? P.xb[kk[u] * BP + f] : Real(0);
                    if constexpr (sizeof(Real) == 4) xl[u][i] = go ? (Real)bf16Value(P.xbLo[kk[u] * BP + f]) : Real(0);
                    else xl[u][i] = Real(0);
                }
                if constexpr (MODEL == M_COX) {
                    const signed char* ef = entryFolds<Real, FMT>(P, e);
                    Real wt = Real(1);
                    if (P.rowFold) wt = ((ef ? (int)ef[u] : P.rowFold[kk[u]]) == heldF[i]) ? Real(0) : Real(1);
                    else if (P.kw) wt = P.kw[kk[u]];
                    ww[u][i] = go ? wt : Real(0);
                }
            }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!val[u]) break;
            VecN<Real, FPL> nvv;
            VecN<unsigned short, FPL> nlv;
#pragma unroll
            for (int i = 0; i < GE::FPL; ++i) {
                if constexpr (FPL > 1) {
                    nvv.v[i] = xv[u][i];
                    if constexpr (sizeof(Real) == 4) nlv.v[i] = bf16Bits((float)xl[u][i]);
                }
                if (dl[i] == 0.0) continue;
                const int f = GE::fit(sl, i);
                const Real old = xv[u][i];
                Real nv = old, nl = xl[u][i];
                addXb<Real>(nv, nl, dl[i], xx[u]);
                if constexpr (FPL > 1) {
                    nvv.v[i] = nv;
                    if constexpr (sizeof(Real) == 4) nlv.v[i] = bf16Bits((float)nl);
                } else {
                    P.xb[kk[u] * BP + f] = nv;
                    if constexpr (sizeof(Real) == 4) P.xbLo[kk[u] * BP + f] = bf16Bits((float)nl);
                }
                if constexpr (GROUPED) {
                    Real diff = dexp(nv) - dexp(old);
                    if constexpr (MODEL == M_COX) diff *= ww[u][i];
                    acc[i] += (double)diff;
                }
            }
            if constexpr (FPL > 1) {
                if (anyLane) {
                    const int64_t q = kk[u] * BP + GE::fit(sl, 0);
                    *reinterpret_cast<VecN<Real, FPL>*>(P.xb + q) = nvv;
                    if constexpr (sizeof(Real) == 4) *reinterpret_cast<VecN<unsigned short, FPL>*>(P.xbLo + q) = nlv;
                }
            }
            if constexpr (GROUPED) {
                if (pn[u] != pp[u]) {   // last entry of a run of equal pid
                    const int p = pp[u];
#pragma unroll
                    for (int i = 0; i < GE::FPL; ++i) {
                        if (dl[i] == 0.0) continue;
                        const int f = GE::fit(sl, i);
                        const size_t q = (size_t)p * BP + f;
                        const double dn = P.denom[q] + acc[i];
                        P.denom[q] = dn;
                        if constexpr (MODEL == M_COX && BP == 1)
                            if (wsum) sweepDelta(P, p, acc[i], wsum);
                        if constexpr (MODEL == M_CLR) {
                            const double nw = P.nW[q];
                            const double r = nw == 0.0 ? 0.0 : nw / dn;
                            P.R1[q] = r;
                            P.R2[q] = nw == 0.0 ? 0.0 : r / dn;
                        }
                        acc[i] = 0.0;
                    }
                }
            }
        }
    }
}

// First half of combinePlain: the butterfly over the lanes of a wave that hold the
// same fit, then wave w's sums in sMon[w * BP + f]. One-workgroup teams of
// independent rows ("solo") add the waves themselves, in combinePlain's order.
template <int BP, class GE>
__device__ __forceinline__ void combineWaves(const Mon (&acc)[GE::FPL], Mon* sMon) {
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int sl = threadIdx.x % GE::SUB;
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) {
        double g = acc[i].g, h = acc[i].h;
#pragma unroll
        for (int off = GE::SUB; off < 64; off <<= 1) {
            g += __shfl_xor(g, off, 64);
            h += __shfl_xor(h, off, 64);
        }
        if (lane < GE::SUB) sMon[wave * BP + GE::fit(sl, i)] = Mon{g, h, 0, 0, 0, 0};
    }
    __syncthreads();
}

// Ordered tree combine of NSG sub-group Mons per fit in LDS; result in sMon[f].
template <int BP>
__device__ __forceinline__ void combineSubgroups(Mon* sMon) {
    using GE = Geo<BP>;
    for (int s = 1; s < GE::NSG; s <<= 1) {
        const int pairs = GE::NSG / (2 * s);
        for (int idx = threadIdx.x; idx < pairs * BP; idx += BLOCK) {
            const int q = idx / BP, f = idx % BP;
            const int a = 2 * s * q, b = a + s;
            sMon[a * BP + f] = monCombine(sMon[a * BP + f], sMon[b * BP + f]);
        }
        __syncthreads();
    }
}

// Plain sums of (g, h) per fit over the sub-groups of a workgroup: a butterfly
// over the lanes of a wave that hold the same fit, then the waves in order (a
// fixed order, so results are reproducible). Result in sMon[f] (g, h).
template <int BP, class GE>
__device__ __forceinline__ void combinePlain(const Mon (&acc)[GE::FPL], Mon* sMon) {
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int sl = threadIdx.x % GE::SUB;
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) {
        double g = acc[i].g, h = acc[i].h;
#pragma unroll
        for (int off = GE::SUB; off < 64; off <<= 1) {
            g += __shfl_xor(g, off, 64);
            h += __shfl_xor(h, off, 64);
        }
        if (lane < GE::SUB) sMon[wave * BP + GE::fit(sl, i)] = Mon{g, h, 0, 0, 0, 0};
    }
    __syncthreads();
    for (int f = threadIdx.x; f < BP; f += BLOCK) {
        Mon m = sMon[f];
        for (int q = 1; q < BLOCK / 64; ++q) {
            m.g += sMon[q * BP + f].g;
            m.h += sMon[q * BP + f].h;
        }
        sMon[f] = m;
    }
    __syncthreads();
}

__device__ __forceinline__ double priorDeltaDev(int type, double variance, double g, double h, double beta) {
    if (type == 0) return -(g / h);
    if (type == 2) return -(g + beta / variance) / (h + 1.0 / variance);
    const double lambda = sqrt(2.0 / variance);
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

__device__ __forceinline__ double logDensityDev(int type, double variance, double x) {
    if (type == 0) return 0.0;
    if (type == 1) {
        const double l = sqrt(2.0 / variance);
        return log(0.5 * l) - l * fabs(x);
    }
    return -0.5 * log(2.0 * M_PI * variance) - 0.5 * x * x / variance;
}

}  // namespace res
}  // namespace ccd

namespace ccd {
namespace res {

__device__ __forceinline__ Mon shflDown(const Mon& v, int d) {
    return Mon{__shfl_down(v.g, d, 64), __shfl_down(v.h, d, 64), __shfl_down(v.l, d, 64),
               __shfl_down(v.m, d, 64), __shfl_down(v.t, d, 64), __shfl_down(v.r, d, 64)};
}
__device__ __forceinline__ SweepRec shflDown(const SweepRec& v, int d) {
    return SweepRec{__shfl_down(v.gl, d, 64), __shfl_down(v.hl, d, 64), __shfl_down(v.af, d, 64),
                    __shfl_down(v.bf, d, 64), __shfl_down(v.xf, d, 64), __shfl_down(v.t1, d, 64),
                    __shfl_down(v.t2, d, 64), __shfl_down(v.f, d, 64), 0};
}
__device__ __forceinline__ Seg2 shflDown(const Seg2& v, int d) {
    Seg2 r;
    r.a = __shfl_down(v.a, d, 64);
    r.b = __shfl_down(v.b, d, 64);
    r.f = __shfl_down(v.f, d, 64);
    r.pad = 0;
    return r;
}

// Ordered combine of n <= 2 * BLOCK items (loaded by ld(i)); every thread gets the
// result. Thread t takes items 2t and 2t+1, each wave combines its threads in a
// pairwise tree of shuffles (lane l with lane l + d: always the next block, so the
// order of the items is kept), and the wave results are combined in order.
template <class T, class Load, class Comb>
__device__ __forceinline__ T treeReduce(int n, Load ld, Comb cb, T* sBuf, T ident) {
    const int i0 = 2 * threadIdx.x;
    T v = ident;
    if (i0 < n) v = ld(i0);
    if (i0 + 1 < n) v = cb(v, ld(i0 + 1));
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = shflDown(v, d);
        if ((lane & (2 * d - 1)) == 0) v = cb(v, o);
    }
    if (lane == 0) sBuf[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T r = sBuf[0];
        for (int q = 1; q < BLOCK / 64; ++q) r = cb(r, sBuf[q]);
        sBuf[WAVES] = r;
    }
    __syncthreads();
    const T r = sBuf[WAVES];
    __syncthreads();
    return r;
}

// Shared scratch of the persistent kernel.
template <int BP>
struct Shared {
    Mon mon[Geo<BP>::NSG * BP];
    union {
        Mon red[MAXG];
        Seg2 seg[MAXG];
        SweepRec rec[WAVES + 1];
        double part[BLOCK * WAVES];
    } u;
    // Cox sweep (one fit per team): segmented denominator total of each wave range
    // (kept exact by the walks, moved by the updates in between) and the walk's
    // per-wave records
    Seg2 waveDen[WAVES];
    Seg2 waveDenNew[WAVES];
    SweepRec wrec[WAVES];
    Seg2 scan[Geo<BP>::NSG * BP];
    double out[BP * 4];
    double beta[BP];
    double hd[BP];
    double carry[BP];
    double carry2[BP];
    // per-fit control (all-read mode: every workgroup keeps an identical copy)
    double last[BP];
    int iter[BP];
    int done[BP];
    int flag[BP];
    int active[BP];   // fits still iterating (identical in every workgroup)
};

// Publish slots, wait, reduce per fit (all-read: every workgroup reduces every fit;
// owner mode: workgroup f % G reduces fit f and broadcasts 4 doubles). On return
// s.out[4f..4f+3] holds fn's output for every fit in every workgroup.
template <typename Real, int BP, class Fn>
__device__ __forceinline__ bool reduceFits(const Params<Real>& P, unsigned& seq, unsigned long long deadline,
                                           Shared<BP>& s, Fn fn, bool ordered = true, int off = 0) {
    if (P.G == 1) {   // a one-workgroup team reduces in LDS only; fits step in parallel
        for (int f = threadIdx.x; f < P.B; f += BLOCK) fn(f, s.mon[f], s.out + 4 * f);
        __syncthreads();
        return true;
    }
    publish(P.epoch, ++seq, P.w);
    if (!waitAll(P.epoch, seq, P.G, P.err, deadline)) return false;
    const double* slot = P.slots + (size_t)(seq & 1) * P.G * BP * SLOTW;
    if (P.allRead && !ordered) {
        // every workgroup sums every fit in one pass: thread t reads the partials
        // w = t / BP, t / BP + BLOCK / BP, ... of fit t % BP; lanes of one fit are
        // combined by a butterfly, then the waves (or parts) in order
        constexpr int TPF = BLOCK / BP;
        const int f = threadIdx.x % BP, part = threadIdx.x / BP, lane = threadIdx.x % 64;
        double g = 0.0, h = 0.0;
        if (f < P.B)
            for (int w = part; w < P.G; w += TPF) {
                const double* q = slot + ((size_t)w * BP + f) * SLOTW + off;
                g += ald(q);
                h += ald(q + 1);
            }
        int rowsOut;
        if constexpr (BP < 64) {
#pragma unroll
            for (int off = BP; off < 64; off <<= 1) {
                g += __shfl_xor(g, off, 64);
                h += __shfl_xor(h, off, 64);
            }
            if (lane < BP) s.u.red[(threadIdx.x / 64) * BP + f] = Mon{g, h, 0, 0, 0, 0};
            rowsOut = BLOCK / 64;
        } else {
            s.u.red[part * BP + f] = Mon{g, h, 0, 0, 0, 0};
            rowsOut = TPF;
        }
        __syncthreads();
        for (int ff = threadIdx.x; ff < P.B; ff += BLOCK) {
            Mon m = s.u.red[ff];
            for (int q = 1; q < rowsOut; ++q) {
                m.g += s.u.red[q * BP + ff].g;
                m.h += s.u.red[q * BP + ff].h;
            }
            fn(ff, m, s.out + 4 * ff);
        }
        __syncthreads();
        return true;
    }
    auto comb = [](const Mon& a, const Mon& b) { return monCombine(a, b); };
    // ordered combine of the G partials of fit f; plain sums (fixed shuffle tree)
    // when the partials carry no segment state
    auto reduceOne = [&](int f) -> Mon {
        if (!ordered) {
            double v[2] = {0.0, 0.0};
            for (int w = threadIdx.x; w < P.G; w += BLOCK) {
                const double* q = slot + ((size_t)w * BP + f) * SLOTW + off;
                v[0] += ald(q);
                v[1] += ald(q + 1);
            }
            blockReduce<2>(v);
            return Mon{v[0], v[1], 0, 0, 0, 0};
        }
        return treeReduce<Mon>(P.G, [&](int w) {
            const double* q = slot + ((size_t)w * BP + f) * SLOTW;
            return Mon{ald(q), ald(q + 1), ald(q + 2), ald(q + 3), ald(q + 4), ald(q + 5)};
        }, comb, s.u.red, monZero());
    };
    if (P.allRead) {
        for (int f = 0; f < P.B; ++f) {
            const Mon m = reduceOne(f);
            if (threadIdx.x == 0) fn(f, m, s.out + 4 * f);
        }
        __syncthreads();
    } else {
        double* bc = P.bcast + (size_t)((seq + 1) & 1) * BP * 4;
        for (int f = P.w; f < P.B; f += P.G) {
            const Mon m = reduceOne(f);
            if (threadIdx.x == 0) {
                double o[4] = {0, 0, 0, 0};
                fn(f, m, o);
                for (int i = 0; i < 4; ++i) ast(&bc[f * 4 + i], o[i]);
            }
        }
        publish(P.epoch, ++seq, P.w);
        if (!waitAll(P.epoch, seq, P.G, P.err, deadline)) return false;
        for (int i = threadIdx.x; i < P.B * 4; i += BLOCK) s.out[i] = ald(&bc[i]);
        __syncthreads();
    }
    return true;
}

// Write this workgroup's per-fit Mons (s.mon[f]) to the slots of the next epoch.
template <typename Real, int BP>
__device__ __forceinline__ void writeSlots(const Params<Real>& P, unsigned seq, const Shared<BP>& s,
                                           bool ordered = true, int off = 0) {
    if (P.G == 1) return;
    double* slot = P.slots + (size_t)((seq + 1) & 1) * P.G * BP * SLOTW;
    for (int f = threadIdx.x; f < BP; f += BLOCK) {
        double* q = slot + ((size_t)P.w * BP + f) * SLOTW + off;
        const Mon& m = s.mon[f];
        ast(q, m.g);
        ast(q + 1, m.h);
        if (ordered) { ast(q + 2, m.l); ast(q + 3, m.m); ast(q + 4, m.t); ast(q + 5, m.r); }
    }
}

// Tagged exchange of the per-coordinate (g, h) partials: FP32 runs of independent
// rows in all-read teams with G x BP <= 1024. Each partial is one 64-bit word whose
// lowest mantissa bit carries the exchange's tag, (t >> 1) & 1 for exchange t, which
// uses slot half t & 1. The stale word a reader may find in that half was written two
// exchanges earlier, with the other tag, so a reader that sees the tag sees the value:
// there is no epoch word, no wait for this workgroup's stores to be acknowledged, and
// the next column's row-id loads stay in flight through the exchange. A half is
// rewritten only after every workgroup has read it: a workgroup writes exchange t + 1
// after it has seen every exchange-t word, each of which was written after its writer
// had summed exchange t - 1. Before every launch the host fills half 0 with all-zero
// bytes and half 1 with all-one bytes, so neither matches its first tag. Readers clear
// the tag bit before summing. A partial of FP32 terms is an exact FP64 sum whose lowest
// mantissa bit is zero unless its terms span more than 2^29, so the value read is the
// value written and results stay bit-identical to the epoch exchange; otherwise it is
// at most one unit in the last place of a double lower. An infinite partial (tag 1
// makes it a NaN pattern) reads back as the same infinity. The once-per-cycle
// objective keeps the epoch exchange on words 2-3.
__device__ __forceinline__ double withTag(double v, unsigned long long tag) {
    return __longlong_as_double((long long)(((unsigned long long)__double_as_longlong(v) & ~1ull) | tag));
}
__device__ __forceinline__ unsigned long long tagOf(double v) {
    return (unsigned long long)__double_as_longlong(v) & 1ull;
}
__device__ __forceinline__ double untag(double v) { return withTag(v, 0ull); }

template <typename Real, int BP>
__device__ __forceinline__ void writeTagged(const Params<Real>& P, unsigned tseq, const Shared<BP>& s) {
    const unsigned t = tseq + 1;
    const unsigned long long tag = (t >> 1) & 1u;
    double* slot = P.slots + (size_t)(t & 1u) * P.G * BP * SLOTW;
    for (int f = threadIdx.x; f < BP; f += BLOCK) {
        double* q = slot + ((size_t)P.w * BP + f) * SLOTW;
        ast(q, withTag(s.mon[f].g, tag));
        ast(q + 1, withTag(s.mon[f].h, tag));
    }
}

// Every workgroup sums every fit (the all-read order of reduceFits: thread t reads the
// partials w = t / BP, t / BP + BLOCK / BP, ... of fit t % BP, then lanes of one fit by
// a butterfly and the waves in order) once all of its words carry the tag.
template <typename Real, int BP, class Fn>
__device__ __forceinline__ bool reduceTagged(const Params<Real>& P, unsigned& tseq, unsigned long long ticks,
                                             Shared<BP>& s, Fn fn) {
    ++tseq;
    const unsigned long long tag = (tseq >> 1) & 1u;
    const double* slot = P.slots + (size_t)(tseq & 1u) * P.G * BP * SLOTW;
    constexpr int TPF = BLOCK / BP;
    constexpr int MW = 1024 / BLOCK;   // G x BP <= 1024: at most 4 workgroups' words per thread
    const int f = threadIdx.x % BP, part = threadIdx.x / BP, lane = threadIdx.x % 64;
    double a[MW], b[MW];
    bool wait[MW];
#pragma unroll
    for (int i = 0; i < MW; ++i) {
        const int w = part + i * TPF;
        wait[i] = f < P.B && w < P.G;
        a[i] = 0.0;
        b[i] = 0.0;
        if (wait[i]) {
            const double* q = slot + ((size_t)w * BP + f) * SLOTW;
            a[i] = ald(q);
            b[i] = ald(q + 1);
        }
    }
    int bad = 0;
    unsigned spins = 0;
    unsigned long long deadline = 0;
    while (true) {
        bool ok = true;
#pragma unroll
        for (int i = 0; i < MW; ++i)
            if (wait[i]) {
                if (tagOf(a[i]) == tag && tagOf(b[i]) == tag) wait[i] = false;
                else ok = false;
            }
        if (ok) break;
        if ((++spins & 63u) == 0) {   // another workgroup's timeout ends this wait too
            if (__hip_atomic_load(P.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { bad = 1; break; }
            if (deadline == 0) deadline = now() + ticks;
            else if (now() > deadline) {
                __hip_atomic_store(P.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bad = 1;
                break;
            }
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int i = 0; i < MW; ++i)
            if (wait[i]) {
                const double* q = slot + ((size_t)(part + i * TPF) * BP + f) * SLOTW;
                a[i] = ald(q);
                b[i] = ald(q + 1);
            }
    }
    if (__syncthreads_or(bad)) return false;
    double g = 0.0, h = 0.0;
#pragma unroll
    for (int i = 0; i < MW; ++i) {
        g += untag(a[i]);
        h += untag(b[i]);
    }
    int rowsOut;
    if constexpr (BP < 64) {
#pragma unroll
        for (int o = BP; o < 64; o <<= 1) {
            g += __shfl_xor(g, o, 64);
            h += __shfl_xor(h, o, 64);
        }
        if (lane < BP) s.u.red[(threadIdx.x / 64) * BP + f] = Mon{g, h, 0, 0, 0, 0};
        rowsOut = BLOCK / 64;
    } else {
        s.u.red[part * BP + f] = Mon{g, h, 0, 0, 0, 0};
        rowsOut = TPF;
    }
    __syncthreads();
    for (int ff = threadIdx.x; ff < P.B; ff += BLOCK) {
        Mon m = s.u.red[ff];
        for (int q = 1; q < rowsOut; ++q) {
            m.g += s.u.red[q * BP + ff].g;
            m.h += s.u.red[q * BP + ff].h;
        }
        fn(ff, m, s.out + 4 * ff);
    }
    __syncthreads();
    return true;
}

// Breslow risk sets of this workgroup's pids, both scans, with the cross-workgroup carries.
// One fit per team: the same rescan with the workgroup's pids split into 4
// contiguous wave ranges, each walked in chunks of 8 consecutive pids per lane
// (a register scan per lane, one wave scan of the lane totals per chunk), then
// carried across waves and, through the slots, across workgroups.
template <typename Real, int BP>
__device__ __forceinline__ bool coxRescanWaves(const Params<Real>& P, unsigned& seq, unsigned long long deadline,
                                               Shared<BP>& s) {
    const int w = P.w;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t q0 = P.pidStart[w], q1 = P.pidStart[w + 1];
    const int64_t cw = (q1 - q0 + WAVES - 1) / WAVES;
    int64_t a = q0 + wave * cw, b = a + cw;
    if (a > q1) a = q1;
    if (b > q1) b = q1;
    const int64_t N = P.N;
    const Seg2 z{0.0, 0.0, 0, 0};
    auto segc = [](const Seg2& x, const Seg2& y) { return segCombine(x, y); };
    // Chunks of CH consecutive pids per lane (64 * CH per wave): each lane scans its
    // pids in registers and ONE wave scan of the lane totals joins them, instead of
    // a wave scan per 64 pids.
    constexpr int CH = 8;
    constexpr int CW = 64 * CH;
    // wave exclusive prefix of the lane totals and the wave's total (DPP scan)
    auto laneJoin = [&](const Seg2& tot, Seg2& excl, Seg2& all) {
        const Seg2 inc = waveSegScanDpp<false>(tot);
        excl = segFromPrevLane(inc);
        all = segLane63(inc);
    };
    // Segment starts come from one byte per pid (pflag[p] = pid p begins a stratum):
    // forward, p > 0 with pflag[p]; backward (descending pids), p == N - 1 or
    // pflag[p + 1].
    const unsigned char* fl = P.pflag;
    // A walk keeps one chunk of loads in flight: the loads of chunk c + 1 are issued
    // (at addresses clamped into the range, masked later) before chunk c is scanned,
    // so a wave waits for memory once per walk rather than once per chunk.
    struct Raw {
        double x[CH], y[CH];
        int f[CH + 1];
    };
    // forward chunk at base: x = denom, y = nW (store walk), f[i] = pflag[p0 + i]
    auto fetchFwd = [&](int64_t base, bool store, Raw& r) {
        const int64_t p0 = base + (int64_t)lane * CH;
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int64_t p = p0 + i < b ? p0 + i : b - 1;
            r.x[i] = P.denom[p];
            if (store) r.y[i] = P.nW[p];
        }
#pragma unroll
        for (int i = 0; i <= CH; ++i) r.f[i] = fl[p0 + i < N ? p0 + i : N - 1];
    };
    // walks [a, b) forward: D = segmented prefix of denom. With store it writes D and
    // also forms, from those D, the wave's total for the backward scan (its pids in
    // descending order: each lane's pids from the top, then the lanes from the top),
    // so the backward totals need no walk of their own.
    auto fwdWalk = [&](Seg2 run, bool store, Seg2& bwdTot) {
        Raw nxt;
        if (a < b) fetchFwd(a, store, nxt);
        for (int64_t base = a; base < b; base += CW) {
            const Raw cur = nxt;
            if (base + CW < b) fetchFwd(base + CW, store, nxt);
            const int64_t p0 = base + (int64_t)lane * CH;
            Seg2 acc = z;
            double inc[CH];
            bool started[CH];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int64_t p = p0 + i;
                const bool ok = p < b;
                const bool st = ok && p > 0 && cur.f[i] != 0;
                acc = segCombine(acc, Seg2{ok ? cur.x[i] : 0.0, 0.0, st ? 1 : 0, 0});
                inc[i] = acc.a;
                started[i] = acc.f != 0;
            }
            Seg2 excl, all;
            laneJoin(acc, excl, all);
            if (store) {
                const Seg2 pre = segCombine(run, excl);
                double dd[CH];
#pragma unroll
                for (int i = 0; i < CH; ++i) {
                    dd[i] = started[i] ? inc[i] : pre.a + inc[i];
                    if (p0 + i < b) P.Dacc[p0 + i] = dd[i];
                }
                Seg2 bacc = z;   // this lane's pids, descending
#pragma unroll
                for (int i = CH - 1; i >= 0; --i) {
                    const int64_t p = p0 + i;
                    if (p >= b) continue;
                    const bool stE = p == N - 1 || cur.f[i + 1] != 0;
                    const double nv = cur.y[i];
                    // one divide: nE/D, then nE/D^2 from it (D may be 0 where nE is)
                    const double inv = nv == 0.0 ? 0.0 : 1.0 / dd[i];
                    const double r = nv * inv;
                    bacc = segCombine(bacc, Seg2{r, r * inv, stE ? 1 : 0, 0});
                }
                // lanes from the top: a scan in descending lane order, total at lane 63
                bwdTot = segCombine(segLane63(waveSegScanDpp<true>(bacc)), bwdTot);   // precedes the earlier chunks
            }
            run = segCombine(run, all);
        }
        return run;
    };
    // backward chunk at top: x = nW, y = D, f[i] = pflag[p + 1] of this lane's pid p0 - i
    auto fetchBwd = [&](int64_t top, Raw& r) {
        const int64_t p0 = top - (int64_t)lane * CH;
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int64_t p = p0 - i >= a ? p0 - i : a;
            r.x[i] = P.nW[p];
            r.y[i] = P.Dacc[p];
            r.f[i] = fl[p + 1 < N ? p + 1 : N - 1];
        }
    };
    // walks [a, b) backward (descending pids): R1, R2 = reverse segmented sums of
    // nE/D and nE/D^2 from the carry `run`
    auto bwdWalk = [&](Seg2 run) {
        Raw nxt;
        if (a < b) fetchBwd(b - 1, nxt);
        for (int64_t top = b - 1; top >= a; top -= CW) {
            const Raw cur = nxt;
            if (top - CW >= a) fetchBwd(top - CW, nxt);
            const int64_t p0 = top - (int64_t)lane * CH;   // this lane's pids: p0, p0 - 1, ...
            Seg2 acc = z;
            double i1[CH], i2[CH];
            bool started[CH];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int64_t p = p0 - i;
                Seg2 v = z;
                if (p >= a) {
                    const double nv = cur.x[i], dv = cur.y[i];
                    const bool stE = p == N - 1 || cur.f[i] != 0;
                    const double inv = nv == 0.0 ? 0.0 : 1.0 / dv;
                    const double r = nv * inv;
                    v = Seg2{r, r * inv, stE ? 1 : 0, 0};
                }
                acc = segCombine(acc, v);
                i1[i] = acc.a;
                i2[i] = acc.b;
                started[i] = acc.f != 0;
            }
            Seg2 excl, all;
            laneJoin(acc, excl, all);
            const Seg2 pre = segCombine(run, excl);
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int64_t p = p0 - i;
                if (p >= a) {
                    P.R1[p] = started[i] ? i1[i] : pre.a + i1[i];
                    P.R2[p] = started[i] ? i2[i] : pre.b + i2[i];
                }
            }
            run = segCombine(run, all);
        }
        return run;
    };
    // exchange of the workgroup totals; carry = combination of the workgroups
    // before this one (after, for the backward pass)
    auto exchange = [&](bool backward, Seg2& carry) -> bool {
        if (threadIdx.x == 0) {
            Seg2 tot = z;
            for (int i = 0; i < WAVES; ++i) tot = segCombine(tot, s.scan[backward ? WAVES - 1 - i : i]);
            double* o = P.slots + ((size_t)((seq + 1) & 1) * P.G + w) * SLOTW;
            ast(o, tot.a); ast(o + 1, tot.b); ast(o + 2, (double)tot.f);
        }
        publish(P.epoch, ++seq, P.w);
        if (!waitAll(P.epoch, seq, P.G, P.err, deadline)) return false;
        const double* sl2 = P.slots + (size_t)(seq & 1) * P.G * SLOTW;
        const int n = backward ? P.G - 1 - w : w;
        carry = treeReduce<Seg2>(n, [&](int q) {
            const double* o = sl2 + (size_t)(backward ? P.G - 1 - q : q) * SLOTW;
            return Seg2{ald(o), ald(o + 1), (int)ald(o + 2), 0};
        }, segc, s.u.seg, z);
        return true;
    };
    // ---- forward: D = segmented prefix of denom (a totals walk, then the store walk,
    // which also forms the backward totals)
    Seg2 none = z;
    Seg2 mine = fwdWalk(z, false, none);
    if (lane == 0) {
        s.scan[wave] = mine;
        s.waveDen[wave] = mine;   // the sweep's exact wave totals
    }
    __syncthreads();
    Seg2 carry;
    if (!exchange(false, carry)) return false;
    Seg2 run = Seg2{carry.a, 0.0, 0, 0};
    for (int i = 0; i < wave; ++i) run = segCombine(run, s.scan[i]);
    __syncthreads();   // every wave has read the forward totals before they are replaced
    Seg2 back = z;
    fwdWalk(run, true, back);
    if (lane == 0) s.scan[wave] = back;
    __syncthreads();
    // ---- backward: R1, R2 = reverse segmented sums of nE/D, nE/D^2
    if (!exchange(true, carry)) return false;
    run = Seg2{carry.a, carry.b, 0, 0};
    for (int i = WAVES - 1; i > wave; --i) run = segCombine(run, s.scan[i]);
    bwdWalk(run);
    __syncthreads();
    return true;
}

template <typename Real, int BP>
__device__ __forceinline__ bool coxRescan(const Params<Real>& P, unsigned& seq, unsigned long long deadline,
                                          Shared<BP>& s) {
    using GE = Geo<BP>;
    const int sg = threadIdx.x / GE::SUB, sl = threadIdx.x % GE::SUB;
    const int w = P.w;
    const int64_t q0 = P.pidStart[w], q1 = P.pidStart[w + 1];
    const int64_t c = (q1 - q0 + GE::NSG - 1) / GE::NSG;
    int64_t p0 = q0 + sg * c, p1 = p0 + c;
    if (p0 > q1) p0 = q1;
    if (p1 > q1) p1 = q1;
    const int64_t N = P.N;
    auto segc = [](const Seg2& a, const Seg2& b) { return segCombine(a, b); };
    const Seg2 z{0.0, 0.0, 0, 0};
    // Exclusive scan of the sub-group aggregates s.scan[q * BP + f] in sub-group
    // order (backward: reverse order), in place; the workgroup's total goes to its
    // slot. One fit per workgroup (BP == 1) scans all 256 aggregates in parallel;
    // otherwise thread f walks the sub-groups of fit f.
    auto scanSubgroups = [&](double* slot, bool backward) {
        if constexpr (GE::NSG == BLOCK) {
            const int q = backward ? BLOCK - 1 - threadIdx.x : threadIdx.x;
            Seg2 total;
            const Seg2 ex = blockSegScanExcl(s.scan[q], total);
            s.scan[q] = ex;
            if (threadIdx.x == 0) {
                double* o = slot + (size_t)w * SLOTW;
                ast(o, total.a); ast(o + 1, total.b); ast(o + 2, (double)total.f);
            }
            __syncthreads();
        } else {
            for (int f = threadIdx.x; f < BP; f += BLOCK) {
                Seg2 run = z;
                for (int i = 0; i < GE::NSG; ++i) {
                    const int q = backward ? GE::NSG - 1 - i : i;
                    const Seg2 v = s.scan[q * BP + f];
                    s.scan[q * BP + f] = run;
                    run = segCombine(run, v);
                }
                double* o = slot + ((size_t)w * BP + f) * SLOTW;
                ast(o, run.a); ast(o + 1, run.b); ast(o + 2, (double)run.f);
            }
        }
    };
    // The pids of a chunk are visited TL at a time: every load of a tile is issued
    // before the first combine, so a chunk costs one memory latency per TL pids.
    constexpr int TL = 16;
    // forward term of pid p: denominator, reset at a new stratum
    auto fwdTiles = [&](int64_t a, int64_t b, int f, auto visit) {
        for (int64_t p = a; p < b; p += TL) {
            double dv[TL];
            int sv[TL + 1];
            sv[0] = p > 0 ? P.pseg[p - 1] : 0;
#pragma unroll
            for (int t = 0; t < TL; ++t) {
                const bool ok = p + t < b;
                dv[t] = ok ? P.denom[(p + t) * BP + f] : 0.0;
                sv[t + 1] = ok ? P.pseg[p + t] : 0;
            }
#pragma unroll
            for (int t = 0; t < TL; ++t)
                if (p + t < b) visit(p + t, Seg2{dv[t], 0.0, (p + t > 0 && sv[t + 1] != sv[t]) ? 1 : 0, 0});
        }
    };
    // backward term of pid p: nE/D and nE/D^2, reset after the last pid of a stratum
    auto bwdTiles = [&](int64_t a, int64_t b, int f, auto visit) {
        for (int64_t p = b - 1; p >= a; p -= TL) {
            double nv[TL], dv[TL];
            int sv[TL + 1];   // sv[t] = pseg[p - t + 1]
            sv[0] = P.pseg[p + 1];
#pragma unroll
            for (int t = 0; t < TL; ++t) {
                const bool ok = p - t >= a;
                nv[t] = ok ? P.nW[(p - t) * BP + f] : 0.0;
                dv[t] = ok ? P.Dacc[(p - t) * BP + f] : 1.0;
                sv[t + 1] = ok ? P.pseg[p - t] : 0;
            }
#pragma unroll
            for (int t = 0; t < TL; ++t) {
                const int64_t q = p - t;
                if (q < a) break;
                const double inv = nv[t] == 0.0 ? 0.0 : 1.0 / dv[t];
                const double r = nv[t] * inv;
                visit(q, Seg2{r, r * inv, (q == N - 1 || sv[t] != sv[t + 1]) ? 1 : 0, 0});
            }
        }
    };
    // ---- forward: D = segmented prefix of denom
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) {
        const int f = GE::fit(sl, i);
        Seg2 agg = z;
        fwdTiles(p0, p1, f, [&](int64_t, const Seg2& v) { agg = segCombine(agg, v); });
        s.scan[sg * BP + f] = agg;
    }
    __syncthreads();
    double* slot = P.slots + (size_t)((seq + 1) & 1) * P.G * BP * SLOTW;
    scanSubgroups(slot, false);
    publish(P.epoch, ++seq, P.w);
    if (!waitAll(P.epoch, seq, P.G, P.err, deadline)) return false;
    {
        const double* sl2 = P.slots + (size_t)(seq & 1) * P.G * BP * SLOTW;
        for (int f = 0; f < BP; ++f) {
            const Seg2 cr = treeReduce<Seg2>(w, [&](int q) {
                const double* o = sl2 + ((size_t)q * BP + f) * SLOTW;
                return Seg2{ald(o), ald(o + 1), (int)ald(o + 2), 0};
            }, segc, s.u.seg, z);
            if (threadIdx.x == 0) s.carry[f] = cr.a;
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) {
        const int f = GE::fit(sl, i);
        Seg2 run = segCombine(Seg2{s.carry[f], 0.0, 0, 0}, s.scan[sg * BP + f]);
        fwdTiles(p0, p1, f, [&](int64_t p, const Seg2& v) {
            run = segCombine(run, v);
            P.Dacc[p * BP + f] = run.a;
        });
    }
    __syncthreads();
    // ---- backward: R1, R2 = reverse segmented sums of nE/D, nE/D^2
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) {
        const int f = GE::fit(sl, i);
        Seg2 agg = z;
        bwdTiles(p0, p1, f, [&](int64_t, const Seg2& v) { agg = segCombine(agg, v); });
        s.scan[sg * BP + f] = agg;
    }
    __syncthreads();
    slot = P.slots + (size_t)((seq + 1) & 1) * P.G * BP * SLOTW;
    scanSubgroups(slot, true);
    publish(P.epoch, ++seq, P.w);
    if (!waitAll(P.epoch, seq, P.G, P.err, deadline)) return false;
    {
        const double* sl2 = P.slots + (size_t)(seq & 1) * P.G * BP * SLOTW;
        const int nAfter = P.G - 1 - w;
        for (int f = 0; f < BP; ++f) {
            const Seg2 cr = treeReduce<Seg2>(nAfter, [&](int q) {
                const double* o = sl2 + ((size_t)(P.G - 1 - q) * BP + f) * SLOTW;
                return Seg2{ald(o), ald(o + 1), (int)ald(o + 2), 0};
            }, segc, s.u.seg, z);
            if (threadIdx.x == 0) { s.carry[f] = cr.a; s.carry2[f] = cr.b; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) {
        const int f = GE::fit(sl, i);
        Seg2 run = segCombine(Seg2{s.carry[f], s.carry2[f], 0, 0}, s.scan[sg * BP + f]);
        bwdTiles(p0, p1, f, [&](int64_t p, const Seg2& v) {
            run = segCombine(run, v);
            P.R1[p * BP + f] = run.a;
            P.R2[p * BP + f] = run.b;
        });
    }
    __syncthreads();
    return true;
}

}  // namespace res
}  // namespace ccd

namespace ccd {
namespace res {

enum ObjOp : int { OBJ_LOGLIK = 0, OBJ_GRADIENT = 1, OBJ_LANGE = 2 };

// This workgroup's share of the objective of every fit, into s.mon[f].g.
template <int N>
struct DVec {
    double v[N];
};

// Row terms of the objective of this workgroup's rows for each fit of the lane, in
// FP64 in every run (the convergence test compares objectives that differ by a
// relative 1e-6 or less; FP32 runs add the low part of xb). It runs once per cycle
// and is kept out of line, so the registers of its FP64 exp do not add to those the
// coordinate loop holds.
template <typename Real, int MODEL, int BP>
__device__ __attribute__((noinline)) DVec<GeoM<BP, MODEL>::FPL> objectiveRows(const Params<Real>& P, int op) {
    using GE = GeoM<BP, MODEL>;
    const int sg = threadIdx.x / GE::SUB, sl = threadIdx.x % GE::SUB;
    const int w = P.w;
    DVec<GE::FPL> acc;
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) acc.v[i] = 0.0;
    const int64_t a = P.rowStart[w], b = P.rowStart[w + 1];
    const int64_t c = (b - a + GE::NSG - 1) / GE::NSG;
    int64_t n0 = a + sg * c, n1 = n0 + c;
    if (n0 > b) n0 = b;
    if (n1 > b) n1 = b;
    for (int64_t k = n0; k < n1; ++k) {
        const Real yk = P.y[k];
        const int yi = (int)yk;
#pragma unroll
        for (int i = 0; i < GE::FPL; ++i) {
            const int f = GE::fit(sl, i);
            if (f >= P.B) continue;
            double x = (double)P.xb[k * BP + f];
            if constexpr (sizeof(Real) == 4) x += (double)bf16Value(P.xbLo[k * BP + f]);
            const double wt = (double)weightOf(P, k, f);
            if (op == OBJ_GRADIENT) {
                acc.v[i] += x * (double)yk * wt;
                continue;
            }
            double cterm = (MODEL == M_COX) ? (yi != 1 ? 0.0 : x) : yi * x;
            if constexpr (MODEL == M_LOGISTIC) cterm -= log(1.0 + exp(x));
            if constexpr (MODEL == M_POISSON) cterm -= exp(x);
            acc.v[i] += cterm * wt;
        }
    }
    return acc;
}

template <typename Real, int MODEL, int BP>
__device__ __forceinline__ void objectivePartial(const Params<Real>& P, int op, Shared<BP>& s) {
    using GE = GeoM<BP, MODEL>;
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
    const int sg = threadIdx.x / GE::SUB, sl = threadIdx.x % GE::SUB;
    const int w = P.w;
    double acc[GE::FPL];
    {   // rows (once per cycle, in a function of its own: see objectiveRows)
        const DVec<GE::FPL> r = objectiveRows<Real, MODEL, BP>(P, op);
#pragma unroll
        for (int i = 0; i < GE::FPL; ++i) acc[i] = r.v[i];
    }
    if constexpr (GROUPED) {
        if (op != OBJ_GRADIENT) {
            const int64_t a = P.pidStart[w], b = P.pidStart[w + 1];
            const int64_t c = (b - a + GE::NSG - 1) / GE::NSG;
            int64_t n0 = a + sg * c, n1 = n0 + c;
            if (n0 > b) n0 = b;
            if (n1 > b) n1 = b;
            const double* den = (MODEL == M_COX) ? P.Dacc : P.denom;
            for (int64_t p = n0; p < n1; ++p) {
#pragma unroll
                for (int i = 0; i < GE::FPL; ++i) {
                    const int f = GE::fit(sl, i);
                    if (f >= P.B) continue;
                    const double nw = P.nW[p * BP + f];
                    if (nw != 0.0) acc[i] -= nw * log(den[p * BP + f]);
                }
            }
        }
    }
    Mon m[GE::FPL];
#pragma unroll
    for (int i = 0; i < GE::FPL; ++i) m[i] = Mon{acc[i], 0, 0, 0, 0, 0};
    combinePlain<BP, GE>(m, s.mon);
}

// Per-pid initial state of this workgroup: nE, denominators, CLR risk ratios.
template <typename Real, int MODEL, int BP>
__device__ __forceinline__ void initPids(const Params<Real>& P) {
    using GE = Geo<BP>;
    const int sg = threadIdx.x / GE::SUB, sl = threadIdx.x % GE::SUB;
    const int w = P.w;
    const int64_t a = P.pidStart[w], b = P.pidStart[w + 1];
    const int64_t c = (b - a + GE::NSG - 1) / GE::NSG;
    int64_t n0 = a + sg * c, n1 = n0 + c;
    if (n0 > b) n0 = b;
    if (n1 > b) n1 = b;
    for (int64_t p = n0; p < n1; ++p) {
#pragma unroll
        for (int i = 0; i < GE::FPL; ++i) {
            const int f = GE::fit(sl, i);
            double dn = 0.0, nw = 0.0;
            if (f < P.B) {
                for (int k = P.pstart[p]; k < P.pstart[p + 1]; ++k) {
                    const Real wt = weightOf(P, (int64_t)k, f);
                    Real e = dexp(P.xb[(int64_t)k * BP + f]);
                    if constexpr (MODEL == M_COX) e *= wt;
                    dn += (double)e;
                    nw += (double)P.y[k] * (double)wt;
                }
            }
            const size_t q = (size_t)p * BP + f;
            P.denom[q] = dn;
            P.nW[q] = nw;
            if constexpr (MODEL == M_COX && BP == 1)
                if (P.nWs) P.nWs[q] = (Real)nw;
            if constexpr (MODEL == M_CLR) {
                const double r = nw == 0.0 ? 0.0 : nw / dn;
                P.R1[q] = r;
                P.R2[q] = nw == 0.0 ? 0.0 : r / dn;
            }
        }
    }
    __syncthreads();
}

// Per-coordinate metadata of column j for workgroup w: two table rows read with
// vector loads (the vector-memory counter waits for them only where they are
// used, one coordinate later; scalar loads would be waited for at the next LDS
// access).
struct Meta {
    int j, fm, ptype, next;
    int64_t nnz, a, b, c0;
};
__device__ __forceinline__ long long ldv(const long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename Real>
__device__ __forceinline__ Meta loadMeta(const Params<Real>& P, int j, int w) {
    const long long* ip = P.info + (size_t)j * 4;
    const long long* rp = P.rng + ((size_t)j * P.G + w) * 2;
    const long long w0 = ldv(ip), w1 = ldv(ip + 1);
    Meta m;
    m.j = j;
    m.fm = (int)(w0 & 0xffffffffll);
    m.ptype = ((w0 >> 32) & 1) ? 0 : P.priorType;
    m.nnz = (int64_t)(unsigned)(w1 & 0xffffffffll);
    m.next = (int)(w1 >> 32);
    m.c0 = ldv(ip + 2);
    m.a = ldv(rp);
    m.b = ldv(rp + 1);
    return m;
}

// The rescan for the team's fit layout (the wave-range walk holds LDS staging, so
// only the one-fit layout instantiates it).
template <typename Real, int BP>
__device__ __forceinline__ bool coxRescanAny(const Params<Real>& P, unsigned& seq, unsigned long long deadline,
                                             Shared<BP>& s) {
    if constexpr (BP == 1) return coxRescanWaves<Real, BP>(P, seq, deadline, s);
    else return coxRescan<Real, BP>(P, seq, deadline, s);
}

// ---------------------------------------------------------------------------
// Breslow Cox, one fit per team: the accumulated sweep. Per coordinate the
// column's numerators w x e^xb (and w x^2 e^xb) are scattered to their pids, one
// forward walk over every pid of the team forms D, N1, N2 and the partial
// gradient records, and one exchange joins the records (with a second exchange for
// the denominator carries only after an update moved the denominators). This is
// the pid-by-pid accumulation of ModelSpecifics.hpp:904-962 (computeGradientAndHessian
// for the accumulated models) spread over the workgroups: nothing per pid is written
// per coordinate, where the telescoped form rewrites the risk-set ratios of every pid.
// ---------------------------------------------------------------------------

// wave range [a, b) of wave `wave` of workgroup w (the split the rescan uses)
template <typename Real>
__device__ __forceinline__ void waveRange(const Params<Real>& P, int wave, int64_t& a, int64_t& b) {
    const int64_t q0 = P.pidStart[P.w], q1 = P.pidStart[P.w + 1];
    const int64_t cw = (q1 - q0 + WAVES - 1) / WAVES;
    a = q0 + wave * cw;
    b = a + cw;
    if (a > q1) a = q1;
    if (b > q1) b = q1;
}

// Adds every thread's denominator changes to the wave totals in a fixed order.
template <int BP>
__device__ __forceinline__ void sweepAddWaveDeltas(Shared<BP>& s, const double (&wsum)[WAVES]) {
#pragma unroll
    for (int i = 0; i < WAVES; ++i) s.u.part[threadIdx.x * WAVES + i] = wsum[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < BLOCK / 64; ++q) v += s.u.part[(q * 64 + lane) * WAVES + wv];
    v = waveSumDpp(v);
    if (lane == 0) s.waveDen[wv].a += v;
    __syncthreads();
}

// Numerators of the entries of a register batch, summed per run of equal pid and
// stored at the pid (clear: store zeros at the same pids).
template <typename Real, int MODEL, int BP, int U>
__device__ __forceinline__ void sweepScatterBatch(const Params<Real>& P, const Batch<Real, MODEL, BP, U>& bt,
                                                  bool two, bool clear) {
    double c1 = 0.0, c2 = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!bt.val[u]) break;
        if (!clear) {
            const Real x = bt.xx[u];
            const Real c = bt.ww[u][0] * dexp(bt.xv[u][0]) * x;
            c1 += (double)c;
            c2 += (double)(c * x);
        }
        const int pn = (u + 1 < U && bt.val[u + 1]) ? bt.pp[u + 1] : -1;
        if (pn != bt.pp[u]) {   // last entry of a run of equal pid
            P.numer[bt.pp[u]] = (Real)c1;
            if (two) P.numer2[bt.pp[u]] = (Real)c2;
            c1 = c2 = 0.0;
        }
    }
}

// The same over a chunk [n0, n1) of any length (runs may cross batches).
template <typename Real, int MODEL, int BP, int FMT, int U>
__device__ __forceinline__ void sweepScatterPass(const Params<Real>& P, int j, int64_t n0, int64_t n1,
                                                 const bool (&on)[1], bool two, bool clear) {
    double c1 = 0.0, c2 = 0.0;
    for (int64_t e = n0; e < n1; e += U) {
        Batch<Real, MODEL, BP, U> bt;
        loadRows<Real, MODEL, BP, FMT, U>(P, j, e, n1, bt);
        if (!clear) loadState<Real, MODEL, BP, U>(P, on, bt, entryFolds<Real, FMT>(P, e));
        const int pAfter = (e + U < n1) ? P.pid[rowAt<Real, FMT>(P, e + U)] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!bt.val[u]) break;
            if (!clear) {
                const Real x = bt.xx[u];
                const Real c = bt.ww[u][0] * dexp(bt.xv[u][0]) * x;
                c1 += (double)c;
                c2 += (double)(c * x);
            }
            const int pn = (u + 1 < U) ? (bt.val[u + 1] ? bt.pp[u + 1] : -1) : pAfter;
            if (pn != bt.pp[u]) {
                P.numer[bt.pp[u]] = (Real)c1;
                if (two) P.numer2[bt.pp[u]] = (Real)c2;
                c1 = c2 = 0.0;
            }
        }
    }
}

// Carry of the denominators into this workgroup: its wave totals joined and
// published (right after an update, so the exchange overlaps the next column's
// scatter), then, once every workgroup has published, the segmented prefix of
// the workgroups before this one.
template <typename Real, int BP>
__device__ __forceinline__ void sweepCarryPublish(const Params<Real>& P, unsigned& seq, Shared<BP>& s) {
    if (threadIdx.x == 0) {
        Seg2 tot{0.0, 0.0, 0, 0};
        for (int i = 0; i < WAVES; ++i) tot = segCombine(tot, s.waveDen[i]);
        double* o = P.slots + ((size_t)((seq + 1) & 1) * P.G + P.w) * SLOTW;
        ast(o, tot.a);
        ast(o + 1, (double)tot.f);
    }
    publish(P.epoch, ++seq, P.w);
}
template <typename Real, int BP>
__device__ __forceinline__ bool sweepCarryWait(const Params<Real>& P, unsigned seq, unsigned long long deadline,
                                               Shared<BP>& s, double& carryD) {
    const Seg2 z{0.0, 0.0, 0, 0};
    if (!waitAll(P.epoch, seq, P.G, P.err, deadline)) return false;
    const double* sl2 = P.slots + (size_t)(seq & 1) * P.G * SLOTW;
    const Seg2 c = treeReduce<Seg2>(P.w, [&](int q) {
        const double* o = sl2 + (size_t)q * SLOTW;
        return Seg2{ald(o), 0.0, (int)ald(o + 1), 0};
    }, [](const Seg2& x, const Seg2& y) { return segCombine(x, y); }, s.u.seg, z);
    carryD = c.a;
    return true;
}

// The walk of one wave over its pid range: D from the carry, N1 and N2 from the
// range start; leaves the wave's record and its exact denominator total in LDS.
// Chunks of CH consecutive pids per lane, one chunk of loads in flight.
template <typename Real, int BP>
__device__ __forceinline__ void sweepWalk(const Params<Real>& P, Shared<BP>& s, double carryD, bool two) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t a, b;
    waveRange(P, wave, a, b);
    const int64_t N = P.N;
    const unsigned char* fl = P.pflag;
    constexpr int CH = 8;
    constexpr int CW = 64 * CH;
    Seg2 dIn{carryD, 0.0, 0, 0};
    for (int i = 0; i < wave; ++i) dIn = segCombine(dIn, s.waveDen[i]);
    const Seg3 z3{0.0, 0.0, 0.0, 0};
    Seg3 run{dIn.a, 0.0, 0.0, 0};   // flag and numerators count from the range start
    Seg2 own{0.0, 0.0, 0, 0};
    double gl = 0.0, hl = 0.0, af = 0.0, bf = 0.0, xf = 0.0;
    // Loads are coalesced (element e of a chunk, pid base + e, by lane e % 64) and
    // each chunk is turned around in this wave's LDS so that a lane then holds CH
    // consecutive pids: 64 consecutive doubles per load instruction instead of one
    // 64-byte block per lane, whose 64 separate cache lines per instruction held the
    // walk at the rate the memory pipeline accepts requests.
    constexpr int SP = CW + CW / CH;   // element e at e + e / CH: strided reads spread over the banks
    __shared__ double sD[WAVES][SP];
    __shared__ Real sW[WAVES][SP], sN[WAVES][SP], sM[WAVES][SP];
    __shared__ unsigned char sF[WAVES][CW];
    auto px = [](int e) { return e + e / CH; };
    struct Raw {   // one chunk of positions (the weights and numerators in the bulk precision, as stored)
        double d[CH];
        Real n[CH], w[CH], m[CH];
        int f[CH];
    };
    auto fetch = [&](int64_t base, Raw& r) {
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int64_t q = base + i * 64 + lane;
            const int64_t p = q < b ? q : b - 1;
            r.d[i] = P.denom[p];
            r.w[i] = P.nWs[p];
            r.n[i] = P.numer[p];
            if (two) r.m[i] = P.numer2[p];
            r.f[i] = fl[p];
        }
    };
    auto waveSync = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    Raw nxt;
    if (a < b) fetch(a, nxt);
    for (int64_t base = a; base < b; base += CW) {
        Raw cur;
#pragma unroll
        for (int i = 0; i < CH; ++i) {   // into LDS in load order
            const int e = i * 64 + lane;
            sD[wave][px(e)] = nxt.d[i];
            sW[wave][px(e)] = nxt.w[i];
            sN[wave][px(e)] = nxt.n[i];
            if (two) sM[wave][px(e)] = nxt.m[i];
            sF[wave][e] = (unsigned char)nxt.f[i];
        }
        if (base + CW < b) fetch(base + CW, nxt);
        waveSync();
#pragma unroll
        for (int i = 0; i < CH; ++i) {   // out in lane order: this lane's CH consecutive pids
            const int e = lane * CH + i;
            cur.d[i] = sD[wave][px(e)];
            cur.w[i] = sW[wave][px(e)];
            cur.n[i] = sN[wave][px(e)];
            if (two) cur.m[i] = sM[wave][px(e)];
            cur.f[i] = sF[wave][e];
        }
        waveSync();   // read before the next chunk is written
        const int64_t p0 = base + (int64_t)lane * CH;
        auto elem = [&](int i) {
            const int64_t p = p0 + i;
            const bool ok = p < b;
            const bool st = ok && p > 0 && cur.f[i] != 0;
            return Seg3{ok ? cur.d[i] : 0.0, ok ? cur.n[i] : 0.0, ok ? (two ? cur.m[i] : cur.n[i]) : 0.0, st ? 1 : 0};
        };
        Seg3 tot = z3;
#pragma unroll
        for (int i = 0; i < CH; ++i) tot = seg3Combine(tot, elem(i));
        const Seg3 inc = waveSeg3ScanDpp(tot);
        const Seg3 excl = seg3FromPrevLane(inc);
        const Seg3 all = seg3Lane63(inc);
        Seg3 acc = seg3Combine(run, excl);
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            acc = seg3Combine(acc, elem(i));
            const double nw = cur.w[i];
            if (p0 + i < b && nw != 0.0) {
                const double inv = 1.0 / acc.d;
                const double ra = nw * inv, rb = ra * inv;
                gl += ra * acc.n1;
                hl += ra * acc.n2 - rb * acc.n1 * acc.n1;
                if (!acc.f) {   // before the range's first stratum start
                    af += ra;
                    bf += rb;
                    xf += rb * acc.n1;
                }
            }
        }
        run = seg3Combine(run, all);
        own = segCombine(own, Seg2{all.d, 0.0, all.f, 0});
    }
    gl = waveSumDpp(gl);
    hl = waveSumDpp(hl);
    af = waveSumDpp(af);
    bf = waveSumDpp(bf);
    xf = waveSumDpp(xf);
    if (lane == 0) {
        s.wrec[wave] = SweepRec{gl, hl, af, bf, xf, run.n1, run.n2, run.f, 0};
        s.waveDenNew[wave] = own;
    }
    __syncthreads();
}

// Joins the wave records into the workgroup's, exchanges, and joins every
// workgroup's in order: the column's accumulated g and h (in every workgroup).
template <typename Real, int BP>
__device__ __forceinline__ bool sweepReduce(const Params<Real>& P, unsigned& seq, unsigned long long deadline,
                                            Shared<BP>& s, double& g, double& h) {
    if (threadIdx.x == 0) {
        SweepRec r = s.wrec[0];
        for (int i = 1; i < WAVES; ++i) r = recCombine(r, s.wrec[i]);
        double* o = P.slots + ((size_t)((seq + 1) & 1) * P.G + P.w) * SLOTW;
        ast(o, r.gl); ast(o + 1, r.hl); ast(o + 2, r.af); ast(o + 3, r.bf);
        ast(o + 4, r.xf); ast(o + 5, r.t1); ast(o + 6, r.t2); ast(o + 7, (double)r.f);
    }
    if (threadIdx.x < WAVES) s.waveDen[threadIdx.x] = s.waveDenNew[threadIdx.x];
    publish(P.epoch, ++seq, P.w);
    if (!waitAll(P.epoch, seq, P.G, P.err, deadline)) return false;
    const double* sl2 = P.slots + (size_t)(seq & 1) * P.G * SLOTW;
    const SweepRec t = treeReduce<SweepRec>(P.G, [&](int q) {
        const double* o = sl2 + (size_t)q * SLOTW;
        return SweepRec{ald(o), ald(o + 1), ald(o + 2), ald(o + 3), ald(o + 4), ald(o + 5), ald(o + 6),
                        (int)ald(o + 7), 0};
    }, [](const SweepRec& x, const SweepRec& y) { return recCombine(x, y); }, s.u.rec, recZero());
    g = t.gl;
    h = t.hl;
    return true;
}

// The persistent findMode kernel (CyclicCoordinateDescent.cpp:909-1216 for every fit).
// Launch = teams of T workgroups; team t runs the fits described by teams[t]
// (its own data and state pointers), exchanging only inside the team.
template <typename Real, int MODEL, int BP>
__global__ __launch_bounds__(BLOCK) void kFindMode(const Params<Real>* __restrict__ teams, int T) {
    using GE = GeoM<BP, MODEL>;
    // one parameter block per workgroup (w filled in by the host): fields are read
    // from the scalar cache where used instead of being held in registers
    const Params<Real>& P = teams[blockIdx.x];
    constexpr bool GROUPED = (MODEL == M_CLR || MODEL == M_COX);
    // Cox with one fit per team runs the accumulated sweep (every workgroup computes
    // the step: one fit per team is always all-read)
    constexpr bool SWEEP = MODEL == M_COX && BP == 1;
    __shared__ Shared<BP> s;
    const int w = P.w;
    const unsigned long long deadline = P.deadlineTicks;   // per-wait limit
    unsigned seq = 0;
    const bool own = P.allRead != 0;   // control state in LDS (all-read) or global (owner mode)
    // FP32 runs of independent rows: per-coordinate partials through tagged slot words
    // (writeTagged); the objective's epoch exchanges then use words 2-3 of each slot
    constexpr bool TAGGABLE = !GROUPED && sizeof(Real) == 4;
    const bool tagged = TAGGABLE && P.tagged != 0;
    const int objOff = tagged ? 2 : 0;
    unsigned tseq = 0;
    for (int f = threadIdx.x; f < BP; f += BLOCK) {
        const bool real = f < P.B;
        s.iter[f] = (P.resume && real) ? P.ctlIter[f] : 0;
        s.done[f] = !real ? 1 : (P.resume ? P.done[f] : 0);
        s.flag[f] = (P.resume && real) ? P.ctlFlag[f] : 0;
        s.last[f] = (P.resume && real) ? P.ctlLast[f] : 0.0;
        s.active[f] = real && !s.done[f] ? 1 : 0;
    }
    __syncthreads();
    // per-fit control pointers (LDS copy in all-read mode; owner's global copy otherwise)
    int* cIter = own ? s.iter : P.ctlIter;
    int* cFlag = own ? s.flag : P.ctlFlag;
    double* cLast = own ? s.last : P.ctlLast;
    int* cDone = own ? s.done : P.done;

    if constexpr (GROUPED) {
        if (!P.resume) {
            initPids<Real, MODEL, BP>(P);
            if constexpr (MODEL == M_COX) {
                // the pass above wrote only this workgroup's pids; the scan reads them all
                if (!coxRescanAny<Real, BP>(P, seq, deadline, s)) return;
            }
        } else if constexpr (SWEEP) {
            if (!coxRescanAny<Real, BP>(P, seq, deadline, s)) return;   // the sweep's wave totals (LDS)
        }
    }
    const int objOp = P.convType == CCD_CONV_GRADIENT ? OBJ_GRADIENT
                    : (P.convType == CCD_CONV_LANGE ? OBJ_LANGE : OBJ_LOGLIK);
    const bool addFixed = (MODEL == M_POISSON) && objOp != OBJ_GRADIENT;
    // log prior of fit f (JointPrior.h:123-131); read after the epoch that follows the
    // last beta write of the cycle, so every workgroup sees the same beta
    auto logPrior = [&](int f) {
        double lp = 0.0;
        for (int jj = 0; jj < P.J; ++jj)
            lp += logDensityDev(P.exclude[jj] ? 0 : P.priorType, P.var[f], ald(&P.beta[(size_t)f * P.J + jj]));
        return lp;
    };
    // initial objective -> lastObj
    if (!P.resume) {
        objectivePartial<Real, MODEL, BP>(P, objOp, s);
        writeSlots<Real, BP>(P, seq, s, false, objOff);
        if (!reduceFits<Real, BP>(P, seq, deadline, s, [&](int f, const Mon& m, double* o) {
                const double obj = m.g + (addFixed ? P.fixedTermFit[f] : 0.0) + (objOp == OBJ_LANGE ? logPrior(f) : 0.0);
                cLast[f] = obj;
                cIter[f] = 0;
                cFlag[f] = 0;
                cDone[f] = 0;
                o[0] = 0.0;
            }, false, objOff))
            return;
    }

    double loc[4] = {0.0, 0.0, 0.0, 0.0};   // counters of workgroup 0, thread 0
    unsigned long long locSteps = 0;
    bool allDone = false;
    int nActive = 0;   // fits still iterating this cycle
    {
        int act = 0;
        for (int f = threadIdx.x; f < P.B; f += BLOCK) act += s.active[f];
        nActive = __syncthreads_count(act != 0);
        allDone = nActive == 0;
    }
    // coordinates in cycle order: non-fixed columns; the metadata of the next one
    // (and the beta/bound of the fits this workgroup steps) is fetched while the
    // current one is in flight
    const int firstJ = P.firstJ;
    // fits whose step this workgroup computes: all (all-read) or the owned ones
    auto mine = [&](int f) { return f < P.B && (own || f % P.G == w); };
    Meta cur = loadMeta(P, firstJ < P.J ? firstJ : 0, w);
    double preB = 0.0, preH = 0.0;   // thread f: beta / bound of fit f at cur.j
    bool havePre = false;
    // chunks of at most UF entries stay in registers from the column pass to the
    // update (teams of up to 4 fits, or 16 independent-row fits in the wide lane
    // geometry; wider teams would hold too many registers)
    constexpr bool FAST = BP <= 4 || (!GROUPED && BP <= 16);
    constexpr int UF = 4;   // entries per sub-group kept in registers
    // compensated xb: the low parts travel with the registers for independent-row
    // models; grouped models, whose update also walks per-pid state, fetch them there
    constexpr bool LOW = sizeof(Real) == 4 && !GROUPED;
    double carryD = 0.0;        // sweep: denominator carry into this workgroup
    bool denomStale = true;     // sweep: the wave totals were recomputed; no carry published
    bool carryPending = false;  // sweep: this workgroup's totals are published, not yet joined
    auto fitsBatch = [&](const Meta& m) {   // every sub-group chunk of m has <= UF entries
        return (m.b - m.a + GE::NSG - 1) / GE::NSG <= UF;
    };
    Batch<Real, MODEL, BP, UF> nb;   // prefetched row ids / values of cur (ungrouped models)
    bool haveRows = false;
    // One-workgroup teams of independent rows ("solo"): lane f of every wave holds the
    // beta and bound of fit f in registers, and every wave adds the wave partials and
    // computes the step itself, so a coordinate takes two barriers (the partials, the
    // update) instead of six. The partials alternate between two halves of s.mon, so a
    // wave that runs ahead to the next coordinate never overwrites what another wave
    // still reads. The sums keep combinePlain's order: results are bit-identical.
    constexpr bool SOLO_OK = !GROUPED && BP <= 4;
    const bool solo = SOLO_OK && P.G == 1;
    const int me = solo ? (int)(threadIdx.x & 63) : (int)threadIdx.x;   // the fit whose beta / bound this thread loads
    double curB = 0.0, curH = 0.0;   // solo: beta / bound of fit `me` at this coordinate
    unsigned par = 0;                // solo: half of s.mon for this coordinate's partials
    bool soloStore = true;           // solo: only wave 0 stores beta and bound
    for (int cycle = 0; !allDone && cycle < P.cycleBudget; ++cycle) {
        while (firstJ < P.J) {
            const int j = cur.j;
            const int fm = cur.fm;
            const int64_t nnzj = cur.nnz;
            const int ptype = cur.ptype;
            if (solo) {
                if (mine(me)) {
                    curB = havePre ? preB : ald(&P.beta[(size_t)me * P.J + j]);
                    curH = havePre ? preH : ald(&P.hDelta[(size_t)me * P.J + j]);
                }
            } else {
                const int f = threadIdx.x;
                if (mine(f)) {
                    s.beta[f] = havePre ? preB : ald(&P.beta[(size_t)f * P.J + j]);
                    s.hd[f] = havePre ? preH : ald(&P.hDelta[(size_t)f * P.J + j]);
                }
            }
            const int jn = cur.next;
            // step of fit f from its reduced column pass: ccdUpdateBeta (:1394-1413),
            // prior (CovariatePrior.h), applyBounds (:1503-1520)
            auto stepWith = [&](int f, const Mon& m, double* o, double b, double hd) {
                double g = 0.0, h = 0.0;
                if (nnzj != 0) {
                    const double xjy = P.xjyW ? P.xjyW[j]
                                     : P.xjyTot[j] - ((P.xjyFold && P.held[f] >= 0) ? P.xjyFold[(size_t)j * P.F + P.held[f]] : 0.0);
                    g = m.g - xjy;
                    h = GROUPED ? m.h - m.l : m.h;
                }
                if (cDone[f]) { o[0] = 0.0; return; }
                if (P.fixJ && P.fixJ[f] == j) { o[0] = 0.0; return; }   // setFixedBeta for this fit only
                if (h < 0.0) { g = 0.0; h = 0.0; }
                double delta = priorDeltaDev(ptype, P.var[f], g, h, b);
                if (delta < -hd) delta = -hd;
                else if (delta > hd) delta = hd;
                hd = fmax(fmax(fabs(delta) * 2.0, hd / 2.0), 1E-3);
                if ((!own || w == 0) && soloStore) {
                    ast(&P.hDelta[(size_t)f * P.J + j], hd);
                    if (delta != 0.0) ast(&P.beta[(size_t)f * P.J + j], b + delta);
                }
                o[0] = delta;
            };
            auto step = [&](int f, const Mon& m, double* o) { stepWith(f, m, o, s.beta[f], s.hd[f]); };
            // solo: every wave steps the fits from its registers (wave 0 stores)
            auto soloStep = [&](const Mon& m) {
                double o4[4] = {0.0, 0.0, 0.0, 0.0};
                soloStore = threadIdx.x < 64;
                stepWith(me, m, o4, curB, curH);
                soloStore = true;
                return o4[0];
            };
            if (!solo) __syncthreads();   // s.beta / s.hd written above
            // prefetch for the next coordinate (a different column, so its beta and
            // bound are final for this cycle)
            const Meta nxt = loadMeta(P, jn, w);
            // not across the end of the cycle: the next cycle's first beta and bound are
            // read after the objective exchange, which orders them after workgroup 0's
            // stores (tagged exchanges carry no such ordering)
            havePre = jn > j;
            if (havePre && mine(me)) {
                preB = ald(&P.beta[(size_t)me * P.J + jn]);
                preH = ald(&P.hDelta[(size_t)me * P.J + jn]);
            }
            // independent-row models: a small next column's row ids and values do not
            // depend on this coordinate; they are fetched once this column's are used
            const bool curRows = haveRows;
            auto prefetchRows = [&]() {
                if constexpr (FAST && !GROUPED) {
                    haveRows = havePre && nxt.nnz != 0 && fitsBatch(nxt);
                    if (haveRows) {
                        const int sgq = threadIdx.x / GE::SUB;
                        int64_t q0, q1;
                        subChunk<Real, F_INTERCEPT, false>(P, nxt.a, nxt.b, sgq, GE::NSG, q0, q1);
                        switch (nxt.fm) {
                        case F_INDICATOR: loadRows<Real, MODEL, BP, F_INDICATOR, UF>(P, jn, q0, q1, nb); break;
                        case F_SPARSE: loadRows<Real, MODEL, BP, F_SPARSE, UF>(P, jn, q0, q1, nb); break;
                        case F_DENSE: loadRows<Real, MODEL, BP, F_DENSE, UF>(P, jn, q0, q1, nb); break;
                        default: loadRows<Real, MODEL, BP, F_INTERCEPT, UF>(P, jn, q0, q1, nb); break;
                        }
                    }
                }
            };
            if (nnzj == 0) {
                // g = h = 0 without a pass (ModelSpecifics.hpp:720-723); every fit this
                // workgroup steps moves its bound; no row changes
                if (solo) {
                    if (mine(me)) (void)soloStep(monZero());
                    prefetchRows();
                } else {
                    for (int f = threadIdx.x; f < P.B; f += BLOCK)
                        if (mine(f)) step(f, monZero(), s.out + 4 * f);
                    prefetchRows();
                    __syncthreads();
                }
            } else if constexpr (SWEEP) {
                const unsigned long long tA = P.timing ? now() : 0ull;
                int64_t n0, n1;   // this thread's run-aligned chunk of the column
                switch (fm) {
                case F_INDICATOR: subChunk<Real, F_INDICATOR, true>(P, cur.a, cur.b, threadIdx.x, BLOCK, n0, n1); break;
                case F_SPARSE: subChunk<Real, F_SPARSE, true>(P, cur.a, cur.b, threadIdx.x, BLOCK, n0, n1); break;
                case F_DENSE: subChunk<Real, F_DENSE, true>(P, cur.a, cur.b, threadIdx.x, BLOCK, n0, n1); break;
                default: subChunk<Real, F_INTERCEPT, true>(P, cur.a, cur.b, threadIdx.x, BLOCK, n0, n1); break;
                }
                const bool on[1] = {P.B > 0 && s.active[0] != 0};
                const bool two = fm == F_SPARSE || fm == F_DENSE;   // x^2 != x
                // a chunk of at most UF entries stays in registers until the update
                const bool fast = __syncthreads_and(n1 - n0 <= UF);
                Batch<Real, MODEL, BP, UF> bt;
                auto scatter = [&](bool clear) {
                    if (fast) {
                        sweepScatterBatch<Real, MODEL, BP, UF>(P, bt, two, clear);
                    } else {
                        switch (fm) {
                        case F_INDICATOR: sweepScatterPass<Real, MODEL, BP, F_INDICATOR, UF>(P, j, n0, n1, on, two, clear); break;
                        case F_SPARSE: sweepScatterPass<Real, MODEL, BP, F_SPARSE, UF>(P, j, n0, n1, on, two, clear); break;
                        case F_DENSE: sweepScatterPass<Real, MODEL, BP, F_DENSE, UF>(P, j, n0, n1, on, two, clear); break;
                        default: sweepScatterPass<Real, MODEL, BP, F_INTERCEPT, UF>(P, j, n0, n1, on, two, clear); break;
                        }
                    }
                };
                if (fast) {
                    switch (fm) {
                    case F_INDICATOR: loadBatch<Real, MODEL, BP, F_INDICATOR, UF>(P, j, n0, n1, on, bt); break;
                    case F_SPARSE: loadBatch<Real, MODEL, BP, F_SPARSE, UF>(P, j, n0, n1, on, bt); break;
                    case F_DENSE: loadBatch<Real, MODEL, BP, F_DENSE, UF>(P, j, n0, n1, on, bt); break;
                    default: loadBatch<Real, MODEL, BP, F_INTERCEPT, UF>(P, j, n0, n1, on, bt); break;
                    }
                }
                scatter(false);
                __syncthreads();   // the numerators are stored before any wave walks
                if (denomStale) {
                    sweepCarryPublish<Real, BP>(P, seq, s);
                    carryPending = true;
                    denomStale = false;
                }
                if (carryPending) {
                    if (!sweepCarryWait<Real, BP>(P, seq, deadline, s, carryD)) return;
                    carryPending = false;
                }
                sweepWalk<Real, BP>(P, s, carryD, two);
                const unsigned long long tB = P.timing ? now() : 0ull;

                double g = 0.0, h = 0.0;
                if (!sweepReduce<Real, BP>(P, seq, deadline, s, g, h)) return;
                if (threadIdx.x == 0) step(0, Mon{g, h, 0.0, 0.0, 0.0, 0.0}, s.out);
                __syncthreads();
                const unsigned long long tC = P.timing ? now() : 0ull;
                scatter(true);   // this workgroup's walk is over: numerators back to zero
                double dl[1] = {s.out[0]};
                if (dl[0] != 0.0) {
                    double wsum[WAVES];
#pragma unroll
                    for (int i = 0; i < WAVES; ++i) wsum[i] = 0.0;
                    if (fast) {
                        applyBatch<Real, MODEL, BP, UF, false>(P, bt, dl, wsum);
                    } else {
                        switch (fm) {
                        case F_INDICATOR: updatePass<Real, MODEL, BP, F_INDICATOR>(P, j, n0, n1, dl, wsum); break;
                        case F_SPARSE: updatePass<Real, MODEL, BP, F_SPARSE>(P, j, n0, n1, dl, wsum); break;
                        case F_DENSE: updatePass<Real, MODEL, BP, F_DENSE>(P, j, n0, n1, dl, wsum); break;
                        default: updatePass<Real, MODEL, BP, F_INTERCEPT>(P, j, n0, n1, dl, wsum); break;
                        }
                    }
                    __syncthreads();
                    sweepAddWaveDeltas<BP>(s, wsum);
                    sweepCarryPublish<Real, BP>(P, seq, s);
                    carryPending = true;
                }
                if (w == 0 && threadIdx.x == 0) {
                    const unsigned long long tD = P.timing ? now() : 0ull;
                    loc[1] += (double)(tB - tA);
                    loc[2] += (double)(tC - tB);
                    loc[3] += (double)(tD - tC);
                    locSteps += 1;
                    // compulsory bytes: the column's ids / values / weights, xb and pid
                    // of its rows and the numerator stores; the walk reads D, nW and
                    // the numerators of every pid; an update reads and writes xb (and
                    // its low part) and the denominators of its pids
                    const double n = (double)nnzj;
                    const double base = n * (((fm == F_INDICATOR || fm == F_SPARSE) ? 4.0 : 0.0) +
                                             (two ? sizeof(Real) : 0.0) + (P.entFold && (fm == F_INDICATOR || fm == F_SPARSE) ? 1.0 : ((P.rowFold || P.kw) ? 4.0 : 0.0)));
                    constexpr double lowB = sizeof(Real) == 4 ? 2.0 : 0.0;
                    // (per pid: D in FP64, nW and the numerators in the bulk precision, a flag byte)
                    const double numB = sizeof(Real) * (two ? 2.0 : 1.0);
                    double by = base + n * (sizeof(Real) + 4.0 + numB) +
                                (double)P.N * (9.0 + sizeof(Real) + numB);
                    if (dl[0] != 0.0) by += base + n * (2.0 * (sizeof(Real) + lowB) + 4.0 + 16.0);
                    loc[0] += by;
                }
            } else {
                const unsigned long long tA = P.timing ? now() : 0ull;
                const int sg = threadIdx.x / GE::SUB, sl = threadIdx.x % GE::SUB;
                int64_t n0, n1;
                switch (fm) {
                case F_INDICATOR: subChunk<Real, F_INDICATOR, GROUPED>(P, cur.a, cur.b, sg, GE::NSG, n0, n1); break;
                case F_SPARSE: subChunk<Real, F_SPARSE, GROUPED>(P, cur.a, cur.b, sg, GE::NSG, n0, n1); break;
                case F_DENSE: subChunk<Real, F_DENSE, GROUPED>(P, cur.a, cur.b, sg, GE::NSG, n0, n1); break;
                default: subChunk<Real, F_INTERCEPT, GROUPED>(P, cur.a, cur.b, sg, GE::NSG, n0, n1); break;
                }
                bool on[GE::FPL];
#pragma unroll
                for (int i = 0; i < GE::FPL; ++i) {
                    const int f = GE::fit(sl, i);
                    on[i] = f < P.B && s.active[f];
                }
                // a chunk of at most U entries stays in registers until the update
                bool fast;
                if constexpr (GROUPED) fast = FAST && __syncthreads_and(n1 - n0 <= UF);
                else fast = FAST && fitsBatch(cur);
                Batch<Real, MODEL, BP, UF> bt;
                Mon acc[GE::FPL];
#pragma unroll
                for (int i = 0; i < GE::FPL; ++i) acc[i] = monZero();
                if (FAST && fast) {
                    if (!GROUPED && curRows) {
                        bt = nb;   // row ids and values fetched during the previous coordinate
                        const bool byEntry = P.entFold && (fm == F_INDICATOR || fm == F_SPARSE);
                        loadState<Real, MODEL, BP, UF, LOW>(P, on, bt, byEntry ? P.entFold + n0 : nullptr);
                    } else {
                        switch (fm) {
                        case F_INDICATOR: loadBatch<Real, MODEL, BP, F_INDICATOR, UF, LOW>(P, j, n0, n1, on, bt); break;
                        case F_SPARSE: loadBatch<Real, MODEL, BP, F_SPARSE, UF, LOW>(P, j, n0, n1, on, bt); break;
                        case F_DENSE: loadBatch<Real, MODEL, BP, F_DENSE, UF, LOW>(P, j, n0, n1, on, bt); break;
                        default: loadBatch<Real, MODEL, BP, F_INTERCEPT, UF, LOW>(P, j, n0, n1, on, bt); break;
                        }
                    }
                    int prevSeg = 0;
                    bool hasPrev = false;
                    if constexpr (GROUPED) {
                        if (n0 < n1 && n0 > cur.c0) {
                            prevSeg = P.seg[fm == F_INDICATOR || fm == F_SPARSE ? P.rows[n0 - 1] : n0 - 1];
                            hasPrev = true;
                        }
                    }
                    accumBatch<Real, MODEL, BP, UF>(bt, on, acc, prevSeg, hasPrev);
                } else {
                    switch (fm) {
                    case F_INDICATOR: gradPass<Real, MODEL, BP, F_INDICATOR>(P, j, n0, n1, cur.c0, acc, on); break;
                    case F_SPARSE: gradPass<Real, MODEL, BP, F_SPARSE>(P, j, n0, n1, cur.c0, acc, on); break;
                    case F_DENSE: gradPass<Real, MODEL, BP, F_DENSE>(P, j, n0, n1, cur.c0, acc, on); break;
                    default: gradPass<Real, MODEL, BP, F_INTERCEPT>(P, j, n0, n1, cur.c0, acc, on); break;
                    }
                }
                prefetchRows();
                if constexpr (GROUPED) {
#pragma unroll
                    for (int i = 0; i < GE::FPL; ++i) s.mon[sg * BP + GE::fit(sl, i)] = acc[i];
                    __syncthreads();
                    combineSubgroups<BP>(s.mon);
                } else if (solo) {
                    combineWaves<BP, GE>(acc, s.mon + par * (BLOCK / 64) * BP);
                } else {
                    combinePlain<BP, GE>(acc, s.mon);
                }
                const unsigned long long tB = P.timing ? now() : 0ull;
                if (solo) {
                    // nothing to exchange: the wave partials are in s.mon
                } else if (tagged) {
                    writeTagged<Real, BP>(P, tseq, s);
                    if (!reduceTagged<Real, BP>(P, tseq, deadline, s, step)) return;
                } else {
                    writeSlots<Real, BP>(P, seq, s, GROUPED);
                    if (!reduceFits<Real, BP>(P, seq, deadline, s, step, GROUPED)) return;
                }
                const unsigned long long tC = P.timing ? now() : 0ull;
                double dl[GE::FPL];
                int nUpd;
                if (solo) {
                    const Mon* wm = s.mon + par * (BLOCK / 64) * BP;
                    par ^= 1u;
                    double dlt = 0.0;
                    if (mine(me)) {
                        Mon m = wm[me];
                        for (int q = 1; q < BLOCK / 64; ++q) {
                            m.g += wm[q * BP + me].g;
                            m.h += wm[q * BP + me].h;
                        }
                        dlt = soloStep(m);
                    }
#pragma unroll
                    for (int i = 0; i < GE::FPL; ++i) dl[i] = __shfl(dlt, GE::fit(sl, i), 64);
                    nUpd = __popcll(__ballot(dlt != 0.0));   // lane f of the wave holds fit f
                } else {
#pragma unroll
                    for (int i = 0; i < GE::FPL; ++i) {
                        const int f = GE::fit(sl, i);
                        dl[i] = f < P.B ? s.out[4 * f] : 0.0;
                    }
                    // fits that moved (B <= BP <= BLOCK: thread f looks at fit f)
                    nUpd = __syncthreads_count(threadIdx.x < P.B && s.out[4 * threadIdx.x] != 0.0);
                }
                if (nUpd > 0) {
                    if (FAST && fast) {
                        applyBatch<Real, MODEL, BP, UF, LOW>(P, bt, dl);
                    } else {
                        switch (fm) {
                        case F_INDICATOR: updatePass<Real, MODEL, BP, F_INDICATOR>(P, j, n0, n1, dl); break;
                        case F_SPARSE: updatePass<Real, MODEL, BP, F_SPARSE>(P, j, n0, n1, dl); break;
                        case F_DENSE: updatePass<Real, MODEL, BP, F_DENSE>(P, j, n0, n1, dl); break;
                        default: updatePass<Real, MODEL, BP, F_INTERCEPT>(P, j, n0, n1, dl); break;
                        }
                    }
                    __syncthreads();
                    if constexpr (MODEL == M_COX) {
                        if (!coxRescanAny<Real, BP>(P, seq, deadline, s)) return;
                    }
                }
                if (w == 0 && threadIdx.x == 0) {
                    const unsigned long long tD = P.timing ? now() : 0ull;
                    loc[1] += (double)(tB - tA);
                    loc[2] += (double)(tC - tB);
                    loc[3] += (double)(tD - tC);
                    locSteps += 1;
                    // compulsory bytes of this coordinate: the column pass reads ids /
                    // values / weights once and xb of every active fit; the update reads
                    // and writes xb of every fit that moved
                    const double nAct = (double)nActive;
                    const double n = (double)nnzj;
                    const double base = n * (((fm == F_INDICATOR || fm == F_SPARSE) ? 4.0 : 0.0) +
                                             ((fm == F_SPARSE || fm == F_DENSE) ? sizeof(Real) : 0.0) +
                                             (P.entFold && (fm == F_INDICATOR || fm == F_SPARSE) ? 1.0 : ((P.rowFold || P.kw) ? 4.0 : 0.0)));
                    // (FP32 runs also move the low part of xb: read with the gathers on the
                    // fast path, read and written by every update)
                    constexpr double lowB = sizeof(Real) == 4 ? 2.0 : 0.0;
                    double by = base + n * nAct * (sizeof(Real) + (fast ? lowB : 0.0)) +
                                (GROUPED ? n * (8.0 + 16.0 * nAct) : 0.0);
                    if (nUpd > 0) {
                        by += base + n * nUpd * 2.0 * (sizeof(Real) + lowB) + (GROUPED ? n * (4.0 + 16.0 * nUpd) : 0.0);
                        if (MODEL == M_COX) by += (double)P.N * P.B * 56.0;
                    }
                    loc[0] += by;
                }
            }
            cur = nxt;
            if (jn <= j) break;   // wrapped: the cycle is complete
        }
        if (solo) __syncthreads();   // the last step's reads of s.mon are done before the objective reuses it
        // convergence test of every fit (performCheckConvergence :839-906)
        const unsigned long long tO = P.timing ? now() : 0ull;
        if constexpr (SWEEP) {   // D of every pid for the objective (and exact wave totals)
            if (carryPending) {   // the last update's exchange, joined so the sequence stays in step
                if (!sweepCarryWait<Real, BP>(P, seq, deadline, s, carryD)) return;
                carryPending = false;
            }
            if (!coxRescanAny<Real, BP>(P, seq, deadline, s)) return;
            denomStale = true;
        }
        objectivePartial<Real, MODEL, BP>(P, objOp, s);
        writeSlots<Real, BP>(P, seq, s, false, objOff);
        if (!reduceFits<Real, BP>(P, seq, deadline, s, [&](int f, const Mon& m, double* o) {
                if (cDone[f]) { o[0] = 1.0; return; }
                const double obj = m.g + (addFixed ? P.fixedTermFit[f] : 0.0) + (objOp == OBJ_LANGE ? logPrior(f) : 0.0);
                cIter[f] += 1;
                double conv;
                bool ill = false;
                if (obj != obj) { conv = 0.0; ill = true; }
                else conv = fabs(obj - cLast[f]) / (fabs(obj) + 1.0);
                cLast[f] = obj;
                int done = 0;
                if (P.tol > 0 && conv < P.tol) { cFlag[f] = ill ? CCD_ILLCONDITIONED : CCD_SUCCESS; done = 1; }
                else if (cIter[f] == P.maxIt) { cFlag[f] = CCD_MAX_ITERATIONS; done = 1; }
                cDone[f] = done;
                o[0] = done;
            }, false, objOff))
            return;
        int notDone = 0;
        for (int f = threadIdx.x; f < P.B; f += BLOCK) {
            const int act = (s.out[4 * f] == 0.0) ? 1 : 0;
            s.active[f] = act;
            notDone |= act;
        }
        nActive = __syncthreads_count(notDone);
        allDone = nActive == 0;
        if (w == 0 && threadIdx.x == 0) {   // flush this cycle's counters (once per cycle, not per coordinate)
            if (P.timing) atomicAdd(&P.stats[4], (double)(now() - tO));
            for (int i = 0; i < 4; ++i) atomicAdd(&P.stats[i], loc[i]);
            atomicAdd(P.steps, locSteps);
            loc[0] = loc[1] = loc[2] = loc[3] = 0.0;
            locSteps = 0;
        }
        if (w == 0 && threadIdx.x == 0) {   // objective pass: y, xb of every fit, per-pid terms
            double nAct = 0;
            for (int f = 0; f < P.B; ++f) nAct += 1;
            atomicAdd(&P.stats[0], (double)P.K * (sizeof(Real) * nAct + sizeof(Real) + (P.rowFold ? 4.0 : 0.0)) +
                                       (GROUPED ? (double)P.N * nAct * 16.0 : 0.0));
        }
    }
    if (own && w == 0) {   // persist the control state for a later launch
        for (int f = threadIdx.x; f < P.B; f += BLOCK) {
            P.ctlIter[f] = s.iter[f];
            P.ctlFlag[f] = s.flag[f];
            P.ctlLast[f] = s.last[f];
            P.done[f] = s.done[f];
        }
    }
    if (w == 0 && threadIdx.x == 0 && !allDone) atomicAnd(P.allDoneOut, 0);
    if (!allDone) return;
    // final log-likelihood of every fit (reported beside beta)
    objectivePartial<Real, MODEL, BP>(P, OBJ_LOGLIK, s);
    writeSlots<Real, BP>(P, seq, s, false, objOff);
    if (!reduceFits<Real, BP>(P, seq, deadline, s, [&](int f, const Mon& m, double* o) {
            o[0] = m.g + ((MODEL == M_POISSON) ? P.fixedTermFit[f] : 0.0);
            o[1] = cIter[f];
            o[2] = cFlag[f];
        }, false, objOff))
        return;
    if (w == 0)
        for (int f = threadIdx.x; f < P.B; f += BLOCK) {
            P.ll[f] = s.out[4 * f];
            P.iters[f] = (int)s.out[4 * f + 1];
            P.flag[f] = (int)s.out[4 * f + 2];
        }
}

}  // namespace res
}  // namespace ccd
