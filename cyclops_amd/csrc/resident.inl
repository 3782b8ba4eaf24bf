// resident.inl -- host side of the device-resident and batched fits
// (kFindMode in resident_kernels.h). Included by engine.hip after EngineT.
#pragma once

#include "resident_kernels.h"

#include <array>
#include <chrono>
#include <map>
#include <numeric>
#include <cstdlib>
#include <cstdio>

namespace ccd {

// Offset of the first entry of column j with row >= rowStart[w], relative to the
// column start: lower_bound per (column, workgroup).
__global__ __launch_bounds__(BLOCK) void kColOff(int J, int G, const int* __restrict__ fmt,
                                                const int64_t* __restrict__ colPtr, const int* __restrict__ rows,
                                                const int* __restrict__ rowStart, int* __restrict__ colOff) {
    const int64_t total = (int64_t)J * (G + 1);
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x; t < total; t += stride) {
        const int j = (int)(t / (G + 1)), w = (int)(t % (G + 1));
        const int f = fmt[j];
        if (!(f == F_INDICATOR || f == F_SPARSE)) { colOff[t] = 0; continue; }
        const int64_t a = colPtr[j], n = colPtr[j + 1] - a;
        const int target = rowStart[w];
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (rows[a + mid] < target) lo = mid + 1;
            else hi = mid;
        }
        colOff[t] = (int)lo;
    }
}

// Absolute entry range of workgroup w in column j: [rng[2(jG+w)], rng[2(jG+w)+1]).
__global__ __launch_bounds__(BLOCK) void kColRange(int J, int G, const int* __restrict__ fmt,
                                                  const int64_t* __restrict__ colPtr, const int* __restrict__ colOff,
                                                  const int* __restrict__ rowStart, long long* __restrict__ rng) {
    const int64_t total = (int64_t)J * G;
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x; t < total; t += stride) {
        const int j = (int)(t / G), w = (int)(t % G);
        const int f = fmt[j];
        if (f == F_INDICATOR || f == F_SPARSE) {
            const int* off = colOff + (size_t)j * (G + 1);
            rng[2 * t] = colPtr[j] + off[w];
            rng[2 * t + 1] = colPtr[j] + off[w + 1];
        } else {
            rng[2 * t] = rowStart[w];
            rng[2 * t + 1] = rowStart[w + 1];
        }
    }
}

// Fold of the row of every INDICATOR/SPARSE entry (the column pass then reads one
// byte per entry in order instead of gathering rowFold by row).
__global__ __launch_bounds__(BLOCK) void kEntryFolds(int64_t n, const int* __restrict__ rows,
                                                    const int* __restrict__ rowFold, signed char* __restrict__ out) {
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < n; e += (int64_t)gridDim.x * BLOCK)
        out[e] = (signed char)rowFold[rows[e]];
}

// sum over column j of x*y, in total and per fold (F <= 32 folds).
template <typename Real>
__global__ __launch_bounds__(BLOCK) void kXjYFold(int J, int64_t K, int F, const int* __restrict__ fmt,
                                                 const int64_t* __restrict__ colPtr,
                                                 const int64_t* __restrict__ densePtr, const int* __restrict__ rows,
                                                 const Real* __restrict__ vals, const Real* __restrict__ dense,
                                                 const Real* __restrict__ y, const int* __restrict__ rowFold,
                                                 double* __restrict__ tot, double* __restrict__ byFold) {
    __shared__ double red[33][WAVES];
    for (int j = blockIdx.x; j < J; j += gridDim.x) {
        double acc[33];
        for (int i = 0; i < 33; ++i) acc[i] = 0.0;
        const int f = fmt[j];
        const bool idx = (f == F_INDICATOR || f == F_SPARSE);
        const int64_t a = idx ? colPtr[j] : 0, b = idx ? colPtr[j + 1] : K;
        for (int64_t e = a + threadIdx.x; e < b; e += BLOCK) {
            const int64_t k = idx ? rows[e] : e;
            double c = (double)y[k];
            if (f == F_SPARSE) c *= (double)vals[e];
            if (f == F_DENSE) c *= (double)dense[densePtr[j] + e];
            acc[32] += c;
            if (rowFold) {
                const int q = rowFold[k];
                for (int i = 0; i < 32; ++i)
                    if (i == q) acc[i] += c;
            }
        }
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        for (int i = 0; i < 33; ++i) {
            double x = acc[i];
            for (int d = 32; d > 0; d >>= 1) x += __shfl_down(x, d, 64);
            if (lane == 0) red[i][wid] = x;
        }
        __syncthreads();
        if (threadIdx.x < 33) {
            double s = 0.0;
            for (int wv = 0; wv < WAVES; ++wv) s += red[threadIdx.x][wv];
            if (threadIdx.x == 32) tot[j] = s;
            else if ((int)threadIdx.x < F && byFold) byFold[(size_t)j * F + threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// xb of every fit from the row-major copy of the INDICATOR/SPARSE columns: thread
// (k, f) sums beta_f[j] * x_kj over row k's entries in column order and stores it
// (DENSE/INTERCEPT columns are added afterwards, column by column).
template <typename Real>
__global__ __launch_bounds__(BLOCK) void kXBetaRows(int64_t K, int B, int Bt, int BP, int J,
                                                   const int64_t* __restrict__ rowPtr,
                                                   const int32_t* __restrict__ rowCol,
                                                   const Real* __restrict__ rowVal,
                                                   const double* __restrict__ beta, Real* __restrict__ xb,
                                                   unsigned short* __restrict__ xbLo) {
    const int64_t total = K * (int64_t)B;
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x; t < total; t += stride) {
        const int64_t k = t / B;
        const int f = (int)(t % B);
        const double* bf = beta + (size_t)f * J;
        double acc = 0.0;
        for (int64_t e = rowPtr[k]; e < rowPtr[k + 1]; ++e) acc += bf[rowCol[e]] * (double)rowVal[e];
        const size_t q = (size_t)(f / Bt) * K * BP + k * BP + f % Bt;
        xb[q] = (Real)acc;
        if (xbLo) xbLo[q] = bf16Bits((float)(acc - (double)xb[q]));   // FP32 runs: the rounding remainder
    }
}

// xb[k*BP + f] += beta_f * x_kj for every fit f (one column).
template <typename Real, int FMT>
__global__ __launch_bounds__(BLOCK) void kAxpyBatch(Col<Real> c, const double* __restrict__ beta, int J, int j,
                                                   int B, int Bt, int BP, int64_t K, Real* __restrict__ xb,
                                                   unsigned short* __restrict__ xbLo) {
    // xb of fit f lives in team f / Bt, lane f % Bt: xb[(f/Bt)*K*BP + row*BP + f%Bt]
    const int64_t total = c.n * (int64_t)B;
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x; t < total; t += stride) {
        const int64_t e = t / B;
        const int f = (int)(t % B);
        const double bf = beta[(size_t)f * J + j];
        if (bf == 0.0) continue;
        const size_t q = (size_t)(f / Bt) * K * BP + rowOf<Real, FMT>(c, e) * BP + f % Bt;
        if (xbLo) {   // FP32 runs: compensated, as the kernel's own updates
            const double sum = (double)xb[q] + (double)bf16Value(xbLo[q]) + bf * (double)xOf<Real, FMT>(c, e);
            xb[q] = (Real)sum;
            xbLo[q] = bf16Bits((float)(sum - (double)xb[q]));
        } else {
            xb[q] += (Real)bf * xOf<Real, FMT>(c, e);
        }
    }
}

static int padFits(int B) {
    int p = 1;
    while (p < B) p *= 4;
    return p;
}

// Device workspace of the persistent kernel, kept across fits of one engine.
template <typename Real>
struct ResWork {
    int G = 0, BP = 0, B = 0, nTeams = 0;
    int64_t N = 0;
    double bytes = 0.0;
    // structure (depends on G)
    int* rowStart = nullptr;
    int* pidStart = nullptr;
    int* colOff = nullptr;
    long long* rng = nullptr;
    long long* info = nullptr;
    int* pid = nullptr;       // Cox: unweighted tie pids
    int* pstart = nullptr;
    int* pseg = nullptr;
    unsigned char* pflag = nullptr;   // Cox: 1 where a pid begins a stratum
    int* waveLast = nullptr;          // Cox: last stratum start of each wave range [G][WAVES]
    // per batch
    Real* xb = nullptr;
    unsigned short* xbLo = nullptr;   // FP32 runs: low part of xb (bfloat16 bits)
    double* predBuf = nullptr;        // held-out predictive pass (grow-only)
    size_t predCap = 0;
    double* timeDev = nullptr;        // Cox survival times (held-out risk sets)
    double *denom = nullptr, *Dacc = nullptr, *R1 = nullptr, *R2 = nullptr, *nW = nullptr;
    Real *numer = nullptr, *numer2 = nullptr, *nWs = nullptr;   // Cox sweep (one fit per team)
    double *beta = nullptr, *hDelta = nullptr, *var = nullptr, *initBound = nullptr;
    unsigned char *exclude = nullptr, *fixed = nullptr;
    double *xjyTot = nullptr, *xjyFold = nullptr, *xjyW = nullptr, *fixedTerm = nullptr;
    int *held = nullptr, *rowFold = nullptr, *fixJ = nullptr;
    signed char* entFold = nullptr;   // [nnz] row fold per INDICATOR/SPARSE entry (allocated on first use)
    int *flag = nullptr, *iters = nullptr, *done = nullptr, *ctlIter = nullptr, *ctlFlag = nullptr;
    double* ctlLast = nullptr;
    double* ll = nullptr;
    unsigned* epoch = nullptr;
    double *slots = nullptr, *bcast = nullptr;
    int* err = nullptr;
    int* allDone = nullptr;
    unsigned long long* steps = nullptr;
    double* stats = nullptr;
    Real* kw = nullptr;
    res::Params<Real>* teams = nullptr;
    // parameter blocks of fused launches led by this engine (grow-only, so repeated
    // fused launches call no hipMalloc / hipFree, which wait for the whole device)
    res::Params<Real>* fusedTeams = nullptr;
    size_t fusedCap = 0;
    std::map<int, std::array<int, 3>> tuned;   // batch size -> {Bt, T, nTeams} picked by timing

    void freeBatch() {
        dfree(xb); dfree(xbLo); dfree(predBuf); predCap = 0; dfree(timeDev); dfree(denom); dfree(Dacc); dfree(R1); dfree(R2); dfree(nW);
        dfree(numer); dfree(numer2); dfree(nWs);
        dfree(beta); dfree(hDelta); dfree(var); dfree(initBound); dfree(exclude); dfree(fixed);
        dfree(xjyTot); dfree(xjyFold); dfree(xjyW); dfree(fixedTerm); dfree(held); dfree(rowFold); dfree(entFold); dfree(fixJ);
        dfree(flag); dfree(iters); dfree(done); dfree(ctlIter); dfree(ctlFlag); dfree(ctlLast); dfree(ll); dfree(epoch); dfree(slots); dfree(bcast);
        dfree(err); dfree(allDone); dfree(steps); dfree(stats); dfree(kw); dfree(teams); dfree(info);
        dfree(fusedTeams);
        fusedCap = 0;
        B = BP = nTeams = 0;
    }
    void freeAll() {
        freeBatch();
        dfree(rowStart); dfree(pidStart); dfree(colOff); dfree(rng); dfree(pid); dfree(pstart); dfree(pseg);
        dfree(pflag); dfree(waveLast);
        G = 0;
    }
};

template <typename Real>
void EngineT<Real>::releaseResident() {
    if (res) {
        auto* W = static_cast<ResWork<Real>*>(res);
        W->freeAll();
        delete W;
        res = nullptr;
    }
}

// Workgroup count of the persistent launch: enough rows per workgroup to amortise
// the per-coordinate grid exchange, never more workgroups than compute units.
// One or two fits are bound by the latency of the per-coordinate exchange, which
// grows with the number of workgroups; batches are bound by bandwidth and use the
// whole chip. Grouped models (CLR, Cox) also walk their per-pid state after every
// update, so they take 4x more workgroups (a single Cox fit 16x).
static int chooseWorkgroups(int64_t K, int B, int device, int model) {
    const bool grouped = model == M_CLR || model == M_COX;
    int cus = 0;
    CCD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    int cap = std::min(res::MAXG, cus);
    if (B <= 2 && threadMaxWorkgroups() > 0) cap = std::min(cap, threadMaxWorkgroups());
    if (const char* e = std::getenv("CCD_RESIDENT_WG")) {
        const int v = std::atoi(e);
        if (v > 0) return std::min(v, cap);
    }
    // One fit of independent rows on up to 100k rows runs fastest on one workgroup,
    // with no per-coordinate grid exchange: 15-30% less per cycle than on 2-8
    // workgroups (LR/PR, 16k-87k rows, 1.4k-8.8k columns). Grouped models keep gaining
    // from more workgroups at these sizes (their risk-set rescan is split over them).
    if (B == 1 && !grouped && K <= 100000) return 1;
    // One Cox fit: every coordinate rescans all risk-set positions, split over the
    // workgroups; about 1k rows per workgroup was fastest (11k-43k rows: 1.4-2.5x
    // less per cycle than 4k rows per workgroup). CLR was flat from 6 to 48 workgroups.
    const int64_t rowsPerWG = (B == 1 && model == M_COX) ? 1024 : (B <= 2 && !grouped) ? 16384 : 4096;
    int64_t g = (K + rowsPerWG - 1) / rowsPerWG;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// Layout of a persistent launch: nTeams teams of T workgroups; team t runs fits
// [t*Bt, min((t+1)*Bt, B)) packed fit-minor (padded to BP lanes). Teams never wait
// on each other; a team of T > 1 workgroups exchanges per coordinate, so all
// nTeams*T workgroups must be co-resident when T > 1.
struct Layout {
    int nTeams = 1, T = 1, Bt = 1;
};

// CCD_BATCH_LAYOUT="Bt,T" overrides the team shape of batches (B > 2).
static Layout chooseLayout(int64_t K, int B, int device, int model) {
    Layout L;
    int bt = 0, t = 0;
    if (const char* e = std::getenv("CCD_BATCH_LAYOUT")) std::sscanf(e, "%d,%d", &bt, &t);
    if (B <= 2) {   // one or two fits: the whole chip in lockstep
        L.Bt = B;
        L.T = chooseWorkgroups(K, B, device, model);
    } else if (bt > 0 && t > 0) {
        L.Bt = std::min(bt, 256);
        L.T = std::min(t, res::MAXG);
    } else if (K <= 200000) {   // short data: independent one-workgroup fits, no exchange
        L.Bt = 1;
        L.T = 1;
    } else {        // teams of 4 fits on about 32k rows per workgroup
        L.Bt = 4;
        L.T = (int)std::min<int64_t>(64, std::max<int64_t>(8, K / 32768));
    }
    L.nTeams = (B + L.Bt - 1) / L.Bt;
    return L;
}

// Fits per persistent launch: lockstep teams must be co-resident, so a launch holds
// what fits on the chip; one-workgroup teams are independent and take any count,
// bounded by the per-team state (xb per team, per-pid arrays for grouped models)
// staying within 16 GiB.
static int chunkFits(const Layout& L, int64_t K, int64_t N, bool grouped, int realBytes) {
    if (L.T > 1) return L.Bt * std::max(1, 2048 / (L.T * 4));
    int BP = 1;
    while (BP < L.Bt) BP *= 4;
    const double perTeam = (double)K * BP * realBytes + (grouped ? (double)N * BP * 40.0 : 0.0);
    const double budget = 16.0 * (1ull << 30);
    const int teams = (int)std::max(1.0, std::min(4096.0, budget / perTeam));
    return teams * L.Bt;
}

// Launch of kFindMode<Real, MODEL, BP> over nTeams teams of T workgroups; defined
// per model in resident_<model>.hip so the instantiations compile in parallel.
template <typename Real, int MODEL>
void launchByBP(int BP, const res::Params<Real>* teams, int nTeams, int T, hipStream_t s);
template <typename Real, int MODEL>
int occupancyByBP(int BP);

template <typename Real>
static int occupancyOf(int model, int BP) {
    switch (model) {
    case M_LOGISTIC: return occupancyByBP<Real, M_LOGISTIC>(BP);
    case M_POISSON: return occupancyByBP<Real, M_POISSON>(BP);
    case M_CLR: return occupancyByBP<Real, M_CLR>(BP);
    default: return occupancyByBP<Real, M_COX>(BP);
    }
}

// Team shapes worth timing for a batch of B fits on K rows: for up to 16 grouped
// fits, one fit per team; one lockstep team of every fit; narrower teams (4, 16
// or 64 fits) of 8 to 128 workgroups; each shape co-resident on the device with
// 2,048 to 65,536 rows per workgroup; plus one-workgroup teams for short data.
template <typename Real>
static std::vector<Layout> layoutCandidates(int model, int64_t K, int B, int device) {
    int cus = 0;
    CCD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    std::vector<Layout> out;
    auto add = [&](int bt, int t) {
        Layout L;
        L.Bt = bt;
        L.T = t;
        L.nTeams = (B + bt - 1) / bt;
        const int cap = occupancyOf<Real>(model, padFits(bt)) * cus;
        if (t > 1 && (int64_t)L.nTeams * t > cap) return;
        for (const Layout& q : out) if (q.Bt == L.Bt && q.T == L.T) return;
        out.push_back(L);
    };
    const bool grouped = model == M_CLR || model == M_COX;
    // Cox: one fit per team, as many workgroups per team as keep every team resident
    // (each team streams its own fit's per-pid state once per coordinate in the
    // accumulated sweep); these probes are slow, so the first one usually decides
    if (model == M_COX) {
        int t = cus * occupancyOf<Real>(model, 1) / std::max(B, 1);
        if (t >= 1) add(1, std::min(t, 512));
    }
    // a few grouped fits: one fit per team first (each team walks its own risk sets
    // with the one-fit rescan)
    if (grouped && B <= 16)
        for (int t : {64, 128})
            if (K / t >= 2048) add(1, t);
    if (B <= 256) add(B, chooseWorkgroups(K, B, device, model));
    for (int bt : {4, 16, 64}) {
        if (bt >= B) continue;
        for (int t : {8, 16, 24, 32, 48, 64, 96, 128})
            if (K / t >= 2048 && K / t <= 65536) add(bt, t);   // 2k..64k rows per workgroup
    }
    if (K <= 200000) { add(1, 1); add(4, 1); }
    return out;
}

// A prepared persistent run: the fits' device state and one parameter block per team.
template <typename Real>
struct ResRun {
    EngineT<Real>* E = nullptr;
    ResWork<Real>* W = nullptr;
    int B = 0, J = 0, BP = 0, nTeams = 0, G = 0, m = 0, rateKHz = 100000;
    std::vector<res::Params<Real>> teamP, blockP;
    // the kernel reads workgroup i's block: its team's parameters with w = i mod G
    void fill() {
        blockP.resize((size_t)nTeams * G);
        for (int t = 0; t < nTeams; ++t)
            for (int w = 0; w < G; ++w) {
                blockP[(size_t)t * G + w] = teamP[t];
                blockP[(size_t)t * G + w].w = w;
            }
    }
};

// Set up fits [0, B) (host arrays indexed by fit) of engine E for persistent
// launches: structure, per-batch buffers, inputs, xb, parameter blocks.
// kwHost: row weights of a single fit.
template <typename Real>
static ResRun<Real> prepareResident(EngineT<Real>& E, const ccd_fit_args& a, int B, const double* var,
                                    const int32_t* held, const int32_t* rowFold, int F, const double* kwHost,
                                    const double* initBound, const double* betas, Layout L,
                                    const int32_t* fixJ = nullptr) {
    auto& D = *E.D;
    hipStream_t s = E.s;
    const int64_t K = D.K;
    const int J = D.J;
    const int m = D.model;
    const bool grouped = (m == M_CLR || m == M_COX);
    if (E.stepped) throw std::runtime_error("a stepped batch is open on this engine: finish it first");
    if (!E.res) E.res = new ResWork<Real>();
    auto& W = *static_cast<ResWork<Real>*>(E.res);
    const int G = L.T;
    const int nTeams = L.nTeams;
    const int BP = padFits(L.Bt);
    if (BP > 256) throw std::invalid_argument("at most 256 fits per team");
    if ((int64_t)nTeams * L.Bt < B || (int64_t)(nTeams - 1) * L.Bt >= B)
        throw std::logic_error("layout does not cover the fits");
    // ---- structure for this team size -------------------------------------------
    if (W.G != G) {
        W.freeAll();
        W.G = G;
        std::vector<int32_t> pid(K), pstart, pseg;
        std::vector<unsigned char> fl;
        std::vector<int> wl;
        int64_t N;
        if (m == M_COX) {   // unweighted Breslow tie groups (ModelSpecifics.hpp:1809-1868 with no weights)
            int cur = 0;
            pid[0] = 0;
            pseg.push_back(D.strat[0]);
            for (int64_t k = 1; k < K; ++k) {
                if (D.strat[k] != D.strat[k - 1]) { ++cur; pseg.push_back(D.strat[k]); }
                else if (!(D.y[k - 1] == 1.0 && D.time[k - 1] == D.time[k] && D.y[k] == 1.0)) { ++cur; pseg.push_back(D.strat[k]); }
                pid[k] = cur;
            }
            N = cur + 1;
            pstart.assign(N + 1, 0);
            for (int64_t k = K - 1; k >= 0; --k) pstart[pid[k]] = (int32_t)k;
            pstart[N] = (int32_t)K;
        } else if (m == M_CLR) {
            N = D.nStrata;
            pstart = D.stratStart;
            pseg.resize(N);
            std::iota(pseg.begin(), pseg.end(), 0);
        } else {
            N = K;
        }
        W.N = N;
        // row ranges: even split, moved to the next pid start for grouped models;
        // Cox splits the pids evenly instead (every coordinate walks every pid, while
        // a column touches a few rows of each workgroup)
        std::vector<int32_t> rs(G + 1), ps(G + 1, 0);
        for (int w = 0; w <= G; ++w) {
            int64_t r = (K * w) / G;
            if (m == M_COX && w > 0 && w < G) {
                const int64_t p = (N * w) / G;
                r = pstart[p];
                ps[w] = (int32_t)p;
            } else if (grouped && w > 0 && w < G) {
                int64_t p = (m == M_CLR) ? ((r < K) ? D.strat[r] : N) : ((r < K) ? pid[r] : N);
                if (p < N && pstart[p] < r) ++p;
                r = (p < N) ? pstart[p] : K;
                ps[w] = (int32_t)p;
            }
            rs[w] = (int32_t)r;
        }
        for (int w = 1; w <= G; ++w) if (rs[w] < rs[w - 1]) rs[w] = rs[w - 1];
        if (grouped) {
            ps[0] = 0;
            ps[G] = (int32_t)N;
            for (int w = 1; w <= G; ++w) if (ps[w] < ps[w - 1]) ps[w] = ps[w - 1];
        }
        double& Bb = W.bytes;
        W.rowStart = dalloc<int>(G + 1, Bb);
        h2d(W.rowStart, rs.data(), G + 1, s);
        W.pidStart = dalloc<int>(G + 1, Bb);
        h2d(W.pidStart, ps.data(), G + 1, s);
        W.colOff = dalloc<int>((size_t)J * (G + 1), Bb);
        kColOff<<<gridFor((int64_t)J * (G + 1)), BLOCK, 0, s>>>(J, G, D.dFmt, D.dColPtr, D.dRows, W.rowStart, W.colOff);
        CCD_HIP(hipGetLastError());
        W.rng = dalloc<long long>((size_t)J * G * 2, Bb);
        kColRange<<<gridFor((int64_t)J * G), BLOCK, 0, s>>>(J, G, D.dFmt, D.dColPtr, W.colOff, W.rowStart, W.rng);
        CCD_HIP(hipGetLastError());
        if (m == M_COX) {
            W.pid = dalloc<int>(K, Bb);
            h2d(W.pid, pid.data(), K, s);
        }
        if (grouped) {
            W.pstart = dalloc<int>(N + 1, Bb);
            h2d(W.pstart, pstart.data(), N + 1, s);
            W.pseg = dalloc<int>(N + 1, Bb);
            pseg.push_back(-1);
            h2d(W.pseg, pseg.data(), N + 1, s);
            if (m == M_COX) {   // fl lives until the synchronisation below
                fl.assign(N, 0);
                for (int64_t p = 0; p < N; ++p) fl[p] = (p == 0 || pseg[p] != pseg[p - 1]) ? 1 : 0;
                W.pflag = dalloc<unsigned char>(N, Bb);
                h2d(W.pflag, fl.data(), N, s);
                // the sweep's wave ranges split each workgroup's pids as the rescan does
                wl.assign((size_t)G * WAVES, -1);
                for (int w = 0; w < G; ++w) {
                    const int64_t q0 = ps[w], q1 = ps[w + 1], cw = (q1 - q0 + WAVES - 1) / WAVES;
                    for (int i = 0; i < WAVES; ++i) {
                        const int64_t a = std::min(q0 + i * cw, q1), b = std::min(a + cw, q1);
                        for (int64_t p = b - 1; p >= a; --p)
                            if (p > 0 && fl[p]) { wl[(size_t)w * WAVES + i] = (int)p; break; }
                    }
                }
                W.waveLast = dalloc<int>((size_t)G * WAVES, Bb);
                h2d(W.waveLast, wl.data(), (size_t)G * WAVES, s);
            }
        }
        CCD_HIP(hipStreamSynchronize(s));
    }
    const int64_t N = W.N;
    // ---- per-batch buffers ------------------------------------------------------
    if (W.B != B || W.BP != BP || W.nTeams != nTeams) {
        W.freeBatch();
        W.B = B;
        W.BP = BP;
        W.nTeams = nTeams;
        double& Bb = W.bytes;
        W.xb = dalloc<Real>((size_t)nTeams * K * BP, Bb);
        if (sizeof(Real) == 4) W.xbLo = dalloc<unsigned short>((size_t)nTeams * K * BP, Bb);
        if (grouped) {
            const size_t pn = (size_t)nTeams * N * BP;
            W.denom = dalloc<double>(pn, Bb);
            W.Dacc = dalloc<double>(pn, Bb);
            W.R1 = dalloc<double>(pn, Bb);
            W.R2 = dalloc<double>(pn, Bb);
            W.nW = dalloc<double>(pn, Bb);
            if (m == M_COX && BP == 1) {
                // per-coordinate scratch and event weights of the sweep in the bulk
                // precision (the walk streams them); the denominators stay FP64
                W.numer = dalloc<Real>(pn, Bb);
                W.numer2 = dalloc<Real>(pn, Bb);
                W.nWs = dalloc<Real>(pn, Bb);
            }
        }
        W.beta = dalloc<double>((size_t)B * J, Bb);
        W.hDelta = dalloc<double>((size_t)B * J, Bb);
        W.var = dalloc<double>(B, Bb);
        W.initBound = dalloc<double>(B, Bb);
        W.info = dalloc<long long>((size_t)J * 4, Bb);
        W.exclude = dalloc<unsigned char>(J, Bb);
        W.fixed = dalloc<unsigned char>(J, Bb);
        W.xjyTot = dalloc<double>(J, Bb);
        W.xjyFold = dalloc<double>((size_t)J * 32, Bb);
        W.xjyW = dalloc<double>(J, Bb);
        W.fixedTerm = dalloc<double>(B, Bb);
        W.held = dalloc<int>(B, Bb);
        W.fixJ = dalloc<int>(B, Bb);
        W.rowFold = dalloc<int>(K, Bb);
        W.flag = dalloc<int>(B, Bb);
        W.iters = dalloc<int>(B, Bb);
        W.done = dalloc<int>(B, Bb);
        W.ctlIter = dalloc<int>(B, Bb);
        W.ctlFlag = dalloc<int>(B, Bb);
        W.ctlLast = dalloc<double>(B, Bb);
        W.ll = dalloc<double>(B, Bb);
        W.epoch = dalloc<unsigned>((size_t)nTeams * G, Bb);
        W.slots = dalloc<double>((size_t)nTeams * 2 * G * BP * res::SLOTW, Bb);
        W.bcast = dalloc<double>((size_t)nTeams * 2 * BP * 4, Bb);
        W.err = dalloc<int>(1, Bb);
        W.allDone = dalloc<int>(1, Bb);
        W.steps = dalloc<unsigned long long>(1, Bb);
        W.stats = dalloc<double>(8, Bb);
        W.kw = dalloc<Real>(K, Bb);
        W.teams = dalloc<res::Params<Real>>((size_t)nTeams * G, Bb);   // one block per workgroup
    }
    // ---- inputs ---------------------------------------------------------------
    std::vector<unsigned char> excl(J, 0), fix(J, 0);
    for (int j = 0; j < J; ++j) {
        if (a.exclude && a.exclude[j]) excl[j] = 1;
        if (a.fixed && a.fixed[j]) fix[j] = 1;
    }
    if (D.hasOffset) fix[0] = 1;
    std::vector<double> hd((size_t)B * J);
    for (int f = 0; f < B; ++f) std::fill(hd.begin() + (size_t)f * J, hd.begin() + (size_t)(f + 1) * J, initBound[f]);
    h2d(W.beta, betas, (size_t)B * J, s);
    h2d(W.hDelta, hd.data(), (size_t)B * J, s);
    h2d(W.var, var, B, s);
    h2d(W.initBound, initBound, B, s);
    h2d(W.exclude, excl.data(), J, s);
    h2d(W.fixed, fix.data(), J, s);
    // column table: fmt | exclude << 32, nnz | next non-fixed column << 32, c0
    int firstJ = 0;
    while (firstJ < J && fix[firstJ]) ++firstJ;
    std::vector<long long> info((size_t)J * 4, 0);
    {
        int nextFree = firstJ;   // next non-fixed column after j, wrapping to firstJ
        for (int j = J - 1; j >= 0; --j) {
            const bool sp = D.fmt[j] == F_INDICATOR || D.fmt[j] == F_SPARSE;
            const int64_t nnz = sp ? D.colPtr[j + 1] - D.colPtr[j] : K;
            info[(size_t)j * 4 + 0] = (long long)(unsigned)D.fmt[j] | ((long long)excl[j] << 32);
            info[(size_t)j * 4 + 1] = (long long)(unsigned)nnz | ((long long)nextFree << 32);
            info[(size_t)j * 4 + 2] = sp ? D.colPtr[j] : 0;
            if (!fix[j]) nextFree = j;
        }
    }
    h2d(W.info, info.data(), (size_t)J * 4, s);
    std::vector<int32_t> hh(B, -1);
    if (held) std::copy(held, held + B, hh.begin());
    h2d(W.held, hh.data(), B, s);
    std::vector<int32_t> fj(B, -1);
    if (fixJ) std::copy(fixJ, fixJ + B, fj.begin());
    h2d(W.fixJ, fj.data(), B, s);
    if (rowFold) h2d(W.rowFold, rowFold, K, s);
    // entry folds (checkBatch keeps fold ids in [0, 32), so each fits a byte)
    const bool entryFolds = rowFold != nullptr && D.nnz > 0;
    if (entryFolds) {
        if (!W.entFold) W.entFold = dalloc<signed char>((size_t)D.nnz, W.bytes);
        kEntryFolds<<<gridFor(D.nnz), BLOCK, 0, s>>>(D.nnz, D.dRows, W.rowFold, W.entFold);
        CCD_HIP(hipGetLastError());
    }
    if (kwHost) h2dConvert(W.kw, kwHost, K, s);
    {   // XjY: total (and per fold), or weighted for a single weighted fit
        const int grid = std::min(J, 4096);
        if (kwHost) {
            kXjY<Real, true><<<grid, BLOCK, 0, s>>>(J, K, D.dFmt, D.dColPtr, D.dDensePtr, D.dRows, D.dVals, D.dDense,
                                                   D.dY, W.kw, W.xjyW);
        }
        kXjYFold<Real><<<grid, BLOCK, 0, s>>>(J, K, F, D.dFmt, D.dColPtr, D.dDensePtr, D.dRows, D.dVals, D.dDense,
                                              D.dY, rowFold ? W.rowFold : nullptr, W.xjyTot,
                                              rowFold ? W.xjyFold : nullptr);
        CCD_HIP(hipGetLastError());
    }
    // Poisson fixed term per fit: sum_k w_fk * (-log y_k!)
    std::vector<double> ft(B, 0.0);
    if (m == M_POISSON) {
        std::vector<double> byFold(F > 0 ? F : 1, 0.0);
        double total = 0.0, totalW = 0.0;
        for (int64_t k = 0; k < K; ++k) {
            double c = 0.0;
            for (int i = 2; i <= (int)D.y[k]; ++i) c -= std::log((double)i);
            total += c;
            if (rowFold) byFold[rowFold[k]] += c;
            if (kwHost) totalW += c * kwHost[k];
        }
        for (int f = 0; f < B; ++f)
            ft[f] = kwHost ? totalW : total - ((rowFold && hh[f] >= 0) ? byFold[hh[f]] : 0.0);
    }
    h2d(W.fixedTerm, ft.data(), B, s);
    // xb = X beta for every fit (column order, as CyclicCoordinateDescent::computeXBeta);
    // team t's block is xb[t*K*BP ...], fit-minor inside the team
    CCD_HIP(hipMemsetAsync(W.xb, 0, (size_t)nTeams * K * BP * sizeof(Real), s));
    if (W.numer) {   // the sweep keeps the numerators at zero between coordinates
        CCD_HIP(hipMemsetAsync(W.numer, 0, (size_t)nTeams * N * sizeof(Real), s));
        CCD_HIP(hipMemsetAsync(W.numer2, 0, (size_t)nTeams * N * sizeof(Real), s));
    }
    if (W.xbLo) CCD_HIP(hipMemsetAsync(W.xbLo, 0, (size_t)nTeams * K * BP * sizeof(unsigned short), s));
    {
        // warm starts with many non-zero sparse columns form the sparse part of xb in
        // one pass over rows; otherwise (and for DENSE/INTERCEPT columns) one axpy per column
        std::vector<char> live(J, 0);
        int liveSparse = 0;
        for (int j = 0; j < J; ++j) {
            for (int f = 0; f < B && !live[j]; ++f) live[j] = betas[(size_t)f * J + j] != 0.0;
            if (live[j] && (D.fmt[j] == F_INDICATOR || D.fmt[j] == F_SPARSE)) ++liveSparse;
        }
        const bool byRows = liveSparse > 32;
        if (byRows) {
            D.buildRowMajor(s);
            kXBetaRows<Real><<<gridFor(K * (int64_t)B), BLOCK, 0, s>>>(K, B, L.Bt, BP, J, D.dRowPtr, D.dRowCol,
                                                                      D.dRowVal, W.beta, W.xb, W.xbLo);
            CCD_HIP(hipGetLastError());
        }
        for (int j = 0; j < J; ++j) {
            const Col<Real> c = E.col(j);
            if (!live[j] || c.n == 0) continue;
            if (byRows && (D.fmt[j] == F_INDICATOR || D.fmt[j] == F_SPARSE)) continue;
            CCD_DISPATCH_FMT(D.fmt[j], (kAxpyBatch<Real, FMT><<<gridFor(c.n * (int64_t)B), BLOCK, 0, s>>>(c, W.beta, J, j, B, L.Bt, BP, K, W.xb, W.xbLo)));
        }
    }
    CCD_HIP(hipGetLastError());
    CCD_HIP(hipMemsetAsync(W.steps, 0, sizeof(unsigned long long), s));
    CCD_HIP(hipMemsetAsync(W.stats, 0, 8 * sizeof(double), s));
    CCD_HIP(hipMemsetAsync(W.done, 0, B * sizeof(int), s));
    CCD_HIP(hipMemsetAsync(W.flag, 0, B * sizeof(int), s));
    CCD_HIP(hipMemsetAsync(W.iters, 0, B * sizeof(int), s));
    CCD_HIP(hipMemsetAsync(W.ctlIter, 0, B * sizeof(int), s));
    // ---- parameters: one block per team -------------------------------------------
    int rateKHz = 0;
    CCD_HIP(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, D.device));
    if (rateKHz <= 0) rateKHz = 100000;
    res::Params<Real> P{};
    P.K = K; P.J = J; P.G = G; P.B = L.Bt; P.N = grouped ? N : K;
    P.fmt = D.dFmt; P.colPtr = D.dColPtr; P.densePtr = D.dDensePtr; P.rows = D.dRows; P.vals = D.dVals;
    P.dense = D.dDense; P.y = D.dY;
    P.seg = D.dStrat;
    P.pid = (m == M_COX) ? W.pid : D.dStrat;
    P.pstart = W.pstart; P.pseg = W.pseg; P.pflag = W.pflag; P.waveLast = W.waveLast;
    P.rowStart = W.rowStart; P.pidStart = W.pidStart; P.colOff = W.colOff;
    P.info = W.info; P.rng = W.rng; P.firstJ = firstJ;
    P.kw = kwHost ? W.kw : nullptr;
    P.rowFold = rowFold ? W.rowFold : nullptr;
    P.entFold = entryFolds ? W.entFold : nullptr;
    P.exclude = W.exclude; P.fixed = W.fixed;
    P.xjyTot = W.xjyTot; P.xjyFold = rowFold ? W.xjyFold : nullptr; P.F = F; P.xjyW = kwHost ? W.xjyW : nullptr;
    P.priorType = a.prior_type; P.convType = a.convergence_type; P.tol = a.tolerance; P.maxIt = a.max_iterations;
    P.err = W.err; P.allDoneOut = W.allDone; P.steps = W.steps; P.stats = W.stats;
    P.deadlineTicks = (unsigned long long)rateKHz * 1000ull * 20ull;   // 20 s for one grid exchange
    P.timing = std::getenv("CCD_PHASES") ? 1 : 0;
    // all-read: every workgroup reduces and steps every fit after one exchange; used
    // when the partials to read per coordinate are few (G x fits), else the owner of
    // each fit steps it and broadcasts the delta (a second exchange)
    P.allRead = (L.Bt <= 2 || G == 1 || (!grouped && (int64_t)G * BP <= 1024)) ? 1 : 0;
    // FP32 runs of independent rows in all-read teams: tagged per-coordinate partials
    // (no epoch exchange per coordinate); the slot halves are reset before every launch
    P.tagged = (sizeof(Real) == 4 && !grouped && P.allRead && G > 1 && (int64_t)G * BP <= 1024) ? 1 : 0;

    std::vector<res::Params<Real>> teamP(nTeams, P);
    for (int t = 0; t < nTeams; ++t) {
        auto& Q = teamP[t];
        const size_t f0 = (size_t)t * L.Bt;
        Q.B = std::min(L.Bt, B - (int)f0);
        Q.xb = W.xb + (size_t)t * K * BP;
        Q.xbLo = W.xbLo ? W.xbLo + (size_t)t * K * BP : nullptr;
        if (grouped) {
            const size_t o = (size_t)t * N * BP;
            Q.denom = W.denom + o; Q.Dacc = W.Dacc + o; Q.R1 = W.R1 + o; Q.R2 = W.R2 + o; Q.nW = W.nW + o;
            Q.numer = W.numer ? W.numer + o : nullptr;
            Q.numer2 = W.numer2 ? W.numer2 + o : nullptr;
            Q.nWs = W.nWs ? W.nWs + o : nullptr;
        }
        Q.beta = W.beta + f0 * J; Q.hDelta = W.hDelta + f0 * J;
        Q.var = W.var + f0; Q.initBound = W.initBound + f0; Q.held = W.held + f0;
        Q.fixJ = fixJ ? W.fixJ + f0 : nullptr;
        Q.fixedTermFit = W.fixedTerm + f0;
        Q.flag = W.flag + f0; Q.done = W.done + f0; Q.iters = W.iters + f0; Q.ll = W.ll + f0;
        Q.ctlIter = W.ctlIter + f0; Q.ctlFlag = W.ctlFlag + f0; Q.ctlLast = W.ctlLast + f0;
        Q.epoch = W.epoch + (size_t)t * G;
        Q.slots = W.slots + (size_t)t * 2 * G * BP * res::SLOTW;
        Q.bcast = W.bcast + (size_t)t * 2 * BP * 4;
    }
    ResRun<Real> R;
    R.E = &E;
    R.W = &W;
    R.B = B;
    R.J = J;
    R.BP = BP;
    R.nTeams = nTeams;
    R.G = G;
    R.m = m;
    R.rateKHz = rateKHz;
    R.teamP = std::move(teamP);
    return R;
}

// Launches of prepared runs. One run launches its own grid; several runs (same
// model, precision and fit padding) share ONE grid, each run's teams keeping their
// own sizes, exchanges and state, and one error and completion flag. Each launch
// runs up to `budget` cycles of every unfinished fit and leaves the state in HBM,
// so where the launches break changes no result.
template <typename Real>
struct Launcher {
    std::vector<ResRun<Real>*> runs;
    res::Params<Real>* dTeams = nullptr;
    bool fused = false, exchanges = false;
    int total = 0, fitsAll = 0, launches = 0, lastBudget = -1;
    std::vector<res::Params<Real>> all;

    explicit Launcher(const std::vector<ResRun<Real>*>& rs) : runs(rs) {
        ResRun<Real>& R0 = *runs[0];
        auto& W0 = *R0.W;
        fused = runs.size() > 1;
        for (auto* R : runs) {
            if (R->m != R0.m || R->BP != R0.BP) throw std::logic_error("fused runs need one kernel instance");
            total += R->nTeams * R->G;
            fitsAll += R->B;
            exchanges = exchanges || R->G > 1;
        }
        dTeams = W0.teams;
        if (fused) {
            for (auto* R : runs) CCD_HIP(hipStreamSynchronize(R->E->s));   // every run's set-up is on the device
            for (auto* R : runs)
                for (auto& Q : R->teamP) { Q.err = W0.err; Q.allDoneOut = W0.allDone; }
            if (W0.fusedCap < (size_t)total) {
                dfree(W0.fusedTeams);
                W0.fusedCap = 0;
                W0.fusedTeams = dalloc<res::Params<Real>>(total, W0.bytes);
                W0.fusedCap = (size_t)total;
            }
            dTeams = W0.fusedTeams;
        }
    }

    void upload(hipStream_t s) {
        ResRun<Real>& R0 = *runs[0];
        if (!fused) {
            R0.fill();
            h2d(R0.W->teams, R0.blockP.data(), R0.blockP.size(), s);
            return;
        }
        all.clear();
        for (auto* R : runs) {
            R->fill();
            all.insert(all.end(), R->blockP.begin(), R->blockP.end());
        }
        h2d(dTeams, all.data(), all.size(), s);
    }

    // One launch of up to `budget` cycles; true once every fit has finished. With
    // `evA`/`evB` the kernel alone is bracketed by those events on the stream.
    bool launch(int budget, hipEvent_t evA = nullptr, hipEvent_t evB = nullptr) {
        ResRun<Real>& R0 = *runs[0];
        auto& W0 = *R0.W;
        hipStream_t s = R0.E->s;
        const int one = 1;
        for (auto* R : runs) {
            CCD_HIP(hipMemsetAsync(R->W->epoch, 0, (size_t)R->nTeams * R->G * sizeof(unsigned), s));
            if (R->teamP[0].tagged) {
                // tagged slot words: half 0 with tag bit 0, half 1 with tag bit 1, so that
                // neither matches the tag of its first exchange in the launch
                const size_t half = (size_t)R->G * R->BP * res::SLOTW;
                for (int t = 0; t < R->nTeams; ++t) {
                    double* base = R->W->slots + (size_t)t * 2 * half;
                    CCD_HIP(hipMemsetAsync(base, 0x00, half * sizeof(double), s));
                    CCD_HIP(hipMemsetAsync(base + half, 0xFF, half * sizeof(double), s));
                }
            }
        }
        CCD_HIP(hipMemsetAsync(W0.err, 0, sizeof(int), s));
        h2d(W0.allDone, &one, 1, s);
        if (launches <= 1 || budget != lastBudget) {   // the first launch starts the fits, later ones resume
            for (auto* R : runs)
                for (auto& Q : R->teamP) { Q.resume = launches > 0 ? 1 : 0; Q.cycleBudget = budget; }
            upload(s);
            lastBudget = budget;
        }
        // teams that exchange need every workgroup of the grid resident at once
        const int nT = fused ? (exchanges ? 1 : total) : R0.nTeams;
        const int T = fused ? (exchanges ? total : 1) : R0.G;
        if (evA) CCD_HIP(hipEventRecord(evA, s));
        switch (R0.m) {
        case M_LOGISTIC: launchByBP<Real, M_LOGISTIC>(R0.BP, dTeams, nT, T, s); break;
        case M_POISSON: launchByBP<Real, M_POISSON>(R0.BP, dTeams, nT, T, s); break;
        case M_CLR: launchByBP<Real, M_CLR>(R0.BP, dTeams, nT, T, s); break;
        default: launchByBP<Real, M_COX>(R0.BP, dTeams, nT, T, s); break;
        }
        CCD_HIP(hipGetLastError());
        if (evB) CCD_HIP(hipEventRecord(evB, s));
        int hs[2] = {0, 0};
        CCD_HIP(hipMemcpyAsync(&hs[0], W0.err, sizeof(int), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipMemcpyAsync(&hs[1], W0.allDone, sizeof(int), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipStreamSynchronize(s));
        if (hs[0]) throw HipError("persistent kernel: grid exchange timed out (workgroups not co-resident?)");
        ++launches;
        return hs[1] != 0;
    }
};

// Launch prepared runs until all their fits have finished. Cycles per launch:
// CCD_CYCLES_PER_LAUNCH, else adapted so a launch lasts about 4 s (the host sees
// progress and errors at launch boundaries): 2 cycles first, then as many as the
// measured cycle time allows, at most 32.
template <typename Real>
static void launchRuns(const std::vector<ResRun<Real>*>& runs, double& deviceMs, int& nLaunches) {
    Launcher<Real> Ln(runs);
    EngineT<Real>& E0 = *runs[0]->E;
    hipStream_t s = E0.s;
    int budget = 2;
    bool fixedBudget = false;
    if (const char* e = std::getenv("CCD_CYCLES_PER_LAUNCH")) { budget = std::max(1, std::atoi(e)); fixedBudget = true; }
    const bool progress = std::getenv("CCD_PROGRESS") != nullptr;
    const auto tWall = std::chrono::steady_clock::now();
    CCD_HIP(hipEventRecord(E0.ev0, s));
    bool done = false;
    int cyclesRun = 0;
    while (!done) {
        const auto tL = std::chrono::steady_clock::now();
        done = Ln.launch(budget);
        cyclesRun += budget;
        if (progress)
            std::fprintf(stderr, "[ccd] launch %d (%d runs, %d workgroups, %d fits): %d cycles, %s (%.1fs)\n",
                         Ln.launches, (int)runs.size(), Ln.total, Ln.fitsAll, cyclesRun,
                         done ? "all fits finished" : "fits still running",
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - tWall).count());
        if (!fixedBudget && !done) {
            const double perCycle = std::chrono::duration<double>(std::chrono::steady_clock::now() - tL).count() / budget;
            budget = (int)std::max(1.0, std::min(32.0, std::floor(4.0 / std::max(perCycle, 1e-6))));
        }
    }
    CCD_HIP(hipEventRecord(E0.ev1, s));
    CCD_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    CCD_HIP(hipEventElapsedTime(&ms, E0.ev0, E0.ev1));
    deviceMs = ms;
    nLaunches = Ln.launches;
}

// Results of a finished run.
template <typename Real>
static void gatherRun(ResRun<Real>& R, double* betas, int* flags, int* iters, double* lls, double& algoBytes,
                      double& steps) {
    auto& W = *R.W;
    const int B = R.B, J = R.J;
    // copies on the engine's own stream: a synchronous hipMemcpy goes through the null
    // stream, whose hardware queue other engines' persistent launches may share, and
    // would wait for them
    hipStream_t s = R.E->s;
    unsigned long long st = 0;
    double st8[8];
    CCD_HIP(hipMemcpyAsync(&st, W.steps, sizeof(st), hipMemcpyDeviceToHost, s));
    CCD_HIP(hipMemcpyAsync(betas, W.beta, (size_t)B * J * sizeof(double), hipMemcpyDeviceToHost, s));
    CCD_HIP(hipMemcpyAsync(flags, W.flag, B * sizeof(int), hipMemcpyDeviceToHost, s));
    CCD_HIP(hipMemcpyAsync(iters, W.iters, B * sizeof(int), hipMemcpyDeviceToHost, s));
    CCD_HIP(hipMemcpyAsync(lls, W.ll, B * sizeof(double), hipMemcpyDeviceToHost, s));
    CCD_HIP(hipMemcpyAsync(st8, W.stats, sizeof(st8), hipMemcpyDeviceToHost, s));
    CCD_HIP(hipStreamSynchronize(s));
    steps = (double)st;
    algoBytes = st8[0];
    if (std::getenv("CCD_PHASES"))
        std::fprintf(stderr, "[ccd] teams=%d T=%d B=%d BP=%d steps=%llu  ms (summed over teams): column pass %.1f  exchange %.1f  update %.1f  objective %.1f\n",
                     R.nTeams, R.G, B, R.BP, st, st8[1] / R.rateKHz, st8[2] / R.rateKHz, st8[3] / R.rateKHz,
                     st8[4] / R.rateKHz);
}

// ---------------------------------------------------------------------------
// Held-out predictive log-likelihood of every fit of a finished run, from the xb
// the run leaves in HBM (getPredictiveLogLikelihood, ModelSpecifics.hpp:652-690,
// with the held-out rows of fit f, rowFold[k] == held[f], as the weights; the terms
// of ModelSpecifics.h: CLR :867-873, LR :1002-1013, Breslow Cox :1243-1256,
// Poisson :1583-1594). One pass for every fit of a team replaces, per fit, the
// reference's setWeights + computeXBeta + computeRemainingStatistics round trip.
// Everything is summed in FP64 from xb (plus its low part in FP32 runs), in a
// fixed order.
// ---------------------------------------------------------------------------
template <typename Real>
__device__ __forceinline__ double xbOf(const Real* xb, const unsigned short* lo, size_t q) {
    double x = (double)xb[q];
    if (lo) x += (double)bf16Value(lo[q]);
    return x;
}

// LR / PR: one group per row; CLR: one group per stratum (rows [gs[g], gs[g+1])).
// Thread (group, lane); block partials per lane: part[block * BP + lane].
template <typename Real, int MODEL>
__global__ __launch_bounds__(BLOCK) void kPredGroups(int64_t nGroups, const int* __restrict__ gs, int BP, int Bt,
                                                    const Real* __restrict__ xb, const unsigned short* __restrict__ lo,
                                                    const Real* __restrict__ y, const int* __restrict__ rowFold,
                                                    const int* __restrict__ held, double* __restrict__ part) {
    __shared__ double sh[BLOCK];
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int64_t g = t / BP;
    const int f = (int)(t % BP);
    double c = 0.0;
    if (g < nGroups && f < Bt) {
        const int hf = held[f];
        if constexpr (MODEL == M_CLR) {
            const int64_t a = gs[g], b = gs[g + 1];
            double den = 0.0;
            for (int64_t k = a; k < b; ++k) den += exp(xbOf(xb, lo, (size_t)k * BP + f));
            const double lden = log(den);
            for (int64_t k = a; k < b; ++k)
                if (rowFold[k] == hf && y[k] != Real(0)) c += (double)y[k] * (xbOf(xb, lo, (size_t)k * BP + f) - lden);
        } else {
            const int64_t k = g;
            if (rowFold[k] == hf) {
                const double x = xbOf(xb, lo, (size_t)k * BP + f);
                if constexpr (MODEL == M_LOGISTIC) c = (double)y[k] * (x - log1p(exp(x)));
                else c = (double)y[k] * x - exp(x);
            }
        }
    }
    sh[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x < BP) {   // groups of the block in order
        double sum = 0.0;
        for (int i = threadIdx.x; i < BLOCK; i += BP) sum += sh[i];
        part[(size_t)blockIdx.x * BP + threadIdx.x] = sum;
    }
}

// Per lane, the block partials summed in block order: out[lane].
__global__ void kPredSum(int64_t nParts, int BP, int Bt, const double* __restrict__ part, double* __restrict__ out) {
    const int f = threadIdx.x;
    if (f >= Bt) return;
    double s = 0.0;
    for (int64_t b = 0; b < nParts; ++b) s += part[(size_t)b * BP + f];
    out[f] = s;
}

// Breslow Cox over held-out risk sets: rows sorted (stratum, time desc, y asc), so the
// risk set of an event run (tied events of one time, the last rows of that time) is
// every row from the stratum's start to the run's end; D = sum of w exp(xb) over it,
// w = held-out indicator. Chunks [cs[c], cs[c+1]) never split an event run.
// Pass 1: per (chunk, lane) the sum after the chunk's last stratum start.
template <typename Real>
__global__ __launch_bounds__(BLOCK) void kPredCoxTail(int nChunks, const int* __restrict__ cs, int BP, int Bt,
                                                     const int* __restrict__ strat, const Real* __restrict__ xb,
                                                     const unsigned short* __restrict__ lo,
                                                     const int* __restrict__ rowFold, const int* __restrict__ held,
                                                     double* __restrict__ tail) {
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int c = (int)(t / BP), f = (int)(t % BP);
    if (c >= nChunks || f >= Bt) return;
    const int hf = held[f];
    double s = 0.0;
    for (int k = cs[c]; k < cs[c + 1]; ++k) {
        if (k > cs[c] && strat[k] != strat[k - 1]) s = 0.0;
        if (rowFold[k] == hf) s += exp(xbOf(xb, lo, (size_t)k * BP + f));
    }
    tail[(size_t)c * BP + f] = s;
}

// Pass 2: the carry into each chunk (the sum since its stratum's start), per lane.
// startsNew[c]: chunk c begins a stratum; inner[c]: a stratum begins inside chunk c.
__global__ void kPredCoxCarry(int nChunks, int BP, int Bt, const unsigned char* __restrict__ startsNew,
                              const unsigned char* __restrict__ inner, const double* __restrict__ tail,
                              double* __restrict__ carry) {
    const int f = threadIdx.x;
    if (f >= Bt) return;
    double cin = 0.0;
    for (int c = 0; c < nChunks; ++c) {
        if (startsNew[c]) cin = 0.0;
        carry[(size_t)c * BP + f] = cin;
        cin = inner[c] ? tail[(size_t)c * BP + f] : cin + tail[(size_t)c * BP + f];
    }
}

// Pass 3: walk the chunk with its carry; each event run adds sum(xb) - n log D.
template <typename Real>
__global__ __launch_bounds__(BLOCK) void kPredCoxWalk(int nChunks, const int* __restrict__ cs, int BP, int Bt,
                                                     const int* __restrict__ strat, const double* __restrict__ time,
                                                     const Real* __restrict__ y, const Real* __restrict__ xb,
                                                     const unsigned short* __restrict__ lo,
                                                     const int* __restrict__ rowFold, const int* __restrict__ held,
                                                     const double* __restrict__ carry, double* __restrict__ part) {
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const int c = (int)(t / BP), f = (int)(t % BP);
    if (c >= nChunks || f >= Bt) return;
    const int hf = held[f];
    double run = carry[(size_t)c * BP + f], sumX = 0.0, nEv = 0.0, out = 0.0;
    const int a = cs[c], b = cs[c + 1];
    for (int k = a; k < b; ++k) {
        if (k > a && strat[k] != strat[k - 1]) run = 0.0;
        const bool w = rowFold[k] == hf;
        const double x = xbOf(xb, lo, (size_t)k * BP + f);
        if (w) run += exp(x);
        if (y[k] != Real(0) && w) { sumX += x; nEv += 1.0; }
        const bool more = k + 1 < b && y[k] != Real(0) && y[k + 1] != Real(0) && time[k + 1] == time[k] &&
                          strat[k + 1] == strat[k];
        if (!more && y[k] != Real(0)) {   // end of an event run
            if (nEv > 0.0) out += sumX - nEv * log(run);
            sumX = 0.0;
            nEv = 0.0;
        }
    }
    part[(size_t)c * BP + f] = out;
}

// Held-out predictive log-likelihood of fits [0, R.B) of a finished run -> pred[f].
template <typename Real>
static void predictHeldOut(ResRun<Real>& R, double* pred) {
    if (!R.teamP[0].rowFold) throw std::logic_error("held-out prediction needs row folds");
    EngineT<Real>& E = *R.E;
    auto& D = *E.D;
    auto& W = *R.W;
    hipStream_t s = E.s;
    const int64_t K = D.K;
    const int BP = R.BP;
    const int m = D.model;
    // groups (LR / PR rows, CLR strata) or Cox chunks
    int64_t nGroups = 0;
    int nChunks = 0;
    std::vector<int> cs;
    std::vector<unsigned char> flags;   // Cox: startsNew[nChunks], inner[nChunks]
    if (m == M_COX) {
        constexpr int64_t CH = 2048;
        cs.push_back(0);
        for (int64_t k = CH; k < K;) {
            // move forward while row k continues an event run of row k - 1
            while (k < K && D.y[k] != 0.0 && D.y[k - 1] != 0.0 && D.time[k] == D.time[k - 1] &&
                   D.strat[k] == D.strat[k - 1])
                ++k;
            if (k < K) cs.push_back((int)k);
            k += CH;
        }
        cs.push_back((int)K);
        nChunks = (int)cs.size() - 1;
        flags.assign(2 * (size_t)nChunks, 0);
        for (int c = 0; c < nChunks; ++c) {
            const int a = cs[c];
            flags[c] = (c == 0 || D.strat[a] != D.strat[a - 1]) ? 1 : 0;
            for (int k = a + 1; k < cs[c + 1]; ++k)
                if (D.strat[k] != D.strat[k - 1]) { flags[nChunks + c] = 1; break; }
        }
    } else {
        nGroups = m == M_CLR ? D.nStrata : K;
    }
    const int64_t nParts = m == M_COX ? nChunks : (nGroups * BP + BLOCK - 1) / BLOCK;
    const size_t need = (size_t)nParts * BP * (m == M_COX ? 3 : 1) + (size_t)BP + (size_t)cs.size() + flags.size();
    if (W.predCap < need) {
        dfree(W.predBuf);
        W.predCap = 0;
        W.predBuf = dalloc<double>(need, W.bytes);
        W.predCap = need;
    }
    double* part = W.predBuf;
    double* outDev = part + (size_t)nParts * BP * (m == M_COX ? 3 : 1);
    int* csDev = reinterpret_cast<int*>(outDev + BP);
    unsigned char* flagDev = reinterpret_cast<unsigned char*>(csDev + cs.size());
    if (m == M_COX) {
        h2d(csDev, cs.data(), cs.size(), s);
        h2d(flagDev, flags.data(), flags.size(), s);
        if (!W.timeDev) {   // FP64 times, so distinct times never tie
            W.timeDev = dalloc<double>(K, W.bytes);
            h2d(W.timeDev, D.time.data(), K, s);
        }
    }
    std::vector<double> h(BP);
    for (int t = 0; t < R.nTeams; ++t) {
        const res::Params<Real>& Q = R.teamP[t];
        const int Bt = Q.B;
        if (m == M_COX) {
            double* tail = part + (size_t)nChunks * BP;
            double* carry = tail + (size_t)nChunks * BP;
            const int grid = (int)(((int64_t)nChunks * BP + BLOCK - 1) / BLOCK);
            kPredCoxTail<Real><<<grid, BLOCK, 0, s>>>(nChunks, csDev, BP, Bt, D.dStrat, Q.xb, Q.xbLo, Q.rowFold,
                                                      Q.held, tail);
            kPredCoxCarry<<<1, std::max(64, BP), 0, s>>>(nChunks, BP, Bt, flagDev, flagDev + nChunks, tail, carry);
            kPredCoxWalk<Real><<<grid, BLOCK, 0, s>>>(nChunks, csDev, BP, Bt, D.dStrat, W.timeDev, D.dY, Q.xb, Q.xbLo,
                                                      Q.rowFold, Q.held, carry, part);
        } else {
            const int grid = (int)nParts;
            switch (m) {
            case M_LOGISTIC:
                kPredGroups<Real, M_LOGISTIC><<<grid, BLOCK, 0, s>>>(nGroups, nullptr, BP, Bt, Q.xb, Q.xbLo, D.dY,
                                                                     Q.rowFold, Q.held, part);
                break;
            case M_POISSON:
                kPredGroups<Real, M_POISSON><<<grid, BLOCK, 0, s>>>(nGroups, nullptr, BP, Bt, Q.xb, Q.xbLo, D.dY,
                                                                    Q.rowFold, Q.held, part);
                break;
            default:
                kPredGroups<Real, M_CLR><<<grid, BLOCK, 0, s>>>(nGroups, D.dStratStart, BP, Bt, Q.xb, Q.xbLo, D.dY,
                                                                Q.rowFold, Q.held, part);
                break;
            }
        }
        kPredSum<<<1, std::max(64, BP), 0, s>>>(nParts, BP, Bt, part, outDev);
        CCD_HIP(hipGetLastError());
        CCD_HIP(hipMemcpyAsync(h.data(), outDev, Bt * sizeof(double), hipMemcpyDeviceToHost, s));
        CCD_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < Bt; ++i) pred[(size_t)t * R.teamP[0].B + i] = h[i];
    }
}

// Run fits [0, B) of engine E as persistent launches until all have finished;
// fills betas/flags/iters/ll.
template <typename Real>
static void runResident(EngineT<Real>& E, const ccd_fit_args& a, int B, const double* var, const int32_t* held,
                        const int32_t* rowFold, int F, const double* kwHost, const double* initBound, double* betas,
                        int* flags, int* iters, double* lls, double& deviceMs, double& algoBytes, double& steps,
                        int& nLaunches, Layout L, const int32_t* fixJ = nullptr, double* pred = nullptr) {
    ResRun<Real> R = prepareResident<Real>(E, a, B, var, held, rowFold, F, kwHost, initBound, betas, L, fixJ);
    std::vector<ResRun<Real>*> one{&R};
    launchRuns<Real>(one, deviceMs, nLaunches);
    gatherRun<Real>(R, betas, flags, iters, lls, algoBytes, steps);
    if (pred) predictHeldOut<Real>(R, pred);   // while the run's xb is still in HBM
}

template <typename Real>
void EngineT<Real>::fitResident(const ccd_fit_args& a, double* beta, ccd_fit_result& r) {
    CCD_HIP(hipSetDevice(D->device));
    const auto t0 = std::chrono::steady_clock::now();
    const int J = D->J;
    std::vector<double> b(beta, beta + J);
    if (D->hasOffset) b[0] = 1.0;
    double initBound = a.initial_bound;
    int count = 0, flag = 0, it = 0;
    double ll = 0.0;
    double var = a.variance;
    for (;;) {
        double ms = 0, bytes = 0, steps = 0;
        int nl = 0;
        runResident<Real>(*this, a, 1, &var, nullptr, nullptr, 0, a.weights, &initBound, b.data(), &flag, &it, &ll,
                          ms, bytes, steps, nl,
                          chooseLayout(D->K, 1, D->device, D->model));
        ++count;
        if (timing) {
            counters.deviceMs += ms;
            counters.launches += nl;
            counters.bytes += bytes;
            counters.coordinates += steps;
        }
        if (flag == CCD_ILLCONDITIONED && count < a.max_bound_count) {   // update() retry (:599-614)
            initBound /= 10.0;
            for (int j = D->hasOffset ? 1 : 0; j < J; ++j) b[j] = 0.0;
        } else {
            break;
        }
    }
    std::copy(b.begin(), b.end(), beta);
    double lp = 0.0;
    for (int j = 0; j < J; ++j)
        lp += priorLogDensity((a.exclude && a.exclude[j]) ? CCD_PRIOR_NONE : a.prior_type, a.variance, b[j]);
    r.return_flag = flag;
    r.iterations = it;
    r.find_mode_calls = count;
    r.log_likelihood = ll;
    r.log_prior = lp;
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Checks the batch arguments, forces the offset column of every fit to 1 and
// returns the number of folds in row_fold.
template <typename Real>
static int checkBatch(const EngineT<Real>& E, const ccd_fit_args& a, const ccd_batch_args& bargs, double* betas) {
    const auto& D = *E.D;
    const int J = D.J;
    const int n = bargs.n_fits;
    int F = 0;
    if (bargs.row_fold) {
        for (int64_t k = 0; k < D.K; ++k) {
            const int q = bargs.row_fold[k];
            if (q < 0 || q >= 32) throw std::invalid_argument("row_fold values must be in [0, 32)");
            F = std::max(F, q + 1);
        }
    }
    if (bargs.held_out_fold && !bargs.row_fold)
        for (int f = 0; f < n; ++f)
            if (bargs.held_out_fold[f] >= 0) throw std::invalid_argument("held_out_fold needs row_fold");
    if (bargs.fixed_index)
        for (int f = 0; f < n; ++f)
            if (bargs.fixed_index[f] < -1 || bargs.fixed_index[f] >= J)
                throw std::invalid_argument("fixed_index must be -1 or a column index");
    for (int f = 0; f < n; ++f) {
        if (a.prior_type != CCD_PRIOR_NONE && !(bargs.variances[f] > 0))
            throw std::invalid_argument("variances must be > 0");
        if (D.hasOffset) betas[(size_t)f * J] = 1.0;
    }
    return F;
}

// Team shape of a whole batch: CCD_BATCH_LAYOUT, else the fastest candidate on one
// timed cycle of these fits (kept per batch size). False when neither applies (the
// caller then uses chooseLayout per chunk).
template <typename Real>
static bool tuneLayout(EngineT<Real>& E, const ccd_fit_args& a, const ccd_batch_args& bargs, const double* betas,
                       int F, Layout& LT) {
    const auto& D = *E.D;
    const int J = D.J;
    const int n = bargs.n_fits;
    if (!(n > 2 && !std::getenv("CCD_BATCH_LAYOUT") &&
          !(std::getenv("CCD_AUTOTUNE") && std::atoi(std::getenv("CCD_AUTOTUNE")) == 0)))
        return false;
    if (!E.res) E.res = new ResWork<Real>();
    auto& W = *static_cast<ResWork<Real>*>(E.res);
    auto it = W.tuned.find(n);
    if (it != W.tuned.end()) {
        LT.Bt = it->second[0];
        LT.T = it->second[1];
        LT.nTeams = it->second[2];
        return true;
    }
    const std::vector<Layout> cand = layoutCandidates<Real>(D.model, D.K, n, D.device);
    if (cand.size() <= 1) return false;
    ccd_fit_args one = a;
    one.max_iterations = 1;
    std::vector<double> v(n), bd(n, a.initial_bound), l(n);
    std::vector<int32_t> h(n);
    std::vector<int> fl(n), it2(n);
    for (int f = 0; f < n; ++f) {
        v[f] = bargs.variances[f];
        h[f] = bargs.held_out_fold ? bargs.held_out_fold[f] : -1;
    }
    double best = 1e300;
    for (const Layout& L : cand) {
        std::vector<double> bb(betas, betas + (size_t)n * J);
        double ms = 0, bytes = 0, steps = 0;
        int nl = 0;
        runResident<Real>(E, one, n, v.data(), h.data(), bargs.row_fold, F, nullptr, bd.data(), bb.data(),
                          fl.data(), it2.data(), l.data(), ms, bytes, steps, nl, L, bargs.fixed_index);
        if (std::getenv("CCD_PROGRESS"))
            std::fprintf(stderr, "[ccd] autotune %d fits: %d fits x %d workgroups per team: %.1f ms per cycle\n", n,
                         L.Bt, L.T, ms);
        if (ms < best) { best = ms; LT = L; }
        if (best > 2000.0) break;   // every shape is slow here: probing costs more than it saves
    }
    W.tuned[n] = {LT.Bt, LT.T, LT.nTeams};
    return true;
}

// Log prior of every fit and the result records of a finished batch.
template <typename Real>
static void batchResults(const EngineT<Real>& E, const ccd_fit_args& a, const double* variances, int n,
                         const double* betas, const int* flags, const int* its, const double* lls, const int* count,
                         double secs, ccd_fit_result* r) {
    const int J = E.D->J;
    for (int f = 0; f < n; ++f) {
        double lp = 0.0;
        for (int j = 0; j < J; ++j)
            lp += priorLogDensity((a.exclude && a.exclude[j]) ? CCD_PRIOR_NONE : a.prior_type, variances[f],
                                  betas[(size_t)f * J + j]);
        r[f].return_flag = flags[f];
        r[f].iterations = its[f];
        r[f].find_mode_calls = count[f];
        r[f].log_likelihood = lls[f];
        r[f].log_prior = lp;
        r[f].seconds = secs;
    }
}

template <typename Real>
void EngineT<Real>::fitBatch(const ccd_fit_args& a, const ccd_batch_args& bargs, double* betas, ccd_fit_result* r,
                             bool autotune, double* predHeldOut) {
    CCD_HIP(hipSetDevice(D->device));
    const auto t0 = std::chrono::steady_clock::now();
    const int J = D->J;
    const int n = bargs.n_fits;
    const int F = checkBatch<Real>(*this, a, bargs, betas);
    std::vector<int> count(n, 0), flags(n, 0), its(n, 0);
    std::vector<double> lls(n, 0.0), bound(n, a.initial_bound);
    std::vector<int> pending(n);
    std::iota(pending.begin(), pending.end(), 0);
    const bool grouped = D->model == M_CLR || D->model == M_COX;
    const int64_t nGroups = D->model == M_CLR ? D->nStrata : D->K;
    Layout LT;
    const bool tunedOk = autotune && tuneLayout<Real>(*this, a, bargs, betas, F, LT);
    while (!pending.empty()) {
        std::vector<int> retry;
        const bool whole = tunedOk && (int)pending.size() == n;
        const Layout L0 = whole ? LT : chooseLayout(D->K, (int)pending.size(), D->device, D->model);
        const size_t chunk = whole ? (size_t)n : (size_t)chunkFits(L0, D->K, nGroups, grouped, (int)sizeof(Real));
        for (size_t c0 = 0; c0 < pending.size(); c0 += chunk) {
            const int B = (int)std::min(chunk, pending.size() - c0);
            const Layout L = whole ? LT : chooseLayout(D->K, B, D->device, D->model);
            std::vector<double> v(B), bd(B), bb((size_t)B * J), l(B);
            std::vector<int32_t> h(B), fj(B);
            std::vector<int> fl(B), it(B);
            for (int i = 0; i < B; ++i) {
                const int f = pending[c0 + i];
                v[i] = bargs.variances[f];
                h[i] = bargs.held_out_fold ? bargs.held_out_fold[f] : -1;
                fj[i] = bargs.fixed_index ? bargs.fixed_index[f] : -1;
                bd[i] = bound[f];
                std::copy(betas + (size_t)f * J, betas + (size_t)(f + 1) * J, bb.begin() + (size_t)i * J);
            }
            double ms = 0, bytes = 0, steps = 0;
            int nl = 0;
            std::vector<double> pr(predHeldOut ? B : 0);
            runResident<Real>(*this, a, B, v.data(), h.data(), bargs.row_fold, F, nullptr, bd.data(), bb.data(),
                              fl.data(), it.data(), l.data(), ms, bytes, steps, nl, L,
                              bargs.fixed_index ? fj.data() : nullptr, predHeldOut ? pr.data() : nullptr);
            if (predHeldOut)
                for (int i = 0; i < B; ++i) predHeldOut[pending[c0 + i]] = pr[i];
            if (timing) {
                counters.deviceMs += ms;
                counters.launches += nl;
                counters.bytes += bytes;
                counters.coordinates += steps;
            }
            for (int i = 0; i < B; ++i) {
                const int f = pending[c0 + i];
                std::copy(bb.begin() + (size_t)i * J, bb.begin() + (size_t)(i + 1) * J, betas + (size_t)f * J);
                flags[f] = fl[i];
                its[f] = it[i];
                lls[f] = l[i];
                count[f] += 1;
                if (fl[i] == CCD_ILLCONDITIONED && count[f] < a.max_bound_count) {
                    bound[f] /= 10.0;
                    for (int j = D->hasOffset ? 1 : 0; j < J; ++j) betas[(size_t)f * J + j] = 0.0;
                    retry.push_back(f);
                }
            }
        }
        pending.swap(retry);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    batchResults<Real>(*this, a, bargs.variances, n, betas, flags.data(), its.data(), lls.data(), count.data(), secs,
                       r);
}

// A batch of ccd_fit_batch advanced by the caller a few cycles at a time: every fit
// in one persistent run (one grid) whose state stays in HBM between steps. The
// arithmetic of each cycle does not depend on where launches break, so the result
// equals ccd_fit_batch's.
template <typename Real>
class SteppedBatchT : public SteppedBatch {
public:
    SteppedBatchT(EngineT<Real>& e, const ccd_fit_args& a, const ccd_batch_args& b, const double* betas0)
        : E(e), n(b.n_fits), J(e.D->J) {
        CCD_HIP(hipSetDevice(E.D->device));
        t0 = std::chrono::steady_clock::now();
        args = a;
        if (a.exclude) { excl.assign(a.exclude, a.exclude + J); args.exclude = excl.data(); }
        if (a.fixed) { fix.assign(a.fixed, a.fixed + J); args.fixed = fix.data(); }
        args.weights = nullptr;
        var.assign(b.variances, b.variances + n);
        held.assign(n, -1);
        if (b.held_out_fold) held.assign(b.held_out_fold, b.held_out_fold + n);
        if (b.row_fold) rowFold.assign(b.row_fold, b.row_fold + E.D->K);
        if (b.fixed_index) fixIdx.assign(b.fixed_index, b.fixed_index + n);
        bargs = ccd_batch_args{n, var.data(), held.data(), rowFold.empty() ? nullptr : rowFold.data(),
                               fixIdx.empty() ? nullptr : fixIdx.data()};
        betas.assign(betas0, betas0 + (size_t)n * J);
        const int F = checkBatch<Real>(E, args, bargs, betas.data());
        Layout L;
        if (!tuneLayout<Real>(E, args, bargs, betas.data(), F, L)) {
            L = chooseLayout(E.D->K, n, E.D->device, E.D->model);
            const bool grouped = E.D->model == M_CLR || E.D->model == M_COX;
            const int64_t nGroups = E.D->model == M_CLR ? E.D->nStrata : E.D->K;
            if (chunkFits(L, E.D->K, nGroups, grouped, (int)sizeof(Real)) < n)
                throw std::invalid_argument("batch too large for one stepped run; use ccd_fit_batch");
        }
        lay = L;
        std::vector<double> bd(n, args.initial_bound);
        run = prepareResident<Real>(E, args, n, var.data(), held.data(), bargs.row_fold, F, nullptr, bd.data(),
                                    betas.data(), L, bargs.fixed_index);
        ptrs.push_back(&run);
        launcher.reset(new Launcher<Real>(ptrs));
        E.stepped = this;
    }
    ~SteppedBatchT() override {
        if (E.stepped == this) E.stepped = nullptr;
    }

    int step(int cycles, int32_t* doneOut) override {
        if (cycles < 1) throw std::invalid_argument("cycles must be >= 1");
        hipStream_t s = E.s;
        if (!finished) {
            finished = launcher->launch(cycles, E.timing ? E.ev0 : nullptr, E.timing ? E.ev1 : nullptr);
            cyclesRun += cycles;
            if (E.timing) {
                float ms = 0.f;
                CCD_HIP(hipEventElapsedTime(&ms, E.ev0, E.ev1));
                unsigned long long st = 0;
                double b0 = 0.0;
                CCD_HIP(hipMemcpyAsync(&st, run.W->steps, sizeof(st), hipMemcpyDeviceToHost, s));
                CCD_HIP(hipMemcpyAsync(&b0, run.W->stats, sizeof(double), hipMemcpyDeviceToHost, s));
                CCD_HIP(hipStreamSynchronize(s));
                E.counters.deviceMs += ms;
                E.counters.launches += 1;
                E.counters.bytes += b0 - lastBytes;
                E.counters.coordinates += (double)(st - lastSteps);
                lastBytes = b0;
                lastSteps = st;
            }
        }
        std::vector<int> dn(n, 1);
        if (!finished) {
            CCD_HIP(hipMemcpyAsync(dn.data(), run.W->done, n * sizeof(int), hipMemcpyDeviceToHost, s));
            CCD_HIP(hipStreamSynchronize(s));
        }
        int running = 0;
        for (int f = 0; f < n; ++f) {
            if (doneOut) doneOut[f] = dn[f] ? 1 : 0;
            running += dn[f] ? 0 : 1;
        }
        return running;
    }

    void finish(double* betasOut, ccd_fit_result* r) override {
        while (!finished) step(32, nullptr);
        std::vector<int> flags(n), its(n), count(n, 1);
        std::vector<double> lls(n);
        double algo = 0.0, steps = 0.0;
        gatherRun<Real>(run, betas.data(), flags.data(), its.data(), lls.data(), algo, steps);
        E.stepped = nullptr;   // the engine's workspace is free again
        // update()'s retry of ill-conditioned fits with a tenth of the bound
        // (CyclicCoordinateDescent.cpp:599-614), as ccd_fit_batch does
        std::vector<int> retry;
        for (int f = 0; f < n; ++f)
            if (flags[f] == CCD_ILLCONDITIONED && args.max_bound_count > 1) retry.push_back(f);
        if (!retry.empty()) {
            const int m = (int)retry.size();
            std::vector<double> v(m), bb((size_t)m * J, 0.0);
            std::vector<int32_t> h(m), fj(m);
            for (int i = 0; i < m; ++i) {
                v[i] = var[retry[i]];
                h[i] = held[retry[i]];
                fj[i] = fixIdx.empty() ? -1 : fixIdx[retry[i]];
            }
            ccd_fit_args a2 = args;
            a2.initial_bound = args.initial_bound / 10.0;
            a2.max_bound_count = args.max_bound_count - 1;
            ccd_batch_args b2{m, v.data(), h.data(), bargs.row_fold, fixIdx.empty() ? nullptr : fj.data()};
            std::vector<ccd_fit_result> r2(m);
            E.fitBatch(a2, b2, bb.data(), r2.data(), false, nullptr);
            for (int i = 0; i < m; ++i) {
                const int f = retry[i];
                std::copy(bb.begin() + (size_t)i * J, bb.begin() + (size_t)(i + 1) * J, betas.begin() + (size_t)f * J);
                flags[f] = r2[i].return_flag;
                its[f] = r2[i].iterations;
                lls[f] = r2[i].log_likelihood;
                count[f] = 1 + r2[i].find_mode_calls;
            }
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::copy(betas.begin(), betas.end(), betasOut);
        batchResults<Real>(E, args, var.data(), n, betas.data(), flags.data(), its.data(), lls.data(), count.data(),
                           secs, r);
    }

    int fits() const override { return n; }
    int cycles() const override { return cyclesRun; }
    void layout(int32_t* fitsPerTeam, int32_t* workgroupsPerTeam, int32_t* teams) const override {
        *fitsPerTeam = lay.Bt;
        *workgroupsPerTeam = lay.T;
        *teams = lay.nTeams;
    }

private:
    EngineT<Real>& E;
    int n, J;
    int cyclesRun = 0;
    bool finished = false;
    ccd_fit_args args{};
    ccd_batch_args bargs{};
    std::vector<unsigned char> excl, fix;
    std::vector<double> var, betas;
    std::vector<int32_t> held, rowFold, fixIdx;
    Layout lay;
    ResRun<Real> run;
    std::vector<ResRun<Real>*> ptrs;
    std::unique_ptr<Launcher<Real>> launcher;
    double lastBytes = 0.0;
    unsigned long long lastSteps = 0;
    std::chrono::steady_clock::time_point t0;
};

template <typename Real>
SteppedBatch* EngineT<Real>::beginStepped(const ccd_fit_args& a, const ccd_batch_args& b, const double* betas0) {
    if (stepped) throw std::runtime_error("a stepped batch is already open on this engine");
    return new SteppedBatchT<Real>(*this, a, b, betas0);
}

// Single fits of several datasets as one persistent grid (CyclicCoordinateDescent::
// update, CyclicCoordinateDescent.cpp:586-615, for every engine at once). Each fit
// keeps the team shape it would have alone, so its result is the same as its own
// resident fit; fits are packed into grids whose workgroups can all be resident.
template <typename Real>
static void fitMultiT(const std::vector<EngineT<Real>*>& engs, const ccd_fit_args* args, double* const* betas,
                      ccd_fit_result* res) {
    const int n = (int)engs.size();
    const auto t0 = std::chrono::steady_clock::now();
    const int device = engs[0]->D->device;
    const int model = engs[0]->D->model;
    const bool grouped = model == M_CLR || model == M_COX;
    for (int i = 0; i < n; ++i) {
        if (engs[i]->D->device != device || engs[i]->D->model != model)
            throw std::invalid_argument("fit_multi needs engines of one model type on one device");
        for (int k = 0; k < i; ++k)
            if (engs[k] == engs[i]) throw std::invalid_argument("fit_multi needs distinct engines");
    }
    CCD_HIP(hipSetDevice(device));
    int cus = 0;
    CCD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int capacity = std::max(1, occupancyOf<Real>(model, 1) * cus);
    std::vector<std::vector<double>> b(n);
    std::vector<double> bound(n), var(n), lls(n, 0.0);
    std::vector<int> count(n, 0), flags(n, 0), its(n, 0);
    std::vector<Layout> lay(n);
    for (int i = 0; i < n; ++i) {
        const int J = engs[i]->D->J;
        b[i].assign(betas[i], betas[i] + J);
        if (engs[i]->D->hasOffset) b[i][0] = 1.0;
        bound[i] = args[i].initial_bound;
        var[i] = args[i].variance;
        lay[i] = chooseLayout(engs[i]->D->K, 1, device, model);   // the shape of its own resident fit
        if (lay[i].T > capacity) throw std::invalid_argument("a fit needs more workgroups than the device holds");
    }
    std::vector<int> pending(n);
    std::iota(pending.begin(), pending.end(), 0);
    while (!pending.empty()) {
        std::vector<int> retry;
        for (size_t g0 = 0; g0 < pending.size();) {
            // next grid: as many pending fits as fit co-resident
            std::vector<int> grp;
            int used = 0;
            while (g0 < pending.size() && (grp.empty() || used + lay[pending[g0]].T <= capacity)) {
                used += lay[pending[g0]].T;
                grp.push_back(pending[g0++]);
            }
            std::vector<ResRun<Real>> runs;
            runs.reserve(grp.size());
            for (int i : grp)
                runs.push_back(prepareResident<Real>(*engs[i], args[i], 1, &var[i], nullptr, nullptr, 0,
                                                     args[i].weights, &bound[i], b[i].data(), lay[i]));
            std::vector<ResRun<Real>*> ptrs;
            for (auto& R : runs) ptrs.push_back(&R);
            double ms = 0.0;
            int nl = 0;
            launchRuns<Real>(ptrs, ms, nl);
            for (size_t q = 0; q < grp.size(); ++q) {
                const int i = grp[q];
                double bytes = 0.0, steps = 0.0;
                gatherRun<Real>(runs[q], b[i].data(), &flags[i], &its[i], &lls[i], bytes, steps);
                ++count[i];
                if (flags[i] == CCD_ILLCONDITIONED && count[i] < args[i].max_bound_count) {   // update() retry
                    bound[i] /= 10.0;
                    for (int j = engs[i]->D->hasOffset ? 1 : 0; j < engs[i]->D->J; ++j) b[i][j] = 0.0;
                    retry.push_back(i);
                }
            }
        }
        pending.swap(retry);
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; ++i) {
        const int J = engs[i]->D->J;
        std::copy(b[i].begin(), b[i].end(), betas[i]);
        double lp = 0.0;
        for (int j = 0; j < J; ++j)
            lp += priorLogDensity((args[i].exclude && args[i].exclude[j]) ? CCD_PRIOR_NONE : args[i].prior_type,
                                  args[i].variance, b[i][j]);
        res[i].return_flag = flags[i];
        res[i].iterations = its[i];
        res[i].find_mode_calls = count[i];
        res[i].log_likelihood = lls[i];
        res[i].log_prior = lp;
        res[i].seconds = secs;
    }
}

void fitMulti(EngineBase* const* engs, int n, const ccd_fit_args* args, double* const* betas, ccd_fit_result* res) {
    if (n < 1) throw std::invalid_argument("n must be >= 1");
    if (engs[0]->precision() == CCD_FP32) {
        std::vector<EngineT<float>*> e(n);
        for (int i = 0; i < n; ++i) {
            e[i] = dynamic_cast<EngineT<float>*>(engs[i]);
            if (!e[i]) throw std::invalid_argument("fit_multi needs engines of one precision");
        }
        fitMultiT<float>(e, args, betas, res);
    } else {
        std::vector<EngineT<double>*> e(n);
        for (int i = 0; i < n; ++i) {
            e[i] = dynamic_cast<EngineT<double>*>(engs[i]);
            if (!e[i]) throw std::invalid_argument("fit_multi needs engines of one precision");
        }
        fitMultiT<double>(e, args, betas, res);
    }
}

}  // namespace ccd
