// kernels.h -- HIP kernels of the per-coordinate CCD engine (gfx950, wave64).
//
// Every kernel here is a pass over one column slice or one row/pid vector; none of
// the work is a dense contraction, so nothing runs on MFMA. The passes are
// HBM/L2-bound gathers and scans: coalesced reads of CSC row indices, gathers of
// per-row state, wave64 shuffle reductions and scans, LDS for cross-wave steps.
#pragma once

#include "common.h"

#include <climits>

namespace ccd {

// ---------------------------------------------------------------------------
// Column view: INDICATOR / SPARSE hold row ids; SPARSE and DENSE hold values.
// ---------------------------------------------------------------------------
template <typename Real>
struct Col {
    const int* rows;
    const Real* vals;
    int64_t n;
};

template <typename Real, int FMT>
__device__ __forceinline__ int64_t rowOf(const Col<Real>& c, int64_t e) {
    if constexpr (FMT == F_INDICATOR || FMT == F_SPARSE) return c.rows[e];
    else return e;
}
template <typename Real, int FMT>
__device__ __forceinline__ Real xOf(const Col<Real>& c, int64_t e) {
    if constexpr (FMT == F_SPARSE || FMT == F_DENSE) return c.vals[e];
    else return Real(1);
}
template <int FMT>
constexpr bool isIndicatorFmt() { return FMT == F_INDICATOR || FMT == F_INTERCEPT; }

// bfloat16 bits of a float, rounded to nearest even (finite inputs), and back: the
// storage of the low part of a compensated FP32 xb.
__host__ __device__ __forceinline__ unsigned short bf16Bits(float x) {
    unsigned u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16Value(unsigned short b) {
    const unsigned u = (unsigned)b << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }

// ---------------------------------------------------------------------------
// Block primitives
// ---------------------------------------------------------------------------
// Segmented sum: value(s) + "a segment starts here" flag. (a,fa) o (b,fb) =
// (fb ? b : a+b, fa|fb) is associative with identity (0, 0).
struct Seg2 {
    double a, b;
    int f;
    int pad;
};
__device__ __forceinline__ Seg2 segCombine(const Seg2& x, const Seg2& y) {
    Seg2 r;
    r.a = y.f ? y.a : x.a + y.a;
    r.b = y.f ? y.b : x.b + y.b;
    r.f = x.f | y.f;
    r.pad = 0;
    return r;
}

__device__ __forceinline__ Seg2 waveSegScanIncl(Seg2 v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        Seg2 p;
        p.a = __shfl_up(v.a, d, 64);
        p.b = __shfl_up(v.b, d, 64);
        p.f = __shfl_up(v.f, d, 64);
        p.pad = 0;
        if (lane >= d) v = segCombine(p, v);
    }
    return v;
}

// The same scan on DPP lane moves (CDNA's row shifts and row broadcasts), which
// stay in the VALU instead of a round trip through the LDS crossbar per step:
// Hillis-Steele inside each row of 16 lanes (row_shr 1, 2, 4, 8), then row 0's
// total into row 1 and row 2's into row 3 (row_bcast:15), then the total of rows
// 0-1 into rows 2-3 (row_bcast:31). Lanes whose source is outside the row or
// masked keep the identity. REV combines in descending lane order (lane 63 then
// holds x63 o x62 o ... o x0).
template <int CTRL, int ROWS>
__device__ __forceinline__ int dppInt(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dppDouble(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = dppInt<CTRL, ROWS>((int)(u & 0xffffffffll));
    const int hi = dppInt<CTRL, ROWS>((int)(u >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL, int ROWS, bool REV>
__device__ __forceinline__ void segDppStep(Seg2& v) {
    const Seg2 p{dppDouble<CTRL, ROWS>(v.a), dppDouble<CTRL, ROWS>(v.b), dppInt<CTRL, ROWS>(v.f), 0};
    v = REV ? segCombine(v, p) : segCombine(p, v);
}
template <bool REV = false>
__device__ __forceinline__ Seg2 waveSegScanDpp(Seg2 v) {
    segDppStep<0x111, 0xf, REV>(v);   // row_shr:1
    segDppStep<0x112, 0xf, REV>(v);   // row_shr:2
    segDppStep<0x114, 0xf, REV>(v);   // row_shr:4
    segDppStep<0x118, 0xf, REV>(v);   // row_shr:8
    segDppStep<0x142, 0xa, REV>(v);   // row_bcast:15 into rows 1 and 3
    segDppStep<0x143, 0xc, REV>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}
// value of the previous lane (identity at lane 0): wave_shr:1
__device__ __forceinline__ Seg2 segFromPrevLane(const Seg2& v) {
    return Seg2{dppDouble<0x138, 0xf>(v.a), dppDouble<0x138, 0xf>(v.b), dppInt<0x138, 0xf>(v.f), 0};
}
// lane 63's value in every lane (through scalar registers)
__device__ __forceinline__ double readLane63(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ Seg2 segLane63(const Seg2& v) {
    return Seg2{readLane63(v.a), readLane63(v.b), __builtin_amdgcn_readlane(v.f, 63), 0};
}

// Plain sum over the 64 lanes on the same DPP sequence; every lane gets the total.
__device__ __forceinline__ double waveSumDpp(double v) {
    v += dppDouble<0x111, 0xf>(v);
    v += dppDouble<0x112, 0xf>(v);
    v += dppDouble<0x114, 0xf>(v);
    v += dppDouble<0x118, 0xf>(v);
    v += dppDouble<0x142, 0xa>(v);
    v += dppDouble<0x143, 0xc>(v);
    return readLane63(v);
}

// Segmented sums of three values (risk-set denominator and the two column
// numerators of the Breslow sweep) with one stratum-start flag.
struct Seg3 {
    double d, n1, n2;
    int f;
};
__device__ __forceinline__ Seg3 seg3Combine(const Seg3& x, const Seg3& y) {
    return Seg3{y.f ? y.d : x.d + y.d, y.f ? y.n1 : x.n1 + y.n1, y.f ? y.n2 : x.n2 + y.n2, x.f | y.f};
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg3DppStep(Seg3& v) {
    const Seg3 p{dppDouble<CTRL, ROWS>(v.d), dppDouble<CTRL, ROWS>(v.n1), dppDouble<CTRL, ROWS>(v.n2),
                 dppInt<CTRL, ROWS>(v.f)};
    v = seg3Combine(p, v);
}
__device__ __forceinline__ Seg3 waveSeg3ScanDpp(Seg3 v) {
    seg3DppStep<0x111, 0xf>(v);
    seg3DppStep<0x112, 0xf>(v);
    seg3DppStep<0x114, 0xf>(v);
    seg3DppStep<0x118, 0xf>(v);
    seg3DppStep<0x142, 0xa>(v);
    seg3DppStep<0x143, 0xc>(v);
    return v;
}
__device__ __forceinline__ Seg3 seg3FromPrevLane(const Seg3& v) {
    return Seg3{dppDouble<0x138, 0xf>(v.d), dppDouble<0x138, 0xf>(v.n1), dppDouble<0x138, 0xf>(v.n2),
                dppInt<0x138, 0xf>(v.f)};
}
__device__ __forceinline__ Seg3 seg3Lane63(const Seg3& v) {
    return Seg3{readLane63(v.d), readLane63(v.n1), readLane63(v.n2), __builtin_amdgcn_readlane(v.f, 63)};
}

// Exclusive segmented scan across the workgroup; `total` receives the
// combination of all threads' values. Ends with a barrier (safe to call again).
__device__ __forceinline__ Seg2 blockSegScanExcl(Seg2 v, Seg2& total) {
    __shared__ Seg2 wtot[WAVES];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const Seg2 inc = waveSegScanIncl(v);
    if (lane == 63) wtot[wid] = inc;
    __syncthreads();
    Seg2 wpre{0.0, 0.0, 0, 0};
    Seg2 all{0.0, 0.0, 0, 0};
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        if (w < wid) wpre = segCombine(wpre, wtot[w]);
        all = segCombine(all, wtot[w]);
    }
    total = all;
    Seg2 ex;
    ex.a = __shfl_up(inc.a, 1, 64);
    ex.b = __shfl_up(inc.b, 1, 64);
    ex.f = __shfl_up(inc.f, 1, 64);
    ex.pad = 0;
    if (lane == 0) ex = Seg2{0.0, 0.0, 0, 0};
    const Seg2 res = segCombine(wpre, ex);
    __syncthreads();
    return res;
}

// Sum of NV doubles over the workgroup; the result is valid in thread 0.
// Fixed reduction tree, so results are bit-reproducible for a fixed launch.
template <int NV>
__device__ __forceinline__ void blockReduce(double (&v)[NV]) {
    __shared__ double red[NV][WAVES];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double x = v[i];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) x += __shfl_down(x, d, 64);
        if (lane == 0) red[i][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < WAVES; ++w) s += red[i][w];
            v[i] = s;
        }
    }
    __syncthreads();
}

// Partials that another workgroup reads are stored and loaded with relaxed
// agent-scope atomics (coherent per word across the XCDs, no whole-cache flush);
// workgroup-scope fences order them around the arrival count.
__device__ __forceinline__ void pst(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pld(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void psti(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int pldi(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until every global store this wave issued has been acknowledged. Agent-scope
// stores are coherent once acknowledged, but stores to different addresses are
// acknowledged in any order, and a workgroup-scope fence does not wait for them:
// without this, a flag or counter written after a workgroup barrier can become
// visible to another workgroup before the data stored ahead of the barrier.
__device__ __forceinline__ void drainStores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Publish-then-count: true in the workgroup that arrives last.
__device__ __forceinline__ bool lastBlockArrive(unsigned* counter) {
    __shared__ int amLast;
    drainStores();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        amLast = (prev == gridDim.x - 1);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return amLast != 0;
}

// ---------------------------------------------------------------------------
// Independent-row models (LR, PR): ModelSpecifics.hpp:1022-1038 with the traits
// Logistic::incrementGradientAndHessian (ModelSpecifics.h:540-568) and
// PoissonRegression::incrementGradientAndHessian (ModelSpecifics.h:1508-1543).
// out[0] = sum(...) - XjY, out[1] = hessian.
// ---------------------------------------------------------------------------
template <typename Real, int MODEL, int FMT, bool W>
__global__ __launch_bounds__(BLOCK) void kGradIndep(Col<Real> c, const Real* __restrict__ ex,
                                                   const Real* __restrict__ kw, Partial* parts,
                                                   unsigned* counter, double xjy, double* out,
                                                   int64_t perBlock) {
    constexpr bool IND = isIndicatorFmt<FMT>();
    constexpr int U = 4;   // entries whose gathers are in flight together
    const int64_t lo = (int64_t)blockIdx.x * perBlock;
    const int64_t hi = lo + perBlock < c.n ? lo + perBlock : c.n;
    double g = 0.0, h = 0.0;
    for (int64_t e0 = lo + threadIdx.x; e0 < hi; e0 += U * BLOCK) {
        int64_t k[U];
        Real x[U], ee[U], w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t e = e0 + (int64_t)u * BLOCK;
            k[u] = e < hi ? rowOf<Real, FMT>(c, e) : -1;
            x[u] = e < hi ? xOf<Real, FMT>(c, e) : Real(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ee[u] = k[u] >= 0 ? ex[k[u]] : Real(0);
            if constexpr (W) w[u] = k[u] >= 0 ? kw[k[u]] : Real(0);
        }
        // accumulated in entry order, as one entry per step would
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k[u] < 0) break;
            Real gi, hh;
            if constexpr (MODEL == M_LOGISTIC) {
                const Real D = Real(1) + ee[u];
                gi = ee[u] * x[u] / D;
                if constexpr (IND) hh = gi * (Real(1) - gi);
                else hh = ee[u] * x[u] * x[u] / D - gi * gi;
            } else {
                gi = ee[u] * x[u];
                if constexpr (IND) hh = gi;
                else hh = gi * x[u];
            }
            if constexpr (W) {
                gi *= w[u];
                hh *= w[u];
            }
            g += (double)gi;
            h += (double)hh;
        }
    }
    double v[2] = {g, h};
    blockReduce<2>(v);
    if (threadIdx.x == 0) {
        pst(&parts[blockIdx.x].g, v[0]);
        pst(&parts[blockIdx.x].h, v[1]);
    }
    if (lastBlockArrive(counter)) {
        double s[2] = {0.0, 0.0};
        for (int b = threadIdx.x; b < (int)gridDim.x; b += BLOCK) {
            s[0] += pld(&parts[b].g);
            s[1] += pld(&parts[b].h);
        }
        blockReduce<2>(s);
        if (threadIdx.x == 0) {
            out[0] = s[0] - xjy;
            out[1] = s[1];
            *counter = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// Grouped models (CLR, Breslow Cox): one pass over the column's entries.
//
// With u_k = w_k e_k x_k, pid p(k), segment s(k) (stratum) and per-pid weights
//   CLR : R1[p] = nE_p / denom_p,            R2[p] = nE_p / denom_p^2
//   Cox : R1[p] = sum_{i>=p, same stratum} nE_i / D_i,   R2[p] = ... nE_i / D_i^2
// the reference's derivatives (ModelSpecifics.hpp:904-962 for Cox,
// :1127-1179 for CLR; traits ModelSpecifics.h:820-837, 1185-1202) are
//   g = sum_k u_k R1[p(k)] - XjY
//   h = sum_k u_k x_k R1[p(k)] - sum_k R2[p(k)] u_k (2 S_k - u_k)
// where S_k is the running sum of u over the column in row order, restarted at
// every segment change (Cox: stratum; CLR: stratum = pid). The second sum is
// sum_p nE_p A_p^2 / D_p^2 written as a telescoping series, so the O(N) risk-set
// sweep of the reference becomes O(nnz_j) per coordinate.
//
// Blocks own contiguous entry ranges; S is scanned inside the block and carried
// across blocks in the final reduction: block w reports its local quadratic term,
// the R2-weighted sum m of entries before its first reset, and its trailing segment
// sum t; the carry into block w is the segmented prefix of the t's.
// ---------------------------------------------------------------------------
template <typename Real, int FMT, bool COXW>
__global__ __launch_bounds__(BLOCK) void kGradGrouped(Col<Real> c, const Real* __restrict__ ex,
                                                     const Real* __restrict__ kw,
                                                     const int* __restrict__ pid,
                                                     const int* __restrict__ seg,
                                                     const double* __restrict__ R1,
                                                     const double* __restrict__ R2, Partial* parts,
                                                     unsigned* counter, double xjy, double* out,
                                                     int64_t perBlock) {
    __shared__ int lastSeg[BLOCK];
    const int64_t lo = (int64_t)blockIdx.x * perBlock;
    const int64_t hi = lo + perBlock < c.n ? lo + perBlock : c.n;
    double G = 0.0, H1 = 0.0, L = 0.0, M = 0.0;
    double carry = 0.0;
    int seen = 0;
    int prevSeg = INT_MIN;
    if (lo > 0 && lo < hi) prevSeg = seg[rowOf<Real, FMT>(c, lo - 1)];

    for (int64_t base = lo; base < hi; base += TILE) {
        const int64_t e0 = base + (int64_t)threadIdx.x * ITEMS;
        double u[ITEMS], xv[ITEMS], s[ITEMS];
        int sg[ITEMS], pp[ITEMS], any[ITEMS];
        int nvalid = 0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int64_t e = e0 + i;
            if (e < hi) {
                const int64_t k = rowOf<Real, FMT>(c, e);
                const Real x = xOf<Real, FMT>(c, e);
                Real uu = ex[k] * x;
                if constexpr (COXW) uu *= kw[k];
                u[i] = (double)uu;
                xv[i] = (double)x;
                sg[i] = seg[k];
                pp[i] = pid[k];
                nvalid = i + 1;
            } else {
                u[i] = 0.0;
                xv[i] = 0.0;
                sg[i] = 0;
                pp[i] = 0;
            }
        }
        lastSeg[threadIdx.x] = nvalid > 0 ? sg[nvalid - 1] : INT_MIN;
        __syncthreads();
        const int prev = threadIdx.x == 0 ? prevSeg : lastSeg[threadIdx.x - 1];
        // thread-local segmented prefix
        double run = 0.0;
        int anyR = 0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (i < nvalid) {
                const int before = (i == 0) ? prev : sg[i - 1];
                if (before != INT_MIN && sg[i] != before) {
                    run = 0.0;
                    anyR = 1;
                }
                run += u[i];
            }
            s[i] = run;
            any[i] = anyR;
        }
        Seg2 agg{run, 0.0, anyR, 0};
        Seg2 tot;
        const Seg2 exc = blockSegScanExcl(agg, tot);
        const Seg2 inc = segCombine(Seg2{carry, 0.0, seen, 0}, exc);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (i < nvalid) {
                const double prefix = any[i] ? s[i] : inc.a + s[i];
                const int p = pp[i];
                const double r1 = R1[p], r2 = R2[p];
                G += u[i] * r1;
                H1 += u[i] * xv[i] * r1;
                L += r2 * u[i] * (2.0 * prefix - u[i]);
                if (!inc.f && !any[i]) M += r2 * u[i];
            }
        }
        const Seg2 nc = segCombine(Seg2{carry, 0.0, seen, 0}, tot);
        carry = nc.a;
        seen = nc.f;
        const int64_t lastE = (base + TILE < hi ? base + TILE : hi) - 1;
        prevSeg = lastSeg[(int)((lastE - base) / ITEMS)];
        __syncthreads();
    }
    double v[4] = {G, H1, L, M};
    blockReduce<4>(v);
    if (threadIdx.x == 0) {
        Partial& P = parts[blockIdx.x];
        pst(&P.g, v[0]);
        pst(&P.h, v[1]);
        pst(&P.q, v[2]);
        pst(&P.m, v[3]);
        pst(&P.t, carry);
        psti(&P.reset, seen);
    }
    if (lastBlockArrive(counter)) {
        const int b = threadIdx.x;
        const bool act = b < (int)gridDim.x;
        Seg2 a{act ? pld(&parts[b].t) : 0.0, 0.0, act ? pldi(&parts[b].reset) : 0, 0};
        Seg2 tot;
        const Seg2 off = blockSegScanExcl(a, tot);
        double s2[3] = {act ? pld(&parts[b].g) : 0.0, act ? pld(&parts[b].h) : 0.0,
                        act ? pld(&parts[b].q) + 2.0 * off.a * pld(&parts[b].m) : 0.0};
        blockReduce<3>(s2);
        if (threadIdx.x == 0) {
            out[0] = s2[0] - xjy;
            out[1] = s2[1] - s2[2];
            *counter = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// updateXBeta (ModelSpecifics.hpp:1524-1586)
// ---------------------------------------------------------------------------
template <typename Real, int FMT>
__global__ __launch_bounds__(BLOCK) void kUpdateIndep(Col<Real> c, Real delta, Real* __restrict__ xb,
                                                     Real* __restrict__ ex) {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < c.n; e += stride) {
        const int64_t k = rowOf<Real, FMT>(c, e);
        const Real v = xb[k] + delta * xOf<Real, FMT>(c, e);
        xb[k] = v;
        ex[k] = dexp(v);
    }
}

// Grouped update: the first entry of each run of equal pid updates the whole run
// and applies one change to denom[pid]; runs are disjoint, so no atomics.
template <typename Real, int FMT, bool COXW, bool CLRR>
__global__ __launch_bounds__(BLOCK) void kUpdateGrouped(Col<Real> c, Real delta, Real* __restrict__ xb,
                                                       Real* __restrict__ ex, const Real* __restrict__ kw,
                                                       const int* __restrict__ pid, double* denom,
                                                       const double* __restrict__ nW, double* R1,
                                                       double* R2) {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < c.n; e += stride) {
        const int p = pid[rowOf<Real, FMT>(c, e)];
        if (e > 0 && pid[rowOf<Real, FMT>(c, e - 1)] == p) continue;
        double acc = 0.0;
        for (int64_t e2 = e; e2 < c.n; ++e2) {
            const int64_t k2 = rowOf<Real, FMT>(c, e2);
            if (e2 > e && pid[k2] != p) break;
            const Real old = ex[k2];
            const Real v = xb[k2] + delta * xOf<Real, FMT>(c, e2);
            xb[k2] = v;
            const Real nv = dexp(v);
            ex[k2] = nv;
            if constexpr (COXW) acc += (double)(kw[k2] * nv) - (double)(kw[k2] * old);
            else acc += (double)nv - (double)old;
        }
        const double dn = denom[p] + acc;
        denom[p] = dn;
        if constexpr (CLRR) {
            const double r = nW[p] / dn;
            R1[p] = r;
            R2[p] = r / dn;
        }
    }
}

template <typename Real, int FMT>
__global__ __launch_bounds__(BLOCK) void kAxpy(Col<Real> c, Real beta, Real* __restrict__ xb) {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x; e < c.n; e += stride)
        xb[rowOf<Real, FMT>(c, e)] += beta * xOf<Real, FMT>(c, e);
}

// ---------------------------------------------------------------------------
// computeRemainingStatistics (ModelSpecifics.hpp:1588-1633)
// ---------------------------------------------------------------------------
template <typename Real>
__global__ __launch_bounds__(BLOCK) void kExp(int64_t K, const Real* __restrict__ xb, Real* __restrict__ ex) {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x; k < K; k += stride) ex[k] = dexp(xb[k]);
}

// Per-pid denominators; rows of pid p are [pstart[p], pstart[p+1]).
template <typename Real, bool W, bool CLRR>
__global__ __launch_bounds__(BLOCK) void kDenomByPid(int64_t N, const int* __restrict__ pstart,
                                                    const Real* __restrict__ xb, Real* __restrict__ ex,
                                                    const Real* __restrict__ kw, double* __restrict__ denom,
                                                    const double* __restrict__ nW, double* R1, double* R2) {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x; p < N; p += stride) {
        double s = 0.0;
        for (int k = pstart[p]; k < pstart[p + 1]; ++k) {
            const Real v = dexp(xb[k]);
            ex[k] = v;
            if constexpr (W) s += (double)(kw[k] * v);
            else s += (double)v;
        }
        denom[p] = s;
        if constexpr (CLRR) {
            const double r = nW[p] / s;
            R1[p] = r;
            R2[p] = r / s;
        }
    }
}

// ---------------------------------------------------------------------------
// Segmented scans over pids for the Breslow risk sets (three phases).
//   forward : D[p]  = sum_{i<=p, same stratum} denom[i]      (ModelSpecifics.hpp:1670-1694)
//   backward: R1[p] = sum_{i>=p, same stratum} nE_i / D_i,  R2 with D_i^2
// ---------------------------------------------------------------------------
struct ScanFwd {
    int64_t n;
    const double* denom;
    const int* pseg;
    double* D;
    __device__ Seg2 load(int64_t i) const {
        return Seg2{denom[i], 0.0, (i > 0 && pseg[i] != pseg[i - 1]) ? 1 : 0, 0};
    }
    __device__ void store(int64_t i, const Seg2& v) const { D[i] = v.a; }
};
struct ScanBwd {
    int64_t n;
    const double* nW;
    const double* D;
    const int* pseg;
    double* R1;
    double* R2;
    __device__ Seg2 load(int64_t i) const {
        const int64_t p = n - 1 - i;
        const double r = nW[p] / D[p];
        return Seg2{r, r / D[p], (p < n - 1 && pseg[p + 1] != pseg[p]) ? 1 : 0, 0};
    }
    __device__ void store(int64_t i, const Seg2& v) const {
        const int64_t p = n - 1 - i;
        R1[p] = v.a;
        R2[p] = v.b;
    }
};

template <class Op>
__global__ __launch_bounds__(BLOCK) void kScanReduce(Op op, Seg2* parts) {
    const int64_t base = (int64_t)blockIdx.x * TILE + (int64_t)threadIdx.x * ITEMS;
    Seg2 acc{0.0, 0.0, 0, 0};
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
        if (base + i < op.n) acc = segCombine(acc, op.load(base + i));
    Seg2 tot;
    blockSegScanExcl(acc, tot);
    if (threadIdx.x == 0) parts[blockIdx.x] = tot;
}

// Exclusive segmented scan of nb block aggregates, in place, one workgroup.
static __global__ __launch_bounds__(BLOCK) void kScanPartials(int64_t nb, Seg2* parts) {
    const int64_t per = (nb + BLOCK - 1) / BLOCK;
    const int64_t lo = (int64_t)threadIdx.x * per;
    Seg2 acc{0.0, 0.0, 0, 0};
    for (int64_t i = lo; i < lo + per && i < nb; ++i) acc = segCombine(acc, parts[i]);
    Seg2 tot;
    Seg2 run = blockSegScanExcl(acc, tot);
    for (int64_t i = lo; i < lo + per && i < nb; ++i) {
        const Seg2 v = parts[i];
        parts[i] = run;
        run = segCombine(run, v);
    }
}

template <class Op>
__global__ __launch_bounds__(BLOCK) void kScanApply(Op op, const Seg2* parts) {
    const int64_t base = (int64_t)blockIdx.x * TILE + (int64_t)threadIdx.x * ITEMS;
    Seg2 v[ITEMS];
    Seg2 acc{0.0, 0.0, 0, 0};
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        v[i] = (base + i < op.n) ? op.load(base + i) : Seg2{0.0, 0.0, 0, 0};
        acc = segCombine(acc, v[i]);
    }
    Seg2 tot;
    Seg2 run = segCombine(parts[blockIdx.x], blockSegScanExcl(acc, tot));
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        run = segCombine(run, v[i]);
        if (base + i < op.n) op.store(base + i, run);
    }
}

// ---------------------------------------------------------------------------
// Row/pid reductions: log-likelihood (ModelSpecifics.hpp:605-650), predictive
// log-likelihood (:652-690), gradient objective (:103-120), XjY (:486-524).
// ---------------------------------------------------------------------------
enum RowOp : int { OP_LOGLIK = 0, OP_PRED = 1, OP_GRADOBJ = 2 };

template <typename Real, int MODEL, int OP, bool W>
__global__ __launch_bounds__(BLOCK) void kRowReduce(int64_t K, int64_t N, const Real* __restrict__ y,
                                                   const Real* __restrict__ xb, const Real* __restrict__ ex,
                                                   const Real* __restrict__ kw, const double* __restrict__ pw,
                                                   const int* __restrict__ pid, const double* __restrict__ nW,
                                                   const double* __restrict__ den, Partial* parts,
                                                   unsigned* counter, double add, double* out) {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    double acc = 0.0;
    const int64_t len = K > N ? K : N;
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < len; i += stride) {
        if (i < K) {
            const double yy = (double)y[i];
            const double x = (double)xb[i];
            if constexpr (OP == OP_GRADOBJ) {
                double c = x * yy;
                if constexpr (W) c *= (double)kw[i];
                acc += c;
            } else if constexpr (OP == OP_LOGLIK) {
                const int yi = (int)y[i];
                double c = (MODEL == M_COX) ? (yi != 1 ? 0.0 : x) : yi * x;
                double w = 1.0;
                if constexpr (W) w = (double)kw[i];
                acc += c * w;
                if constexpr (MODEL == M_LOGISTIC) acc -= w * log(1.0 + (double)ex[i]);
                if constexpr (MODEL == M_POISSON) acc -= w * (double)ex[i];
            } else {   // OP_PRED, weights pw
                const double w = pw[i];
                if constexpr (MODEL == M_LOGISTIC) acc += yy * w * (x - log(1.0 + (double)ex[i]));
                else if constexpr (MODEL == M_POISSON) acc += (yy * x - exp(x)) * w;
                else if constexpr (MODEL == M_CLR) acc += yy * w * (x - log(den[pid[i]]));
                else acc += (w == 0.0) ? 0.0 : yy * w * (x - log(den[pid[i]]));
            }
        }
        if constexpr (OP == OP_LOGLIK && (MODEL == M_CLR || MODEL == M_COX)) {
            if (i < N) acc -= nW[i] * log(den[i]);
        }
    }
    double v[1] = {acc};
    blockReduce<1>(v);
    if (threadIdx.x == 0) pst(&parts[blockIdx.x].g, v[0]);
    if (lastBlockArrive(counter)) {
        double s[1] = {0.0};
        for (int b = threadIdx.x; b < (int)gridDim.x; b += BLOCK) s[0] += pld(&parts[b].g);
        blockReduce<1>(s);
        if (threadIdx.x == 0) {
            out[0] = s[0] + add;
            *counter = 0u;
        }
    }
}

// XjY[j] = sum over column j of x * y * (w): one workgroup per column.
template <typename Real, bool W>
__global__ __launch_bounds__(BLOCK) void kXjY(int J, int64_t K, const int* __restrict__ fmt,
                                             const int64_t* __restrict__ colPtr,
                                             const int64_t* __restrict__ densePtr,
                                             const int* __restrict__ rows, const Real* __restrict__ vals,
                                             const Real* __restrict__ dense, const Real* __restrict__ y,
                                             const Real* __restrict__ kw, double* XjY) {
    for (int j = blockIdx.x; j < J; j += gridDim.x) {
        const int f = fmt[j];
        double acc = 0.0;
        if (f == F_INDICATOR || f == F_SPARSE) {
            for (int64_t e = colPtr[j] + threadIdx.x; e < colPtr[j + 1]; e += BLOCK) {
                const int k = rows[e];
                double c = (double)y[k];
                if (f == F_SPARSE) c *= (double)vals[e];
                if (W) c *= (double)kw[k];
                acc += c;
            }
        } else {
            const Real* dv = (f == F_DENSE) ? dense + densePtr[j] : nullptr;
            for (int64_t k = threadIdx.x; k < K; k += BLOCK) {
                double c = (double)y[k];
                if (dv) c *= (double)dv[k];
                if (W) c *= (double)kw[k];
                acc += c;
            }
        }
        double v[1] = {acc};
        blockReduce<1>(v);
        if (threadIdx.x == 0) XjY[j] = v[0];
    }
}

template <typename Real, int MODEL>
__global__ __launch_bounds__(BLOCK) void kPredEst(int64_t K, const Real* __restrict__ xb, double* out) {
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    for (int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x; k < K; k += stride) {
        const double x = (double)xb[k];
        if constexpr (MODEL == M_LOGISTIC) {
            const double t = exp(x);
            out[k] = t / (t + 1.0);
        } else if constexpr (MODEL == M_POISSON) {
            out[k] = exp(x);
        } else {
            out[k] = x;
        }
    }
}

}  // namespace ccd
