// fisher.inl -- Fisher information of a set of columns (included by engine.hip).
//
// ModelSpecifics::computeFisherInformationImpl (ModelSpecifics.hpp:1269-1352) over the
// pairs CyclicCoordinateDescent::computeFisherInformation asks for
// (CyclicCoordinateDescent.cpp:1288-1300), with unit row weights:
//
//   I[a][b] = sum_k w_k x_{k,a} x_{k,b}  -  sum_n C[n][a] C[n][b] / denom_n^2
//
// w_k is the model's increment (ModelSpecifics.h: Logistic :521-535 p - p^2,
// PoissonRegression :1497-1507 e_k, Survival :488-497 e_k / denom of k's group, for
// CLR strata and Cox tie groups); the second term exists for grouped models only
// (hasStrataCrossTerms), with C[n][a] = sum over rows k of group n of x_{k,a} e_k
// (getSubjectSpecificHessianIterator, :1245-1267) and denom_n the group's sum of e_k
// (computeRemainingStatistics, :1588-1612; for Cox the per-pid sum, not the risk set).
//
// On the device both terms are weighted Gram matrices of a dense row-major block:
// the selected columns are scattered into Z[rows][m] (a row chunk at a time, so Z
// stays within ZBYTES), the grouped models fold Z into C[N][m], and kGram forms the
// upper 64 x 64 tiles of Z^T diag(w) Z in FP64 over row splits whose partial sums
// kGramReduce adds in a fixed order: the result is the same on every run.

namespace ccd {
namespace fis {

constexpr int GT = 64;                       // Gram tile edge
constexpr int GR = 32;                       // rows staged in LDS per step
constexpr size_t ZBYTES = size_t(256) << 20; // row-chunk buffer

__device__ __forceinline__ int64_t lowerBound(const int32_t* r, int64_t n, int64_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (r[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Z[(k - r0) * m + a] = x_{k, idx[a]} for rows k in [r0, r1); Z is zero beforehand.
// blockIdx.y = a; the sorted row ids of a sparse column are searched for the chunk.
template <typename Real>
__global__ __launch_bounds__(BLOCK) void kScatter(int m, const int32_t* __restrict__ idx, const int32_t* __restrict__ fmt,
                                                  const int64_t* __restrict__ colPtr, const int64_t* __restrict__ densePtr,
                                                  const int32_t* __restrict__ rows, const Real* __restrict__ vals,
                                                  const Real* __restrict__ dense, int64_t r0, int64_t r1,
                                                  double* __restrict__ Z) {
    const int a = blockIdx.y;
    const int j = idx[a];
    const int f = fmt[j];
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (f == F_INDICATOR || f == F_SPARSE) {
        const int64_t c0 = colPtr[j];
        const int64_t n = colPtr[j + 1] - c0;
        const int64_t e0 = lowerBound(rows + c0, n, r0), e1 = lowerBound(rows + c0, n, r1);
        for (int64_t e = e0 + t; e < e1; e += stride)
            Z[(int64_t)(rows[c0 + e] - r0) * m + a] = (f == F_SPARSE) ? (double)vals[c0 + e] : 1.0;
    } else {
        const Real* v = (f == F_DENSE) ? dense + densePtr[j] : nullptr;
        for (int64_t k = r0 + t; k < r1; k += stride) Z[(k - r0) * m + a] = v ? (double)v[k] : 1.0;
    }
}

// Row weight of the first term (incrementFisherInformation, unit weights).
template <typename Real, int MODEL>
__global__ __launch_bounds__(BLOCK) void kRowWeight(int64_t K, const Real* __restrict__ ex, const double* __restrict__ denom,
                                                    const int32_t* __restrict__ group, double* __restrict__ w) {
    for (int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x; k < K; k += (int64_t)gridDim.x * BLOCK) {
        const double e = (double)ex[k];
        if constexpr (MODEL == M_LOGISTIC) {
            const double g = e / (1.0 + e);
            w[k] = g - g * g;
        } else if constexpr (MODEL == M_POISSON) {
            w[k] = e;
        } else {
            w[k] = e / denom[group[k]];
        }
    }
}

// 1 / denom_n^2 of each group (the cross term's weight).
__global__ __launch_bounds__(BLOCK) void kGroupWeight(int64_t N, const double* __restrict__ denom, double* __restrict__ w) {
    for (int64_t n = (int64_t)blockIdx.x * BLOCK + threadIdx.x; n < N; n += (int64_t)gridDim.x * BLOCK)
        w[n] = 1.0 / (denom[n] * denom[n]);
}

// C[n * m + a] += sum over rows k of group n inside [r0, r1) of e_k Z[(k - r0) * m + a],
// for groups n in [n0, n1); groups are contiguous row ranges [gstart[n], gstart[n+1]).
template <typename Real>
__global__ __launch_bounds__(BLOCK) void kGroupSums(int64_t n0, int64_t n1, int m, const int32_t* __restrict__ gstart,
                                                    int64_t r0, int64_t r1, const Real* __restrict__ ex,
                                                    const double* __restrict__ Z, double* __restrict__ Cm) {
    const int64_t total = (n1 - n0) * m;
    for (int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x; t < total; t += (int64_t)gridDim.x * BLOCK) {
        const int64_t n = n0 + t / m;
        const int a = (int)(t % m);
        const int64_t k0 = max((int64_t)gstart[n], r0), k1 = min((int64_t)gstart[n + 1], r1);
        double s = 0.0;
        for (int64_t k = k0; k < k1; ++k) s += (double)ex[k] * Z[(k - r0) * m + a];
        Cm[n * m + a] += s;
    }
}

// Partial Gram of the tile pair blockIdx.x (ti <= tj over nt tiles of GT columns) on
// row split blockIdx.y: part[y][i][j] = sum_r w_r Z[r][i] Z[r][j] for i in tile ti,
// j in tile tj. 256 threads as 16 x 16; thread (tx, ty) owns the 4 x 4 outputs
// i = ty + 16a, j = tx + 16b, so each staged row costs 8 LDS reads for 16 FMAs.
__global__ __launch_bounds__(256) void kGram(int64_t R, int m, int nt, const double* __restrict__ Z,
                                             const double* __restrict__ w, int64_t rowsPerSplit,
                                             double* __restrict__ part) {
    __shared__ double zi[GR][GT + 1];
    __shared__ double zj[GR][GT + 1];
    // tile pair of this block: row-major over the upper triangle
    int p = blockIdx.x, ti = 0;
    while (p >= nt - ti) { p -= nt - ti; ++ti; }
    const int tj = ti + p;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t rs0 = (int64_t)blockIdx.y * rowsPerSplit;
    const int64_t rs1 = min(R, rs0 + rowsPerSplit);
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    // staging: thread t loads column c = t % GT of rows t / GT + 4q (GR x GT per matrix)
    const int lc = threadIdx.x % GT, lr = threadIdx.x / GT;
    const int ci = ti * GT + lc, cj = tj * GT + lc;
    for (int64_t rb = rs0; rb < rs1; rb += GR) {
#pragma unroll
        for (int q = 0; q < GR / (256 / GT); ++q) {
            const int rr = lr + q * (256 / GT);
            const int64_t r = rb + rr;
            const bool in = r < rs1;
            zi[rr][lc] = (in && ci < m) ? w[r] * Z[r * m + ci] : 0.0;
            zj[rr][lc] = (in && cj < m) ? Z[r * m + cj] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < GR; ++rr) {
            double u[4], v[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) u[a] = zi[rr][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) v[b] = zj[rr][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] += u[a] * v[b];
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int i = ti * GT + ty + 16 * a;
        if (i >= m) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = tj * GT + tx + 16 * b;
            if (j < m) part[((size_t)blockIdx.y * m + i) * m + j] = acc[a][b];
        }
    }
}

// out[i][j] += sign * sum over splits of the partial of (min(i,j), max(i,j)): every
// such entry lies in an upper tile, and reading one entry for both halves keeps the
// result exactly symmetric (inside a diagonal tile (i,j) and (j,i) are rounded apart).
__global__ __launch_bounds__(BLOCK) void kGramReduce(int m, int splits, const double* __restrict__ part, double sign,
                                                     double* __restrict__ out) {
    const int64_t mm = (int64_t)m * m;
    for (int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x; t < mm; t += (int64_t)gridDim.x * BLOCK) {
        int i = (int)(t / m), j = (int)(t % m);
        if (i > j) { const int x = i; i = j; j = x; }
        double s = 0.0;
        for (int p = 0; p < splits; ++p) s += part[((size_t)p * m + i) * m + j];
        out[t] += sign * s;
    }
}

}  // namespace fis

// out += sign * Z^T diag(w) Z over R rows (Z row-major [R][m]).
static void gramAccumulate(int64_t R, int m, const double* Z, const double* w, double sign, double* out,
                           double* part, int splits, hipStream_t s) {
    if (R <= 0) return;
    const int nt = (m + fis::GT - 1) / fis::GT;
    const int pairs = nt * (nt + 1) / 2;
    const int64_t rps = (R + splits - 1) / splits;
    const int used = (int)((R + rps - 1) / rps);
    fis::kGram<<<dim3(pairs, used), 256, 0, s>>>(R, m, nt, Z, w, rps, part);
    CCD_HIP(hipGetLastError());
    fis::kGramReduce<<<gridFor((int64_t)m * m), BLOCK, 0, s>>>(m, used, part, sign, out);
    CCD_HIP(hipGetLastError());
}

template <typename Real>
void EngineT<Real>::fisherInformation(const double* beta, const int32_t* idx, int m, double* out) {
    CCD_HIP(hipSetDevice(D->device));
    const int J = D->J;
    for (int a = 0; a < m; ++a)
        if (idx[a] < 0 || idx[a] >= J) throw std::invalid_argument("column index out of range");
    if (m == 0) return;
    if (m > 16384) throw std::invalid_argument("at most 16384 columns in one information matrix");
    const int64_t K = D->K;
    const int model = D->model;
    const bool grouped = model == M_CLR || model == M_COX;
    setWeights(nullptr, false);
    computeXBeta(beta);
    remainingStats(false);
    const int64_t N = grouped ? Ncur : K;
    const int32_t* dGroup = model == M_CLR ? D->dStrat : (model == M_COX ? pidArr : nullptr);
    const int32_t* dGStart = model == M_CLR ? D->dStratStart : (model == M_COX ? pstart : nullptr);
    // host copy of each row's group, to find the groups a row chunk touches
    const std::vector<int32_t>* hGroup = model == M_CLR ? &D->strat : (model == M_COX ? &hPid : nullptr);

    const int64_t chunk = std::max<int64_t>(fis::GR, (int64_t)(fis::ZBYTES / sizeof(double)) / m);
    const int64_t rowsZ = std::min(K, chunk);
    const int nt = (m + fis::GT - 1) / fis::GT;
    const int pairs = nt * (nt + 1) / 2;
    // row splits: enough blocks to fill the device, at least 4 staging steps each
    const int splits = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(1, 2048 / pairs),
                                                                   (std::max(rowsZ, N) + 4 * fis::GR - 1) / (4 * fis::GR)));
    double b = 0.0;
    int32_t* dIdx = dalloc<int32_t>(m, b);
    double* acc = dalloc<double>((size_t)m * m, b);
    double* Z = dalloc<double>((size_t)rowsZ * m, b);
    double* wRow = dalloc<double>(K, b);
    double* part = dalloc<double>((size_t)splits * m * m, b);
    double* Cm = grouped ? dalloc<double>((size_t)N * m, b) : nullptr;
    double* wN = grouped ? dalloc<double>(N, b) : nullptr;
    h2d(dIdx, idx, m, s);
    CCD_HIP(hipMemsetAsync(acc, 0, (size_t)m * m * sizeof(double), s));
    if (grouped) CCD_HIP(hipMemsetAsync(Cm, 0, (size_t)N * m * sizeof(double), s));
    switch (model) {
    case M_LOGISTIC: fis::kRowWeight<Real, M_LOGISTIC><<<gridFor(K), BLOCK, 0, s>>>(K, ex, denom, dGroup, wRow); break;
    case M_POISSON: fis::kRowWeight<Real, M_POISSON><<<gridFor(K), BLOCK, 0, s>>>(K, ex, denom, dGroup, wRow); break;
    default: fis::kRowWeight<Real, M_CLR><<<gridFor(K), BLOCK, 0, s>>>(K, ex, denom, dGroup, wRow); break;
    }
    CCD_HIP(hipGetLastError());
    for (int64_t r0 = 0; r0 < K; r0 += rowsZ) {
        const int64_t r1 = std::min(K, r0 + rowsZ);
        CCD_HIP(hipMemsetAsync(Z, 0, (size_t)(r1 - r0) * m * sizeof(double), s));
        const int gx = (int)std::min<int64_t>(64, (r1 - r0 + BLOCK - 1) / BLOCK);
        fis::kScatter<Real><<<dim3(gx, m), BLOCK, 0, s>>>(m, dIdx, D->dFmt, D->dColPtr, D->dDensePtr, D->dRows,
                                                          D->dVals, D->dDense, r0, r1, Z);
        CCD_HIP(hipGetLastError());
        gramAccumulate(r1 - r0, m, Z, wRow + r0, 1.0, acc, part, splits, s);
        if (grouped) {
            const int64_t n0 = (*hGroup)[r0], n1 = (int64_t)(*hGroup)[r1 - 1] + 1;
            fis::kGroupSums<Real><<<gridFor((n1 - n0) * m), BLOCK, 0, s>>>(n0, n1, m, dGStart, r0, r1, ex, Z, Cm);
            CCD_HIP(hipGetLastError());
        }
    }
    if (grouped) {
        fis::kGroupWeight<<<gridFor(N), BLOCK, 0, s>>>(N, denom, wN);
        CCD_HIP(hipGetLastError());
        gramAccumulate(N, m, Cm, wN, -1.0, acc, part, splits, s);
    }
    CCD_HIP(hipMemcpyAsync(out, acc, (size_t)m * m * sizeof(double), hipMemcpyDeviceToHost, s));
    CCD_HIP(hipStreamSynchronize(s));
    dfree(dIdx); dfree(acc); dfree(Z); dfree(wRow); dfree(part); dfree(Cm); dfree(wN);
}

}  // namespace ccd
