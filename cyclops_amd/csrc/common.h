// common.h -- shared definitions of the MI355X CCD engine (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace ccd {

// Reference enumerations (src/cyclops/Types.h, CompressedDataMatrix.h:36-38).
enum Model : int { M_POISSON = 2, M_LOGISTIC = 3, M_CLR = 4, M_COX = 8 };
enum Fmt : int { F_DENSE = 0, F_SPARSE = 1, F_INDICATOR = 2, F_INTERCEPT = 3 };

// Launch geometry: 256-thread workgroups = 4 wave64 wavefronts.
constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int ITEMS = 4;                 // consecutive entries per thread in ordered passes
constexpr int TILE = BLOCK * ITEMS;
constexpr int MAX_RED_BLOCKS = BLOCK;    // reduction kernels: one partial per thread of the last block
constexpr int MAX_GRAD_BLOCKS = 4 * BLOCK;   // column gradient passes of the plain path

// Per-block partial of a column reduction.
//   g, h : gradient / Hessian numerators
//   q    : segment-local quadratic term (grouped models)
//   m    : R2-weighted sum of entries before the block's first segment reset
//   t    : running segment sum at the block's end (carry out)
//   reset: the block contains a segment reset
struct Partial {
    double g, h, q, m, t;
    int reset;
    int pad;
};

struct HipError : std::runtime_error {
    explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define CCD_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t err_ = (call);                                                        \
        if (err_ != hipSuccess)                                                          \
            throw ::ccd::HipError(std::string(#call) + ": " + hipGetErrorString(err_) + \
                                  " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

}  // namespace ccd
