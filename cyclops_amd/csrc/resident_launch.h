// resident_launch.h -- launchers of the persistent findMode kernel. Each
// resident_<model>.hip includes this and instantiates one model, so the
// heavy kernel instantiations compile in parallel translation units.
#pragma once

#include "resident_kernels.h"

#include <string>

namespace ccd {

template <typename Real, int MODEL, int BP>
static void launchFindMode(const res::Params<Real>* teams, int nTeams, int T, hipStream_t s) {
    if (T > 1) {   // workgroups of a team spin on each other: all must be resident at once
        int dev = 0, cus = 0, perCU = 0;
        CCD_HIP(hipGetDevice(&dev));
        CCD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        CCD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, res::kFindMode<Real, MODEL, BP>, BLOCK, 0));
        if ((int64_t)nTeams * T > (int64_t)perCU * cus)
            throw std::invalid_argument("persistent launch: " + std::to_string(nTeams * T) +
                                        " workgroups cannot be co-resident (" + std::to_string(perCU) +
                                        " per compute unit)");
    }
    hipLaunchKernelGGL((res::kFindMode<Real, MODEL, BP>), dim3(nTeams * T), dim3(BLOCK), 0, s, teams, T);
}

template <typename Real, int MODEL>
void launchByBP(int BP, const res::Params<Real>* teams, int nTeams, int T, hipStream_t s) {
    switch (BP) {
    case 1: launchFindMode<Real, MODEL, 1>(teams, nTeams, T, s); break;
    case 4: launchFindMode<Real, MODEL, 4>(teams, nTeams, T, s); break;
    case 16: launchFindMode<Real, MODEL, 16>(teams, nTeams, T, s); break;
    case 64: launchFindMode<Real, MODEL, 64>(teams, nTeams, T, s); break;
    case 256: launchFindMode<Real, MODEL, 256>(teams, nTeams, T, s); break;
    default: throw std::invalid_argument("unsupported fit padding");
    }
}

// Workgroups of kFindMode<Real, MODEL, BP> that one compute unit holds at once.
template <typename Real, int MODEL>
int occupancyByBP(int BP) {
    int perCU = 0;
    switch (BP) {
    case 1: CCD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, res::kFindMode<Real, MODEL, 1>, BLOCK, 0)); break;
    case 4: CCD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, res::kFindMode<Real, MODEL, 4>, BLOCK, 0)); break;
    case 16: CCD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, res::kFindMode<Real, MODEL, 16>, BLOCK, 0)); break;
    case 64: CCD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, res::kFindMode<Real, MODEL, 64>, BLOCK, 0)); break;
    case 256: CCD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, res::kFindMode<Real, MODEL, 256>, BLOCK, 0)); break;
    default: throw std::invalid_argument("unsupported fit padding");
    }
    return perCU;
}

}  // namespace ccd

#define CCD_INSTANTIATE_LAUNCH(MODEL)                                                                   \
    namespace ccd {                                                                                     \
    template void launchByBP<float, MODEL>(int, const res::Params<float>*, int, int, hipStream_t);     \
    template void launchByBP<double, MODEL>(int, const res::Params<double>*, int, int, hipStream_t);   \
    template int occupancyByBP<float, MODEL>(int);                                                      \
    template int occupancyByBP<double, MODEL>(int);                                                     \
    }
