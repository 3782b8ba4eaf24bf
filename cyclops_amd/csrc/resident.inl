// resident.inl -- device-resident and batched fits (persistent kernels).
#pragma once

namespace ccd {

template <typename Real>
void EngineT<Real>::releaseResident() {}

template <typename Real>
void EngineT<Real>::fitResident(const ccd_fit_args&, double*, ccd_fit_result&) {
    throw std::logic_error("resident mode not built");
}

template <typename Real>
void EngineT<Real>::fitBatch(const ccd_fit_args&, const ccd_batch_args&, double*, ccd_fit_result*) {
    throw std::logic_error("batched mode not built");
}

}  // namespace ccd
