// LODE spherical expansion: host orchestration (see kernels_lode.hip).
#include "calculator.hpp"

namespace rascaline_amd {

template <typename T>
mts_tensormap_t* Calculator::run_lode(const HostBatch&, const Gradients&, hipStream_t) {
    throw Error::invalid_parameter("lode_spherical_expansion is not available yet");
}

template mts_tensormap_t* Calculator::run_lode<float>(const HostBatch&, const Gradients&, hipStream_t);
template mts_tensormap_t* Calculator::run_lode<double>(const HostBatch&, const Gradients&, hipStream_t);

}  // namespace rascaline_amd
