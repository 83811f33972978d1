// Batched MFMA integrator (placeholder until the MFMA kernel lands).
#include "engine.hpp"

namespace cubnm {

bool batched_supported(int, int) { return false; }

hipError_t launch_integrate_batched(int, const IntegrateArgs&, const void*, hipStream_t, int*) {
    return hipErrorNotSupported;
}

}  // namespace cubnm
