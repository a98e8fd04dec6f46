// Batched chemistry production stage, BDF integrator (SourceSystemCVODE semantics).
#include <hip/hip_runtime.h>

#include "chem_device.h"
#include "kernels.h"

namespace emx {

int launch_bdf_stage(const DevMech&, const DevOpts&, const double*, const double*, const double*,
                     const double*, double, int, const int*, unsigned int*, double*, int*, int*,
                     int*, double*, int, int, hipStream_t) {
    return (int)hipErrorNotSupported;
}

}  // namespace emx
