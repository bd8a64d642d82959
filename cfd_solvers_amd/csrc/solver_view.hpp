// solver_view.hpp -- what the device-resident coupling modules need to know about a flow solver:
// its device, stream and lattice extents, and the width of its device views (4: float fields).
#pragma once

#include "../../include/cfd_lbm/c_api.h"

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cfd_hip {

struct FlowInfo {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t nx = 0, ny = 0, nz = 0, cells = 0;
    size_t real = 4;  // bytes per value of the macroscopic and acceleration views
};

// Throws std::invalid_argument for a null handle.
FlowInfo flow_info(cfd_lbm_solver* s);

}  // namespace cfd_hip
