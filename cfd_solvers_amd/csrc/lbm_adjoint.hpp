// lbm_adjoint.hpp -- host-side interface to the adjoint kernels in lbm_adjoint.hip.
#pragma once

#include "lbm_kernels.hpp"

namespace cfd_hip {

// Pressure-drop objective j(f) = c_in * sum_{x = x_in} rho + c_out * sum_{x = x_out} rho over fluid
// cells (c_in = 1 / (3 |A| window), c_out = -1 / (3 |B| window)), and the FP64 gradient sinks.
struct AdjointArgs {
    int x_in, x_out;
    double c_in, c_out;
    int add_objective;   // adjoint_step: add dj/df_t for this step's state
    double* grad_gamma;  // per cell, accumulated
    double* grad_peak;   // per cell (inlet cells), accumulated
};

struct AdjointOps {
    virtual ~AdjointOps() = default;
    // lam <- (d step / d f)^T lam at the forward state fstate (+ dj/df if a.add_objective).
    virtual void step(const LaunchState& st, bool odd, bool bounded, const void* fstate, void* lam,
                      const AdjointArgs& a) const = 0;
    // lam <- dj(f)/df at parity odd.
    virtual void seed(const LaunchState& st, bool odd, bool bounded, void* lam, const AdjointArgs& a) const = 0;
    // *acc += weight * j(f) (one block, fixed order).
    virtual void objective(const LaunchState& st, bool odd, bool bounded, const void* f, const AdjointArgs& a,
                           double weight, double* acc) const = 0;
    virtual void zero(const LaunchState& st, double* v, uint64_t n) const = 0;
};

const AdjointOps* adjoint_ops(int q, int storage);

}  // namespace cfd_hip
