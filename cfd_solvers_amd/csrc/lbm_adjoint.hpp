// lbm_adjoint.hpp -- host-side interface to the adjoint kernels in lbm_adjoint.hip.
#pragma once

#include "lbm_kernels.hpp"

namespace cfd_hip {

enum ObjectiveKind : int { kPressureDrop = 0, kDrag = 1 };

// The objective j(f) of one state, and the FP64 gradient sinks.
//  * kPressureDrop: j = c_in * sum_{x = x_in} rho + c_out * sum_{x = x_out} rho over fluid cells
//    (c_in = 1 / (3 |A| window), c_out = -1 / (3 |B| window)).
//  * kDrag: momentum exchange on a body of wall cells, j = c_drag * sum over fluid cells n next to
//    the body and populations d that arrive at n from a body cell (the halfway bounce-back of the
//    population that left n towards it) of -2 c_d[axis] f_d(n), c_drag = 1 / window.
struct AdjointArgs {
    int kind;
    int x_in, x_out;
    double c_in, c_out;
    const uint32_t* drag_cells;  // fluid cells with at least one link into the body
    const uint32_t* drag_links;  // per listed cell: bit d-1 set if population d arrives from the body
    uint32_t drag_count;
    int drag_axis;
    double c_drag;
    double* scratch;      // per-block partial sums: objective_scratch_doubles() of them
    int add_objective;    // adjoint_step: add dj/df_t for this step's state
    double* grad_gamma;   // per cell, accumulated
    double* grad_peak;    // per cell (inlet cells), accumulated
};

struct AdjointOps {
    virtual ~AdjointOps() = default;
    // lam <- (d step / d f)^T lam at the forward state fstate (+ dj/df if a.add_objective).
    virtual void step(const LaunchState& st, bool odd, bool bounded, const void* fstate, void* lam,
                      const AdjointArgs& a) const = 0;
    // lam <- dj(f)/df at parity odd (lam is overwritten).
    virtual void seed(const LaunchState& st, bool odd, bool bounded, void* lam, const AdjointArgs& a) const = 0;
    // *acc += weight * j(f), deterministic (fixed summation order).
    virtual void objective(const LaunchState& st, bool odd, bool bounded, const void* f, const AdjointArgs& a,
                           double weight, double* acc) const = 0;
    virtual void zero(const LaunchState& st, double* v, uint64_t n) const = 0;
    // *out = sum of v[0..n) in a fixed order: kSumBlocks per-block partials (in `partial`, which
    // holds kSumBlocks doubles), then one block over them.
    virtual void sum(const LaunchState& st, const double* v, uint64_t n, double* partial, double* out) const = 0;
};

constexpr uint32_t kSumBlocks = 1024;

const AdjointOps* adjoint_ops(int q, int storage);

// Doubles of AdjointArgs::scratch the objective reductions need on a lattice (drag_count: cells of
// the drag list, 0 if none).
inline uint64_t objective_scratch_doubles(const Geometry& g, uint32_t drag_count) {
    const uint64_t plane = (2ULL * g.ny * g.nz + 255U) / 256U, drag = (drag_count + 255ULL) / 256U;
    return plane > drag ? plane : drag;
}

}  // namespace cfd_hip
