// lbm_kernels.hpp -- host-side interface to the kernel templates in lbm_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "lbm_device.hpp"

namespace cfd_hip {

struct LaunchState {
    void* f;                 // population planes
    Geometry geo;
    const void* physics;     // Physics<float> or Physics<double> (host copy, passed by value)
    hipStream_t stream;
    uint32_t cells;
    uint32_t step_lds;       // dynamic LDS bytes per bounded step-kernel block (occupancy cap; 0 = none)
    uint32_t step_lds_periodic;  // the same for the periodic step
    uint32_t adjoint_lds;    // the same for the transposed step
};

// One instance per (lattice, storage); every launch is asynchronous on st.stream.
struct KernelOps {
    virtual ~KernelOps() = default;
    // One time step from src into dst (src == dst: in place).
    virtual void step(const LaunchState& st, bool odd, bool bounded, bool forced, const void* src,
                      void* dst) const = 0;
    virtual void initialize(const LaunchState& st, const void* init_args, const uint8_t* flags) const = 0;
    virtual void upload(const LaunchState& st, bool odd, const void* logical) const = 0;
    virtual void bounce_back(const LaunchState& st, bool odd, bool to_bounce_back, const uint8_t* flags) const = 0;
    virtual void freeze(const LaunchState& st, bool odd, const uint8_t* flags, const uint8_t* old_flags,
                        void* solid_rho) const = 0;
    virtual void segments(const LaunchState& st, const uint8_t* flags, uint32_t seg_x, uint32_t* out) const = 0;
    // Marks the segments of `segments` (spans already written) that hold a cell with gamma != 0.
    virtual void porous_segments(const LaunchState& st, const void* gamma, uint32_t seg_x,
                                 uint32_t* segments) const = 0;
    virtual void thaw(const LaunchState& st, bool odd, const uint8_t* old_flags, const uint8_t* flags,
                      const void* solid_rho) const = 0;
    virtual void macroscopic(const LaunchState& st, bool odd, bool bounded, void* out) const = 0;
    virtual void download(const LaunchState& st, bool odd, bool bounded, void* out) const = 0;
    // out: q*cells Real, slot (d, n) of the in-place array as a full population, order d*cells + n
    virtual void slots(const LaunchState& st, void* out) const = 0;
    virtual void diagnostics(const LaunchState& st, bool odd, bool bounded, double* partial, double* out) const = 0;
    // ids: q*stride uint32 scratch; on return it holds, per slot, the identifier d*cells + n of the
    // population stored there (0xFFFFFFFF if none); load: q*cells identifiers of the slots read.
    virtual void index_map(const LaunchState& st, bool odd, bool bounded, const uint8_t* flags,
                           const uint32_t* links, uint32_t* ids, int64_t* load) const = 0;
    // z-slab halo messages: elements of one message (crossing pairs x nx x ny), and the pack
    // (unpack = false) or unpack of the up and down messages after a step of parity odd.
    virtual size_t halo_elems(const LaunchState& st) const = 0;
    virtual void halo(const LaunchState& st, bool odd, bool unpack, void* up, void* down) const = 0;
    // out: per-cell link descriptor; vel: (q-1+9)/10 words per cell of wall-velocity indices
    virtual void links(const LaunchState& st, const uint8_t* flags, uint32_t* out, uint32_t* vel) const = 0;
};

const KernelOps* kernel_ops(int q, int storage);

size_t init_args_size(int storage);
void fill_init_args(int storage, void* out, int mode, double rho, double ux, double uy, double uz, double amplitude,
                    double noise, uint64_t seed, const void* tables, const void* fields);

}  // namespace cfd_hip
