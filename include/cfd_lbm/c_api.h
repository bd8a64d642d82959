/*
 * cfd_lbm/c_api.h -- C ABI of the MI355X (gfx950) esoteric-pull lattice-Boltzmann backend.
 *
 * This is the drop-in boundary for the reference's LBM hot path.  Each entry point replaces
 * a member of the reference's solver classes, cited as path:line in the reference tree:
 *   EsotericPullSolver<D>      include/cfd/solvers/lbm/esoteric_pull.hpp:104-278
 *   EsotericPullSyclSolver<D>  include/cfd/solvers/lbm/esoteric_pull_sycl.hpp:18-475
 * and follows the reference's C-ABI convention (include/cfd/python/c_api.h:1-15,
 * src/python/c_api.cpp:8-15):
 *   - every function returns int: 0 = ok, 1 = std::exception (message copied),
 *     2 = unknown exception;
 *   - the last two parameters are a caller-owned error buffer and its capacity; the
 *     message is NUL-terminated and truncated to fit;
 *   - arrays are caller-owned host pointers with the element counts stated below.
 *
 * Cells are numbered n = (z*ny + y)*nx + x (x fastest).  Logical population arrays are
 * direction-major: value (d, n) at index d*cells + n.  Each solver owns its device memory and
 * one HIP stream; calls on one solver are not thread-safe (the reference's host caller is
 * single-threaded too).  Work is asynchronous on that stream until a call that returns host
 * data or cfd_lbm_synchronize().
 */
#ifndef CFD_LBM_C_API_H
#define CFD_LBM_C_API_H

#include <stddef.h>
#include <stdint.h>

#define CFD_LBM_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cfd_lbm_solver cfd_lbm_solver;

/* Lattice ids are the velocity-set sizes q (descriptors.hpp:31-146). */
enum { CFD_LBM_D2Q9 = 9, CFD_LBM_D3Q19 = 19, CFD_LBM_D3Q27 = 27 };

/* Population storage.  Collision arithmetic is FP32 except for FP64 and FP64X.  FP32 follows the
 * reference's arithmetic bit for bit; the shifted storages hold the deviation f - w_d and collide
 * on deviations (rho = 1 + sum f'), which keeps the non-equilibrium part at FP32 precision at low
 * velocities.  FP64 keeps the reference's FP32 lattice-weight constants (as its double solver
 * does), whose sum exceeds 1 by 1.49e-8 (D3Q19) or 7.45e-9 (D2Q9, D3Q27); FP64X uses the weights'
 * exact values and has no reference counterpart (a reference for precision studies). */
enum {
    CFD_LBM_STORE_FP32 = 0,  /* 4 B / population (the reference's storage and arithmetic)  */
    CFD_LBM_STORE_FP16S = 1, /* 2 B / population: binary16 of (f - w_d) * 4096, RNE        */
    CFD_LBM_STORE_FP64 = 2,  /* 8 B / population: verification build of the same kernels   */
    CFD_LBM_STORE_FP32S = 3, /* 4 B / population: f - w_d in FP32                          */
    CFD_LBM_STORE_FP64X = 4  /* 8 B / population: FP64 with the exact lattice weights       */
};

/* uint8 cell flags (cfd_lbm_set_flags). */
enum {
    CFD_LBM_FLUID = 0x00,
    CFD_LBM_VELOCITY_INLET = 0x01, /* x = 0 plane: parabolic inlet, Zou-He density          */
    CFD_LBM_PRESSURE_OUTLET = 0x02,/* x = nx-1 plane: fixed density, Zou-He normal velocity */
    CFD_LBM_WALL = 0x80            /* halfway bounce-back; bits 0-2 select a wall velocity  */
};

CFD_LBM_API const char* cfd_lbm_version(void);

/* EsotericPullSolver(InPlaceLbmConfig) -- esoteric_pull.hpp:109-115,292-302.  Throws (returns 1)
 * if tau <= 0.5, nx or ny < 2, nz < 2 for 3-D lattices.  nz is forced to 1 for D2Q9.  The
 * lattice starts at rest with rho = 1 (initialize_uniform). */
CFD_LBM_API int cfd_lbm_create(int lattice, size_t nx, size_t ny, size_t nz, float tau, float ax, float ay,
                               float az, int storage, cfd_lbm_solver** out, char* error, size_t error_capacity);
/* ~EsotericPullSyclSolver -- esoteric_pull_sycl.hpp:35-45: waits, frees, swallows async errors. */
CFD_LBM_API int cfd_lbm_destroy(cfd_lbm_solver* solver, char* error, size_t error_capacity);

/* initialize_uniform -- esoteric_pull.hpp:117-119. */
CFD_LBM_API int cfd_lbm_initialize_uniform(cfd_lbm_solver* s, float rho, float ux, float uy, float uz, char* error,
                                           size_t error_capacity);
/* initialize_taylor_green -- esoteric_pull.hpp:121-137, plus seeded uniform noise of amplitude
 * noise_amplitude on each velocity component (SplitMix64 of (seed, cell, component)).
 * noise_amplitude = 0 reproduces the reference exactly. */
CFD_LBM_API int cfd_lbm_initialize_taylor_green(cfd_lbm_solver* s, float amplitude, float noise_amplitude,
                                                uint64_t seed, char* error, size_t error_capacity);
/* initialize(initializer) -- esoteric_pull.hpp:312-333 with caller-supplied fields (cells each). */
CFD_LBM_API int cfd_lbm_initialize_fields(cfd_lbm_solver* s, const float* rho, const float* ux, const float* uy,
                                          const float* uz, char* error, size_t error_capacity);
/* Replaces the populations with caller-supplied logical values (q*cells) at the current parity. */
CFD_LBM_API int cfd_lbm_upload_populations(cfd_lbm_solver* s, const float* logical, char* error,
                                           size_t error_capacity);
/* logical_populations -- esoteric_pull.hpp:257-267 (q*cells).  Wall cells report w_d * rho_wall. */
CFD_LBM_API int cfd_lbm_download_populations(cfd_lbm_solver* s, float* logical, char* error,
                                             size_t error_capacity);

/* raw_storage -- esoteric_pull.hpp:276-278: the in-place slot array (q*cells) in the reference's order
 * d*cells + n, at the current parity, each slot decoded to a full float population (FP16S/FP32S
 * shifted states add w_d back; FP64 rounds to float).  On a periodic FP32 lattice it equals the
 * reference's raw storage after the same steps.  Lattices with flags hold this build's in-place
 * halfway bounce-back slots. */
CFD_LBM_API int cfd_lbm_download_raw_slots(cfd_lbm_solver* s, float* slots, char* error, size_t error_capacity);

/* set_solid_mask / clear_solid_mask -- esoteric_pull.hpp:163-175, generalised to uint8 flags
 * (cells values; NULL clears).  Mask value 1 in the reference maps to CFD_LBM_WALL. */
CFD_LBM_API int cfd_lbm_set_flags(cfd_lbm_solver* s, const uint8_t* flags, char* error, size_t error_capacity);
/* The solver's flags (cells bytes; all CFD_LBM_FLUID when none are set) -- the flag form of
 * solid_mask() (esoteric_pull.hpp:177-178).  Host copy, no device transfer. */
CFD_LBM_API int cfd_lbm_get_flags(cfd_lbm_solver* s, uint8_t* flags, char* error, size_t error_capacity);
/* Wall velocity for flag CFD_LBM_WALL|index (index 1..7); src/lbm/d2q9.cpp:109-115,217-223. */
CFD_LBM_API int cfd_lbm_set_wall_velocity(cfd_lbm_solver* s, int index, float ux, float uy, float uz, char* error,
                                          size_t error_capacity);
/* Partial bounce-back porosity gamma in [0,1] per cell (cells values; NULL clears).  On a bounded
 * lattice the steps blend only in the 64-cell row segments that hold a nonzero gamma; the call
 * marks them (one pass over gamma), and so does every later change of the flags. */
CFD_LBM_API int cfd_lbm_set_porosity(cfd_lbm_solver* s, const float* gamma, char* error, size_t error_capacity);
/* set_porosity from DEVICE memory (cells floats on the solver's device), checked on the device: an
 * error, and the field in use unchanged, unless every value lies in [0, 1].  No reference
 * counterpart (the reference has no porosity field). */
CFD_LBM_API int cfd_lbm_set_porosity_device(cfd_lbm_solver* s, const float* gamma, char* error,
                                            size_t error_capacity);
/* Inlet peak velocity and outlet density for CFD_LBM_VELOCITY_INLET / CFD_LBM_PRESSURE_OUTLET cells. */
CFD_LBM_API int cfd_lbm_set_inlet_outlet(cfd_lbm_solver* s, float u_peak, float rho_out, char* error,
                                         size_t error_capacity);
/* set_local_body_acceleration / clear -- esoteric_pull.hpp:144-159 (cells each; NULL clears). */
CFD_LBM_API int cfd_lbm_set_local_acceleration(cfd_lbm_solver* s, const float* ax, const float* ay,
                                               const float* az, char* error, size_t error_capacity);

/* step(count) -- esoteric_pull.hpp:139-142 / esoteric_pull_sycl.hpp:87-211: one kernel launch per
 * time step, enqueued on the solver's stream. */
CFD_LBM_API int cfd_lbm_step(cfd_lbm_solver* s, uint64_t count, char* error, size_t error_capacity);
/* step(count) bracketed by HIP events on the solver's stream; *milliseconds = elapsed GPU time. */
CFD_LBM_API int cfd_lbm_step_timed(cfd_lbm_solver* s, uint64_t count, float* milliseconds, char* error,
                                   size_t error_capacity);
/* wait() -- esoteric_pull_sycl.hpp:213-216. */
CFD_LBM_API int cfd_lbm_synchronize(cfd_lbm_solver* s, char* error, size_t error_capacity);

/* download_macroscopic -- esoteric_pull_sycl.hpp:309-325 (exactly 4*cells floats cross PCIe).
 * Fluid u = m/rho + a/2; wall cells report (rho_wall, 0, 0, 0) -- esoteric_pull.hpp:180-209. */
CFD_LBM_API int cfd_lbm_download_macroscopic(cfd_lbm_solver* s, float* rho, float* ux, float* uy, float* uz,
                                             char* error, size_t error_capacity);
/* Same in double precision (exact for FP64 and FP64X storage). */
CFD_LBM_API int cfd_lbm_download_macroscopic_f64(cfd_lbm_solver* s, double* rho, double* ux, double* uy, double* uz,
                                                 char* error, size_t error_capacity);
/* refresh_device_macroscopic -- esoteric_pull_sycl.hpp:229-307: device pointers to rho, ux, uy, uz
 * (cells each, element type float, or double for FP64 and FP64X storage), valid until the next refresh. */
CFD_LBM_API int cfd_lbm_refresh_device_macroscopic(cfd_lbm_solver* s, void** rho, void** ux, void** uy, void** uz,
                                                   char* error, size_t error_capacity);
/* total_mass -- esoteric_pull_sycl.hpp:327-339, accumulated in FP64 (the reference uses FP32). */
CFD_LBM_API int cfd_lbm_total_mass(cfd_lbm_solver* s, double* mass, char* error, size_t error_capacity);
/* mass / kinetic_energy / max_speed -- esoteric_pull.hpp:211-255 in one device pass (FP64 sums;
 * wall cells excluded from energy and speed). */
CFD_LBM_API int cfd_lbm_diagnostics(cfd_lbm_solver* s, double* mass, double* kinetic_energy, double* max_speed,
                                    char* error, size_t error_capacity);

CFD_LBM_API int cfd_lbm_time_step(cfd_lbm_solver* s, uint64_t* time_step, char* error, size_t error_capacity);
CFD_LBM_API int cfd_lbm_cells(cfd_lbm_solver* s, size_t* cells, char* error, size_t error_capacity);
/* population_bytes -- esoteric_pull.hpp:272-274 (q*cells*bytes per population). */
CFD_LBM_API int cfd_lbm_population_bytes(cfd_lbm_solver* s, size_t* bytes, char* error, size_t error_capacity);
/* resident_bytes -- esoteric_pull_sycl.hpp:368-372 (all device allocations of this solver). */
CFD_LBM_API int cfd_lbm_resident_bytes(cfd_lbm_solver* s, size_t* bytes, char* error, size_t error_capacity);
/* device_name -- esoteric_pull_sycl.hpp:362-364. */
CFD_LBM_API int cfd_lbm_device_name(cfd_lbm_solver* s, char* name, size_t name_capacity, char* error,
                                    size_t error_capacity);
/* queue() -- esoteric_pull_sycl.hpp:358-359: the solver's hipStream_t. */
CFD_LBM_API int cfd_lbm_stream(cfd_lbm_solver* s, void** stream, char* error, size_t error_capacity);
/* transfer_stats / reset_transfer_stats -- esoteric_pull_sycl.hpp:373-374,
 * include/cfd/core/device_residency.hpp:11-42: {h2d, d2h, d2d bytes, synchronizations}. */
CFD_LBM_API int cfd_lbm_transfer_stats(cfd_lbm_solver* s, uint64_t* stats4, int reset, char* error,
                                       size_t error_capacity);

/* InPlaceLbmConfig::smagorinsky_les (esoteric_pull.hpp:28, closure at :72-97, used at :422-425):
 * local relaxation time from the non-equilibrium momentum flux.  FP32, FP64 or FP64X storage. */
CFD_LBM_API int cfd_lbm_set_smagorinsky(cfd_lbm_solver* s, int enabled, char* error, size_t error_capacity);

/* Objective kinds of the adjoint entry points. */
enum { CFD_LBM_OBJECTIVE_PRESSURE_DROP = 0, CFD_LBM_OBJECTIVE_DRAG = 1 };

/* Pressure-drop objective at the current state (FP64): j = mean rho/3 over the fluid cells of plane
 * x = x_in minus the same over plane x = x_out.  No reference counterpart (the reference has no
 * objectives; docs/RELEASE_0_10_9.md:39).  Any storage. */
CFD_LBM_API int cfd_lbm_objective_pressure_drop(cfd_lbm_solver* s, int x_in, int x_out, double* value, char* error,
                                                size_t error_capacity);
/* Registers the body of the drag objective: body[cells] != 0 marks wall cells that belong to it
 * (NULL clears).  The drag is the momentum exchange of halfway bounce-back on the links from fluid
 * cells into the body, along `axis` (0 x, 1 y, 2 z).  set_flags clears the registration.  No
 * reference counterpart (the reference computes no forces on bodies). */
CFD_LBM_API int cfd_lbm_set_drag_body(cfd_lbm_solver* s, const uint8_t* body, int axis, char* error,
                                      size_t error_capacity);
/* Drag on the registered body at the current state (FP64): sum over fluid cells n and populations d
 * arriving from a body cell of -2 c_d[axis] f_d(n). */
CFD_LBM_API int cfd_lbm_objective_drag(cfd_lbm_solver* s, double* value, char* error, size_t error_capacity);
/* The objective of the adjoint entry points without the adjoint (no reference counterpart): runs
 * `steps` steps in place from the current state and returns J, the mean of the objective over the
 * last `window` states (objective_kind and x_in, x_out as for cfd_lbm_adjoint_run_binomial), equal
 * bit for bit to the J those entry points return from the same state.  For callers that need J
 * alone (line searches, finite differences, evaluation under torch.no_grad).  The state is left
 * at time + steps. */
CFD_LBM_API int cfd_lbm_objective_run(cfd_lbm_solver* s, int objective_kind, uint64_t steps, uint64_t window, int x_in,
                                      int x_out, double* objective, char* error, size_t error_capacity);
/* Discrete adjoint (no reference counterpart: the reference has no adjoint).  Runs `steps` steps
 * from the current state, J = mean of the pressure drop over the last `window` states, then the
 * transposed steps back with a snapshot every `checkpoint_every` steps (two-level: each step is
 * recomputed once; memory ceil(steps/every) + every - 1 states); returns J, dJ/dgamma per cell
 * (cells doubles, may be NULL) and dJ/du_peak.  The state is left at time + steps.  Unforced
 * lattices with walls at rest, PBB, inlet and outlet; any storage (FP16: the adjoint of the FP32
 * collision at the stored states, lambda in FP32). */
CFD_LBM_API int cfd_lbm_adjoint_run(cfd_lbm_solver* s, uint64_t steps, uint64_t window, uint64_t checkpoint_every,
                                    int x_in, int x_out, double* objective, double* grad_gamma, double* grad_u_peak,
                                    char* error, size_t error_capacity);
/* The same with binomial checkpointing: `snapshots` free state buffers besides the start state
 * (0: as many as fit in device memory, 1/16 of it kept free), each step recomputed at most r times
 * where C(snapshots + r, r) >= steps, with the minimal number of forward steps.  objective_kind
 * CFD_LBM_OBJECTIVE_PRESSURE_DROP (planes x_in, x_out) or CFD_LBM_OBJECTIVE_DRAG (the registered
 * body; x_in, x_out unused).  forward_steps (may be NULL) returns the forward steps executed, the
 * first sweep included. */
CFD_LBM_API int cfd_lbm_adjoint_run_binomial(cfd_lbm_solver* s, int objective_kind, uint64_t steps, uint64_t window,
                                             uint64_t snapshots, int x_in, int x_out, double* objective,
                                             double* grad_gamma, double* grad_u_peak, uint64_t* forward_steps,
                                             char* error, size_t error_capacity);
/* cfd_lbm_adjoint_run_binomial with dJ/dgamma written into DEVICE memory (grad_gamma: cells
 * doubles on the solver's device, 8-byte aligned, required) instead of host memory, so that a
 * device-resident caller (torch tensors) makes no host round trip.  No reference counterpart. */
CFD_LBM_API int cfd_lbm_adjoint_run_binomial_device(cfd_lbm_solver* s, int objective_kind, uint64_t steps,
                                                    uint64_t window, uint64_t snapshots, int x_in, int x_out,
                                                    double* objective, double* grad_gamma, double* grad_u_peak,
                                                    uint64_t* forward_steps, char* error, size_t error_capacity);
/* The binomial plan without a solver (host only, no device call): forward steps of a run of
 * `steps` steps with `snapshots` buffers from the closed form r*l - C(s+r, r-1) + 1, the
 * repetition number r, and the same count obtained by walking the plan. */
CFD_LBM_API int cfd_lbm_checkpoint_plan(uint64_t steps, uint64_t snapshots, uint64_t* forward_steps,
                                        uint64_t* repetitions, uint64_t* walked_steps, char* error,
                                        size_t error_capacity);
/* GPU time of the last adjoint run's forward phase and reverse (recompute + transposed) phase. */
CFD_LBM_API int cfd_lbm_adjoint_timing(cfd_lbm_solver* s, float* forward_ms, float* reverse_ms, char* error,
                                       size_t error_capacity);

/* z-slab decomposition (include/cfd/solvers/lbm/distributed_pull.hpp:326-395 exchanges the
 * crossing populations of include/cfd/distributed/selective_halo.hpp:15-21 between bricks).  A
 * lattice cut into z-slabs runs one solver per slab, each created with its slab's nz and no flags
 * (periodic in z: the z wrap puts the slots a face cell shares with the next slab into the slab's
 * own opposite face layer).  After every step each slab packs an up message (to the slab above)
 * and a down message (to the slab below), the messages are exchanged (RCCL send/recv between
 * GPUs, or device copies on one GPU), and each slab unpacks the up message of the slab below and
 * the down message of the slab above.  The slabs together then hold bit for bit the state of the
 * whole periodic lattice.  A message is 5 (D3Q19) or 9 (D3Q27) planes of nx*ny populations in
 * the storage's width; buffers are device memory on the solver's device.  A slab that is its own
 * neighbour (one slab) receives its own messages, which changes nothing. */
CFD_LBM_API int cfd_lbm_halo_message_bytes(cfd_lbm_solver* s, size_t* bytes, char* error, size_t error_capacity);
CFD_LBM_API int cfd_lbm_halo_pack(cfd_lbm_solver* s, void* up, void* down, char* error, size_t error_capacity);
CFD_LBM_API int cfd_lbm_halo_unpack(cfd_lbm_solver* s, const void* from_below, const void* from_above, char* error,
                                    size_t error_capacity);

/* Checkpoint / restart.  The file starts with the CFDCHK3 layout of DistributedPullBlock::
 * save_checkpoint (include/cfd/solvers/lbm/distributed_pull.hpp:176-203) for one brick covering
 * the lattice: magic "CFDCHK3", version 1, endian marker, q, global/begin/extent (u64), time step,
 * tau and acceleration (f32), then q x cells float populations.  Periodic lattices store the
 * two-grid block's outgoing populations f(d, x + c_d), so the reference's load_checkpoint
 * (:205-245) restores them; lattices with flags store the logical populations.  A trailer
 * ("CFDESO1") records the parity of the time step, the body's convention and the storage, and
 * with raw != 0 the device state itself, which restores every storage (FP16S included) and the
 * wall densities bit for bit.  load_checkpoint reads either kind of file; the flags, wall
 * velocities and other settings must be set as when the file was written. */
CFD_LBM_API int cfd_lbm_save_checkpoint(cfd_lbm_solver* s, const char* path, int raw, char* error,
                                        size_t error_capacity);
CFD_LBM_API int cfd_lbm_load_checkpoint(cfd_lbm_solver* s, const char* path, char* error, size_t error_capacity);

/* voxelize_surface -- src/lbm/voxelize.cpp:5 (include/cfd/solvers/lbm/voxelize.hpp:9), run on the GPU
 * as the reference's device variant does (src/sycl/voxelize_sycl.cpp:13-68).  triangles: count x 9
 * doubles (a.x a.y a.z b.x b.y b.z c.x c.y c.z, the layout of cfd::io::Triangle3, host memory);
 * minimum, maximum: 3 doubles each (VoxelGrid3D); mask: nx*ny*nz bytes (host), cell (i,j,k) at
 * (k*ny + j)*nx + i, 1 where the +x ray from the cell centre crosses the surface an odd number
 * of times.  Double precision in the reference's CPU operation order: the mask equals
 * voxelize_surface's bit for bit.  Errors: "invalid voxelization inputs" as voxelize.cpp:5, and
 * non-finite vertex coordinates. */
CFD_LBM_API int cfd_lbm_voxelize_surface(const double* triangles, size_t count, size_t nx, size_t ny, size_t nz,
                                         const double* minimum, const double* maximum, uint8_t* mask, char* error,
                                         size_t error_capacity);
/* cfd_lbm_voxelize_surface without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  triangles (count x 9 doubles, 8-byte aligned) and mask
 * (nx*ny*nz bytes) are caller-owned DEVICE memory; stream is a hipStream_t (NULL: the null stream).
 * The working memory is a static device array of 17 MiB per GPU, held until the process ends; it
 * limits the grid to CFD_LBM_VOXELIZE_MAX_TILES tiles of 8 x 8 x-lines, ((ny+7)/8)*((nz+7)/8), and
 * the surface to CFD_LBM_VOXELIZE_MAX_PAIRS (triangle, tile) pairs; larger inputs are rejected with
 * an error.  The call returns when the mask is written.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other threads, other streams) race.  The mask equals
 * cfd_lbm_voxelize_surface's bit for bit; it records no timing. */
#define CFD_LBM_VOXELIZE_MAX_TILES 65536u
#define CFD_LBM_VOXELIZE_MAX_PAIRS 4194304u
CFD_LBM_API int cfd_lbm_voxelize_surface_no_alloc(const double* triangles, size_t count, size_t nx, size_t ny,
                                                  size_t nz, const double* minimum, const double* maximum,
                                                  uint8_t* mask, void* stream, char* error, size_t error_capacity);
/* GPU time of the calling thread's last voxelisation: binning of triangles to line tiles, the
 * line kernel, and the number of (triangle, tile) pairs binned. */
CFD_LBM_API int cfd_lbm_voxelize_timing(double* bin_ms, double* lines_ms, uint64_t* pairs, char* error,
                                        size_t error_capacity);
/* transform_surface -- src/lbm/voxelize.cpp:6 (voxelize.hpp:10): each vertex rotated about `axis`
 * through the origin by `angle` radians (Rodrigues; an axis shorter than 1e-30 leaves it), then
 * translated.  Host arithmetic in the reference's operation order; out may equal triangles. */
CFD_LBM_API int cfd_lbm_transform_surface(const double* triangles, size_t count, const double* translation,
                                          const double* axis, double angle, double* out, char* error,
                                          size_t error_capacity);
/* set_solid_mask(voxelize_surface(surface, grid)) as examples/real_lbm3d_model.cpp:146-151 does,
 * device-resident: the solver's nx, ny, nz with the grid's minimum and maximum; solid cells become
 * CFD_LBM_WALL (at rest).  keep_flags != 0 keeps the current flags of the other cells (inlets,
 * outlets, channel walls); 0 makes them fluid, as set_solid_mask does. */
CFD_LBM_API int cfd_lbm_set_solid_surface(cfd_lbm_solver* s, const double* triangles, size_t count,
                                          const double* minimum, const double* maximum, int keep_flags, char* error,
                                          size_t error_capacity);

/* ---------------------------------------------------------------------------------------------
 * Device-resident coupling (include/cfd/solvers/lbm/resident_multiphysics_sycl.hpp:54-408).
 * Each module lives on the flow solver's device and stream, on the flow's nx x ny x nz lattice,
 * periodic in x, y and z, and exchanges fields with the flow only through device pointers: the
 * macroscopic views of cfd_lbm_refresh_device_macroscopic and the acceleration planes of
 * cfd_lbm_device_local_acceleration.  The views must be float (FP32, FP16S or FP32S storage).
 * Destroy modules before their flow solver.  Field pointers passed in are device pointers of
 * cells floats each. */
typedef struct cfd_lbm_thermal cfd_lbm_thermal;
typedef struct cfd_lbm_surface cfd_lbm_surface;
typedef struct cfd_lbm_particles cfd_lbm_particles;
typedef struct cfd_lbm_flow_diagnostics cfd_lbm_flow_diagnostics;

/* device_local_acceleration(clear) -- esoteric_pull_sycl.hpp:349-358: the solver's per-cell
 * acceleration planes on the device (allocated zeroed on first use; clear != 0 zeroes them); every
 * later step applies them (Guo forcing).  Views are `real`-typed (float except FP64 storage). */
CFD_LBM_API int cfd_lbm_device_local_acceleration(cfd_lbm_solver* s, int clear, void** ax, void** ay, void** az,
                                                  char* error, size_t error_capacity);
/* Copies between a device view and host memory on the solver's stream (synchronous).  No reference
 * counterpart: SYCL's shared views are readable through queue.memcpy. */
CFD_LBM_API int cfd_lbm_device_download(cfd_lbm_solver* s, const void* device, void* host, size_t bytes, char* error,
                                        size_t error_capacity);
CFD_LBM_API int cfd_lbm_device_upload(cfd_lbm_solver* s, void* device, const void* host, size_t bytes, char* error,
                                      size_t error_capacity);

/* ResidentThermalD3Q7Sycl (resident_multiphysics_sycl.hpp:54-198): D3Q7 passive scalar, tau =
 * 0.5 + 4 diffusivity, equilibrium w T (1 + 4 c.u), two lattices. */
CFD_LBM_API int cfd_lbm_thermal_create(cfd_lbm_solver* flow, float diffusivity, cfd_lbm_thermal** out, char* error,
                                       size_t error_capacity);
CFD_LBM_API int cfd_lbm_thermal_destroy(cfd_lbm_thermal* t, char* error, size_t error_capacity);
/* initialize_uniform :82-91 */
CFD_LBM_API int cfd_lbm_thermal_initialize_uniform(cfd_lbm_thermal* t, float temperature, char* error,
                                                   size_t error_capacity);
/* Equilibrium at rest of a host temperature field (cells values); no reference counterpart. */
CFD_LBM_API int cfd_lbm_thermal_initialize_fields(cfd_lbm_thermal* t, const float* temperature, char* error,
                                                  size_t error_capacity);
/* step(flow view, count) :93-125 */
CFD_LBM_API int cfd_lbm_thermal_step(cfd_lbm_thermal* t, const float* ux, const float* uy, const float* uz,
                                     uint64_t count, char* error, size_t error_capacity);
/* refresh_device_scalar :127-136 (device pointer to cells floats, valid until the next refresh) */
CFD_LBM_API int cfd_lbm_thermal_refresh_device_scalar(cfd_lbm_thermal* t, void** scalar, char* error,
                                                      size_t error_capacity);
/* The temperature field on the host (cells floats). */
CFD_LBM_API int cfd_lbm_thermal_download_scalar(cfd_lbm_thermal* t, float* temperature, char* error,
                                                size_t error_capacity);
/* add_boussinesq_acceleration :140-155: a += -beta (T - t_ref) g */
CFD_LBM_API int cfd_lbm_thermal_add_boussinesq(cfd_lbm_thermal* t, float* ax, float* ay, float* az, float t_ref,
                                               float beta, double gx, double gy, double gz, char* error,
                                               size_t error_capacity);
/* total_scalar :157-167, accumulated in FP64 (the reference sums in float) */
CFD_LBM_API int cfd_lbm_thermal_total_scalar(cfd_lbm_thermal* t, double* total, char* error, size_t error_capacity);
/* resident_bytes :169-172, transfer_stats :173-174 ({h2d, d2h, d2d bytes, synchronisations}) */
CFD_LBM_API int cfd_lbm_thermal_resident_bytes(cfd_lbm_thermal* t, size_t* bytes, char* error, size_t error_capacity);
CFD_LBM_API int cfd_lbm_thermal_transfer_stats(cfd_lbm_thermal* t, uint64_t* stats4, int reset, char* error,
                                               size_t error_capacity);

/* ResidentFreeSurfaceSycl3D (resident_multiphysics_sycl.hpp:200-324): fill fraction in [0,1]. */
CFD_LBM_API int cfd_lbm_surface_create(cfd_lbm_solver* flow, cfd_lbm_surface** out, char* error,
                                       size_t error_capacity);
CFD_LBM_API int cfd_lbm_surface_destroy(cfd_lbm_surface* s, char* error, size_t error_capacity);
/* initialize_uniform :220-224, initialize_half_space_x :225-236 */
CFD_LBM_API int cfd_lbm_surface_initialize_uniform(cfd_lbm_surface* s, float fill, char* error,
                                                   size_t error_capacity);
CFD_LBM_API int cfd_lbm_surface_initialize_half_space_x(cfd_lbm_surface* s, float interface_x, char* error,
                                                        size_t error_capacity);
/* Host copies of the fill field (cells floats); no reference counterpart. */
CFD_LBM_API int cfd_lbm_surface_upload_fill(cfd_lbm_surface* s, const float* fill, char* error,
                                            size_t error_capacity);
CFD_LBM_API int cfd_lbm_surface_download_fill(cfd_lbm_surface* s, float* fill, char* error, size_t error_capacity);
/* advect(flow view, dt) :238-263: first-order upwind fluxes, clamped to [0,1] */
CFD_LBM_API int cfd_lbm_surface_advect(cfd_lbm_surface* s, const float* ux, const float* uy, const float* uz,
                                       float dt, char* error, size_t error_capacity);
/* fill_device :265-266 */
CFD_LBM_API int cfd_lbm_surface_fill_device(cfd_lbm_surface* s, void** fill, char* error, size_t error_capacity);
/* refresh_geometry :268-297: unit normals grad(phi)/|grad(phi)| and curvature -div(n) (device
 * pointers to cells floats each; any may be NULL) */
CFD_LBM_API int cfd_lbm_surface_refresh_geometry(cfd_lbm_surface* s, void** normal_x, void** normal_y,
                                                 void** normal_z, void** curvature, char* error,
                                                 size_t error_capacity);
/* add_surface_tension_acceleration :299-317: a += (sigma kappa / density) grad(phi) */
CFD_LBM_API int cfd_lbm_surface_add_surface_tension(cfd_lbm_surface* s, float* ax, float* ay, float* az,
                                                    float surface_tension, float density, char* error,
                                                    size_t error_capacity);
/* volume :319-323, accumulated in FP64 */
CFD_LBM_API int cfd_lbm_surface_volume(cfd_lbm_surface* s, double* volume, char* error, size_t error_capacity);
CFD_LBM_API int cfd_lbm_surface_transfer_stats(cfd_lbm_surface* s, uint64_t* stats4, int reset, char* error,
                                               size_t error_capacity);

/* ResidentParticlesSycl3D (resident_multiphysics_sycl.hpp:326-380).  Particles cross the ABI as 8
 * doubles each in the layout of ImmersedBoundaryParticle (include/cfd/solvers/lbm/particles.hpp:
 * 13-18): position x y z, velocity x y z, mass, response time; they are stored as float. */
CFD_LBM_API int cfd_lbm_particles_create(cfd_lbm_solver* flow, size_t count, cfd_lbm_particles** out, char* error,
                                         size_t error_capacity);
CFD_LBM_API int cfd_lbm_particles_destroy(cfd_lbm_particles* p, char* error, size_t error_capacity);
/* upload :337-344, download :368-376 */
CFD_LBM_API int cfd_lbm_particles_upload(cfd_lbm_particles* p, const double* particles, size_t count, char* error,
                                         size_t error_capacity);
CFD_LBM_API int cfd_lbm_particles_download(cfd_lbm_particles* p, double* particles, size_t count, char* error,
                                           size_t error_capacity);
/* advance(flow view, force view, dt, two_way, gravity) :346-366: trilinear fluid velocity, Stokes
 * drag, explicit Euler, periodic positions; two_way spreads -drag to the 8 cells (atomically). */
CFD_LBM_API int cfd_lbm_particles_advance(cfd_lbm_particles* p, const float* ux, const float* uy, const float* uz,
                                          float* ax, float* ay, float* az, float dt, int two_way, double gx,
                                          double gy, double gz, char* error, size_t error_capacity);
CFD_LBM_API int cfd_lbm_particles_transfer_stats(cfd_lbm_particles* p, uint64_t* stats4, int reset, char* error,
                                                 size_t error_capacity);

/* ResidentFlowDiagnosticsSycl3D (resident_multiphysics_sycl.hpp:382-408): Q-criterion
 * (|Omega|^2 - |S|^2)/2 by central differences; q receives a device pointer (cells floats). */
CFD_LBM_API int cfd_lbm_flow_diagnostics_create(cfd_lbm_solver* flow, cfd_lbm_flow_diagnostics** out, char* error,
                                                size_t error_capacity);
CFD_LBM_API int cfd_lbm_flow_diagnostics_destroy(cfd_lbm_flow_diagnostics* d, char* error, size_t error_capacity);
CFD_LBM_API int cfd_lbm_q_criterion(cfd_lbm_flow_diagnostics* d, const float* ux, const float* uy, const float* uz,
                                    void** q, char* error, size_t error_capacity);

/* Debug: the slots (d'*cells + n') read (load) and written (store) for every (d, n) by the
 * step kernel at the given parity, computed by the kernel's own index code (q*cells each).
 * Wall cells report -1.  Periodic lattices follow esoteric_pull.hpp:349-382. */
CFD_LBM_API int cfd_lbm_index_map(cfd_lbm_solver* s, int parity, int64_t* load_slots, int64_t* store_slots,
                                  char* error, size_t error_capacity);

#ifdef __cplusplus
}
#endif

#endif /* CFD_LBM_C_API_H */
