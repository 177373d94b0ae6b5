/* The explicit compressible-Euler residual loop on one MI355X (gfx950), behind a plain C ABI.
 *
 * A drop-in for the reference's device stepper: make_device_stepper<P>
 * (physics/include/autan/physics/euler_device.hpp:49-54, physics/src/euler_device.cu:8-23) and the
 * DeviceStepper<P> methods step / evaluate / residual / set_second_order / upload / download
 * (physics/include/autan/physics/euler_device.hpp:22-41), implemented by CudaStepper<P> in
 * physics/src/device_stepper.cuh:200-529. Where the reference's stepper reads the host Euler<P>
 * (data(), stencil(), gradient_weights(), scheme(), planar(); device_stepper.cuh:209-311), this
 * ABI takes the same arrays: mesh topology, packed FP64 geometry, the gradient stencil with its
 * weights, boundary codes and data, and the volume source. Conserved states and residuals cross
 * the boundary as FP64 structure-of-arrays, component k of cell c at k * n_cells + c.
 *
 * Conventions (the reference's C ABI precedent, physics/mms/capi.cpp:126-347): every call
 * returning int returns 0 on success and otherwise puts a message in err[0 .. err_len) that starts
 * with the reference's error code where one exists (E1101 for a device failure). The value tells
 * the two exception kinds of the reference's stepper apart: 1 when the input is refused
 * (std::invalid_argument), 2 for a device or runtime failure (std::runtime_error); mxeuler.hpp
 * rethrows them as those exceptions. Handles are opaque. Calls are synchronous: each returns once
 * its results are on the host. One handle is driven by one host thread.
 */
#ifndef MXEULER_H
#define MXEULER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision policies (codegen/include/autan/kernels/runtime.hpp:73-99). */
enum { MXEULER_FP64 = 0, MXEULER_FP32_ACC64 = 1 };
/* Numerical fluxes (physics/include/autan/physics/euler.hpp:27). */
enum { MXEULER_ROE = 0, MXEULER_HLLC = 1, MXEULER_AUSM_UP = 2 };

/* Scheme and Gas (physics/include/autan/physics/euler.hpp:39-71): p_ref and the limiter scales
 * (K / length)^3 scale_k^2 are derived from these as physics/src/euler.cpp:238-245 derives them. */
typedef struct {
  int32_t flux;         /* MXEULER_ROE, MXEULER_HLLC or MXEULER_AUSM_UP */
  int32_t second_order; /* linear reconstruction with least-squares gradients */
  int32_t limiter;      /* Venkatakrishnan limiter */
  int32_t planar;       /* the loops without warped-face terms; refused unless all_planar */
  double gamma;
  double gas_constant;
  double mach_ref;  /* AUSM+-up cut-off Mach number, > 0 */
  double venkat_k;
  double length;
  double scale[5];     /* W units: rho, u, v, w, p */
  double reference[5]; /* U_ref (the freestream), reference[0] > 0 */
} mxeuler_scheme;

/* Mesh topology in the reference's canonical order (core/include/autan/core/mesh.hpp:277-282):
 * interior faces first with owner < neighbour, each cell's faces ascending in cell_faces. */
typedef struct {
  int32_t n_cells;
  int32_t n_faces;
  int32_t n_interior;
  const int32_t* owner;             /* n_faces */
  const int32_t* neighbour;         /* n_interior */
  const int64_t* cell_face_offsets; /* n_cells + 1 */
  const int32_t* cell_faces;        /* cell_face_offsets[n_cells] */
} mxeuler_topology;

/* Packed geometry (kernels::PackedGeometry, codegen/include/autan/kernels/geometry.hpp): FP64,
 * component-major. r_left / r_right are x_f - x_c of the owner / neighbour in FP64 (the operands
 * of the kernels' rel(x_f, x_c)); on a periodic face r_right uses the neighbour's image. */
typedef struct {
  const double* face_normal; /* 3 x n_faces, unit, owner -> neighbour */
  const double* face_area;   /* n_faces */
  const double* r_left;      /* 3 x n_faces */
  const double* r_right;     /* 3 x n_faces (interior faces read) */
  const double* face_moment; /* 9 x n_faces; may be NULL when every face is proven planar */
  const double* cell_volume; /* n_cells */
  int32_t all_planar;        /* every face moment is an exact zero */
} mxeuler_geometry;

/* The gradient stencil (kernels::CellStencil): entries of cell c at offsets[c] .. offsets[c+1],
 * a cell index or -(f + 1) for boundary face f; weights 3 x n_entries (Euler::gradient_weights). */
typedef struct {
  const int64_t* offsets;
  const int32_t* entries;
  const double* weights;
} mxeuler_stencil;

typedef struct mxeuler_stepper mxeuler_stepper;

/* make_device_stepper: uploads everything; bc_kind holds one code (BoundaryKind 0..5, whole
 * numbers as FP64) and bc five values (5 x n_boundary) per boundary face; source is 5 x n_cells
 * per unit volume, or NULL for zero. stream is a hipStream_t to run on, or NULL for the
 * library's own stream. second_order and planar are those of the scheme. */
int mxeuler_create(int32_t policy, const mxeuler_scheme* scheme, const mxeuler_topology* topology,
                   const mxeuler_geometry* geometry, const mxeuler_stencil* stencil, const double* bc_kind,
                   const double* bc, const double* source, void* stream, mxeuler_stepper** out, char* err,
                   size_t err_len);
void mxeuler_destroy(mxeuler_stepper* s);

/* DeviceStepper::upload / download: the FP64 conserved state u (5 x n_cells). */
int mxeuler_upload(mxeuler_stepper* s, const double* u, char* err, size_t err_len);
int mxeuler_download(mxeuler_stepper* s, double* u, char* err, size_t err_len);

/* DeviceStepper::residual: R(u) into res (both 5 x n_cells, FP64); the stepper's state is kept. */
int mxeuler_residual(mxeuler_stepper* s, const double* u, double* res, char* err, size_t err_len);

/* Euler::residual_and_jacobian (physics/src/euler.cpp:983-991; kernels/euler.py:434-440, 854-860,
 * 1051-1057): R(u) into res as mxeuler_residual, and the approximate Jacobian's parts at u, with the
 * reconstruction's gradients and limiter held fixed, in FP64:
 * dwdu (25 x n_cells) dW/dq of each cell, q the stored state, entry (i, j) at (5 i + j) n_cells + c;
 * face_jac (50 x n_interior) dF/dW_L then dF/dW_R of each interior face (owner to neighbour,
 * area-weighted), entry (k, i, j) at (25 k + 5 i + j) n_interior + f; bnd_jac (25 x n_boundary)
 * dF/dW_L of each boundary face (through its boundary state; slip and wall faces the pressure
 * flux's, as the reference's Jacobian loops take them). The residual's Jacobian is then, per cell, sum over its faces of
 * +-J_side D_cell (Euler::assemble, physics/src/euler.cpp:993-1053). Planar meshes, one case. */
int mxeuler_residual_and_jacobian(mxeuler_stepper* s, const double* u, double* res, double* dwdu, double* face_jac,
                                  double* bnd_jac, char* err, size_t err_len);

/* Device-resident upload, download and residual (the hipGraph path's variant of the three calls
 * above): u and res are device memory on the stepper's GPU (5 x n_cells FP64). The work runs on
 * the stepper's stream after everything `stream` holds at the call, and `stream` receives an
 * event that orders its later work after the copies (no host wait; `stream` may be the stepper's
 * own, or NULL for the null stream). */
int mxeuler_upload_device(mxeuler_stepper* s, const double* u, void* stream, char* err, size_t err_len);
int mxeuler_download_device(mxeuler_stepper* s, double* u, void* stream, char* err, size_t err_len);
int mxeuler_residual_device(mxeuler_stepper* s, const double* u, double* res, void* stream, char* err,
                            size_t err_len);

/* DeviceStepper::step: one SSP-RK3 step of the uploaded state. residuals receives the RMS norms
 * of R(u) at the start of the step (sqrt(sum_c R_kc^2 / N), summed exactly); time_step the step
 * used: dt when has_dt, else min_c CFL V_c / lambda_c (with local, each cell its own).
 * invalid_cell receives the least cell whose new state has a non-finite or non-positive density
 * or pressure, else the least with a non-finite residual, else -1. */
int mxeuler_step(mxeuler_stepper* s, double cfl, int32_t has_dt, double dt, int32_t local, double residuals[5],
                 double* time_step, int32_t* invalid_cell, char* err, size_t err_len);

/* DeviceStepper::evaluate: R(u) of the uploaded state, its norms and validity. */
int mxeuler_evaluate(mxeuler_stepper* s, double residuals[5], int32_t* invalid_cell, char* err, size_t err_len);

/* DeviceStepper::set_second_order: gradients and limiter zero without reconstruction, limiter 1
 * when second order is unlimited. */
int mxeuler_set_second_order(mxeuler_stepper* s, int32_t on, char* err, size_t err_len);

/* Euler::set_boundary_data and Euler::set_source (physics/include/autan/physics/euler.hpp:127,
 * 217): new boundary data (5 x n_boundary) and volume source (5 x n_cells, or NULL for zero). */
int mxeuler_set_boundary_data(mxeuler_stepper* s, const double* bc, char* err, size_t err_len);
int mxeuler_set_source(mxeuler_stepper* s, const double* source, char* err, size_t err_len);

/* Euler::spectral_radius after a residual (n_cells, widened to FP64). */
int mxeuler_spectral_radius(mxeuler_stepper* s, double* lambda, char* err, size_t err_len);

/* The arrays of the last residual, widened to FP64, for tests: prim (5 x n_cells, p gauge),
 * grad (15 x n_cells, row (3 i + j)), phi (5 x n_cells), prim_b (5 x n_boundary). Any may be NULL. */
int mxeuler_arrays(mxeuler_stepper* s, double* prim, double* grad, double* phi, double* prim_b, char* err,
                   size_t err_len);

/* Timing of the device work of `steps` steps (no host transfers between them): the average
 * milliseconds per step and per launch of each kernel, measured with HIP events on the stepper's
 * stream. names receives a ';'-separated list of the kernel names in the order of ms_per_kernel
 * (up to max_kernels). Returns the number of kernels through *n_kernels. */
int mxeuler_profile(mxeuler_stepper* s, int32_t steps, double cfl, int32_t local, double* ms_per_step,
                    double* ms_per_kernel, int32_t* launches_per_step, int32_t max_kernels, int32_t* n_kernels,
                    char* names, size_t names_len, char* err, size_t err_len);

/* Device math probe for the tests: sqrt, division, exp and log of x (and x / y) on the device in
 * FP64 (policy 0) or FP32 (policy 1), n values each, widened to FP64. */
int mxeuler_probe_math(int32_t policy, const double* x, const double* y, int64_t n, double* out_sqrt,
                       double* out_div, double* out_exp, double* out_log, char* err, size_t err_len);

/* HIP STREAM (bench/stream/stream.cu's measure of peak bandwidth): n FP64 values per array, the
 * best of `repeats` runs of copy, scale, add and triad; gbps and ms receive 4 values each. */
int mxeuler_stream(int64_t n, int32_t repeats, double* gbps, double* ms, char* err, size_t err_len);

/* Euler::boundary_pressure (physics/src/euler.cpp:1187-1223) for the current state: the residual of
 * u is evaluated first, then for each listed boundary face (global face index) the pressure force
 * p_f S + phi_p M^T grad p in FP64, force 3 x n SoA. */
int mxeuler_face_force(mxeuler_stepper* s, const int32_t* faces, int64_t n, double* force, char* err,
                       size_t err_len);
/* mxeuler_face_force without allocating: no counterpart in the reference. For callers that cannot
 * allocate device memory (tight monitor loops, code that forbids hipMalloc). The face list and the
 * forces are staged in a static device array that each GPU holds until the process ends, at most
 * MXEULER_STATIC_FORCE_FACES faces per call (more is refused, status 1). Every call on one GPU shares
 * that array: calls that overlap on one GPU (two threads, or steppers on different streams) race.
 * Same results as mxeuler_face_force. */
#define MXEULER_STATIC_FORCE_FACES 131072
int mxeuler_face_force_noalloc(mxeuler_stepper* s, const int32_t* faces, int64_t n, double* force, char* err,
                               size_t err_len);
/* EulerPhysics::loads (physics/src/euler_physics.cpp:527-562; wall_force :930-948): exact sums over
 * the listed boundary faces of the face force minus p_inf S_f, and of its moment about `center`.
 * area_vector and face_centroid (3 x n SoA, FP64) are the caller's geometry of the listed faces. */
int mxeuler_loads(mxeuler_stepper* s, const int32_t* faces, int64_t n, const double* area_vector,
                  const double* face_centroid, const double center[3], double p_inf, double force[3],
                  double moment[3], char* err, size_t err_len);
/* Chooses between kernel variants that give the same bits (no counterpart in the reference: the
 * autotuner's handle). Keys: "face_stage" (0/1, fp32+acc64): the fused face kernel stages its
 * block's cells in LDS; "graph" (0/1): steps without a fixed time step replay a recorded HIP graph
 * of their launches; "grad_occupancy" (0, 7 or 8) and "face_occupancy" (0, 5 or 8): the gradient and
 * face kernels built for at least that many waves per SIMD (0: the compiler's choice; 5 for
 * non-planar meshes only); "face_block" (256, 512 or 1024): the most faces a face block holds (the
 * blocks are rebuilt); "grad_nontemporal" (0/1, default 1): the gradient kernel loads its stencil
 * stream non-temporally. An unknown key or a variant this stepper cannot run is an error. */
int mxeuler_tune(mxeuler_stepper* s, const char* key, int32_t value, char* err, size_t err_len);
/* EulerPhysics::check_conservation (physics/src/euler_physics.cpp:1750-1784; docs/design/euler.md,
 * Conservation check) on the current state: per equation k, exact sums over the cells and faces,
 * sums[k] E = sum_c R^F_c - sum_b F_b, sums[5 + k] D = sum |F| over the incidences, sums[10 + k] the
 * boundary fluxes, sums[15 + k] sum_c R^F_c, with R^F the residual without its source. */
int mxeuler_conservation(mxeuler_stepper* s, double sums[20], char* err, size_t err_len);
/* EulerPhysics::residual_scales (physics/src/euler_physics.cpp:440-477): Phi_k = f_k sqrt(sum_i
 * A_i^2 / N) with the freestream's density, speed and pressure. */
int mxeuler_residual_scales(mxeuler_stepper* s, double density, double speed, double pressure, double scales[5],
                            char* err, size_t err_len);

/* ---- Steady runs (EulerPhysics::device_step and the runtime's steady loop) -------------------
 * A steady run on local time steps: an optional first-order start (physics/include/autan/physics/
 * first_order_start.hpp:22-137; physics/src/euler_physics.cpp:1495-1571) and the convergence rule
 * of case-schema.md#time (runtime/src/run.cpp:2262-2389; runtime/include/autan/runtime/
 * round_off.hpp:14-160). The stepper is reached through callbacks, so the same run drives the device
 * stepper (mxeuler_steady) or any other with the same three calls. */
enum { MXSTEADY_CONVERGED = 0, MXSTEADY_BUDGET = 1, MXSTEADY_DIVERGED = 2 };
enum { MXSTEADY_END_NONE = 0, MXSTEADY_END_DROP = 1, MXSTEADY_END_STALL = 2, MXSTEADY_END_CAP = 3,
       MXSTEADY_END_LEVEL = 4 };

typedef struct mxsteady_ops {
  void* ctx;
  /* DeviceStepper::step(cfl, nullopt, local = true): the norms of R(u) before the update. */
  int (*step)(void* ctx, double cfl, double residuals[5], int32_t* invalid_cell);
  /* DeviceStepper::evaluate(): the norms of R(u) and the first invalid cell (-1: none). */
  int (*evaluate)(void* ctx, double residuals[5], int32_t* invalid_cell);
  int (*set_second_order)(void* ctx, int32_t on);
} mxsteady_ops;

typedef struct mxsteady_options {
  int64_t max_iterations;  /* >= 1 (case default 10000) */
  double residual_orders;  /* > 0 (case default 8) */
  double cfl;              /* the device path's steady CFL, 0.8 (euler_physics.cpp:94) */
  int32_t level_rule;      /* 1 for fp32+acc64: level-crossing start, round-off plateau convergence */
  int32_t start_order;     /* 1: first-order start, 2: second order from the first step */
  const double* scales;    /* Phi_k (mxeuler_residual_scales) for the fp32+acc64 floors, or NULL */
  const double* reference; /* 5 residuals to drop from (a warm start's source), or NULL: iteration 1 */
} mxsteady_options;

typedef struct mxsteady_result {
  int32_t status;            /* MXSTEADY_* */
  int32_t first_order_end;   /* MXSTEADY_END_* */
  int32_t first_order_steps; /* provisional iterations */
  int32_t diverged_cell;     /* -1: none */
  int32_t plateau_mask;      /* bit e: equation e converged on its round-off plateau */
  int32_t pad;
  int64_t iterations;        /* the last completed iteration */
  double reference[5];       /* the residuals the run dropped from */
  double residuals[5];       /* the last reported residuals */
} mxsteady_result;

/* history: 5 reported residuals per iteration (or NULL); provisional: 1 per iteration (or NULL). */
int mxsteady_run(const mxsteady_ops* ops, const mxsteady_options* options, mxsteady_result* result,
                 double* history, int32_t* provisional, char* err, size_t err_len);
/* mxsteady_run over the device stepper's state (upload it first). */
int mxeuler_steady(mxeuler_stepper* s, const mxsteady_options* options, mxsteady_result* result, double* history,
                   int32_t* provisional, char* err, size_t err_len);

/* ---- The steady implicit path (EulerPhysics::steady_step, physics/src/euler_physics.cpp:1056-1250,
 * with linear_solver = "gauss-seidel") ------------------------------------------------------------
 * Pseudo-transient continuation (lambda / CFL I + dR/dU) du = -R on the approximate Jacobian, the CFL
 * number following the residual (switched evolution relaxation from CFL 1000, a penalty halved after
 * a shortened, rejected or rising step, 1 <= CFL <= 1e6), each system solved by GCR (at most 10
 * directions, to 0.1; a solve leaving more than 0.9 of its residual rejects the step) preconditioned
 * by 4 symmetric multicolour block Gauss-Seidel sweeps, and the largest fraction of du (at most 1)
 * that keeps every cell's density and pressure above half their values. */

/* R(u) of the stepper's state at the order set, its norms (record_norms) and first invalid cell (as
 * mxeuler_evaluate), the spectral radius and the Jacobian's parts (Euler::residual_and_jacobian,
 * physics/src/euler.cpp:983-991, warped faces' terms included), kept on the device for the solve. */
int mxeuler_implicit_evaluate(mxeuler_stepper* s, double norms[5], int32_t* invalid_cell, char* err, size_t err_len);
/* One solve and update from the last mxeuler_implicit_evaluate: report = {GCR directions taken,
 * residual drop of the solve, alpha (0: no update), failed (0/1), singular diagonal block (0/1)}. */
int mxeuler_implicit_solve(mxeuler_stepper* s, double cfl, double report[5], char* err, size_t err_len);
/* The colour of every cell in the sweeps (BlockMatrix's greedy distance-1 colouring in cell order). */
int mxeuler_implicit_colours(mxeuler_stepper* s, int32_t* colours, char* err, size_t err_len);

typedef struct mximplicit_ops {
  void* ctx;
  /* R(u) at the order set: norms and first invalid cell (-1: none); the history's second order. */
  int (*evaluate)(void* ctx, double residuals[5], int32_t* invalid_cell);
  int (*set_second_order)(void* ctx, int32_t on);
  /* R(u) and the Jacobian at the order set: norms of R(u), first invalid cell (-1: none). */
  int (*jacobian)(void* ctx, double residuals[5], int32_t* invalid_cell);
  /* One solve and update at this CFL number; report as mxeuler_implicit_solve's. */
  int (*solve)(void* ctx, double cfl, double report[5]);
  /* The cell whose density or pressure the full step lowers the most (optional, may be NULL). */
  int (*worst_cell)(void* ctx, int32_t* cell);
} mximplicit_ops;

/* The steady implicit run over any stepper's calls: the runtime's loop and convergence rule
 * (runtime/src/run.cpp:2262-2389) with the first-order start, options as mxsteady_run's (cfl
 * unused); stats receives 4 values per iteration (CFL, GCR directions, linear drop, alpha) or is
 * NULL. */
int mximplicit_run(const mximplicit_ops* ops, const mxsteady_options* options, mxsteady_result* result,
                   double* history, int32_t* provisional, double* stats, char* err, size_t err_len);
/* mximplicit_run over the device stepper's state (upload it first; one case). */
int mxeuler_implicit_steady(mxeuler_stepper* s, const mxsteady_options* options, mxsteady_result* result,
                            double* history, int32_t* provisional, double* stats, char* err, size_t err_len);

/* ---- Batches: several meshes (cases) stepped together in one index space ------------------------
 * No counterpart in the reference (a parameter sweep on one GPU). The topology, geometry, stencil
 * and boundary arrays are the union of the cases': case k owns cells [case_first[k], case_first[k]
 * + case_cells[k]); the rest of its range up to case_first[k + 1] are filler cells without faces
 * or stencil entries; every case_first a multiple of 256, case_first[n_cases] = n_cells. Faces
 * never join two cases. schemes[k] is case k's scheme: flux, order, limiter, planarity and gas
 * agree across cases. The state arrays (upload, download, residual) cover the whole index space. */
int mxeuler_create_batch(int32_t policy, int32_t n_cases, const mxeuler_scheme* schemes, const int64_t* case_first,
                         const int64_t* case_cells, const mxeuler_topology* topology, const mxeuler_geometry* geometry,
                         const mxeuler_stencil* stencil, const double* bc_kind, const double* bc, const double* source,
                         void* stream, mxeuler_stepper** out, char* err, size_t err_len);
/* The number of cases of a stepper (1 for mxeuler_create's). */
int mxeuler_cases(const mxeuler_stepper* s);
/* One SSP-RK3 step of every active case: residuals 5 x n_cases (case-major), the time step and the
 * first invalid cell (relative to the case's first cell, -1: none) per case. */
int mxeuler_batch_step(mxeuler_stepper* s, double cfl, int32_t local, double* residuals, double* time_steps,
                       int32_t* invalid_cells, char* err, size_t err_len);
/* Per case: bit 0 active (an inactive case is not stepped and keeps its state), bit 1 second order
 * (a case turned to first order has its gradients cleared). */
int mxeuler_batch_flags(mxeuler_stepper* s, const int32_t* flags, char* err, size_t err_len);
/* Every case's steady run at once (mxeuler_steady's rules per case): case_scales (5 per case, the
 * fp32+acc64 floors' Phi_k; NULL: none) and case_references (5 per case, the residuals to drop
 * from; NULL: each case's iteration 1) replace options->scales and options->reference; results per
 * case; history (5 per iteration, case k from history + 5 k max_iterations) may be NULL. */
int mxeuler_batch_steady(mxeuler_stepper* s, const mxsteady_options* options, const double* case_scales,
                         const double* case_references, mxsteady_result* results, double* history, char* err,
                         size_t err_len);
/* Times `steps` steps of every kernel variant from the current state (restored afterwards), keeps
 * the fastest and describes the trials in `report` ("best face_stage=..,graph=..;..."). No
 * counterpart in the reference: the project's autotuner. */
int mxeuler_autotune(mxeuler_stepper* s, int32_t steps, double cfl, int32_t local, char* report, size_t report_len,
                     char* err, size_t err_len);

/* Version string of the library. */
const char* mxeuler_version(void);

#ifdef __cplusplus
}
#endif

#endif
