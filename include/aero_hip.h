/* aero_hip.h -- C ABI of the MI355X (gfx950) FP64 Euler residual backend.
 *
 * This is the drop-in boundary for AeroCFD's `--backend hip`.  It replaces the role of the CUDA
 * backend interface cfd::cuda_backend (src/cfd_core/include/cfd_core/cuda_backend.hpp:10-53 in
 * AstaWasTaken/Aero-CFD-Solver v0.1.1); each entry point below names the reference function it
 * stands in for.  Conventions:
 *   - plain C types only; every call returns 0 on success or a non-zero AERO_E* code, and fills the
 *     caller buffer `err` (length `errlen`) with "context: hipGetErrorString(...)"-style text, the
 *     way the reference fills `std::string* error_message` (cuda/euler_backend.cu:11-19);
 *   - host arrays use the reference layout (AoS [cell][5] conserved / residual, xyz-interleaved face
 *     normals and centres) and the reference's cell / face numbering, in FP64;
 *   - a handle owns all its device memory and one HIP stream; a handle is not thread-safe, separate
 *     handles are;  calls taking host pointers are synchronous on return.
 */
#ifndef AERO_HIP_H
#define AERO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { AERO_OK = 0, AERO_EINVAL = 1, AERO_EHIP = 2, AERO_ENODEV = 3, AERO_ENOMEM = 4, AERO_ENONFINITE = 5 };
enum { AERO_FLUX_RUSANOV = 0, AERO_FLUX_HLLC = 1 };           /* EulerFluxScheme, euler_flux.hpp:24-27 */
enum { AERO_LIMITER_MINMOD = 0, AERO_LIMITER_VENKAT = 1 };    /* LimiterType, reconstruction.hpp:10-13 */
enum { AERO_BC_INTERIOR = 0, AERO_BC_FARFIELD = 1, AERO_BC_SLIP_WALL = 2 }; /* mesh_geometry.hpp:9-13 */
enum { AERO_GRAD_GREEN_GAUSS = 0, AERO_GRAD_LEAST_SQUARES = 1 };

/* ------------------------------------------------------------------------------------------------
 * Mesh description in the reference layout (UnstructuredMesh, cfd_core/mesh.hpp:14-36), FP64.
 * Face normals are unit and point out of the owner; boundary faces have neighbor == -1.
 * ---------------------------------------------------------------------------------------------- */
typedef struct aero_mesh_desc {
  int32_t num_cells;
  int32_t num_faces;
  const int32_t* face_owner;    /* [F] */
  const int32_t* face_neighbor; /* [F], -1 on boundary faces */
  const int32_t* face_bc;       /* [F], AERO_BC_* (build_euler_boundary_types, mesh_geometry.cpp:82-104) */
  const double* face_normal;    /* [3F] xyz */
  const double* face_center;    /* [3F] xyz */
  const double* face_area;      /* [F] */
  const double* cell_volume;    /* [N] */
  const double* cell_center;    /* [3N] xyz */
  int32_t wall_start;           /* contiguous "wall" patch range (forces.cpp:152-175) */
  int32_t wall_count;
} aero_mesh_desc;

/* Device layout options for aero_hip_mesh_create. */
enum { AERO_ACCUM_AUTO = 0, AERO_ACCUM_GATHER = 1, AERO_ACCUM_COLOUR = 2 };
typedef struct aero_hip_options {
  int32_t renumber;      /* 1: renumber cells into compact blocks (recursive coordinate bisection); 0: keep order */
  int32_t block_cells;   /* cells per workgroup block: 128, 256 or 512; 0 = chosen by the layout probe below */
  int32_t gradient;      /* AERO_GRAD_* */
  /* Face->cell accumulation of residual evaluations (no reference counterpart; the reference's serial loop
   * sums in face order, residual_assembly.cpp:97-181):
   *  AERO_ACCUM_GATHER: per-block face slots staged in LDS, each cell summing its faces in reference face
   *    order (the reference's summation order: results bit for bit the FP64 oracle's);
   *  AERO_ACCUM_COLOUR: faces coloured so that no two faces of a colour share a cell, one launch per
   *    colour updating both cells in place (contributions arrive in colour order: the same sums to
   *    rounding, within 1e-12 of the face-flux scale);
   *  AERO_ACCUM_AUTO: gather, unless block_cells = 0.
   * With block_cells = 0 the layout probe times one residual evaluation (HLLC + venkat, second order, on a
   * smooth synthetic state) in every candidate -- gather at 128, 256 and 512 cells per block and, unless
   * accumulation = GATHER, the coloured layout -- and keeps the fastest (the gather with the largest blocks when
   * it is within 2 % of the fastest), cached per mesh for the process
   * (aero_hip_layout_probe reports the times).  Stage updates (aero_hip_stage, _iterate, _run_case and the
   * batches) always use the gather layout at the chosen block size. */
  int32_t accumulation;
} aero_hip_options;

/* Per-evaluation scalars of EulerResidualAssemblyConfig (solvers/euler/residual_assembly.hpp:12-31). */
typedef struct aero_residual_config {
  double farfield[5];        /* conservative freestream state (rho, rhou, rhov, rhow, rhoE) */
  double gamma;
  int32_t flux_scheme;       /* AERO_FLUX_* */
  int32_t limiter;           /* AERO_LIMITER_* */
  int32_t second_order;      /* 0: first order (gradients zero) */
  int32_t low_mach_fix;
  double low_mach_cutoff;
  double stab_mach_floor;
  int32_t precond_farfield_bc;
  int32_t precond_enabled;
  double precond_mach_ref, precond_mach_min, precond_beta_min, precond_beta_max;
  int32_t all_speed_fix;
  double all_speed_cutoff, all_speed_f_min, all_speed_ramp;
  /* Beyond the reference: 1 = MUSCL with the handle's frozen limiter ratios (aero_hip_freeze_limiter),
   * q + ratio * a instead of q + phi(a, b); makes the residual smooth so stalled limiter cycles
   * converge and the adjoint is well posed.  0 = the reference's limiter. */
  int32_t limiter_frozen;
} aero_residual_config;

/* Counters of EulerResidualAssemblyDiagnostics (residual_assembly.hpp:33-41). */
typedef struct aero_residual_diag {
  int32_t first_order_fallback;
  int32_t diffusive_fallback;
  int32_t first_failure_face; /* lowest failing face index in the reference numbering, -1 if none */
} aero_residual_diag;

/* The driver configuration: EulerAirfoilCaseConfig (solvers/euler_solver.hpp:14-74) minus the mesh
 * generator settings, plus the resolved AERO_ENABLE_EXPERIMENTAL_PRECOND gate. */
typedef struct aero_case_config {
  int32_t iterations, min_iterations, cfl_ramp_iters;
  double gamma, gas_constant, mach, aoa_deg, p_inf, t_inf, rho_inf, cfl_start, cfl_max;
  double residual_reduction_target, force_stability_tol;
  int32_t force_stability_window;
  double force_mean_drift_tol;
  int32_t startup_first_order_iters, rk_stages, local_time_stepping;
  int32_t precond_on;
  double precond_mach_ref, precond_mach_min, precond_beta_min, precond_beta_max;
  int32_t precond_farfield_bc;
  double stab_floor_start, stab_floor_target, stab_floor_k_start, stab_floor_k_target;
  int32_t stab_ramp_iters;
  double x_ref, y_ref, chord;
  int32_t flux_scheme, limiter;
  int32_t low_mach_fix;
  double mach_cutoff;
  int32_t all_speed_flux_fix;
  double all_speed_mach_cutoff, all_speed_f_min;
  int32_t all_speed_ramp_start_iter, all_speed_ramp_iters;
  int32_t all_speed_staged_controller, all_speed_stage_a_min_iters, all_speed_stage_a_max_iters;
  double all_speed_stage_f_min_high, all_speed_stage_f_min_mid, all_speed_stage_f_min_low;
  double all_speed_pjump_wall_target, all_speed_pjump_spike_factor;
  int32_t all_speed_pjump_hold_iters, all_speed_freeze_iters;
  double all_speed_cfl_drop_factor, all_speed_cfl_restore_factor, all_speed_cfl_min_scale;
  int32_t aoa0_symmetry_enforce, aoa0_symmetry_enforce_interval;
  int32_t precond_experimental_enabled;
  /* Beyond the reference: at the start of this (second-order) iteration the limiter ratios are
   * captured from the state and frozen for the rest of the run; < 0 = never (the reference). */
  int32_t limiter_freeze_iter;
} aero_case_config;

/* Per-iteration history row layout (residuals.csv / forces.csv columns, euler_solver.cpp:1161-1203). */
enum {
  AERO_H_ITER = 0, AERO_H_L1, AERO_H_L2, AERO_H_LINF, AERO_H_RATIO, AERO_H_CL, AERO_H_CD, AERO_H_CM,
  AERO_H_RHO_L2, AERO_H_CFL, AERO_H_FO_FALLBACK, AERO_H_DIFF_FALLBACK, AERO_HIST_FIELDS
};

/* Outputs of a full run (EulerRunResult, euler_solver.hpp:86-148, the fields the parity contract uses).
 * All arrays are caller-owned; any pointer may be NULL to skip it. */
typedef struct aero_run_output {
  int32_t hist_capacity;   /* rows available in `history` */
  int32_t n_iter;          /* rows written */
  int32_t nonfinite;       /* non-finite trap fired (euler_solver.cpp:908-946) */
  double* history;         /* [hist_capacity][AERO_HIST_FIELDS] */
  double* conserved;       /* [N][5] final state */
  double* residual;        /* [N][5] last stage residual */
  double* spectral_radius; /* [N] */
  double* forces;          /* [8]: fx, fy, lift, drag, moment, cl, cd, cm (gauge) */
  double* wall_cp;         /* [wall_count] Cp in wall-patch order */
  double device_seconds;   /* wall time of the iteration loop */
  aero_residual_config* final_config; /* may be NULL: the residual configuration of the last iteration
                                         (ramped floors and all-speed weights, effective preconditioner) */
  double* diagnostics;     /* may be NULL: [hist_capacity][AERO_DIAG_FIELDS] per-iteration diagnostics, the
                              state- and controller-dependent columns of euler_diagnostics.csv (below) */
} aero_run_output;

/* Columns of aero_run_output.diagnostics, one row per history row: the per-iteration quantities of the
 * reference's euler_diagnostics.csv (euler_solver.cpp:717-738, :921-1139, :1161-1203) that depend on the
 * state or the all-speed controller, reduced on the device from the iteration's final state.  The other
 * columns of that file follow from the history rows and the mesh (the runner writes the whole file). */
enum {
  AERO_D_MAX_WALL_UN = 0,     /* max |u . n| of the wall-face owner cells */
  AERO_D_MAX_FF_DP,           /* max |p - p_inf| over farfield-face owner cells */
  AERO_D_MAX_FF_DRHO,         /* max |rho - rho_inf| */
  AERO_D_MAX_FF_DU,           /* max |u - u_inf| */
  AERO_D_MAX_FF_DV,           /* max |v - v_inf| */
  AERO_D_WALL_CP_MIN, AERO_D_WALL_CP_MAX, AERO_D_WALL_CP_NONFINITE,
  AERO_D_FX_GAUGE, AERO_D_FY_GAUGE, AERO_D_FX_ABS, AERO_D_FY_ABS,  /* compute_pressure_force_diagnostics */
  AERO_D_CL_ABS, AERO_D_CD_ABS, AERO_D_CM_ABS,                      /* integrate_pressure_forces, absolute p */
  AERO_D_LIFT, AERO_D_DRAG,                                         /* gauge */
  AERO_D_PJUMP_ND,            /* mean (p_o - p_n)^2 over interior faces / (rho_inf a_inf^2)^2 */
  AERO_D_PJUMP_WALL_ND,       /* the same over interior faces touching a wall-adjacent cell */
  AERO_D_PFLUCT_MAX,          /* max |p - p_inf| / (rho_inf a_inf^2) over all cells */
  AERO_D_PFLUCT_WALL,         /* ... over wall-adjacent cells */
  AERO_D_PFLUCT_FARFIELD,     /* max_ff_dp / (rho_inf a_inf^2) */
  AERO_D_AS_STAGE,            /* all-speed controller stage: 0 single, 1 A, 2 B */
  AERO_D_AS_FIX, AERO_D_AS_FMIN, AERO_D_AS_RAMP,  /* effective all-speed fix, f_min, ramp weight */
  AERO_D_AS_FREEZE,           /* freeze counter */
  AERO_D_AS_ACTION,           /* controller action code (euler_solver.cpp:1070-1139) */
  AERO_D_CFL_SCALE,           /* adaptive CFL scale after this iteration */
  AERO_D_FIRST_FAILURE_FACE,  /* first face that fell back to first order in the iteration's first stage with
                                 one (-1 if none) */
  AERO_DIAG_FIELDS
};

typedef struct aero_hip_mesh aero_hip_mesh; /* opaque device mesh + state + stream */

/* Replaces cfd::cuda_backend::cuda_runtime_available / cuda_euler_available (cuda_backend.hpp:36-37):
 * returns 1 when a gfx950 device is visible, else 0 (with the reason in err). */
int aero_hip_available(char* err, size_t errlen);

/* Replaces cfd::cuda_backend::init_euler_device_buffers (cuda_backend.hpp:44-46): uploads the mesh
 * (AoS FP64 -> SoA FP64, optional renumbering, per-block face tables) and allocates state buffers.
 * AERO_EINVAL for meshes of 2^28 cells or face slots and more (the kernels gather at 32-bit byte
 * offsets). */
int aero_hip_mesh_create(const aero_mesh_desc* desc, const aero_hip_options* opts, aero_hip_mesh** out,
                         char* err, size_t errlen);

/* Replaces cfd::cuda_backend::free_euler_device_buffers (cuda_backend.hpp:47). NULL is a no-op. */
void aero_hip_mesh_destroy(aero_hip_mesh* mesh);

/* Replaces cfd::cuda_backend::euler_residual_cuda (cuda_backend.hpp:49-52): one residual evaluation
 * -- primitive conversion, gradient (computed on the device), limiter, MUSCL, flux, face->cell
 * accumulation and spectral radius -- from host conserved[N][5] to host residual[N][5] and
 * spectral_radius[N], in the reference's numbering.  diag may be NULL. */
int aero_hip_residual(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* conserved,
                      double* residual, double* spectral_radius, aero_residual_diag* diag, char* err,
                      size_t errlen);

/* Device-resident variants for timed loops: load the state once, then evaluate the residual on the
 * handle's stream without host traffic (asynchronous), then read results back. */
int aero_hip_set_state(aero_hip_mesh* mesh, const double* conserved, char* err, size_t errlen);
int aero_hip_residual_async(aero_hip_mesh* mesh, const aero_residual_config* cfg, char* err, size_t errlen);
int aero_hip_get_residual(aero_hip_mesh* mesh, double* residual, double* spectral_radius, aero_residual_diag* diag,
                          char* err, size_t errlen);
int aero_hip_get_gradients(aero_hip_mesh* mesh, double* gradients /* [N][5][2] reference layout */, char* err,
                           size_t errlen);
/* The handle's current conserved state, [N][5] in the reference numbering. */
int aero_hip_get_state(aero_hip_mesh* mesh, double* conserved, char* err, size_t errlen);
/* The layout probe's timings for this handle (block_cells = 0): ms per residual evaluation with gather at
 * 128, 256 and 512 cells per block and with the coloured layout; -1 where a candidate was not probed
 * (explicit block_cells, or not buildable). */
int aero_hip_layout_probe(aero_hip_mesh* mesh, double* ms4);
/* The handle's HIP stream (hipStream_t), for event timing by the caller. */
void* aero_hip_stream(aero_hip_mesh* mesh);

/* Per-kernel timing of the next `max_evals` aero_hip_residual_async calls with HIP events recorded on
 * the handle's stream around the gradient and the flux kernel; _end synchronises and returns the summed
 * gradient / flux kernel milliseconds, the first-to-last event span and the number of timed calls.
 * (Measurement support; the CUDA backend has no counterpart.) */
int aero_hip_timing_begin(aero_hip_mesh* mesh, int32_t max_evals, char* err, size_t errlen);
int aero_hip_timing_end(aero_hip_mesh* mesh, double* gradient_ms, double* flux_ms, double* span_ms,
                        int32_t* n_evals, char* err, size_t errlen);

/* Device layout summary: cells, faces, blocks, cells per block, face slots, max slots per block, -,
 * workgroups launched.  Returns AERO_EINVAL on NULL. */
int aero_hip_layout(aero_hip_mesh* mesh, int64_t* out8);

/* One RK stage on the device-resident state: residual, low-Mach preconditioned residual,
 * dt/V and the SSP-RK3 stage update with enforce_physical_state (euler_solver.cpp:802-900).
 * The stage-0 call snapshots U^n.  last_stage additionally produces the residual-norm partials and the
 * non-finite flag; write_residual keeps R and the spectral radius for aero_hip_get_residual. */
int aero_hip_stage(aero_hip_mesh* mesh, const aero_residual_config* cfg, int32_t stage, int32_t last_stage,
                   double cfl, int32_t local_time_stepping, int32_t write_residual, char* err, size_t errlen);

/* Frozen limiter (no reference counterpart): captures, at `conserved` (NULL = the handle's state) and
 * the gradients of that state, the ratio phi(a, b) / a of every interior face side and primitive
 * variable (0 where a == 0) for the residual configuration's limiter.  Residual calls with
 * limiter_frozen = 1 use them until the next capture.  ratios (may be NULL): [8][F] copy, rows
 * owner-side rho, u, v, p then neighbour-side rho, u, v, p, faces in the reference order. */
int aero_hip_freeze_limiter(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* conserved,
                            double* ratios, char* err, size_t errlen);
/* Sets the handle's frozen limiter ratios ([8][F], the layout above), e.g. from a batch case. */
int aero_hip_set_limiter_ratios(aero_hip_mesh* mesh, const double* ratios, char* err, size_t errlen);

/* n pseudo-time iterations on the handle's state with a fixed residual configuration and CFL (local
 * time stepping, no ramps, no monitor; the stage body of euler_solver.cpp:802-900), e.g. to continue a
 * converged solution after a small parameter change.  rho_l2 (may be NULL): [n] density-residual L2
 * (the last stage's residual, as the driver's norms) after each iteration.  No reference counterpart. */
int aero_hip_iterate(aero_hip_mesh* mesh, const aero_residual_config* cfg, int32_t n_iters, int32_t rk_stages,
                     double cfl, double* rho_l2, char* err, size_t errlen);

/* aero_hip_iterate without any allocation: no hipMalloc / hipFree or other allocating call, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid it).  No
 * counterpart in the reference.  The per-iteration density-residual sums are staged in a static device
 * array of AERO_NOALLOC_MAX_ITERS doubles (512 KiB) that each GPU holds until the process ends; with
 * rho_l2 != NULL, n_iters above that limit is rejected with AERO_EINVAL.  Every call on one GPU shares
 * that GPU's array, so calls that overlap on one GPU (from several host threads or streams) race; calls
 * in sequence do not (each call writes every entry it reads back).  Same results as aero_hip_iterate. */
#define AERO_NOALLOC_MAX_ITERS 65536
int aero_hip_iterate_noalloc(aero_hip_mesh* mesh, const aero_residual_config* cfg, int32_t n_iters,
                             int32_t rk_stages, double cfl, double* rho_l2, char* err, size_t errlen);

/* FreestreamReference (post/forces.hpp:9-17). */
typedef struct aero_freestream_ref {
  double rho_inf, p_inf, aoa_deg, speed_inf, chord, x_ref, y_ref;
} aero_freestream_ref;

/* Pressure forces on the wall patch from the owner cells' pressure, integrate_pressure_forces
 * (post/forces.cpp:250-294; gauge = 1 subtracts p_inf, as the driver's coefficients do), summed over the
 * wall faces in face order: forces[8] = fx, fy, lift, drag, moment, cl, cd, cm.  wall_cp (may be NULL):
 * [wall_count] Cp per wall face in wall-patch order (extract_wall_cp, forces.cpp:177-206, which then
 * orders the faces along the surface).  conserved: [N][5] or NULL for the handle's state. */
int aero_hip_forces(aero_hip_mesh* mesh, const double* conserved, const aero_freestream_ref* ref, int32_t gauge,
                    double gamma, double* forces, double* wall_cp, char* err, size_t errlen);

/* Columns of the per-face scaling record (AcousticScalingInfo, euler_solver.cpp:392-416, and the
 * dissipation_debug.csv row, :1412-1437). */
enum {
  AERO_FS_MACH = 0,        /* face Mach, mean of the two sides' |v| / max(a, 1e-8) */
  AERO_FS_UN_ABS,          /* mean |u.n| */
  AERO_FS_SOUND,           /* mean speed of sound */
  AERO_FS_BETA,            /* face preconditioner beta (1 when the preconditioner is off) */
  AERO_FS_LOW_MACH,        /* low-Mach dissipation scale (1 when low_mach_fix is off) */
  AERO_FS_ACOUSTIC_BASE,   /* acoustic_scale_base */
  AERO_FS_STAB,            /* stabilization_scale */
  AERO_FS_ACOUSTIC,        /* acoustic_scale (the one the flux uses) */
  AERO_FS_ALL_SPEED,       /* all-speed HLLC pressure scale */
  AERO_FS_SMAX_PHYSICAL,   /* max(|u_n| + a) */
  AERO_FS_SMAX_DISSIPATION,/* max of the preconditioned spectral radii at acoustic_scale */
  AERO_FS_WALL_MASS_FLUX,  /* slip walls: mass component of the wall flux; 0 elsewhere */
  AERO_FS_FIELDS
};

/* The post-run scaling diagnostics of run_euler_airfoil_case (euler_solver.cpp:1217-1437), evaluated on
 * the device at conserved ([N][5], or NULL for the handle's state) under cfg (the residual configuration
 * of the last iteration, as aero_run_output.final_config returns it; the limiter fields are unused):
 * face_scales[AERO_FS_FIELDS][F] per face in the reference's face numbering, with the exterior state the
 * interior one at a slip wall and the freestream at a farfield (:1270-1279); cell_scales (may be NULL)
 * [2][N]: cell preconditioner beta, then local Mach, in the reference's cell numbering (:1229-1246).
 * The host writes dissipation_debug.csv, wall_flux.csv and run.log's beta / Mach / acoustic-scale
 * statistics from these. */
int aero_hip_face_scales(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* conserved,
                         double* face_scales, double* cell_scales, char* err, size_t errlen);

/* The full pseudo-time run of run_euler_airfoil_case (euler_solver.cpp:463-1208) with the stage
 * loop, norms and forces on the device and the ramp / CFL / convergence control on the host. */
int aero_hip_run_case(aero_hip_mesh* mesh, const aero_case_config* cfg, aero_run_output* out, char* err,
                      size_t errlen);

/* ------------------------------------------------------------------------------------------------
 * Batched solves (no reference counterpart: the reference sweep runs cases one after another,
 * python/cfd/sweep.py:14-59).  Cases share device geometries (aero_hip_mesh handles, all built with the
 * same block_cells) and advance together, one launch per RK stage over every (case, block); the driver's
 * per-iteration control (CFL ramp and adaptation, floor / all-speed ramps, norms, forces) runs on the
 * device, so an iteration is a single replayable hipGraph.  A case stops at the density-residual drop
 * rho_l2 <= rho_drop * rho_l2(iteration 0), at a non-finite state, or at max_iters.
 * ---------------------------------------------------------------------------------------------- */
typedef struct aero_hip_batch aero_hip_batch;

typedef struct aero_batch_case {
  int32_t mesh_index; /* index into the meshes passed to aero_hip_batch_create */
  double mach;
  double aoa_deg;
  /* Per-case numerics switches, so that one batch can mix low-Mach preconditioned cases with plain
   * ones: 0 takes the case configuration's setting, 1 turns it on, -1 off.  precond = 1 turns the
   * low-Mach preconditioner on without the experimental gate (the caller has passed it). */
  int32_t precond;
  int32_t all_speed_flux_fix;
  /* NULL (the batch's base configuration) or this case's own: freestream reference values, CFL,
   * stabilisation-floor and all-speed schedules, preconditioner and low-Mach settings.  It must agree
   * with the base on flux scheme, limiter, rk_stages, startup iterations, freeze iteration, local time
   * stepping and gamma; its mach and aoa_deg are ignored (the fields above apply). */
  const aero_case_config* config;
} aero_batch_case;

typedef struct aero_batch_result {
  int32_t iterations, converged, failed;
  double l2, rho_l2, rho_ratio, cl, cd, cm;
} aero_batch_result;

/* `base` supplies everything but Mach and AoA (flux, limiter, CFL schedule, startup, precond, ...) for
 * cases without their own configuration; rk_stages 1 or 3, local time stepping only.  hist_capacity rows of (l2, rho_l2, cfl, linf) are kept
 * per case. */
int aero_hip_batch_create(aero_hip_mesh* const* meshes, int32_t n_meshes, const aero_batch_case* cases,
                          int32_t n_cases, const aero_case_config* base, int32_t hist_capacity,
                          aero_hip_batch** out, char* err, size_t errlen);
void aero_hip_batch_destroy(aero_hip_batch* batch);
/* Runs up to max_iters iterations (fixed_iters = 1: ignore the convergence test); the host synchronises
 * every iters_per_sync iterations; use_graph = 1 replays captured hipGraphs.  results: [n_cases]. */
int aero_hip_batch_run(aero_hip_batch* batch, int32_t max_iters, double rho_drop, int32_t fixed_iters,
                       int32_t iters_per_sync, int32_t use_graph, aero_batch_result* results, double* seconds,
                       char* err, size_t errlen);
int aero_hip_batch_get_state(aero_hip_batch* batch, int32_t case_index, double* conserved /* [N][5] */, char* err,
                             size_t errlen);
int aero_hip_batch_history(aero_hip_batch* batch, int32_t case_index, double* hist /* [cap][4] */, int32_t cap,
                           char* err, size_t errlen);
/* The frozen limiter ratios a case captured ([8][F] of its mesh; zeros before the freeze iteration).
 * Fails when the batch was created without a freeze iteration. */
int aero_hip_batch_limiter_ratios(aero_hip_batch* batch, int32_t case_index, double* ratios, char* err,
                                  size_t errlen);
/* Warm starts for batched continuations (no reference counterpart), e.g. the perturbed solutions of
 * finite-difference sensitivities started from a converged state: set_state makes conserved ([N][5] of the
 * case's mesh, reference numbering) the state the case's next iteration starts from.  set_limiter_ratios
 * installs frozen limiter ratios ([8][F], aero_hip_freeze_limiter's layout; held at 2^-15) for the case
 * and marks the whole batch frozen from its next iteration, so no capture happens at limiter_freeze_iter:
 * install them for every case.  It needs a batch created with limiter_freeze_iter >= 0.  A continuation
 * batch typically sets startup_first_order_iters = 0, cfl_ramp_iters = 0 and stab_ramp_iters = 0 so that
 * its first iteration already uses the converged run's final configuration. */
int aero_hip_batch_set_state(aero_hip_batch* batch, int32_t case_index, const double* conserved, char* err,
                             size_t errlen);
int aero_hip_batch_set_limiter_ratios(aero_hip_batch* batch, int32_t case_index, const double* ratios, char* err,
                                      size_t errlen);
/* Case groups (no reference counterpart): the batch iterates its active cases in groups of about equal
 * hot data (state buffers, gradients, frozen ratios: 160 B per cell + 16 B per face slot), each group
 * running a sync interval's iterations in turn, so that a group's data stays in the last-level cache between
 * its kernels.  The number of groups is ceil(hot bytes / group_bytes), at most one per group_items work
 * items (workgroups) and one per case.  0 restores a default: group_bytes 256 MiB (the size of MI355X's
 * Infinity Cache), group_items 4096 (about 16 workgroups per CU).  Results are bitwise the
 * same for any grouping.  *n_groups (may be NULL) receives the number of groups of the active cases. */
int aero_hip_batch_set_groups(aero_hip_batch* batch, int64_t group_bytes, int32_t group_items, int32_t* n_groups,
                              char* err, size_t errlen);

/* ------------------------------------------------------------------------------------------------
 * Discrete tangent and adjoint of the residual (no reference counterpart: the reference has no
 * derivative code).  Derivatives are exact derivatives of the residual the other entry points
 * compute -- the same device code instantiated on dual numbers -- taken on the branch each value
 * follows (limiter cases, HLLC fallbacks, floors).  A NULL `conserved` means the handle's current
 * device state (for example the state aero_hip_run_case ended on); otherwise it becomes that state.
 * Host arrays use the reference layout and numbering; the rhow column is ignored on input and zero on
 * output.
 * ---------------------------------------------------------------------------------------------- */
enum { AERO_PARAM_NONE = 0, AERO_PARAM_AOA = 1, AERO_PARAM_MACH = 2 };

/* A direction in parameter space: derivatives of the freestream-dependent options and, optionally,
 * of the geometry (reference layout; any pointer may be NULL for zero). */
typedef struct aero_tangent_seed {
  double d_farfield[5];
  double d_stab_floor, d_mach_ref, d_mach_min, d_beta_min, d_beta_max;
  const double* d_face_normal; /* [3F] */
  const double* d_face_center; /* [3F] */
  const double* d_face_area;   /* [F] */
  const double* d_cell_center; /* [3N] */
  const double* d_cell_volume; /* [N] */
} aero_tangent_seed;

/* J = w_cl CL + w_cd CD + w_cm CM of the gauge wall pressure forces (forces.cpp:208-294), with the
 * reference values that define the coefficients. */
typedef struct aero_objective {
  double w_cl, w_cd, w_cm;
  double aoa_deg, q_inf, p_inf, chord, x_ref, y_ref, gamma;
} aero_objective;

/* Controls and results of a linear solve.  The pseudo-time iteration (rk_stages 1 or 3, cfl) runs at
 * most max_iters iterations; with gmres_restart = m > 0 it is instead accelerated by restarted GMRES(m)
 * on its fixed-point form x = G x + c, one pseudo-time iteration per matrix-vector product (this
 * converges where the plain iteration diverges on a few unstable modes, as at limit-cycling states).
 * Stop when the residual's L2 norm is <= tol times its first value (checked every check_every
 * iterations of the plain iteration; every Arnoldi step under GMRES, where the norm is that of the
 * fixed-point residual). */
typedef struct aero_linear_solve {
  int32_t rk_stages;
  double cfl;
  int32_t max_iters;
  double tol;
  int32_t check_every;
  int32_t gmres_restart;
  int32_t gmres_sweeps;   /* pseudo-time iterations per GMRES matrix-vector product (>= 1) */
  int32_t iterations;     /* out */
  double res0, res;       /* out: first and last linear residual norms */
  double seconds;         /* out: device time of the iterations */
} aero_linear_solve;

/* The residual configuration of iteration `iteration` of the case (the driver's ramps evaluated
 * there, euler_solver.cpp:463-797), the derivative of its freestream-dependent fields with respect to
 * `param` (AERO_PARAM_*; angle of attack per degree), the objective reference values (weights zero) and
 * d q_inf / d param.  The staged all-speed controller is rejected (its schedule depends on the run). */
int aero_case_linearization(const aero_case_config* cfg, int32_t iteration, int32_t param,
                            aero_residual_config* residual_cfg, aero_tangent_seed* seed, aero_objective* objective,
                            double* d_q_inf, char* err, size_t errlen);

/* Residual R and its directional derivative dR = R_U dU + R_p . seed (dU or seed may be NULL). */
int aero_hip_residual_jvp(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* conserved,
                          const double* dU, const aero_tangent_seed* seed, double* residual /* or NULL */, double* dR,
                          char* err, size_t errlen);
/* out = R_U^T psi, through the stored face Jacobians. */
int aero_hip_residual_vjp(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* conserved,
                          const double* psi, double* out, char* err, size_t errlen);

/* Device-pointer forms of aero_hip_residual (euler_residual_cuda, cuda_backend.hpp:49-52) and
 * aero_hip_residual_vjp, for callers whose arrays already live on the handle's GPU (the torch layer):
 * the same layouts ([N][5] conserved / residual / psi in the reference numbering, spectral radius [N],
 * which may be NULL), as device pointers; the arrays never cross to the host.  Work runs on the handle's
 * stream: the caller makes its inputs ready first (e.g. synchronises its own stream); the call returns
 * when the outputs are written.  They set the handle's state to d_conserved. */
int aero_hip_residual_device(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* d_conserved,
                             double* d_residual, double* d_spectral_radius, aero_residual_diag* diag, char* err,
                             size_t errlen);
int aero_hip_residual_vjp_device(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* d_conserved,
                                 const double* d_psi, double* d_out, char* err, size_t errlen);
/* Objective value, dJ/dU [N][5] (may be NULL) and the explicit partials: dJ/d(aoa, per degree) and
 * dJ/dq_inf at fixed state, and the derivative along the geometry of `seed` (may be NULL). */
int aero_hip_objective(aero_hip_mesh* mesh, const aero_objective* obj, const double* conserved, double* J,
                       double* dJdU, double* dJ_daoa, double* dJ_dq, const aero_tangent_seed* seed, double* dJ_seed,
                       char* err, size_t errlen);
/* Adjoint: psi with R_U^T psi = rhs, by the transposed pseudo-time iteration (same stages, CFL and
 * local time steps as the primal, frozen at the state).  history: [max_iters] residual norms or NULL. */
int aero_hip_adjoint_solve(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* conserved,
                           const double* rhs, aero_linear_solve* ctl, double* psi, double* history, char* err,
                           size_t errlen);
/* Tangent: v with R_U v = -R_p . seed, by the linearised pseudo-time iteration. */
int aero_hip_tangent_solve(aero_hip_mesh* mesh, const aero_residual_config* cfg, const double* conserved,
                           const aero_tangent_seed* seed, aero_linear_solve* ctl, double* v, double* history, char* err,
                           size_t errlen);

/* One solve of a batched adjoint solve. */
typedef struct aero_adjoint_case {
  aero_hip_mesh* mesh;             /* linearised at its state, or at `conserved` when given */
  const aero_residual_config* cfg;
  const double* conserved;         /* [N][5] or NULL */
  const double* rhs;               /* [N][5] dJ/dU */
  double* psi;                     /* [N][5] out: R_U^T psi = rhs */
  double cfl;                      /* > 0: this solve's pseudo-time CFL; 0: ctl->cfl */
  int32_t iterations, converged, failed; /* out */
  double res0, res;                /* out: first and last linear residual L2 norms */
} aero_adjoint_case;

/* Many adjoint solves advanced together (e.g. CL and CD of every case of a sweep, on meshes of uneven
 * size): one launch per transposed RK stage over every (solve, block of 256 cells), per-solve stopping
 * on the device (the tests of aero_hip_adjoint_solve, checked after every iteration), iterations replayed
 * from a hipGraph.  Solves on the same handle share its face Jacobians and must share cfg and
 * conserved; all handles need the same gradient kind.  ctl gives rk_stages, cfl (unless a solve sets
 * its own), max_iters and tol (gmres_restart must be 0); ctl->iterations and ctl->seconds receive the largest iteration count and
 * the wall time of the iterations.  The host checks progress every iters_per_sync iterations and then
 * calls `progress` (may be NULL) with the iterations run and the solves still active; a non-zero return
 * stops the iterations there.  A solve whose residual turns non-finite stops with `failed` set while the
 * others go on (the call itself succeeds).  Each solve equals aero_hip_adjoint_solve with
 * check_every = 1 bit for bit.  No reference counterpart. */
typedef int (*aero_progress_fn)(void* user, int32_t iterations, int32_t active);
int aero_hip_adjoint_solve_batch(aero_adjoint_case* cases, int32_t n_cases, aero_linear_solve* ctl,
                                 int32_t iters_per_sync, aero_progress_fn progress, void* user, char* err,
                                 size_t errlen);

/* ------------------------------------------------------------------------------------------------
 * Mesh generation (host side).  The NACA-4 O-grid of make_airfoil_ogrid_mesh
 * (mesh/airfoil_mesh.cpp:395-544) with the reference's triangles, plus quad and mixed
 * (quad wall block + triangles) cells, seeded interior-node jitter and RCM cell renumbering.
 * precision 4 reproduces the reference's FP32 generator; precision 8 computes in FP64.
 * ---------------------------------------------------------------------------------------------- */
enum { AERO_CELLS_TRI = 0, AERO_CELLS_QUAD = 1, AERO_CELLS_MIXED = 2 };

typedef struct aero_mesh_gen_config {
  char naca_code[8];
  double chord;
  int32_t num_circumferential, num_radial;
  double farfield_radius, radial_stretch;
  int32_t cell_type;     /* AERO_CELLS_* */
  int32_t quad_layers;   /* AERO_CELLS_MIXED: wall-normal layers kept as quads */
  double jitter;         /* interior-node jitter as a fraction of local spacing (0 = none) */
  uint64_t seed;
  int32_t renumber_rcm;  /* 1: reverse Cuthill-McKee cell order before faces are built */
} aero_mesh_gen_config;

typedef struct aero_mesh_arrays {
  int32_t precision; /* 4 or 8: element size of the floating arrays */
  int32_t num_cells, num_faces, num_points;
  int32_t* face_owner;
  int32_t* face_neighbor;
  int32_t* face_bc;
  int32_t* face_vertices; /* [2F] */
  void* face_normal;      /* [3F] */
  void* face_center;      /* [3F] */
  void* face_area;        /* [F] */
  void* cell_volume;      /* [N] */
  void* cell_center;      /* [3N] */
  void* points;           /* [3P] */
  int32_t* cell_offsets;  /* [N] running end offsets into cell_connectivity (VTU style) */
  int32_t* cell_connectivity;
  int32_t wall_start, wall_count, farfield_start, farfield_count;
} aero_mesh_arrays;

int aero_mesh_generate(const aero_mesh_gen_config* cfg, int32_t precision, aero_mesh_arrays* out, char* err,
                       size_t errlen);
void aero_mesh_free(aero_mesh_arrays* arrays);

/* Binary mesh file (no reference counterpart: the reference regenerates its mesh on every run,
 * euler_solver.cpp:474; the file holds the UnstructuredMesh fields of mesh.hpp:14-36 plus this ABI's
 * boundary codes and patch ranges).  Layout: the magic "AEROMSH1", int32 precision (4 or 8), num_cells,
 * num_faces, num_points, connectivity length, wall_start, wall_count, farfield_start, farfield_count, then
 * the arrays of aero_mesh_arrays in declaration order (floating arrays at the file's precision), in the
 * host's byte order.  aero_mesh_read allocates `out` (free with aero_mesh_free) and rejects a file whose
 * indices, counts or patch ranges are out of range, so its meshes are safe to hand to the device. */
int aero_mesh_write(const aero_mesh_arrays* mesh, const char* path, char* err, size_t errlen);
int aero_mesh_read(const char* path, aero_mesh_arrays* out, char* err, size_t errlen);

/* Shape parameters (no reference counterpart): Hicks-Henne bumps b_i(x) = sin^3(pi x^(ln 0.5 / ln h_i)),
 * h_i = linspace(h_min, h_max, bumps_per_surface) in x/c, on the upper surface (parameters
 * 0..n-1) then the lower surface (n..2n-1); amplitude a_i moves that surface's wall nodes by a_i b_i(x)
 * in +y, and every other node by the inverse-distance-weighted average (weights |r|^-idw_exponent) of
 * the wall and farfield node displacements (farfield nodes fixed). */
typedef struct aero_shape_config {
  int32_t bumps_per_surface;
  double h_min, h_max;
  double idw_exponent; /* 3 when <= 0 */
} aero_shape_config;

/* Derivatives of the deformed geometry with respect to each amplitude (caller-allocated, any NULL). */
typedef struct aero_geometry_tangents {
  int32_t n_params;       /* 2 * bumps_per_surface */
  double* d_points;       /* [n_params][3P] */
  double* d_face_normal;  /* [n_params][3F] */
  double* d_face_center;  /* [n_params][3F] */
  double* d_face_area;    /* [n_params][F] */
  double* d_cell_center;  /* [n_params][3N] */
  double* d_cell_volume;  /* [n_params][N] */
} aero_geometry_tangents;

/* Deforms `base` (precision 8, from aero_mesh_generate) by the 2 * bumps_per_surface amplitudes into
 * `out` (same topology; free with aero_mesh_free) and fills `tangents` when not NULL.  Fails if a cell
 * would invert. */
int aero_mesh_deform(const aero_mesh_arrays* base, const aero_shape_config* shape, const double* amplitudes,
                     aero_mesh_arrays* out, aero_geometry_tangents* tangents, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif /* AERO_HIP_H */
