// aero_internal.hpp -- structures shared by the gfx950 kernels and the host side of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "euler_device.hpp"

namespace aero_dev {

// Slot metadata bits: boundary-condition code (AERO_BC_*) and whether this block owns the face's
// owner cell (so per-face counters are taken once although boundary-straddling faces are slots of
// two blocks).
constexpr int kMetaBcMask = 3;
constexpr int kMetaPrimary = 4;
// A face slot's neighbour and bc / primary bits packed in one int (cells < 2^28).
__host__ __device__ __forceinline__ int slot_nb(int nm) { return nm >> 3; }
__host__ __device__ __forceinline__ int slot_meta(int nm) { return nm & 7; }

// Columns of the per-face scaling record (k_face_scales; AERO_FS_* in the public header).
enum : int {
  kFsMach = 0, kFsUnAbs, kFsSound, kFsBeta, kFsLowMach, kFsAcousticBase, kFsStab, kFsAcoustic, kFsAllSpeed,
  kFsSmaxPhysical, kFsSmaxDissipation, kFsWallMassFlux, kFsFields
};

// Device view of a mesh in internal (block) order.  Passed by value to kernels.
struct DevMesh {
  int N;         // cells
  int nfaces;    // faces (reference numbering)
  int ld;        // leading dimension of SoA cell arrays (== N)
  int nblk;      // blocks of BC cells
  int BC;        // cells per block == threads per workgroup
  int smax;      // max slots of any block (LDS sizing)
  int grid;      // workgroups launched (>= nblk when XCD-remapped)
  int xcd_remap;
  int grad;      // AERO_GRAD_*: 0 Green-Gauss, 1 weighted least squares
  int nslots;    // face slots S over all blocks
  const int* slot_off;  // [nblk + 1]
  const int* s_o;       // [S] owner (internal id)
  const int* s_nm;      // [S] neighbour (internal id, or -1) * 8 + (bc | primary): slot_nb, slot_meta
  const int* s_face;    // [S] reference face index
  const double* s_nx;   // [S]
  const double* s_ny;
  const double* s_A;
  const double* s_fx;   // face centre
  const double* s_fy;
  const int* c_off;     // [N + 1] per-cell entry offsets
  const unsigned short* c_ent;  // (local slot << 1) | is_neighbour, increasing reference face index
  const double* vol;    // [N]
  const double* ccx;    // [N] cell centre
  const double* ccy;
};

// Face-coloured layout (the alternative to the per-block gather that the layout probe can choose for
// residual evaluations): faces grouped by colour so that no two faces of one colour share a cell, and
// within a colour ordered by owner cell.  One launch per colour adds each face's F A to its owner and
// subtracts it from its neighbour directly, without floating-point atomics; a cell's contributions then
// arrive in colour order instead of reference face order (the same sums to rounding).
struct ColourFaces {
  const int* o;       // [F] owner (internal id), colour-major
  const int* n;       // [F] neighbour (internal id) or -1
  const int* meta;    // [F] bc code
  const int* face;    // [F] reference face index
  const double* nx;   // [F]
  const double* ny;
  const double* A;
  const double* fx;   // face centre
  const double* fy;
};

struct StageArgs {
  int stage;
  int last_stage;
  int write_res;
  double cfl;
  const double* Un;  // state at the start of the iteration
  double* Unext;     // output state
  double* partials;  // [nblk][4] norm partials (last stage)
};

struct WallArgs {
  int n;
  const int* owner;  // internal cell id, wall-patch order
  const double* nx;
  const double* ny;
  const double* A;
  const double* fx;
  const double* fy;
  double p_inf, q_inf, x_ref, y_ref;
};

// ---- batched execution ---------------------------------------------------------------------------
// One case of a batch: its own state ring and work buffers over a shared geometry.
struct BatchCase {
  int geom;          // index into the geometry table
  int nblk;          // blocks of that geometry
  int N;             // cells of that geometry
  double* U[3];      // state ring
  double* G;
  double* R;
  double* lam;
  double* partials;  // [nblk][4]
  int* diag;         // [4]
  FaceOpts fo;       // this case's face options (freestream, preconditioner, low-Mach and all-speed
                     // settings: euler_solver.cpp:341-355, :543-561); floor, as_w and frozen set per iteration
  double floor_start, floor_target;         // stabilisation-floor schedule (euler_solver.cpp:141-157)
  double cfl_start, cfl_max;                // CFL ramp (euler_solver.cpp:69-79)
  int cfl_ramp, floor_ramp, as_on, as_len;  // ramp lengths and the all-speed ramp start
  WallArgs wall;     // wall faces of the geometry + this case's reference values
  double* lim_ratio; // frozen limiter ratios [8][F] (beyond the reference) or null
  unsigned short* lim_slot;  // the same per face slot of the geometry [S][8] in units of 2^-15 (the flux
                             // kernel's coalesced copy)
};

// Device-resident control state of one case, advanced by k_batch_control after every iteration with the
// driver's per-iteration logic (euler_solver.cpp:771-797, :948-958).
struct CaseCtl {
  int active;     // still iterating
  int converged;  // reached the density-residual drop
  int failed;     // non-finite trap
  int iter;       // iterations completed
  double cfl, floor_cur, as_w;  // parameters of the next iteration
  double cfl_scale, prev_l2, l2_0, rho_l2_0;
  double l1, l2, linf, rho_l2, ratio, rho_ratio;
  double sums[9];  // k_iter_reduce output of the last iteration
};

struct BatchParams {
  int ncases;
  int max_iters;
  int fixed_iters;   // 1: ignore the convergence test (run exactly max_iters)
  double rho_drop;   // converged when rho_l2 <= rho_drop * rho_l2(iteration 0)
  double* hist;      // [ncases][hist_cap][4]: l2, rho_l2, cfl, cl-sum proxy (may be null)
  int hist_cap;
};

// Buffer selectors of the batched kernels (BatchStage cur / base / out, and the `cur` arguments).
// sel >= 0 names a ring buffer directly: RK3 ends every iteration in its start buffer, so every case's
// state stays in one buffer.  RK1 ping-pongs between buffers 0 and 1, and a case that has stopped no longer
// advances, so its buffer follows its own iteration count: kRingParity is buffer iter & 1 (the state after
// `iter` iterations), kRingParityNext the other one (the output of the running iteration).
constexpr int kRingParity = -1;
constexpr int kRingParityNext = -2;

struct BatchStage {
  int stage, last, cur, base, out;
};

// ---- tangent / adjoint -----------------------------------------------------------------------------
// Faces in reference order with internal cell ids, and per-cell face lists (face << 1 | is_neighbour)
// in increasing face index: the summation order of the residual.
struct AdjMesh {
  int N, F, ld;
  int grad;              // AERO_GRAD_*
  const int* fo;         // [F] owner (internal id)
  const int* fn;         // [F] neighbour (internal id) or -1
  const int* fbc;        // [F] AERO_BC_*
  const double* fnx;     // [F]
  const double* fny;
  const double* fA;
  const double* ffx;     // face centre
  const double* ffy;
  const double* fwdx;    // [F] least-squares face weights w dx, w dy (interior faces)
  const double* fwdy;
  const int* cf_off;     // [N + 1]
  const int* cf_ent;
  const double* vol;     // [N]
  const double* ccx;
  const double* ccy;
  const double* H;       // [3][N] least-squares inverse normal matrix (hxx, hxy, hyy); zero when singular
  int wall_start, wall_count;
  // Per-cell face entries in tiles of kEntryTile cells: cell c (group c / kEntryTile, lane
  // c % kEntryTile) keeps its i-th face, in increasing reference face index, in slot
  // (e_tile[group] + i) * kEntryTile + lane, so a wavefront of consecutive cells reads consecutive
  // addresses.  A group has as many tiles as its cells' largest face count.
  const int* e_tile;     // [groups + 1] first tile of each group
  const int* e_nb;       // [slots] other cell << 1 | side (side 1: this cell is the face's neighbour); -2 boundary
  const double* e_gx;    // [slots] gradient-transpose face weights: nx A (Green-Gauss) or w dx (least squares)
  const double* e_gy;
  const double* e_dx;    // [slots] face centre - this cell's centre (the reconstruction vector)
  const double* e_dy;
  const int* f_slot;     // [2][F] owner-side and neighbour-side slot of each face, -1 where none
};

// Cells per entry tile (one workgroup of the adjoint kernels).
constexpr int kEntryTile = 256;
// Stored face-Jacobian doubles per slot, tile-major [tile][kJacCols][kEntryTile]: column d * 4 + k is
// d(F_k A) / d(U_d) of this side's conserved state, 16 + d * 4 + k is d(F_k A) / d(a_d) of this side's
// reconstruction increment a_d = g_d . (face centre - cell centre) of primitive d (rho, u, v, p).
constexpr int kJacCols = 32;

// Geometry tangents (any pointer may be null = zero): faces in reference order, cells internal.
struct GeoTan {
  const double* dnx;
  const double* dny;
  const double* dA;
  const double* dfx;
  const double* dfy;
  const double* dccx;
  const double* dccy;
  const double* dvol;
};

// One solve of a batched adjoint solve: the linearisation (face Jacobians, state, spectral radius,
// options) of its mesh handle and its own right-hand side and iteration buffers ([4][N] / [8][N], ld N).
// Whether a linearisation's face Jacobians take the factored form of k_face_jacobian_w (d(F A)/dW per
// side, with the frozen ratios and the cells' primitive-map Jacobians applied in the gather): the frozen
// limiter, or first order.  With the limiter live, a side's face state depends on the other cell too.
__host__ __device__ inline bool jacobian_factored(bool second, bool frozen) { return !second || frozen; }

struct AdjBatchSolve {
  AdjMesh am;
  const double* J;
  int fact;          // J in the factored form (jacobian_factored)
  const double* U;
  const double* lam;
  FaceOpts o;
  double cfl;
  const double* rhs;
  double* x[3];      // iterate ring
  double* psi;       // P^T D x of the current stage input
  double* Gbar;
  double* direct;
  double* partials;  // [nparts] block partials of the stage-0 squared residual
  int nparts;
};

struct AdjBatchCtl {
  int active, converged, failed, iter;
  double res0, res;
};

}  // namespace aero_dev

namespace aero_host {
// diag (optional): residual counters [4] the next flux launch accumulates, reset by the gradient kernel.
hipError_t launch_gradient(const aero_dev::DevMesh& m, const double* U, double* G, double g, hipStream_t s,
                           int* diag = nullptr);
hipError_t launch_flux(const aero_dev::DevMesh& m, const aero_dev::FaceOpts& o, int scheme, int limiter, bool second,
                       int mode, const double* U, const double* G, double* R, double* lam, int* diag,
                       const aero_dev::StageArgs& st, hipStream_t s);
// The residual R, lam of U on the face-coloured layout (colour_off: [ncolours + 1] host offsets into cf):
// zeroes R and lam, then one launch per colour.  G: the gradients (second order).
hipError_t launch_flux_colour(const aero_dev::DevMesh& m, const aero_dev::ColourFaces& cf, const int* colour_off,
                              int ncolours, const aero_dev::FaceOpts& o, int scheme, int limiter, bool second,
                              const double* U, const double* G, double* R, double* lam, int* diag, hipStream_t s);
// Caller device arrays in the reference layout ([N][5], reference numbering; perm: internal -> reference)
// to / from the internal SoA [4][N]; lam (internal order) goes to lam_ref (reference order) when both are set.
hipError_t launch_aos_to_internal(int N, const int* perm, const double* aos5, double* soa, hipStream_t s);
hipError_t launch_internal_to_aos(int N, const int* perm, const double* soa, double* aos5, const double* lam,
                                  double* lam_ref, hipStream_t s);
// Frozen limiter ratios [8][F] at U with gradients G (k_limiter_capture).
// Frozen limiter ratios [8][F] -> per face slot [S][8] in units of 2^-15 (k_ratio_slots).
hipError_t launch_ratio_slots(const aero_dev::DevMesh& m, double* ratios, unsigned short* slots, hipStream_t s);
hipError_t launch_face_scales(const aero_dev::DevMesh& m, const aero_dev::FaceOpts& o, const double* U, double* faces,
                              double* cells, hipStream_t s);
hipError_t launch_limiter_capture(const aero_dev::DevMesh& m, int limiter, const double* U, const double* G, double g,
                                  double* ratios, hipStream_t s);
hipError_t launch_global_update(const aero_dev::DevMesh& m, const aero_dev::FaceOpts& o, const double* U,
                                const double* R, const double* lam, double cfl, double* dt_scratch, int* diag,
                                const aero_dev::StageArgs& st, hipStream_t s);
hipError_t launch_norm_partials(const aero_dev::DevMesh& m, const double* R, double* partials, hipStream_t s);
hipError_t launch_iter_reduce(const aero_dev::DevMesh& m, const double* partials, const aero_dev::WallArgs& wa,
                              const double* U, double g, double* out, double* cp, hipStream_t s);
hipError_t launch_batch_gradient(int BC, int grad, const aero_dev::DevMesh* geoms, const aero_dev::BatchCase* cases,
                                 const aero_dev::CaseCtl* ctl, const int2* work, int nwork, int smax, double g, int cur,
                                 hipStream_t s);
// plain: every case's face options satisfy plain_options (the flux kernels' PLAIN instantiation).
hipError_t launch_batch_flux(int BC, int scheme, int limiter, bool second, bool frozen, bool plain,
                             const aero_dev::DevMesh* geoms, const aero_dev::BatchCase* cases,
                             const aero_dev::CaseCtl* ctl, const int2* work, int nwork, int smax,
                             const aero_dev::BatchStage& bs, hipStream_t s);
// Frozen limiter capture of every active case at its state U[cur] (gradients in G).
hipError_t launch_batch_limiter_capture(int BC, int limiter, const aero_dev::DevMesh* geoms,
                                        const aero_dev::BatchCase* cases, const aero_dev::CaseCtl* ctl,
                                        const int2* work, int nwork, double g, int cur, hipStream_t s);
// Per-iteration control of the n cases listed in list[0..n) (device array of case indices).
hipError_t launch_batch_reduce(const aero_dev::BatchCase* cases, aero_dev::CaseCtl* ctl, const int* list, int n,
                               double g, int cur, hipStream_t s);
hipError_t launch_batch_control(const aero_dev::BatchCase* cases, aero_dev::CaseCtl* ctl, const int* list, int n,
                                const aero_dev::BatchParams& p, hipStream_t s);
hipError_t launch_batch_init_diag(const aero_dev::BatchCase* cases, const int* list, int n, hipStream_t s);

// tangent / adjoint (aero_adjoint.hip)
hipError_t launch_tan_gradient(const aero_dev::AdjMesh& am, const double* U, const double* dU,
                               const aero_dev::GeoTan& gt, double g, double* G, double* dG, hipStream_t s);
hipError_t launch_tan_face(const aero_dev::AdjMesh& am, int scheme, int limiter, bool second, const double* U,
                           const double* dU, const double* G, const double* dG, const aero_dev::GeoTan& gt,
                           const aero_dev::FaceOptsT<aero_dev::Dual>& o, double* fa, hipStream_t s);
hipError_t launch_tan_gather(const aero_dev::AdjMesh& am, const double* fa, double* R, double* dR, hipStream_t s);
hipError_t launch_face_jacobian(const aero_dev::AdjMesh& am, int scheme, int limiter, bool second, const double* U,
                                const double* G, const aero_dev::FaceOpts& o, double* J, hipStream_t s);
// fact: J was built in the factored form (jacobian_factored of the linearisation).
hipError_t launch_adj_product(const aero_dev::AdjMesh& am, const double* J, bool fact, const double* U,
                              const double* psi, double g, double* Gbar, double* out, hipStream_t s);
// One stage of the transposed pseudo-time iteration (k_adj_gather + k_adj_grad_update): reads psi, writes
// xout, the block partials (may be null) and the next stage's psi.
hipError_t launch_adj_stage(const aero_dev::AdjMesh& am, const double* J, bool fact, const double* U,
                            const aero_dev::FaceOpts& o, int stage, int rk3, const double* x0, const double* xs,
                            const double* rhs, const double* lam, double cfl, double* Gbar, double* direct,
                            double* xout, double* partials, double* psi, hipStream_t s);
hipError_t launch_objective_gradient(const aero_dev::AdjMesh& am, const double* U, double g, const double* coef,
                                     double* out, hipStream_t s);
// Batched transposed stage over (solve a, solve b or -1, block) work items (two solves of one
// linearisation share an item); the iterate ring indices select x[] of each solve.  Then the per-solve control (residual norm, iteration count, stopping tests).
// fact_mode: 0 every solve's face Jacobians unfactored, 1 every solve's factored, 2 mixed.
hipError_t launch_adj_batch_stage(int grad, int fact_mode, const aero_dev::AdjBatchSolve* solves,
                                  const aero_dev::AdjBatchCtl* ctl, const int4* work, int nwork, int stage, int rk3,
                                  int ix0, int ixs, int ixo, int want_partials, hipStream_t s);
hipError_t launch_adj_batch_control(const aero_dev::AdjBatchSolve* solves, aero_dev::AdjBatchCtl* ctl, int nsolves,
                                    double tol, int max_iters, hipStream_t s);
hipError_t launch_adj_scale(const aero_dev::AdjMesh& am, const double* lam, double cfl, const double* U,
                            const aero_dev::FaceOpts& o, const double* phi, double* psi, hipStream_t s);
hipError_t launch_lin_update(const aero_dev::AdjMesh& am, int stage, int rk3, int forward, const double* x0,
                             const double* xs, const double* r, const double* rhs, const double* lam, double cfl,
                             const double* U, const aero_dev::FaceOpts& o, double* xout, double* partials,
                             hipStream_t s);
hipError_t launch_partials_sum(int nparts, const double* partials, double* out, hipStream_t s);
// The sum written to hist[*counter], counter incremented on the device (graph-replayable).
hipError_t launch_partials_sum_counted(int nparts, const double* partials, double* hist, int* counter, int cap,
                                      hipStream_t s);
// y[0..n) = v (a kernel, so it may sit inside a captured graph).
hipError_t launch_fill(long long n, double v, double* y, hipStream_t s);
// Krylov vector operations on flat vectors of n doubles; the dot products use kDotBlocks (512) block
// partials per vector, summed in a fixed order.
hipError_t launch_multidot(long long n, const double* V, int nv, const double* w, double* partials, double* out,
                           hipStream_t s);
hipError_t launch_combine(long long n, double a, const double* V, int nv, const double* c, double* y, hipStream_t s);
hipError_t launch_axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s);
}  // namespace aero_host
