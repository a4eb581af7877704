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

// Device view of a mesh in internal (block) order.  Passed by value to kernels.
struct DevMesh {
  int N;         // cells
  int ld;        // leading dimension of SoA cell arrays (== N)
  int nblk;      // blocks of BC cells
  int BC;        // cells per block == threads per workgroup
  int smax;      // max slots of any block (LDS sizing)
  int grid;      // workgroups launched (>= nblk when XCD-remapped)
  int xcd_remap;
  const int* slot_off;  // [nblk + 1]
  const int* s_o;       // [S] owner (internal id)
  const int* s_n;       // [S] neighbour (internal id) or -1
  const int* s_meta;    // [S] bc | primary
  const int* s_face;    // [S] reference face index
  const double* s_nx;   // [S]
  const double* s_ny;
  const double* s_A;
  const double* s_fx;   // face centre
  const double* s_fy;
  const int* c_off;     // [N + 1] per-cell entry offsets
  const int* c_ent;     // (local slot << 1) | is_neighbour, increasing reference face index
  const double* vol;    // [N]
  const double* ccx;    // [N] cell centre
  const double* ccy;
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

}  // namespace aero_dev

namespace aero_host {
hipError_t launch_gradient(const aero_dev::DevMesh& m, const double* U, double* G, double g, hipStream_t s);
hipError_t launch_flux(const aero_dev::DevMesh& m, const aero_dev::FaceOpts& o, int scheme, int limiter, bool second,
                       int mode, const double* U, const double* G, double* R, double* lam, int* diag,
                       const aero_dev::StageArgs& st, hipStream_t s);
hipError_t launch_global_update(const aero_dev::DevMesh& m, const aero_dev::FaceOpts& o, const double* U,
                                const double* R, const double* lam, double cfl, double* dt_scratch, int* diag,
                                const aero_dev::StageArgs& st, hipStream_t s);
hipError_t launch_norm_partials(const aero_dev::DevMesh& m, const double* R, double* partials, hipStream_t s);
hipError_t launch_iter_reduce(const aero_dev::DevMesh& m, const double* partials, const aero_dev::WallArgs& wa,
                              const double* U, double g, double* out, double* cp, hipStream_t s);
}  // namespace aero_host
