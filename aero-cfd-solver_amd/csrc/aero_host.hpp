// aero_host.hpp -- the device mesh handle behind the opaque aero_hip_mesh of include/aero_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "aero_hip.h"
#include "aero_internal.hpp"

struct AdjointData;  // tangent / adjoint workspace, built on first use (aero_adjoint.cpp)

struct aero_hip_mesh {
  int N = 0, F = 0, BC = 256, nblk = 0, S = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<int> perm;   // internal -> reference cell
  int* d_perm = nullptr;   // the same on the device (the device-pointer entry points)
  std::vector<int> iperm;  // reference -> internal cell
  std::vector<void*> allocs;
  aero_dev::DevMesh dm{};
  double* U[3] = {nullptr, nullptr, nullptr};
  int cur = 0, base = 0;
  double* G = nullptr;
  double* R = nullptr;
  double* lam = nullptr;
  int* diag = nullptr;
  double* partials = nullptr;
  double* dt_scratch = nullptr;
  double* iter_out = nullptr;
  double* cp = nullptr;
  aero_dev::WallArgs wall{};
  double* h_iter = nullptr;  // pinned [64]: iteration scalars [0..9], counters (int) at [10..12], diagnostics [16..]
  std::vector<hipEvent_t> tev;  // per-kernel timing events (3 per residual evaluation)
  int tcap = 0, tcount = 0;
  // Host copies in the reference numbering (unit normals as renormalised at setup), for host-side
  // diagnostics and the tangent / adjoint setup.
  std::vector<double> h_nx, h_ny, h_area, h_fx, h_fy, h_vol, h_ccx, h_ccy;
  std::vector<int> h_owner, h_neighbor, h_bc;  // h_bc: AERO_BC_* per face
  int wall_start = 0, wall_count = 0;
  AdjointData* adj = nullptr;
  double* lim_ratio = nullptr;  // frozen limiter ratios [8][F] (aero_hip_freeze_limiter)
  unsigned short* lim_slot = nullptr;  // the same per face slot [S][8] in units of 2^-15 (flux kernels)
  bool lim_valid = false;
  // face-coloured layout of the residual evaluations (AERO_ACCUM_COLOUR)
  bool colour = false;
  int ncolours = 0;
  std::vector<int> colour_off;  // [ncolours + 1] offsets into cfaces
  aero_dev::ColourFaces cfaces{};
  double probe_ms[4] = {-1.0, -1.0, -1.0, -1.0};  // layout probe (block_cells = 0): gather 128/256/512, colour
};

namespace aero_hostutil {
// The per-evaluation options of an aero_residual_config.
aero_dev::FaceOpts face_opts(const aero_residual_config* c);
// face_opts plus the handle's frozen limiter ratios when the config asks for them; fails (AERO_EINVAL,
// message in err) when none were captured.
int mesh_face_opts(const aero_hip_mesh* m, const aero_residual_config* c, aero_dev::FaceOpts* o, char* err,
                   size_t errlen);
// Host [N][5] conserved (reference numbering) <-> device SoA [4][N] (internal numbering).
int to_internal_state(aero_hip_mesh* m, const double* conserved, std::vector<double>* soa);
void from_internal4(const aero_hip_mesh* m, const std::vector<double>& soa, double* aos5);
// Releases the tangent / adjoint workspace (aero_adjoint.cpp).
void adjoint_free(aero_hip_mesh* m);
}  // namespace aero_hostutil
