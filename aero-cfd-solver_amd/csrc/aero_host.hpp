// aero_host.hpp -- the device mesh handle behind the opaque aero_hip_mesh of include/aero_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "aero_internal.hpp"

struct aero_hip_mesh {
  int N = 0, F = 0, BC = 256, nblk = 0, S = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<int> perm;   // internal -> reference cell
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
  double* h_iter = nullptr;  // pinned: 9 doubles + 4 ints
  std::vector<hipEvent_t> tev;  // per-kernel timing events (3 per residual evaluation)
  int tcap = 0, tcount = 0;
  std::vector<double> h_nx, h_ny, h_area, h_fc;  // for host-side diagnostics
  std::vector<int> h_owner, h_neighbor;
  int wall_start = 0, wall_count = 0;
};
