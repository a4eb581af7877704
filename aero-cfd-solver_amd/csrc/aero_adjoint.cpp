// aero_adjoint.cpp -- host side of the tangent / adjoint entry points of include/aero_hip.h.
//
// The workspace (faces in reference order with internal cell ids, per-cell face lists, least-squares
// weights, face Jacobians and iteration buffers) is built on the first derivative call on a handle and
// lives until the handle is destroyed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "aero_hip.h"
#include "aero_host.hpp"
#include "aero_internal.hpp"

using aero_dev::AdjBatchCtl;
using aero_dev::AdjBatchSolve;
using aero_dev::AdjMesh;
using aero_dev::Dual;
using aero_dev::FaceOpts;
using aero_dev::FaceOptsT;
using aero_dev::GeoTan;
using aero_dev::kEntryTile;
using aero_dev::kJacCols;
using namespace aero_hostutil;

struct AdjointData {
  AdjMesh am{};
  std::vector<void*> allocs;
  int nparts = 0;
  size_t nslots = 0;       // face-entry slots (kEntryTile-cell tiles)
  double* G2 = nullptr;    // [8][N] gradient values of the tangent evaluation
  double* dG = nullptr;    // [8][N]
  double* dU = nullptr;    // [4][N]
  double* fa = nullptr;    // [8][F]
  double* R2 = nullptr;    // [4][N]
  double* dR = nullptr;    // [4][N]
  double* J = nullptr;     // [tiles][kJacCols][kEntryTile] face Jacobians per face-entry slot
  bool jfact = false;      // J in the factored form (jacobian_factored of its linearisation)
  double* Gbar = nullptr;  // [8][N]
  double* out = nullptr;   // [4][N]
  double* psi = nullptr;   // [4][N]
  double* x[3] = {nullptr, nullptr, nullptr};  // iteration ring [4][N] each
  double* rhs = nullptr;   // [4][N]
  double* partials = nullptr;
  double* hist = nullptr;  // [hist_cap] squared residual norms
  int hist_cap = 0;
  int* hist_ctr = nullptr;  // device-side index into hist
  double* coef = nullptr;  // [wall_count]
  double* geo = nullptr;   // [5F + 3N] geometry tangents
  // GMRES(m): Krylov basis [m + 1][4N], work vectors, dot-product scratch
  double* V = nullptr;
  int v_cap = 0;
  double* gx = nullptr;
  double* gc = nullptr;
  double* gw = nullptr;
  double* dots = nullptr;      // [m + 1]
  double* dot_parts = nullptr;  // [512][m + 1]
};

namespace {

void set_err(char* err, size_t errlen, const std::string& msg) {
  if (err && errlen > 0) {
    std::strncpy(err, msg.c_str(), errlen - 1);
    err[errlen - 1] = '\0';
  }
}

#define AERO_TRY(expr, ctx)                                                              \
  do {                                                                                   \
    const hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                              \
      set_err(err, errlen, std::string(ctx) + ": " + hipGetErrorString(e_));             \
      return AERO_EHIP;                                                                  \
    }                                                                                    \
  } while (0)

template <class T> hipError_t aalloc(AdjointData* a, T** p, size_t n) {
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) {
    a->allocs.push_back(q);
    *p = static_cast<T*>(q);
  }
  return e;
}

template <class T> hipError_t aupload(AdjointData* a, T** p, const std::vector<T>& v) {
  hipError_t e = aalloc(a, p, v.size());
  if (e == hipSuccess && !v.empty()) e = hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

// Builds the workspace of handle m (adjoint_setup releases it again when any step fails).
int adjoint_setup_steps(aero_hip_mesh* m, char* err, size_t errlen) {
  const int N = m->N, F = m->F;
  AdjointData* a = new AdjointData();
  m->adj = a;
  std::vector<int> fo(F), fn(F);
  std::vector<double> wdx(F, 0.0), wdy(F, 0.0);
  for (int f = 0; f < F; ++f) {
    fo[f] = m->iperm[m->h_owner[f]];
    const int n = m->h_neighbor[f];
    fn[f] = n >= 0 ? m->iperm[n] : -1;
    if (n >= 0) {  // least-squares face weights, as the gradient kernel forms them
      const int o = m->h_owner[f];
      const double dx = m->h_ccx[n] - m->h_ccx[o];
      const double dy = m->h_ccy[n] - m->h_ccy[o];
      const double w = 1.0 / (dx * dx + dy * dy);
      wdx[f] = w * dx;
      wdy[f] = w * dy;
    }
  }
  // Per internal cell: its faces in increasing reference index.
  std::vector<int> cf_off(N + 1, 0);
  for (int f = 0; f < F; ++f) {
    ++cf_off[fo[f] + 1];
    if (fn[f] >= 0) ++cf_off[fn[f] + 1];
  }
  for (int c = 0; c < N; ++c) cf_off[c + 1] += cf_off[c];
  std::vector<int> cf_ent(cf_off[N]);
  {
    std::vector<int> fill(cf_off.begin(), cf_off.end() - 1);
    for (int f = 0; f < F; ++f) {
      cf_ent[fill[fo[f]]++] = f << 1;
      if (fn[f] >= 0) cf_ent[fill[fn[f]]++] = (f << 1) | 1;
    }
  }
  // Least-squares inverse normal matrices, accumulated per cell in face order like the gradient kernel.
  std::vector<double> H(3 * static_cast<size_t>(N), 0.0);
  for (int c = 0; c < N; ++c) {
    double aa = 0.0, bb = 0.0, cc = 0.0;
    for (int e = cf_off[c]; e < cf_off[c + 1]; ++e) {
      const int f = cf_ent[e] >> 1;
      if (fn[f] < 0) continue;
      const int o = m->h_owner[f], n = m->h_neighbor[f];
      const double dx = m->h_ccx[n] - m->h_ccx[o];
      const double dy = m->h_ccy[n] - m->h_ccy[o];
      aa += wdx[f] * dx;
      bb += wdx[f] * dy;
      cc += wdy[f] * dy;
    }
    const double det = aa * cc - bb * bb;
    const double tr = aa + cc;
    if (det > 1.0e-12 * (tr * tr)) {
      const double inv = 1.0 / det;
      H[c] = cc * inv;
      H[N + c] = -bb * inv;
      H[2 * static_cast<size_t>(N) + c] = aa * inv;
    }
  }
  // Face entries in tiles of kEntryTile cells (the AdjMesh e_* layout): per group of cells as many tiles as
  // its largest face count; per slot the other cell and side, the gradient-transpose face weights and the
  // reconstruction vector; per face the slots of its two sides.
  const int ngroups = (N + kEntryTile - 1) / kEntryTile;
  std::vector<int> e_tile(ngroups + 1, 0);
  for (int b = 0; b < ngroups; ++b) {
    int deg = 0;
    for (int c = b * kEntryTile; c < std::min(N, (b + 1) * kEntryTile); ++c)
      deg = std::max(deg, cf_off[c + 1] - cf_off[c]);
    e_tile[b + 1] = e_tile[b] + deg;
  }
  if (static_cast<long long>(e_tile[ngroups]) * kEntryTile > 0x7fffffffLL) {
    set_err(err, errlen, "adjoint setup: too many face entries for 32-bit slot indices");
    return AERO_EINVAL;
  }
  const size_t nslots = static_cast<size_t>(e_tile[ngroups]) * kEntryTile;
  std::vector<int> e_nb(nslots, -2), f_slot(2 * static_cast<size_t>(F), -1);
  std::vector<double> e_gx(nslots, 0.0), e_gy(nslots, 0.0), e_dx(nslots, 0.0), e_dy(nslots, 0.0);
  for (int c = 0; c < N; ++c) {
    const int rc = m->perm[c];
    for (int e = cf_off[c]; e < cf_off[c + 1]; ++e) {
      const int f = cf_ent[e] >> 1, side = cf_ent[e] & 1;
      const size_t sl =
          static_cast<size_t>(e_tile[c / kEntryTile] + (e - cf_off[c])) * kEntryTile + c % kEntryTile;
      const int other = side ? fo[f] : fn[f];
      e_nb[sl] = other >= 0 ? (other << 1) | side : -2;
      if (m->dm.grad == 1) {
        e_gx[sl] = wdx[f];
        e_gy[sl] = wdy[f];
      } else {
        e_gx[sl] = m->h_nx[f] * m->h_area[f];
        e_gy[sl] = m->h_ny[f] * m->h_area[f];
      }
      e_dx[sl] = m->h_fx[f] - m->h_ccx[rc];
      e_dy[sl] = m->h_fy[f] - m->h_ccy[rc];
      f_slot[static_cast<size_t>(side) * F + f] = static_cast<int>(sl);
    }
  }
  a->nslots = nslots;
  int *d_etile = nullptr, *d_enb = nullptr, *d_fslot = nullptr;
  double *d_egx = nullptr, *d_egy = nullptr, *d_edx = nullptr, *d_edy = nullptr;
  AERO_TRY(aupload(a, &d_etile, e_tile), "adjoint setup");
  AERO_TRY(aupload(a, &d_enb, e_nb), "adjoint setup");
  AERO_TRY(aupload(a, &d_fslot, f_slot), "adjoint setup");
  AERO_TRY(aupload(a, &d_egx, e_gx), "adjoint setup");
  AERO_TRY(aupload(a, &d_egy, e_gy), "adjoint setup");
  AERO_TRY(aupload(a, &d_edx, e_dx), "adjoint setup");
  AERO_TRY(aupload(a, &d_edy, e_dy), "adjoint setup");
  const size_t n4 = 4 * static_cast<size_t>(N), n8 = 8 * static_cast<size_t>(N);
  int *d_fo = nullptr, *d_fn = nullptr, *d_bc = nullptr, *d_off = nullptr, *d_ent = nullptr;
  double *d_nx = nullptr, *d_ny = nullptr, *d_A = nullptr, *d_fx = nullptr, *d_fy = nullptr, *d_wx = nullptr,
         *d_wy = nullptr, *d_H = nullptr;
  AERO_TRY(aupload(a, &d_fo, fo), "adjoint setup");
  AERO_TRY(aupload(a, &d_fn, fn), "adjoint setup");
  AERO_TRY(aupload(a, &d_bc, m->h_bc), "adjoint setup");
  AERO_TRY(aupload(a, &d_nx, m->h_nx), "adjoint setup");
  AERO_TRY(aupload(a, &d_ny, m->h_ny), "adjoint setup");
  AERO_TRY(aupload(a, &d_A, m->h_area), "adjoint setup");
  AERO_TRY(aupload(a, &d_fx, m->h_fx), "adjoint setup");
  AERO_TRY(aupload(a, &d_fy, m->h_fy), "adjoint setup");
  AERO_TRY(aupload(a, &d_wx, wdx), "adjoint setup");
  AERO_TRY(aupload(a, &d_wy, wdy), "adjoint setup");
  AERO_TRY(aupload(a, &d_off, cf_off), "adjoint setup");
  AERO_TRY(aupload(a, &d_ent, cf_ent), "adjoint setup");
  AERO_TRY(aupload(a, &d_H, H), "adjoint setup");
  AdjMesh& am = a->am;
  am.N = N;
  am.F = F;
  am.ld = m->dm.ld;
  am.grad = m->dm.grad;
  am.fo = d_fo;
  am.fn = d_fn;
  am.fbc = d_bc;
  am.fnx = d_nx;
  am.fny = d_ny;
  am.fA = d_A;
  am.ffx = d_fx;
  am.ffy = d_fy;
  am.fwdx = d_wx;
  am.fwdy = d_wy;
  am.cf_off = d_off;
  am.cf_ent = d_ent;
  am.vol = m->dm.vol;
  am.ccx = m->dm.ccx;
  am.ccy = m->dm.ccy;
  am.H = d_H;
  am.e_tile = d_etile;
  am.e_nb = d_enb;
  am.e_gx = d_egx;
  am.e_gy = d_egy;
  am.e_dx = d_edx;
  am.e_dy = d_edy;
  am.f_slot = d_fslot;
  am.wall_start = m->wall_start;
  am.wall_count = m->wall_count;
  a->nparts = (N + 255) / 256;
  AERO_TRY(aalloc(a, &a->G2, n8), "adjoint setup");
  AERO_TRY(aalloc(a, &a->dG, n8), "adjoint setup");
  AERO_TRY(aalloc(a, &a->dU, n4), "adjoint setup");
  AERO_TRY(aalloc(a, &a->fa, 8 * static_cast<size_t>(F)), "adjoint setup");
  AERO_TRY(aalloc(a, &a->R2, n4), "adjoint setup");
  AERO_TRY(aalloc(a, &a->dR, n4), "adjoint setup");
  AERO_TRY(aalloc(a, &a->Gbar, n8), "adjoint setup");
  AERO_TRY(aalloc(a, &a->out, n4), "adjoint setup");
  AERO_TRY(aalloc(a, &a->psi, n4), "adjoint setup");
  for (int i = 0; i < 3; ++i) AERO_TRY(aalloc(a, &a->x[i], n4), "adjoint setup");
  AERO_TRY(aalloc(a, &a->rhs, n4), "adjoint setup");
  AERO_TRY(aalloc(a, &a->partials, static_cast<size_t>(a->nparts)), "adjoint setup");
  AERO_TRY(aalloc(a, &a->hist_ctr, 1), "adjoint setup");
  AERO_TRY(aalloc(a, &a->coef, static_cast<size_t>(std::max(m->wall_count, 1))), "adjoint setup");
  AERO_TRY(aalloc(a, &a->geo, 5 * static_cast<size_t>(F) + 3 * static_cast<size_t>(N)), "adjoint setup");
  return AERO_OK;
}

int adjoint_setup(aero_hip_mesh* m, char* err, size_t errlen) {
  if (m->adj) return AERO_OK;
  const int rc = adjoint_setup_steps(m, err, errlen);
  if (rc) adjoint_free(m);
  return rc;
}

// The stored face Jacobians (kJacCols doubles per face-entry slot, the largest adjoint buffer), on demand.
int ensure_jacobian_buffers(aero_hip_mesh* m, char* err, size_t errlen) {
  AdjointData* a = m->adj;
  if (!a->J) AERO_TRY(aalloc(a, &a->J, kJacCols * a->nslots), "face Jacobian buffer");
  return AERO_OK;
}

int ensure_history(aero_hip_mesh* m, int n, char* err, size_t errlen) {
  AdjointData* a = m->adj;
  if (a->hist_cap >= n) return AERO_OK;
  if (a->hist) {
    (void)hipFree(a->hist);
    a->allocs.erase(std::find(a->allocs.begin(), a->allocs.end(), static_cast<void*>(a->hist)));
    a->hist = nullptr;
  }
  AERO_TRY(aalloc(a, &a->hist, static_cast<size_t>(n)), "history buffer");
  a->hist_cap = n;
  return AERO_OK;
}

int check_cfg(const aero_residual_config* cfg, const char* who, char* err, size_t errlen) {
  if (!cfg) {
    set_err(err, errlen, std::string(who) + ": null config");
    return AERO_EINVAL;
  }
  if (cfg->flux_scheme < 0 || cfg->flux_scheme > 1 || cfg->limiter < 0 || cfg->limiter > 1) {
    set_err(err, errlen, std::string(who) + ": unknown flux scheme or limiter");
    return AERO_EINVAL;
  }
  return AERO_OK;
}

// Loads `conserved` as the handle's state when given.
int load_state(aero_hip_mesh* m, const double* conserved, char* err, size_t errlen) {
  if (!conserved) return AERO_OK;
  return aero_hip_set_state(m, conserved, err, errlen);
}

int upload_internal(aero_hip_mesh* m, const double* aos5, double* dst, char* err, size_t errlen) {
  std::vector<double> soa;
  to_internal_state(m, aos5, &soa);
  AERO_TRY(hipMemcpyAsync(dst, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, m->stream),
           "upload");
  AERO_TRY(hipStreamSynchronize(m->stream), "upload");
  return AERO_OK;
}

int download_internal(aero_hip_mesh* m, const double* src, double* aos5, char* err, size_t errlen) {
  std::vector<double> soa(4 * static_cast<size_t>(m->N));
  AERO_TRY(hipMemcpyAsync(soa.data(), src, soa.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream),
           "download");
  AERO_TRY(hipStreamSynchronize(m->stream), "download");
  from_internal4(m, soa, aos5);
  return AERO_OK;
}

FaceOptsT<Dual> dual_opts(const FaceOpts& o, const aero_tangent_seed* seed) {
  FaceOptsT<Dual> d;
  const double z[5] = {0, 0, 0, 0, 0};
  const double* df = seed ? seed->d_farfield : z;
  d.ff.r = Dual(o.ff.r, df[0]);
  d.ff.ru = Dual(o.ff.ru, df[1]);
  d.ff.rv = Dual(o.ff.rv, df[2]);
  d.ff.E = Dual(o.ff.E, df[4]);
  d.g = o.g;
  d.low_mach_fix = o.low_mach_fix;
  d.low_mach_cutoff = o.low_mach_cutoff;
  d.stab_floor = Dual(o.stab_floor, seed ? seed->d_stab_floor : 0.0);
  d.char_bc = o.char_bc;
  d.precond = o.precond;
  d.mach_ref = Dual(o.mach_ref, seed ? seed->d_mach_ref : 0.0);
  d.mach_min = Dual(o.mach_min, seed ? seed->d_mach_min : 0.0);
  d.beta_min = Dual(o.beta_min, seed ? seed->d_beta_min : 0.0);
  d.beta_max = Dual(o.beta_max, seed ? seed->d_beta_max : 0.0);
  d.as_fix = o.as_fix;
  d.as_cut = o.as_cut;
  d.as_fmin = o.as_fmin;
  d.as_w = o.as_w;
  d.frozen = o.frozen;
  d.frozen_slots = o.frozen_slots;
  return d;
}

// Geometry tangents of `seed` on the device (cell arrays in internal order), or all-null.
int geometry_tangent(aero_hip_mesh* m, const aero_tangent_seed* seed, GeoTan* gt, char* err, size_t errlen) {
  *gt = GeoTan{};
  if (!seed) return AERO_OK;
  const bool any = seed->d_face_normal || seed->d_face_center || seed->d_face_area || seed->d_cell_center ||
                   seed->d_cell_volume;
  if (!any) return AERO_OK;
  const int N = m->N, F = m->F;
  std::vector<double> h(5 * static_cast<size_t>(F) + 3 * static_cast<size_t>(N), 0.0);
  double* nx = h.data();
  double* ny = nx + F;
  double* A = ny + F;
  double* fx = A + F;
  double* fy = fx + F;
  double* cx = fy + F;
  double* cy = cx + N;
  double* vol = cy + N;
  for (int f = 0; f < F; ++f) {
    if (seed->d_face_normal) {
      nx[f] = seed->d_face_normal[3 * f];
      ny[f] = seed->d_face_normal[3 * f + 1];
    }
    if (seed->d_face_area) A[f] = seed->d_face_area[f];
    if (seed->d_face_center) {
      fx[f] = seed->d_face_center[3 * f];
      fy[f] = seed->d_face_center[3 * f + 1];
    }
  }
  for (int c = 0; c < N; ++c) {
    const int r = m->perm[c];
    if (seed->d_cell_center) {
      cx[c] = seed->d_cell_center[3 * r];
      cy[c] = seed->d_cell_center[3 * r + 1];
    }
    if (seed->d_cell_volume) vol[c] = seed->d_cell_volume[r];
  }
  double* g = m->adj->geo;
  AERO_TRY(hipMemcpyAsync(g, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, m->stream), "geometry tangent");
  AERO_TRY(hipStreamSynchronize(m->stream), "geometry tangent");
  gt->dnx = g;
  gt->dny = g + F;
  gt->dA = g + 2 * static_cast<size_t>(F);
  gt->dfx = g + 3 * static_cast<size_t>(F);
  gt->dfy = g + 4 * static_cast<size_t>(F);
  gt->dccx = g + 5 * static_cast<size_t>(F);
  gt->dccy = gt->dccx + N;
  gt->dvol = gt->dccy + N;
  return AERO_OK;
}

// dR = R_U dU + R_p seed at the handle's state (dU on the device, may be null), into a->dR (and the
// residual into a->R2).
int enqueue_jvp(aero_hip_mesh* m, const aero_residual_config* cfg, const double* dU, const GeoTan& gt,
                const FaceOptsT<Dual>& o, char* err, size_t errlen) {
  AdjointData* a = m->adj;
  const double* U = m->U[m->cur];
  const size_t n8 = 8 * static_cast<size_t>(m->N);
  if (cfg->second_order) {
    AERO_TRY(aero_host::launch_tan_gradient(a->am, U, dU, gt, cfg->gamma, a->G2, a->dG, m->stream), "k_tan_gradient");
  } else {  // fill kernels rather than memsets: this sequence may be captured in a graph
    AERO_TRY(aero_host::launch_fill(static_cast<long long>(n8), 0.0, a->G2, m->stream), "zero gradient");
    AERO_TRY(aero_host::launch_fill(static_cast<long long>(n8), 0.0, a->dG, m->stream), "zero gradient");
  }
  AERO_TRY(aero_host::launch_tan_face(a->am, cfg->flux_scheme, cfg->limiter, cfg->second_order != 0, U, dU, a->G2,
                                      a->dG, gt, o, a->fa, m->stream),
           "k_tan_face");
  AERO_TRY(aero_host::launch_tan_gather(a->am, a->fa, a->R2, a->dR, m->stream), "k_tan_gather");
  return AERO_OK;
}

// Primal gradient + face Jacobians at the handle's state.
int enqueue_jacobian(aero_hip_mesh* m, const aero_residual_config* cfg, char* err, size_t errlen) {
  AdjointData* a = m->adj;
  int rc = ensure_jacobian_buffers(m, err, errlen);
  if (rc) return rc;
  const double* U = m->U[m->cur];
  FaceOpts o;
  rc = mesh_face_opts(m, cfg, &o, err, errlen);
  if (rc) return rc;
  if (cfg->second_order)
    AERO_TRY(aero_host::launch_gradient(m->dm, U, m->G, o.g, m->stream), "k_gradient");
  AERO_TRY(aero_host::launch_face_jacobian(a->am, cfg->flux_scheme, cfg->limiter, cfg->second_order != 0, U, m->G, o,
                                           a->J, m->stream),
           "k_face_jacobian");
  a->jfact = aero_dev::jacobian_factored(cfg->second_order != 0, o.frozen != nullptr);
  return AERO_OK;
}

// Spectral radius sums lambda at the handle's state (one residual evaluation), into m->lam.
int enqueue_spectral_radius(aero_hip_mesh* m, const aero_residual_config* cfg, char* err, size_t errlen) {
  return aero_hip_residual_async(m, cfg, err, errlen);
}

Dual dcos(const Dual& a) { return Dual(std::cos(a.v), -std::sin(a.v) * a.d); }
Dual dsin(const Dual& a) { return Dual(std::sin(a.v), std::cos(a.v) * a.d); }
Dual dclampd(const Dual& v, double lo, double hi) { return aero_dev::dclamp(v, Dual(lo), Dual(hi)); }

}  // namespace

namespace aero_hostutil {
void adjoint_free(aero_hip_mesh* m) {
  if (!m || !m->adj) return;
  for (void* p : m->adj->allocs) (void)hipFree(p);
  delete m->adj;
  m->adj = nullptr;
}
}  // namespace aero_hostutil

extern "C" {

int aero_case_linearization(const aero_case_config* c, int32_t iter, int32_t param, aero_residual_config* rc,
                            aero_tangent_seed* seed, aero_objective* obj, double* d_q_inf, char* err, size_t errlen) {
  if (!c || !rc) {
    set_err(err, errlen, "aero_case_linearization: null argument");
    return AERO_EINVAL;
  }
  if (param < AERO_PARAM_NONE || param > AERO_PARAM_MACH) {
    set_err(err, errlen, "aero_case_linearization: unknown parameter");
    return AERO_EINVAL;
  }
  if (c->all_speed_staged_controller) {
    set_err(err, errlen, "aero_case_linearization: the staged all-speed controller has no fixed schedule");
    return AERO_EINVAL;
  }
  using aero_dev::dmax;
  using aero_dev::dmin;
  const double g = c->gamma;
  const Dual M(c->mach, param == AERO_PARAM_MACH ? 1.0 : 0.0);
  const Dual aoa(c->aoa_deg, param == AERO_PARAM_AOA ? 1.0 : 0.0);
  const bool precond = c->precond_on && c->precond_experimental_enabled;
  // build_precondition_config (euler_solver.cpp:341-355)
  const Dual minf = dmax(M, Dual(1.0e-4));
  const Dual auto_bmin = dmax(Dual(1.0e-4), dmin(Dual(1.0e-3), 0.1 * minf));
  const Dual mach_ref = dmax(c->precond_mach_ref > 0.0 ? Dual(c->precond_mach_ref) : minf, Dual(1.0e-4));
  const Dual mach_min = dmax(Dual(c->precond_mach_min > 0.0 ? c->precond_mach_min : 1.0e-6), Dual(1.0e-6));
  const Dual beta_min = dmax(c->precond_beta_min > 0.0 ? Dual(c->precond_beta_min) : auto_bmin, Dual(1.0e-4));
  const Dual beta_max = dmax(Dual(c->precond_beta_max), beta_min);
  // stabilisation floor schedule (:95-106, :141-157)
  const Dual k0 = c->stab_floor_k_start >= 0.0 ? Dual(c->stab_floor_k_start) : Dual(c->stab_floor_start) / minf;
  const Dual k1 = c->stab_floor_k_target >= 0.0 ? Dual(c->stab_floor_k_target) : Dual(c->stab_floor_target) / minf;
  const Dual floor_start = dclampd(dmax(k0, Dual(0.0)) * minf, 0.0, 1.0);
  const Dual floor_target = dclampd(dmax(k1, Dual(0.0)) * minf, 0.0, 1.0);
  const int stab_ramp = std::max(c->stab_ramp_iters, 0);
  Dual floor_cur;
  {
    const Dual s0 = dclampd(floor_start, 0.0, 1.0), s1 = dclampd(floor_target, 0.0, 1.0);
    if (stab_ramp <= 0) {
      floor_cur = s1;
    } else {
      const double xr = static_cast<double>(std::max(iter, 0)) / static_cast<double>(std::max(stab_ramp, 1));
      const double x = xr < 0.0 ? 0.0 : (xr > 1.0 ? 1.0 : xr);
      floor_cur = s0 + x * (s1 - s0);
    }
  }
  // freestream (:543-561)
  double rho_inf = c->rho_inf;
  if (rho_inf <= 0.0) rho_inf = c->p_inf / (c->gas_constant * c->t_inf);
  if (rho_inf <= 0.0) {
    set_err(err, errlen, "Invalid freestream density. Set rho_inf or p_inf/T_inf.");
    return AERO_EINVAL;
  }
  const double pi = 3.14159265358979323846;
  const Dual alpha = aoa * (pi / 180.0);
  const double a_inf = std::sqrt(g * c->p_inf / rho_inf);
  const Dual v_inf = dmax(M, Dual(0.0)) * a_inf;
  const Dual u_inf = v_inf * dcos(alpha), vv_inf = v_inf * dsin(alpha);
  const double r = std::max(rho_inf, 1.0e-8);
  const Dual ke = 0.5 * r * (u_inf * u_inf + vv_inf * vv_inf + 0.0 * 0.0);
  const Dual ff[5] = {Dual(r), r * u_inf, r * vv_inf, Dual(r * 0.0), std::max(c->p_inf, 1.0e-8) / (g - 1.0) + ke};
  std::memset(rc, 0, sizeof(*rc));
  for (int k = 0; k < 5; ++k) rc->farfield[k] = ff[k].v;
  rc->gamma = g;
  rc->flux_scheme = c->flux_scheme;
  rc->limiter = c->limiter;
  rc->second_order = iter >= std::max(c->startup_first_order_iters, 0) ? 1 : 0;
  rc->limiter_frozen = (rc->second_order && c->limiter_freeze_iter >= 0 && iter >= c->limiter_freeze_iter) ? 1 : 0;
  rc->low_mach_fix = c->low_mach_fix;
  rc->low_mach_cutoff = c->mach_cutoff;
  rc->stab_mach_floor = floor_cur.v;
  rc->precond_farfield_bc = c->precond_farfield_bc;
  rc->precond_enabled = precond ? 1 : 0;
  rc->precond_mach_ref = mach_ref.v;
  rc->precond_mach_min = mach_min.v;
  rc->precond_beta_min = beta_min.v;
  rc->precond_beta_max = beta_max.v;
  rc->all_speed_fix = c->all_speed_flux_fix;
  rc->all_speed_cutoff = c->all_speed_mach_cutoff;
  {
    const double fmin = c->all_speed_f_min;
    rc->all_speed_f_min = fmin < 1.0e-6 ? 1.0e-6 : (fmin > 1.0 ? 1.0 : fmin);
    const int as_on = std::max(c->all_speed_ramp_start_iter, 0);
    const int as_len = std::max(c->all_speed_ramp_iters, 0);
    if (as_len <= 0) {
      rc->all_speed_ramp = iter >= as_on ? 1.0 : 0.0;
    } else {
      const double t = static_cast<double>(iter - as_on) / static_cast<double>(std::max(as_len, 1));
      rc->all_speed_ramp = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    }
  }
  if (seed) {
    std::memset(seed, 0, sizeof(*seed));
    for (int k = 0; k < 5; ++k) seed->d_farfield[k] = ff[k].d;
    seed->d_stab_floor = floor_cur.d;
    seed->d_mach_ref = mach_ref.d;
    seed->d_mach_min = mach_min.d;
    seed->d_beta_min = beta_min.d;
    seed->d_beta_max = beta_max.d;
  }
  const Dual q_inf = 0.5 * rho_inf * v_inf * v_inf + 1.0e-12;
  if (obj) {
    std::memset(obj, 0, sizeof(*obj));
    obj->aoa_deg = c->aoa_deg;
    obj->q_inf = q_inf.v;
    obj->p_inf = c->p_inf;
    obj->chord = c->chord;
    obj->x_ref = c->x_ref;
    obj->y_ref = c->y_ref;
    obj->gamma = g;
  }
  if (d_q_inf) *d_q_inf = q_inf.d;
  return AERO_OK;
}

int aero_hip_residual_jvp(aero_hip_mesh* m, const aero_residual_config* cfg, const double* conserved,
                          const double* dU, const aero_tangent_seed* seed, double* residual, double* dR, char* err,
                          size_t errlen) {
  if (!m || !dR) {
    set_err(err, errlen, "aero_hip_residual_jvp: null argument");
    return AERO_EINVAL;
  }
  int rc = check_cfg(cfg, "aero_hip_residual_jvp", err, errlen);
  if (!rc) rc = adjoint_setup(m, err, errlen);
  if (!rc) rc = load_state(m, conserved, err, errlen);
  if (rc) return rc;
  AdjointData* a = m->adj;
  if (dU) {
    rc = upload_internal(m, dU, a->dU, err, errlen);
    if (rc) return rc;
  }
  GeoTan gt;
  FaceOpts fo;
  rc = mesh_face_opts(m, cfg, &fo, err, errlen);
  if (!rc) rc = geometry_tangent(m, seed, &gt, err, errlen);
  if (!rc) rc = enqueue_jvp(m, cfg, dU ? a->dU : nullptr, gt, dual_opts(fo, seed), err, errlen);
  if (!rc && residual) rc = download_internal(m, a->R2, residual, err, errlen);
  if (!rc) rc = download_internal(m, a->dR, dR, err, errlen);
  return rc;
}

int aero_hip_residual_vjp(aero_hip_mesh* m, const aero_residual_config* cfg, const double* conserved,
                          const double* psi, double* out, char* err, size_t errlen) {
  if (!m || !psi || !out) {
    set_err(err, errlen, "aero_hip_residual_vjp: null argument");
    return AERO_EINVAL;
  }
  int rc = check_cfg(cfg, "aero_hip_residual_vjp", err, errlen);
  if (!rc) rc = adjoint_setup(m, err, errlen);
  if (!rc) rc = load_state(m, conserved, err, errlen);
  if (!rc) rc = enqueue_jacobian(m, cfg, err, errlen);
  if (rc) return rc;
  AdjointData* a = m->adj;
  rc = upload_internal(m, psi, a->psi, err, errlen);
  if (rc) return rc;
  AERO_TRY(aero_host::launch_adj_product(a->am, a->J, a->jfact, m->U[m->cur], a->psi, cfg->gamma, a->Gbar, a->out,
                                         m->stream),
           "adjoint product");
  return download_internal(m, a->out, out, err, errlen);
}

int aero_hip_residual_vjp_device(aero_hip_mesh* m, const aero_residual_config* cfg, const double* d_conserved,
                                 const double* d_psi, double* d_out, char* err, size_t errlen) {
  if (!m || !d_conserved || !d_psi || !d_out) {
    set_err(err, errlen, "aero_hip_residual_vjp_device: null argument");
    return AERO_EINVAL;
  }
  int rc = check_cfg(cfg, "aero_hip_residual_vjp_device", err, errlen);
  if (!rc) rc = adjoint_setup(m, err, errlen);
  if (rc) return rc;
  m->cur = m->base = 0;
  AERO_TRY(aero_host::launch_aos_to_internal(m->N, m->d_perm, d_conserved, m->U[0], m->stream),
           "aero_hip_residual_vjp_device: state");
  rc = enqueue_jacobian(m, cfg, err, errlen);
  if (rc) return rc;
  AdjointData* a = m->adj;
  AERO_TRY(aero_host::launch_aos_to_internal(m->N, m->d_perm, d_psi, a->psi, m->stream),
           "aero_hip_residual_vjp_device: psi");
  AERO_TRY(aero_host::launch_adj_product(a->am, a->J, a->jfact, m->U[m->cur], a->psi, cfg->gamma, a->Gbar, a->out,
                                         m->stream),
           "adjoint product");
  AERO_TRY(aero_host::launch_internal_to_aos(m->N, m->d_perm, a->out, d_out, nullptr, nullptr, m->stream),
           "aero_hip_residual_vjp_device: output");
  AERO_TRY(hipStreamSynchronize(m->stream), "aero_hip_residual_vjp_device: sync");
  return AERO_OK;
}

int aero_hip_objective(aero_hip_mesh* m, const aero_objective* obj, const double* conserved, double* J, double* dJdU,
                       double* dJ_daoa, double* dJ_dq, const aero_tangent_seed* seed, double* dJ_seed, char* err,
                       size_t errlen) {
  if (!m || !obj) {
    set_err(err, errlen, "aero_hip_objective: null argument");
    return AERO_EINVAL;
  }
  int rc = adjoint_setup(m, err, errlen);
  if (!rc) rc = load_state(m, conserved, err, errlen);
  if (rc) return rc;
  AdjointData* a = m->adj;
  const int N = m->N;
  // Wall owner pressures from the device state (forces.cpp:220-246 uses the owner cell's p).
  std::vector<double> soa(4 * static_cast<size_t>(N));
  AERO_TRY(hipMemcpyAsync(soa.data(), m->U[m->cur], soa.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream),
           "aero_hip_objective: state");
  AERO_TRY(hipStreamSynchronize(m->stream), "aero_hip_objective: state");
  const double pi = 3.14159265358979323846;
  const double gam = obj->gamma;
  if (!(gam > 1.0)) {
    set_err(err, errlen, "aero_hip_objective: gamma must be > 1");
    return AERO_EINVAL;
  }
  auto pressure = [&](int ic) {
    const aero_dev::C4 s{soa[ic], soa[N + ic], soa[2 * static_cast<size_t>(N) + ic], soa[3 * static_cast<size_t>(N) + ic]};
    const double rho = s.r < 1.0e-8 ? 1.0e-8 : s.r;
    const double inv = 1.0 / rho;
    const double u = s.ru * inv, v = s.rv * inv;
    const double ke = 0.5 * rho * (u * u + v * v);
    const double p = (gam - 1.0) * (s.E - ke);
    return p < 1.0e-8 ? 1.0e-8 : p;
  };
  const Dual alpha = Dual(obj->aoa_deg, 1.0) * (pi / 180.0);
  const Dual ca = dcos(alpha), sa = dsin(alpha);
  const double chord = std::max(obj->chord, 1.0e-8);
  const double sref = chord * 1.0;
  const Dual q(obj->q_inf, 1.0);
  // geometry tangents along `seed` (wall faces only enter)
  const bool geo = seed && (seed->d_face_normal || seed->d_face_area || seed->d_face_center);
  Dual fx(0.0), fy(0.0), mom(0.0);  // derivative part: along the geometry seed
  std::vector<double> coef(std::max(m->wall_count, 1), 0.0);
  for (int i = 0; i < m->wall_count; ++i) {
    const int f = m->wall_start + i;
    const int o = m->h_owner[f];
    const double p = pressure(m->iperm[o]);
    const double gp = p - obj->p_inf;
    const Dual A(m->h_area[f], geo && seed->d_face_area ? seed->d_face_area[f] : 0.0);
    const Dual nx(m->h_nx[f], geo && seed->d_face_normal ? seed->d_face_normal[3 * f] : 0.0);
    const Dual ny(m->h_ny[f], geo && seed->d_face_normal ? seed->d_face_normal[3 * f + 1] : 0.0);
    const Dual x = Dual(m->h_fx[f], geo && seed->d_face_center ? seed->d_face_center[3 * f] : 0.0) - obj->x_ref;
    const Dual y = Dual(m->h_fy[f], geo && seed->d_face_center ? seed->d_face_center[3 * f + 1] : 0.0) - obj->y_ref;
    const Dual nA0 = -(nx * A), nA1 = -(ny * A);
    fx = fx - gp * nA0;
    fy = fy - gp * nA1;
    const Dual dfx = -gp * nA0, dfy = -gp * nA1;
    mom = mom + (x * dfy - y * dfx);
    // d/dp of (CL, CD, CM) through this face
    const double dfx_dp = -nA0.v, dfy_dp = -nA1.v;
    const double dlift = -dfx_dp * sa.v + dfy_dp * ca.v;
    const double ddrag = dfx_dp * ca.v + dfy_dp * sa.v;
    const double dmom = x.v * dfy_dp - y.v * dfx_dp;
    coef[i] = (obj->w_cl * dlift + obj->w_cd * ddrag) / (q.v * sref) + obj->w_cm * dmom / (q.v * sref * chord);
  }
  // Coefficients with derivatives along aoa (through ca, sa) and along the geometry seed.
  const Dual drag_a = fx.v * ca + fy.v * sa, lift_a = -fx.v * sa + fy.v * ca;
  const double J0 = (obj->w_cl * lift_a.v + obj->w_cd * drag_a.v) / (q.v * sref) +
                    obj->w_cm * mom.v / (q.v * sref * chord);
  if (J) *J = J0;
  if (dJ_daoa) *dJ_daoa = (obj->w_cl * lift_a.d + obj->w_cd * drag_a.d) / (q.v * sref);
  if (dJ_dq) *dJ_dq = -J0 / q.v;
  if (dJ_seed) {
    const double dlift = -fx.d * sa.v + fy.d * ca.v;
    const double ddrag = fx.d * ca.v + fy.d * sa.v;
    *dJ_seed = (obj->w_cl * dlift + obj->w_cd * ddrag) / (q.v * sref) + obj->w_cm * mom.d / (q.v * sref * chord);
  }
  if (dJdU) {
    AERO_TRY(hipMemcpyAsync(a->coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice, m->stream),
             "aero_hip_objective: coefficients");
    AERO_TRY(hipMemsetAsync(a->out, 0, 4 * static_cast<size_t>(N) * sizeof(double), m->stream), "aero_hip_objective");
    AERO_TRY(aero_host::launch_objective_gradient(a->am, m->U[m->cur], gam, a->coef, a->out, m->stream),
             "k_objective_gradient");
    rc = download_internal(m, a->out, dJdU, err, errlen);
  }
  return rc;
}

}  // extern "C"

namespace {

// One pseudo-time iteration of the linearised problem on the iterate ring a->x[]: x[start] holds the
// input; on return *start_io indexes the output.  forcing = false drops the right-hand side (adjoint) or
// the parameter derivative (tangent), leaving the linear map G.  partials_out (or null) receives the
// squared norm of the stage-0 linear residual.
int one_iteration(aero_hip_mesh* m, const aero_residual_config* cfg, int forward, bool forcing, const GeoTan& gt,
                  const FaceOptsT<Dual>& od, const aero_linear_solve* ctl, int* start_io, bool record, char* err,
                  size_t errlen) {
  AdjointData* a = m->adj;
  const FaceOpts o = face_opts(cfg);
  const int rk3 = ctl->rk_stages == 3 ? 1 : 0;
  const int nst = rk3 ? 3 : 1;
  const double* U = m->U[m->cur];
  int start = *start_io;
  const double* x0 = a->x[start];
  const double* xs = x0;
  static const GeoTan no_geo{};
  FaceOptsT<Dual> od_plain = od;  // without the parameter derivatives when forcing is off
  if (!forcing) {
    for (Dual* d : {&od_plain.ff.r, &od_plain.ff.ru, &od_plain.ff.rv, &od_plain.ff.E, &od_plain.stab_floor,
                    &od_plain.mach_ref, &od_plain.mach_min, &od_plain.beta_min, &od_plain.beta_max})
      d->d = 0.0;
  }
  for (int st = 0; st < nst; ++st) {
    double* xo = rk3 ? a->x[st == 0 ? (start + 1) % 3 : (st == 1 ? (start + 2) % 3 : start)] : a->x[1 - start];
    const double* r;
    if (forward) {
      const int rc = enqueue_jvp(m, cfg, xs, forcing ? gt : no_geo, od_plain, err, errlen);
      if (rc) return rc;
      r = a->dR;
    } else {  // psi = P^T D xs is current (set by the previous stage, or by the caller before stage 0)
      const bool want = st == 0 && record;
      AERO_TRY(aero_host::launch_adj_stage(a->am, a->J, a->jfact, U, o, st, rk3, x0, xs, forcing ? a->rhs : nullptr, m->lam,
                                           ctl->cfl, a->Gbar, a->out, xo, want ? a->partials : nullptr, a->psi,
                                           m->stream),
               "adjoint stage");
      if (want)
        AERO_TRY(aero_host::launch_partials_sum_counted(a->nparts, a->partials, a->hist, a->hist_ctr, a->hist_cap,
                                                        m->stream),
                 "k_partials_sum");
      xs = xo;
      continue;
    }
    const bool want = st == 0 && record;
    AERO_TRY(aero_host::launch_lin_update(a->am, st, rk3, forward, x0, xs, r, (forward || !forcing) ? nullptr : a->rhs,
                                          m->lam, ctl->cfl, U, o, xo, want ? a->partials : nullptr, m->stream),
             "k_lin_update");
    if (want)
      AERO_TRY(aero_host::launch_partials_sum_counted(a->nparts, a->partials, a->hist, a->hist_ctr, a->hist_cap,
                                                      m->stream),
               "k_partials_sum");
    xs = xo;
  }
  if (!rk3) start = 1 - start;
  *start_io = start;
  return AERO_OK;
}

// Plain pseudo-time iteration.  On return a->x[*start_out] holds the final iterate.
int plain_iterate(aero_hip_mesh* m, const aero_residual_config* cfg, int forward, const GeoTan& gt,
                  const FaceOptsT<Dual>& od, aero_linear_solve* ctl, double* history, int* start_out, char* err,
                  size_t errlen) {
  AdjointData* a = m->adj;
  const int maxit = std::max(ctl->max_iters, 1);
  const int every = std::max(ctl->check_every, 1);
  int rc = ensure_history(m, maxit, err, errlen);
  if (rc) return rc;
  const size_t n4 = 4 * static_cast<size_t>(m->N);
  AERO_TRY(hipMemsetAsync(a->x[0], 0, n4 * sizeof(double), m->stream), "linear solve init");
  AERO_TRY(hipMemsetAsync(a->hist_ctr, 0, sizeof(int), m->stream), "linear solve init");
  // One graph per replay unit: an iteration under RK3 (its ring ends where it starts), two under RK1
  // (which ping-pongs), so the launches of every iteration replay from the same graph.
  const int unit = ctl->rk_stages == 3 ? 1 : 2;
  hipGraphExec_t gx = nullptr;
  {
    hipGraph_t graph = nullptr;
    int s0 = 0;
    AERO_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    for (int u = 0; u < unit && !rc; ++u) rc = one_iteration(m, cfg, forward, true, gt, od, ctl, &s0, true, err, errlen);
    const hipError_t ec = hipStreamEndCapture(m->stream, &graph);
    hipError_t ei = hipSuccess;
    if (!rc && ec == hipSuccess) ei = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);  // on every path, the failed ones included
    if (rc) return rc;
    AERO_TRY(ec, "hipStreamEndCapture");
    AERO_TRY(ei, "hipGraphInstantiate");
  }
  struct GraphGuard {
    hipGraphExec_t g;
    ~GraphGuard() {
      if (g) (void)hipGraphExecDestroy(g);
    }
  } guard{gx};
  if (!forward)  // the transposed stages read psi = P^T D x of their input
    AERO_TRY(aero_host::launch_adj_scale(a->am, m->lam, ctl->cfl, m->U[m->cur], face_opts(cfg), a->x[0], a->psi,
                                         m->stream),
             "k_adj_scale");
  int start = 0;  // after whole units the ring is back at buffer 0
  int done = 0;
  double res0 = -1.0, res = 0.0;
  std::vector<double> h(maxit);
  const auto t0 = std::chrono::steady_clock::now();
  while (done < maxit) {
    int batch = std::min(std::max(every, unit), maxit - done);
    batch = std::max(unit, (batch / unit) * unit);
    if (done + batch > maxit) batch = maxit - done;
    for (int it = 0; it + unit <= batch; it += unit) AERO_TRY(hipGraphLaunch(gx, m->stream), "hipGraphLaunch");
    for (int it = (batch / unit) * unit; it < batch; ++it) {  // a last odd RK1 iteration
      rc = one_iteration(m, cfg, forward, true, gt, od, ctl, &start, true, err, errlen);
      if (rc) return rc;
    }
    AERO_TRY(hipMemcpyAsync(h.data() + done, a->hist + done, batch * sizeof(double), hipMemcpyDeviceToHost, m->stream),
             "linear solve history");
    AERO_TRY(hipStreamSynchronize(m->stream), "linear solve sync");
    done += batch;
    for (int i = done - batch; i < done; ++i) h[i] = std::sqrt(h[i]);
    if (res0 < 0.0) res0 = h[0];
    res = h[done - 1];
    if (!std::isfinite(res)) break;
    if (res <= ctl->tol * res0) break;
  }
  const auto t1 = std::chrono::steady_clock::now();
  ctl->iterations = done;
  ctl->res0 = res0;
  ctl->res = res;
  ctl->seconds = std::chrono::duration<double>(t1 - t0).count();
  if (history) std::memcpy(history, h.data(), done * sizeof(double));
  *start_out = start;
  if (!std::isfinite(res)) {
    set_err(err, errlen, "linear solve: non-finite residual (the linearised pseudo-time iteration diverged; "
                         "use gmres_restart > 0)");
    return AERO_ENONFINITE;
  }
  return AERO_OK;
}

int ensure_krylov(aero_hip_mesh* m, int mdim, char* err, size_t errlen) {
  AdjointData* a = m->adj;
  const size_t n4 = 4 * static_cast<size_t>(m->N);
  if (a->v_cap < mdim + 1) {
    for (double** p : {&a->V, &a->dots, &a->dot_parts}) {
      if (*p) {
        (void)hipFree(*p);
        a->allocs.erase(std::find(a->allocs.begin(), a->allocs.end(), static_cast<void*>(*p)));
        *p = nullptr;
      }
    }
    AERO_TRY(aalloc(a, &a->V, n4 * (mdim + 1)), "Krylov basis");
    AERO_TRY(aalloc(a, &a->dots, static_cast<size_t>(mdim + 1)), "Krylov scratch");
    AERO_TRY(aalloc(a, &a->dot_parts, static_cast<size_t>(512) * (mdim + 1)), "Krylov scratch");
    a->v_cap = mdim + 1;
  }
  if (!a->gx) AERO_TRY(aalloc(a, &a->gx, n4), "Krylov vectors");
  if (!a->gc) AERO_TRY(aalloc(a, &a->gc, n4), "Krylov vectors");
  if (!a->gw) AERO_TRY(aalloc(a, &a->gw, n4), "Krylov vectors");
  return AERO_OK;
}

// y = G^k x (forcing = false) or the k-fold iteration from x with the right-hand side (forcing = true),
// k = gmres_sweeps, through the iterate ring.
int apply_iteration(aero_hip_mesh* m, const aero_residual_config* cfg, int forward, bool forcing, const GeoTan& gt,
                    const FaceOptsT<Dual>& od, const aero_linear_solve* ctl, const double* x, double* y, char* err,
                    size_t errlen) {
  AdjointData* a = m->adj;
  const size_t bytes = 4 * static_cast<size_t>(m->N) * sizeof(double);
  if (x)
    AERO_TRY(hipMemcpyAsync(a->x[0], x, bytes, hipMemcpyDeviceToDevice, m->stream), "iterate in");
  else
    AERO_TRY(hipMemsetAsync(a->x[0], 0, bytes, m->stream), "iterate in");
  if (!forward)
    AERO_TRY(aero_host::launch_adj_scale(a->am, m->lam, ctl->cfl, m->U[m->cur], face_opts(cfg), a->x[0], a->psi,
                                         m->stream),
             "k_adj_scale");
  int start = 0;
  for (int k = 0; k < std::max(ctl->gmres_sweeps, 1); ++k) {
    const int rc = one_iteration(m, cfg, forward, forcing, gt, od, ctl, &start, false, err, errlen);
    if (rc) return rc;
  }
  AERO_TRY(hipMemcpyAsync(y, a->x[start], bytes, hipMemcpyDeviceToDevice, m->stream), "iterate out");
  return AERO_OK;
}

int dots_to_host(aero_hip_mesh* m, long long n, const double* V, int nv, const double* w, double* host, char* err,
                 size_t errlen) {
  AdjointData* a = m->adj;
  AERO_TRY(aero_host::launch_multidot(n, V, nv, w, a->dot_parts, a->dots, m->stream), "k_multidot");
  AERO_TRY(hipMemcpyAsync(host, a->dots, nv * sizeof(double), hipMemcpyDeviceToHost, m->stream), "dots");
  AERO_TRY(hipStreamSynchronize(m->stream), "dots");
  return AERO_OK;
}

// Restarted GMRES(mdim) on (I - G) x = c, the fixed point of gmres_sweeps pseudo-time iterations
// x -> G x + c;
// classical Gram-Schmidt with one re-orthogonalisation, Givens rotations for the residual estimate.
// The solution is left in a->gx.
int gmres_iterate(aero_hip_mesh* m, const aero_residual_config* cfg, int forward, const GeoTan& gt,
                  const FaceOptsT<Dual>& od, aero_linear_solve* ctl, double* history, char* err, size_t errlen) {
  AdjointData* a = m->adj;
  const int mdim = ctl->gmres_restart;
  int rc = ensure_krylov(m, mdim, err, errlen);
  if (rc) return rc;
  const long long n = 4LL * m->N;
  const size_t bytes = static_cast<size_t>(n) * sizeof(double);
  const int maxit = std::max(ctl->max_iters, 1);
  std::vector<double> hist;
  const auto t0 = std::chrono::steady_clock::now();
  AERO_TRY(hipMemsetAsync(a->gx, 0, bytes, m->stream), "gmres init");
  rc = apply_iteration(m, cfg, forward, true, gt, od, ctl, nullptr, a->gc, err, errlen);  // c = G 0 + c
  if (rc) return rc;
  double res0 = -1.0, res = 0.0;
  int total = 0;
  std::vector<double> H(static_cast<size_t>(mdim + 1) * mdim), cs(mdim), sn(mdim), gv(mdim + 1), h(mdim + 1),
    h2(mdim + 1), y(mdim);
  bool done = false;
  while (!done) {
    // r = c - (x - G x) = G x + c - x
    rc = apply_iteration(m, cfg, forward, false, gt, od, ctl, a->gx, a->gw, err, errlen);
    if (rc) return rc;
    AERO_TRY(aero_host::launch_axpby(n, -1.0, a->gx, 1.0, a->gw, m->stream), "gmres residual");
    AERO_TRY(aero_host::launch_axpby(n, 1.0, a->gc, 1.0, a->gw, m->stream), "gmres residual");
    double beta2 = 0.0;
    rc = dots_to_host(m, n, a->gw, 1, a->gw, &beta2, err, errlen);
    if (rc) return rc;
    const double beta = std::sqrt(beta2);
    if (res0 < 0.0) res0 = beta;
    res = beta;
    if (!std::isfinite(beta)) break;
    if (beta <= ctl->tol * res0 || total >= maxit) break;
    AERO_TRY(aero_host::launch_axpby(n, 1.0 / beta, a->gw, 0.0, a->V, m->stream), "gmres basis");
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int jend = -1;
    for (int j = 0; j < mdim; ++j) {
      double* vj = a->V + static_cast<size_t>(j) * n;
      rc = apply_iteration(m, cfg, forward, false, gt, od, ctl, vj, a->gw, err, errlen);
      if (rc) return rc;
      AERO_TRY(aero_host::launch_axpby(n, 1.0, vj, -1.0, a->gw, m->stream), "gmres operator");  // w = v - G v
      rc = dots_to_host(m, n, a->V, j + 1, a->gw, h.data(), err, errlen);
      if (rc) return rc;
      for (int i = 0; i <= j; ++i) h2[i] = -h[i];
      AERO_TRY(hipMemcpyAsync(a->dots, h2.data(), (j + 1) * sizeof(double), hipMemcpyHostToDevice, m->stream), "h");
      AERO_TRY(aero_host::launch_combine(n, 1.0, a->V, j + 1, a->dots, a->gw, m->stream), "gram-schmidt");
      rc = dots_to_host(m, n, a->V, j + 1, a->gw, h2.data(), err, errlen);  // re-orthogonalise
      if (rc) return rc;
      for (int i = 0; i <= j; ++i) {
        h[i] += h2[i];
        h2[i] = -h2[i];
      }
      AERO_TRY(hipMemcpyAsync(a->dots, h2.data(), (j + 1) * sizeof(double), hipMemcpyHostToDevice, m->stream), "h");
      AERO_TRY(aero_host::launch_combine(n, 1.0, a->V, j + 1, a->dots, a->gw, m->stream), "gram-schmidt");
      double hn2 = 0.0;
      rc = dots_to_host(m, n, a->gw, 1, a->gw, &hn2, err, errlen);
      if (rc) return rc;
      const double hn = std::sqrt(hn2);
      // Givens rotations on column j
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * h[i] + sn[i] * h[i + 1];
        h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
        h[i] = t;
      }
      const double rr = std::hypot(h[j], hn);
      cs[j] = rr > 0.0 ? h[j] / rr : 1.0;
      sn[j] = rr > 0.0 ? hn / rr : 0.0;
      h[j] = rr;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      for (int i = 0; i <= j; ++i) H[static_cast<size_t>(i) * mdim + j] = h[i];
      ++total;
      res = std::abs(gv[j + 1]);
      hist.push_back(res);
      jend = j;
      if (!std::isfinite(res) || res <= ctl->tol * res0 || total >= maxit || hn == 0.0) break;
      AERO_TRY(aero_host::launch_axpby(n, 1.0 / hn, a->gw, 0.0, a->V + static_cast<size_t>(j + 1) * n, m->stream),
               "gmres basis");
    }
    if (jend < 0) break;
    for (int i = jend; i >= 0; --i) {  // back substitution
      double t = gv[i];
      for (int k = i + 1; k <= jend; ++k) t -= H[static_cast<size_t>(i) * mdim + k] * y[k];
      y[i] = t / H[static_cast<size_t>(i) * mdim + i];
    }
    AERO_TRY(hipMemcpyAsync(a->dots, y.data(), (jend + 1) * sizeof(double), hipMemcpyHostToDevice, m->stream), "y");
    AERO_TRY(aero_host::launch_combine(n, 1.0, a->V, jend + 1, a->dots, a->gx, m->stream), "gmres update");
    if (!std::isfinite(res) || total >= maxit) done = true;
  }
  AERO_TRY(hipStreamSynchronize(m->stream), "gmres sync");
  const auto t1 = std::chrono::steady_clock::now();
  ctl->iterations = total;
  ctl->res0 = res0;
  ctl->res = res;
  ctl->seconds = std::chrono::duration<double>(t1 - t0).count();
  if (history) std::memcpy(history, hist.data(), std::min<size_t>(hist.size(), maxit) * sizeof(double));
  if (!std::isfinite(res)) {
    set_err(err, errlen, "linear solve: non-finite residual in GMRES");
    return AERO_ENONFINITE;
  }
  return AERO_OK;
}

// Dispatch: plain iteration or GMRES; returns the device pointer of the solution.
int linear_solve(aero_hip_mesh* m, const aero_residual_config* cfg, int forward, const GeoTan& gt,
                 const FaceOptsT<Dual>& od, aero_linear_solve* ctl, double* history, const double** x_out, char* err,
                 size_t errlen) {
  if (ctl->gmres_restart > 0) {
    const int rc = gmres_iterate(m, cfg, forward, gt, od, ctl, history, err, errlen);
    *x_out = m->adj->gx;
    return rc;
  }
  int start = 0;
  const int rc = plain_iterate(m, cfg, forward, gt, od, ctl, history, &start, err, errlen);
  *x_out = m->adj->x[start];
  return rc;
}

int check_ctl(const aero_linear_solve* ctl, const char* who, char* err, size_t errlen) {
  if (!ctl || (ctl->rk_stages != 1 && ctl->rk_stages != 3) || !(ctl->cfl > 0.0) || ctl->max_iters <= 0 ||
      ctl->gmres_restart < 0 || ctl->gmres_restart > 200) {
    set_err(err, errlen, std::string(who) + ": rk_stages must be 1 or 3, cfl > 0, max_iters > 0, 0 <= gmres_restart <= 200");
    return AERO_EINVAL;
  }
  return AERO_OK;
}

}  // namespace

extern "C" {

int aero_hip_adjoint_solve(aero_hip_mesh* m, const aero_residual_config* cfg, const double* conserved,
                           const double* rhs, aero_linear_solve* ctl, double* psi, double* history, char* err,
                           size_t errlen) {
  if (!m || !rhs || !psi) {
    set_err(err, errlen, "aero_hip_adjoint_solve: null argument");
    return AERO_EINVAL;
  }
  int rc = check_cfg(cfg, "aero_hip_adjoint_solve", err, errlen);
  if (!rc) rc = check_ctl(ctl, "aero_hip_adjoint_solve", err, errlen);
  if (!rc) rc = adjoint_setup(m, err, errlen);
  if (!rc) rc = load_state(m, conserved, err, errlen);
  if (!rc) rc = enqueue_spectral_radius(m, cfg, err, errlen);
  if (!rc) rc = enqueue_jacobian(m, cfg, err, errlen);
  if (rc) return rc;
  AdjointData* a = m->adj;
  rc = upload_internal(m, rhs, a->rhs, err, errlen);
  if (rc) return rc;
  const double* x = nullptr;
  FaceOpts fo;
  rc = mesh_face_opts(m, cfg, &fo, err, errlen);
  if (rc) return rc;
  rc = linear_solve(m, cfg, 0, GeoTan{}, dual_opts(fo, nullptr), ctl, history, &x, err, errlen);
  if (rc) return rc;
  AERO_TRY(aero_host::launch_adj_scale(a->am, m->lam, ctl->cfl, m->U[m->cur], face_opts(cfg), x, a->psi, m->stream),
           "k_adj_scale");
  return download_internal(m, a->psi, psi, err, errlen);
}

int aero_hip_tangent_solve(aero_hip_mesh* m, const aero_residual_config* cfg, const double* conserved,
                           const aero_tangent_seed* seed, aero_linear_solve* ctl, double* v, double* history, char* err,
                           size_t errlen) {
  if (!m || !v) {
    set_err(err, errlen, "aero_hip_tangent_solve: null argument");
    return AERO_EINVAL;
  }
  int rc = check_cfg(cfg, "aero_hip_tangent_solve", err, errlen);
  if (!rc) rc = check_ctl(ctl, "aero_hip_tangent_solve", err, errlen);
  if (!rc) rc = adjoint_setup(m, err, errlen);
  if (!rc) rc = load_state(m, conserved, err, errlen);
  if (!rc) rc = enqueue_spectral_radius(m, cfg, err, errlen);
  if (rc) return rc;
  GeoTan gt;
  rc = geometry_tangent(m, seed, &gt, err, errlen);
  if (rc) return rc;
  const double* x = nullptr;
  FaceOpts fo;
  rc = mesh_face_opts(m, cfg, &fo, err, errlen);
  if (rc) return rc;
  rc = linear_solve(m, cfg, 1, gt, dual_opts(fo, seed), ctl, history, &x, err, errlen);
  if (rc) return rc;
  return download_internal(m, x, v, err, errlen);
}

int aero_hip_adjoint_solve_batch(aero_adjoint_case* cases, int32_t n, aero_linear_solve* ctl, int32_t iters_per_sync,
                                 aero_progress_fn progress, void* user, char* err, size_t errlen) {
  const char* who = "aero_hip_adjoint_solve_batch";
  if (!cases || n <= 0) {
    set_err(err, errlen, std::string(who) + ": no solves");
    return AERO_EINVAL;
  }
  int rc = check_ctl(ctl, who, err, errlen);
  if (rc) return rc;
  if (ctl->gmres_restart > 0) {
    set_err(err, errlen, std::string(who) + ": batched solves use the plain iteration (gmres_restart = 0)");
    return AERO_EINVAL;
  }
  // Solves on one handle share its linearisation: the first solve on a handle sets it up.
  std::vector<int> first(n, -1);
  for (int i = 0; i < n; ++i) {
    const aero_adjoint_case& c = cases[i];
    if (!c.mesh || !c.rhs || !c.psi) {
      set_err(err, errlen, std::string(who) + ": solve " + std::to_string(i) + " has a null mesh, rhs or psi");
      return AERO_EINVAL;
    }
    rc = check_cfg(c.cfg, who, err, errlen);
    if (rc) return rc;
    if (!(c.cfl >= 0.0)) {
      set_err(err, errlen, std::string(who) + ": solve " + std::to_string(i) + " has a negative CFL");
      return AERO_EINVAL;
    }
    if (c.mesh->dm.grad != cases[0].mesh->dm.grad) {
      set_err(err, errlen, std::string(who) + ": all meshes need the same gradient kind");
      return AERO_EINVAL;
    }
    for (int j = 0; j < i; ++j)
      if (cases[j].mesh == c.mesh) {
        if (std::memcmp(cases[j].cfg, c.cfg, sizeof(aero_residual_config)) != 0 || cases[j].conserved != c.conserved) {
          set_err(err, errlen, std::string(who) + ": solves " + std::to_string(j) + " and " + std::to_string(i) +
                                 " share a mesh handle but not its configuration and state");
          return AERO_EINVAL;
        }
        first[i] = j;
        break;
      }
  }
  for (int i = 0; i < n; ++i) {
    if (first[i] >= 0) continue;
    aero_hip_mesh* m = cases[i].mesh;
    rc = adjoint_setup(m, err, errlen);
    if (!rc) rc = load_state(m, cases[i].conserved, err, errlen);
    if (!rc) rc = enqueue_spectral_radius(m, cases[i].cfg, err, errlen);
    if (!rc) rc = enqueue_jacobian(m, cases[i].cfg, err, errlen);
    if (rc) return rc;
    AERO_TRY(hipStreamSynchronize(m->stream), "adjoint linearisation");
  }
  // Per-solve buffers: x ring (3 x 4N), psi (4N), Gbar (8N), direct (4N), rhs (4N), block partials.
  std::vector<size_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const size_t N = cases[i].mesh->N;
    off[i + 1] = off[i] + 32 * N + (N + 255) / 256;
  }
  struct Res {
    void* buf = nullptr;
    void* tab = nullptr;
    void* ctl = nullptr;
    void* work = nullptr;
    hipStream_t st = nullptr;
    hipGraphExec_t gx = nullptr;
    ~Res() {
      if (st) (void)hipStreamSynchronize(st);
      if (gx) (void)hipGraphExecDestroy(gx);
      for (void* p : {buf, tab, ctl, work})
        if (p) (void)hipFree(p);
      if (st) (void)hipStreamDestroy(st);
    }
  } r;
  AERO_TRY(hipStreamCreateWithFlags(&r.st, hipStreamNonBlocking), who);
  AERO_TRY(hipMalloc(&r.buf, off[n] * sizeof(double)), "adjoint batch buffers");
  double* base = static_cast<double*>(r.buf);
  std::vector<AdjBatchSolve> hs(n);
  std::vector<AdjBatchCtl> hc(n);
  const int rk3 = ctl->rk_stages == 3 ? 1 : 0;
  for (int i = 0; i < n; ++i) {
    aero_hip_mesh* m = cases[i].mesh;
    const size_t N = m->N;
    double* p = base + off[i];
    AdjBatchSolve& s = hs[i];
    s.am = m->adj->am;
    s.J = m->adj->J;
    s.fact = m->adj->jfact ? 1 : 0;
    s.U = m->U[m->cur];
    s.lam = m->lam;
    s.o = face_opts(cases[i].cfg);
    s.cfl = cases[i].cfl > 0.0 ? cases[i].cfl : ctl->cfl;
    for (int k = 0; k < 3; ++k) s.x[k] = p + 4 * N * k;
    s.psi = p + 12 * N;
    s.Gbar = p + 16 * N;
    s.direct = p + 24 * N;
    double* rhs = p + 28 * N;
    s.rhs = rhs;
    s.partials = p + 32 * N;
    s.nparts = static_cast<int>((N + 255) / 256);
    std::vector<double> soa;
    to_internal_state(m, cases[i].rhs, &soa);
    AERO_TRY(hipMemcpyAsync(rhs, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, r.st), "rhs upload");
    AERO_TRY(aero_host::launch_fill(static_cast<long long>(4 * N), 0.0, s.x[0], r.st), "adjoint batch init");
    AERO_TRY(aero_host::launch_fill(static_cast<long long>(4 * N), 0.0, s.psi, r.st), "adjoint batch init");
    AERO_TRY(hipStreamSynchronize(r.st), "rhs upload");
    hc[i] = AdjBatchCtl{1, 0, 0, 0, 0.0, 0.0};
  }
  // Work items: solves on one handle with the same CFL go in pairs (a, b), the rest alone (a, -1).  Per
  // handle, block-major / item-minor (the items of a block read the same face Jacobians), then spread so
  // that each XCD (workgroup index mod 8) receives a contiguous run.
  std::vector<int4> list;
  for (int i = 0; i < n; ++i) {
    if (first[i] >= 0) continue;
    std::vector<int> mine;
    for (int j = i; j < n; ++j)
      if (cases[j].mesh == cases[i].mesh) mine.push_back(j);
    std::vector<int2> items;
    std::vector<char> used(mine.size(), 0);
    for (size_t a = 0; a < mine.size(); ++a) {
      if (used[a]) continue;
      used[a] = 1;
      int partner = -1;
      for (size_t b = a + 1; b < mine.size(); ++b)
        if (!used[b] && hs[mine[b]].cfl == hs[mine[a]].cfl) {
          partner = mine[b];
          used[b] = 1;
          break;
        }
      items.push_back(make_int2(mine[a], partner));
    }
    for (int blk = 0; blk < hs[i].nparts; ++blk)
      for (const int2& it : items) list.push_back(make_int4(it.x, it.y, blk, 0));
  }
  const int W = static_cast<int>(list.size());
  const int per = (W + 7) / 8;
  std::vector<int4> work(static_cast<size_t>(per) * 8, make_int4(-1, -1, -1, 0));
  for (int j = 0; j < static_cast<int>(work.size()); ++j) {
    const int src = (j & 7) * per + (j >> 3);
    if (src < W) work[j] = list[src];
  }
  const int nwork = static_cast<int>(work.size());
  AERO_TRY(hipMalloc(&r.tab, n * sizeof(AdjBatchSolve)), "adjoint batch table");
  AERO_TRY(hipMalloc(&r.ctl, n * sizeof(AdjBatchCtl)), "adjoint batch control");
  AERO_TRY(hipMalloc(&r.work, work.size() * sizeof(int4)), "adjoint batch work list");
  auto* d_tab = static_cast<AdjBatchSolve*>(r.tab);
  auto* d_ctl = static_cast<AdjBatchCtl*>(r.ctl);
  auto* d_work = static_cast<int4*>(r.work);
  AERO_TRY(hipMemcpy(d_tab, hs.data(), n * sizeof(AdjBatchSolve), hipMemcpyHostToDevice), "adjoint batch table");
  AERO_TRY(hipMemcpy(d_ctl, hc.data(), n * sizeof(AdjBatchCtl), hipMemcpyHostToDevice), "adjoint batch control");
  AERO_TRY(hipMemcpy(d_work, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice), "adjoint batch work");
  // One replay unit: four iterations under RK3 (its ring ends where it starts), two under RK1 (whose ring
  // alternates).  Several iterations per graph amortise the launch of the graph; a solve that stops inside
  // a unit is skipped by the rest of it, so iteration counts and results do not depend on the unit.
  const int grad = cases[0].mesh->dm.grad;
  const int unit = rk3 ? 4 : 2;
  int n_fact = 0;
  for (const AdjBatchSolve& s : hs) n_fact += s.fact ? 1 : 0;
  const int fact_mode = n_fact == 0 ? 0 : (n_fact == n ? 1 : 2);
  {
    hipGraph_t graph = nullptr;
    AERO_TRY(hipStreamBeginCapture(r.st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    hipError_t e = hipSuccess;
    for (int u = 0; u < unit && e == hipSuccess; ++u) {
      if (rk3) {
        const int ixs[3] = {0, 1, 2}, ixo[3] = {1, 2, 0};
        for (int st = 0; st < 3 && e == hipSuccess; ++st)
          e = aero_host::launch_adj_batch_stage(grad, fact_mode, d_tab, d_ctl, d_work, nwork, st, 1, 0, ixs[st],
                                                ixo[st], st == 0 ? 1 : 0, r.st);
      } else {
        e = aero_host::launch_adj_batch_stage(grad, fact_mode, d_tab, d_ctl, d_work, nwork, 0, 0, u, u, 1 - u, 1, r.st);
      }
      if (e == hipSuccess) e = aero_host::launch_adj_batch_control(d_tab, d_ctl, n, ctl->tol, ctl->max_iters, r.st);
    }
    const hipError_t ec = hipStreamEndCapture(r.st, &graph);
    hipError_t ei = hipSuccess;
    if (e == hipSuccess && ec == hipSuccess) ei = hipGraphInstantiate(&r.gx, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);  // on every path, the failed ones included
    AERO_TRY(e, "adjoint batch capture");
    AERO_TRY(ec, "hipStreamEndCapture");
    AERO_TRY(ei, "hipGraphInstantiate");
  }
  const int every = std::max(iters_per_sync, unit);
  const auto t0 = std::chrono::steady_clock::now();
  long long launched = 0;
  const long long limit = static_cast<long long>(std::max(ctl->max_iters, 1)) + unit;
  for (;;) {
    for (int k = 0; k < every; k += unit) AERO_TRY(hipGraphLaunch(r.gx, r.st), "hipGraphLaunch");
    launched += every;
    AERO_TRY(hipMemcpyAsync(hc.data(), d_ctl, n * sizeof(AdjBatchCtl), hipMemcpyDeviceToHost, r.st), "control");
    AERO_TRY(hipStreamSynchronize(r.st), "adjoint batch sync");
    int active = 0;
    for (const auto& k : hc) active += k.active ? 1 : 0;
    if (!active || launched >= limit) break;
    if (progress && progress(user, static_cast<int32_t>(launched), active) != 0) break;
  }
  const auto t1 = std::chrono::steady_clock::now();
  int max_it = 0;
  for (int i = 0; i < n; ++i) {
    aero_hip_mesh* m = cases[i].mesh;
    const AdjBatchSolve& s = hs[i];
    const AdjBatchCtl& k = hc[i];
    const int fin = rk3 ? 0 : (k.iter & 1);
    AERO_TRY(aero_host::launch_adj_scale(s.am, s.lam, s.cfl, s.U, s.o, s.x[fin], s.direct, r.st), "k_adj_scale");
    std::vector<double> soa(4 * static_cast<size_t>(m->N));
    AERO_TRY(hipMemcpyAsync(soa.data(), s.direct, soa.size() * sizeof(double), hipMemcpyDeviceToHost, r.st), "psi");
    AERO_TRY(hipStreamSynchronize(r.st), "psi");
    from_internal4(m, soa, cases[i].psi);
    cases[i].iterations = k.iter;
    cases[i].converged = k.converged;
    cases[i].failed = k.failed;
    cases[i].res0 = k.res0;
    cases[i].res = k.res;
    max_it = std::max(max_it, k.iter);
  }
  ctl->iterations = max_it;
  ctl->res0 = 0.0;
  ctl->res = 0.0;
  ctl->seconds = std::chrono::duration<double>(t1 - t0).count();
  return AERO_OK;
}

}  // extern "C"
