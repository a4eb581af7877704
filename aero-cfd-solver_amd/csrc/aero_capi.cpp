// aero_capi.cpp -- host side of the C ABI declared in include/aero_hip.h: device mesh setup,
// residual / stage entry points and the device-resident pseudo-time driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "aero_hip.h"
#include "aero_host.hpp"
#include "aero_internal.hpp"

using aero_dev::DevMesh;
using aero_dev::FaceOpts;
using aero_dev::StageArgs;
using aero_dev::WallArgs;
using namespace aero_hostutil;

namespace {

void set_err(char* err, size_t errlen, const std::string& msg) {
  if (err && errlen > 0) {
    std::strncpy(err, msg.c_str(), errlen - 1);
    err[errlen - 1] = '\0';
  }
}

#define AERO_HIP_TRY(expr, ctx)                                                         \
  do {                                                                                  \
    const hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                             \
      set_err(err, errlen, std::string(ctx) + ": " + hipGetErrorString(e_));            \
      return AERO_EHIP;                                                                 \
    }                                                                                   \
  } while (0)

__global__ void k_init_diag(int* d) {
  if (threadIdx.x == 0) {
    d[0] = 0;
    d[1] = 0;
    d[2] = INT_MAX;
    d[3] = 0;
    d[4] = INT_MAX;
  }
}

// aoa = 0 symmetry enforcement on the state (euler_solver.cpp:380-410).
__global__ void k_aoa0_symmetry(int N, int ld, double* U, double g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  aero_dev::C4 s{U[c], U[ld + c], U[2 * ld + c], U[3 * ld + c]};
  aero_dev::P4 p = aero_dev::c2p(s, g);
  p.v = 0.0;
  const aero_dev::C4 y = aero_dev::p2c(p, g);
  U[c] = y.r;
  U[ld + c] = y.ru;
  U[2 * ld + c] = y.rv;
  U[3 * ld + c] = y.E;
}

// Wall pressure-jump energy of the staged all-speed controller (euler_solver.cpp:1026-1047): over the
// interior faces that touch a wall-adjacent cell, in increasing face index, sum (p_o - p_n)^2 of the
// cells' primitive pressures.  The per-face terms are formed in parallel; one thread then adds them in
// face order, the serial loop's summation order.  out[0] = the sum.  One workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_pjump_sum(int n, const int2* __restrict__ faces, const double* __restrict__ U,
                                                   int ld, double g, double* __restrict__ terms,
                                                   double* __restrict__ out) {
  for (int i = threadIdx.x; i < n; i += 256) {
    const int2 f = faces[i];
    const aero_dev::C4 so{U[f.x], U[ld + f.x], U[2 * ld + f.x], U[3 * ld + f.x]};
    const aero_dev::C4 sn{U[f.y], U[ld + f.y], U[2 * ld + f.y], U[3 * ld + f.y]};
    const double dp = aero_dev::c2p(so, g).p - aero_dev::c2p(sn, g).p;
    terms[i] = dp * dp;
  }
  __threadfence_block();
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc += terms[i];
    out[0] = acc;
  }
}

// The first failing face of an iteration's first stage with one (residual_assembly.hpp:33-41 per stage, the
// driver keeping the first stage's, euler_solver.cpp:870-900): after each stage d[4] latches d[2] once and
// d[2] is cleared for the next stage.
__global__ void k_latch_first_failure(int* d) {
  if (threadIdx.x == 0) {
    if (d[4] == INT_MAX && d[2] != INT_MAX) d[4] = d[2];
    d[2] = INT_MAX;
  }
}

// Per-iteration diagnostics of the reference driver's euler_diagnostics.csv (euler_solver.cpp:921-1060) from
// the iteration's final state: two passes, block partials then one fixed-order reduction.  Slots:
enum : int {
  kDgMaxUn = 0, kDgFfDp, kDgFfDrho, kDgFfDu, kDgFfDv,  // max |u.n| at wall owners; farfield-owner deviations
  kDgCpMin, kDgCpMax, kDgCpNonfinite,                   // wall Cp
  kDgMomentAbs,                                         // moment of the absolute-pressure wall forces
  kDgPjAll, kDgPjWall,                                  // sum (p_o - p_n)^2, all / wall-adjacent interior faces
  kDgPfMax, kDgPfWall,                                  // max |p - p_inf|, all / wall-adjacent cells
  kDgSlots
};
struct DiagArgs {
  int N, ld, nint;
  const int2* faces;                // interior faces (owner, neighbour), internal ids
  const unsigned char* wall_adj;    // [N] internal order: 1 for wall-face owner cells
  aero_dev::WallArgs wall;          // wall faces + reference values
  int nff;
  const int* ff_owner;              // farfield-face owner cells, internal ids
  double g, rho_inf, u_inf, v_inf;
};

__device__ __forceinline__ bool dg_is_min(int k) { return k == kDgCpMin; }
__device__ __forceinline__ bool dg_is_sum(int k) {
  return k == kDgCpNonfinite || k == kDgMomentAbs || k == kDgPjAll || k == kDgPjWall;
}
__device__ __forceinline__ double dg_init(int k) {
  return k == kDgCpMin ? 1.79769313486231570815e+308 : (k == kDgCpMax ? -1.79769313486231570815e+308 : 0.0);
}
__device__ __forceinline__ double dg_join(int k, double a, double b) {
  if (dg_is_sum(k)) return a + b;
  if (dg_is_min(k)) return aero_dev::dmin(a, b);
  return aero_dev::dmax(a, b);
}

__device__ __forceinline__ double dg_pressure(const double* U, int ld, int c, double g) {
  return aero_dev::c2p(aero_dev::C4{U[c], U[ld + c], U[2 * ld + c], U[3 * ld + c]}, g).p;
}

__device__ void dg_block_reduce(double (*v)[kDgSlots], double* acc) {
  const int t = threadIdx.x;
  for (int k = 0; k < kDgSlots; ++k) v[t][k] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < kDgSlots; ++k) v[t][k] = dg_join(k, v[t][k], v[t + w][k]);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_diag_partials(DiagArgs a, const double* __restrict__ U,
                                                       double* __restrict__ partials) {
  __shared__ double v[256][kDgSlots];
  double acc[kDgSlots];
  for (int k = 0; k < kDgSlots; ++k) acc[k] = dg_init(k);
  const double pinf = a.wall.p_inf;
  const int stride = gridDim.x * 256;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < a.N; i += stride) {
    const double dp = fabs(dg_pressure(U, a.ld, i, a.g) - pinf);
    acc[kDgPfMax] = aero_dev::dmax(acc[kDgPfMax], dp);
    if (a.wall_adj[i]) acc[kDgPfWall] = aero_dev::dmax(acc[kDgPfWall], dp);
  }
  for (int i = blockIdx.x * 256 + threadIdx.x; i < a.nint; i += stride) {
    const int2 f = a.faces[i];
    const double dp = dg_pressure(U, a.ld, f.x, a.g) - dg_pressure(U, a.ld, f.y, a.g);
    const double e = dp * dp;
    acc[kDgPjAll] += e;
    if (a.wall_adj[f.x] || a.wall_adj[f.y]) acc[kDgPjWall] += e;
  }
  for (int i = blockIdx.x * 256 + threadIdx.x; i < a.wall.n; i += stride) {
    const int c = a.wall.owner[i];
    const aero_dev::P4 q = aero_dev::c2p(aero_dev::C4{U[c], U[a.ld + c], U[2 * a.ld + c], U[3 * a.ld + c]}, a.g);
    acc[kDgMaxUn] = aero_dev::dmax(acc[kDgMaxUn], fabs(q.u * a.wall.nx[i] + q.v * a.wall.ny[i]));
    const double cp = (q.p - pinf) / a.wall.q_inf;
    if (isfinite(cp)) {
      acc[kDgCpMin] = aero_dev::dmin(acc[kDgCpMin], cp);
      acc[kDgCpMax] = aero_dev::dmax(acc[kDgCpMax], cp);
    } else {
      acc[kDgCpNonfinite] += 1.0;
    }
    const double nA0 = -a.wall.nx[i] * a.wall.A[i], nA1 = -a.wall.ny[i] * a.wall.A[i];
    const double dfx = -q.p * nA0, dfy = -q.p * nA1;
    acc[kDgMomentAbs] += (a.wall.fx[i] - a.wall.x_ref) * dfy - (a.wall.fy[i] - a.wall.y_ref) * dfx;
  }
  for (int i = blockIdx.x * 256 + threadIdx.x; i < a.nff; i += stride) {
    const int c = a.ff_owner[i];
    const aero_dev::P4 q = aero_dev::c2p(aero_dev::C4{U[c], U[a.ld + c], U[2 * a.ld + c], U[3 * a.ld + c]}, a.g);
    acc[kDgFfDp] = aero_dev::dmax(acc[kDgFfDp], fabs(q.p - pinf));
    acc[kDgFfDrho] = aero_dev::dmax(acc[kDgFfDrho], fabs(q.r - a.rho_inf));
    acc[kDgFfDu] = aero_dev::dmax(acc[kDgFfDu], fabs(q.u - a.u_inf));
    acc[kDgFfDv] = aero_dev::dmax(acc[kDgFfDv], fabs(q.v - a.v_inf));
  }
  dg_block_reduce(v, acc);
  if (threadIdx.x < kDgSlots) partials[blockIdx.x * kDgSlots + threadIdx.x] = v[0][threadIdx.x];
}

__global__ __launch_bounds__(256) void k_diag_final(int nparts, const double* __restrict__ partials,
                                                    double* __restrict__ out) {
  __shared__ double v[256][kDgSlots];
  double acc[kDgSlots];
  for (int k = 0; k < kDgSlots; ++k) acc[k] = dg_init(k);
  for (int b = threadIdx.x; b < nparts; b += 256)
    for (int k = 0; k < kDgSlots; ++k) acc[k] = dg_join(k, acc[k], partials[b * kDgSlots + k]);
  dg_block_reduce(v, acc);
  if (threadIdx.x < kDgSlots) out[threadIdx.x] = v[0][threadIdx.x];
}

// Recursive coordinate bisection of cell centres into blocks of BC cells (every block full except the
// last); cells inside a block keep increasing reference order.
void rcb(std::vector<int>& idx, int lo, int hi, const double* cc, int BC) {
  const int n = hi - lo;
  if (n <= BC) {
    std::sort(idx.begin() + lo, idx.begin() + hi);
    return;
  }
  const int nb = (n + BC - 1) / BC;
  const int left = (nb / 2) * BC;
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (int i = lo; i < hi; ++i) {
    const double x = cc[3 * idx[i]], y = cc[3 * idx[i] + 1];
    x0 = std::min(x0, x);
    x1 = std::max(x1, x);
    y0 = std::min(y0, y);
    y1 = std::max(y1, y);
  }
  const int ax = (x1 - x0 >= y1 - y0) ? 0 : 1;
  std::nth_element(idx.begin() + lo, idx.begin() + lo + left, idx.begin() + hi, [&](int a, int b) {
    const double va = cc[3 * a + ax], vb = cc[3 * b + ax];
    return va < vb || (va == vb && a < b);
  });
  rcb(idx, lo, lo + left, cc, BC);
  rcb(idx, lo + left, hi, cc, BC);
}


}  // namespace


namespace {

template <class T> hipError_t dev_alloc(aero_hip_mesh* m, T** p, size_t n) {
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) {
    m->allocs.push_back(q);
    *p = static_cast<T*>(q);
  }
  return e;
}

template <class T> hipError_t upload(aero_hip_mesh* m, T** p, const std::vector<T>& v) {
  hipError_t e = dev_alloc(m, p, v.size());
  if (e != hipSuccess) return e;
  if (!v.empty()) e = hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

// The frozen-ratio buffers of a handle: face order [8][F] (FP64) and the flux kernels' slot order [S][8] (16-bit).
int ensure_ratio_buffers(aero_hip_mesh* m, const char* who, char* err, size_t errlen) {
  void** bufs[2] = {reinterpret_cast<void**>(&m->lim_ratio), reinterpret_cast<void**>(&m->lim_slot)};
  const size_t bytes[2] = {8 * static_cast<size_t>(m->F) * sizeof(double),
                           8 * static_cast<size_t>(m->S) * sizeof(unsigned short)};
  for (int k = 0; k < 2; ++k) {
    if (*bufs[k]) continue;
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, bytes[k]);
    if (e != hipSuccess) {
      set_err(err, errlen, std::string(who) + ": hipMalloc: " + hipGetErrorString(e));
      return AERO_EHIP;
    }
    m->allocs.push_back(q);
    *bufs[k] = q;
  }
  return AERO_OK;
}

void free_events(aero_hip_mesh* m) {
  for (hipEvent_t e : m->tev) (void)hipEventDestroy(e);
  m->tev.clear();
  m->tcap = m->tcount = 0;
}

void free_all(aero_hip_mesh* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  adjoint_free(m);
  free_events(m);
  for (void* p : m->allocs) (void)hipFree(p);
  m->allocs.clear();
  if (m->h_iter) (void)hipHostFree(m->h_iter);
  m->h_iter = nullptr;
  if (m->stream) (void)hipStreamDestroy(m->stream);
  m->stream = nullptr;
}

bool env_on(const char* name) {  // euler_solver.cpp:38-45
  const char* v = std::getenv(name);
  if (!v) return false;
  const std::string t(v);
  return t == "1" || t == "true" || t == "TRUE" || t == "on" || t == "ON";
}

}  // namespace

namespace aero_hostutil {

FaceOpts face_opts(const aero_residual_config* c) {
  FaceOpts o;
  o.ff = aero_dev::C4{c->farfield[0], c->farfield[1], c->farfield[2], c->farfield[4]};
  o.g = c->gamma;
  o.low_mach_fix = c->low_mach_fix;
  o.low_mach_cutoff = c->low_mach_cutoff;
  o.stab_floor = c->stab_mach_floor;
  o.char_bc = c->precond_farfield_bc;
  o.precond = c->precond_enabled;
  o.mach_ref = c->precond_mach_ref;
  o.mach_min = c->precond_mach_min;
  o.beta_min = c->precond_beta_min;
  o.beta_max = c->precond_beta_max;
  o.as_fix = c->all_speed_fix;
  o.as_cut = c->all_speed_cutoff;
  o.as_fmin = c->all_speed_f_min;
  o.as_w = c->all_speed_ramp;
  o.frozen = nullptr;
  o.frozen_slots = nullptr;
  return o;
}

int mesh_face_opts(const aero_hip_mesh* m, const aero_residual_config* c, FaceOpts* o, char* err, size_t errlen) {
  *o = face_opts(c);
  if (c->limiter_frozen) {
    if (!m->lim_valid) {
      set_err(err, errlen, "limiter_frozen = 1 but no limiter ratios were captured (aero_hip_freeze_limiter)");
      return AERO_EINVAL;
    }
    o->frozen = m->lim_ratio;
    o->frozen_slots = m->lim_slot;
  }
  return AERO_OK;
}

int to_internal_state(aero_hip_mesh* m, const double* conserved, std::vector<double>* soa) {
  soa->assign(static_cast<size_t>(m->N) * 4, 0.0);
  for (int c = 0; c < m->N; ++c) {
    const double* u = conserved + 5 * static_cast<size_t>(m->perm[c]);
    (*soa)[c] = u[0];
    (*soa)[m->N + c] = u[1];
    (*soa)[2 * static_cast<size_t>(m->N) + c] = u[2];
    (*soa)[3 * static_cast<size_t>(m->N) + c] = u[4];
  }
  return 0;
}

void from_internal4(const aero_hip_mesh* m, const std::vector<double>& soa, double* aos5) {
  for (int c = 0; c < m->N; ++c) {
    double* o = aos5 + 5 * static_cast<size_t>(m->perm[c]);
    o[0] = soa[c];
    o[1] = soa[m->N + c];
    o[2] = soa[2 * static_cast<size_t>(m->N) + c];
    o[3] = 0.0;
    o[4] = soa[3 * static_cast<size_t>(m->N) + c];
  }
}

}  // namespace aero_hostutil

namespace {

int mesh_create_layout(const aero_mesh_desc* d, const aero_hip_options* opts, int accum, aero_hip_mesh** out,
                       char* err, size_t errlen);

// The face-coloured layout (AERO_ACCUM_COLOUR): faces in internal-owner order coloured greedily so that no
// two faces of a colour share a cell, then grouped by colour (owner order kept inside a colour).
int build_colour_layout(aero_hip_mesh* m, const aero_mesh_desc* d, char* err, size_t errlen) {
  const int N = m->N, F = m->F;
  std::vector<int> order(F);
  for (int f = 0; f < F; ++f) order[f] = f;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return m->iperm[d->face_owner[a]] < m->iperm[d->face_owner[b]]; });
  std::vector<uint32_t> used(N, 0u);
  std::vector<int> colour(F, 0);
  int nc = 0;
  for (int f : order) {
    const int o = m->iperm[d->face_owner[f]];
    const int n = d->face_neighbor[f] >= 0 ? m->iperm[d->face_neighbor[f]] : -1;
    const uint32_t taken = used[o] | (n >= 0 ? used[n] : 0u);
    if (taken == 0xffffffffu) {
      set_err(err, errlen, "aero_hip_mesh_create: a cell has more than 32 faces (face colouring)");
      return AERO_EINVAL;
    }
    const int c = __builtin_ctz(~taken);
    colour[f] = c;
    used[o] |= 1u << c;
    if (n >= 0) used[n] |= 1u << c;
    nc = std::max(nc, c + 1);
  }
  m->colour_off.assign(nc + 1, 0);
  for (int f = 0; f < F; ++f) ++m->colour_off[colour[f] + 1];
  for (int c = 0; c < nc; ++c) m->colour_off[c + 1] += m->colour_off[c];
  std::vector<int> fill(m->colour_off.begin(), m->colour_off.end() - 1), co(F), cn(F), cmeta(F), cface(F);
  std::vector<double> nx(F), ny(F), A(F), fx(F), fy(F);
  for (int f : order) {
    const int i = fill[colour[f]]++;
    co[i] = m->iperm[d->face_owner[f]];
    cn[i] = d->face_neighbor[f] >= 0 ? m->iperm[d->face_neighbor[f]] : -1;
    cmeta[i] = m->h_bc[f];
    cface[i] = f;
    nx[i] = m->h_nx[f];
    ny[i] = m->h_ny[f];
    A[i] = d->face_area[f];
    fx[i] = m->h_fx[f];
    fy[i] = m->h_fy[f];
  }
  aero_dev::ColourFaces& cf = m->cfaces;
  int *p_o, *p_n, *p_meta, *p_face;
  double *p_nx, *p_ny, *p_A, *p_fx, *p_fy;
  for (const hipError_t e : {upload(m, &p_o, co), upload(m, &p_n, cn), upload(m, &p_meta, cmeta),
                             upload(m, &p_face, cface), upload(m, &p_nx, nx), upload(m, &p_ny, ny), upload(m, &p_A, A),
                             upload(m, &p_fx, fx), upload(m, &p_fy, fy)})
    if (e != hipSuccess) {
      set_err(err, errlen, std::string("aero_hip_mesh_create: colour layout: ") + hipGetErrorString(e));
      return AERO_EHIP;
    }
  cf = aero_dev::ColourFaces{p_o, p_n, p_meta, p_face, p_nx, p_ny, p_A, p_fx, p_fy};
  m->ncolours = nc;
  m->colour = true;
  return AERO_OK;
}

// The layout probe behind block_cells = 0: the residual (HLLC + venkat, second order, on a smooth
// synthetic state) timed with HIP events in each candidate layout -- gather with 128, 256 and 512 cells per
// block, and the face-coloured layout -- and the fastest kept (the largest-block gather within 2 % of it), per mesh (keyed by a hash of its arrays and
// options) for the life of the process.
struct Tuned {
  int BC = 256, accum = AERO_ACCUM_GATHER;
  double ms[4] = {-1.0, -1.0, -1.0, -1.0};  // gather 128, 256, 512, colour (-1: not probed)
};

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

uint64_t mesh_fingerprint(const aero_mesh_desc* d, const aero_hip_options* opts, int accum) {
  uint64_t h = 1469598103934665603ull;
  const int32_t head[7] = {d->num_cells, d->num_faces, d->wall_start, d->wall_count, opts->renumber, opts->gradient,
                           accum};
  h = fnv1a(h, head, sizeof(head));
  const size_t F = static_cast<size_t>(d->num_faces), N = static_cast<size_t>(d->num_cells);
  h = fnv1a(h, d->face_owner, F * sizeof(int32_t));
  h = fnv1a(h, d->face_neighbor, F * sizeof(int32_t));
  if (d->face_bc) h = fnv1a(h, d->face_bc, F * sizeof(int32_t));
  h = fnv1a(h, d->face_normal, 3 * F * sizeof(double));
  h = fnv1a(h, d->face_center, 3 * F * sizeof(double));
  h = fnv1a(h, d->face_area, F * sizeof(double));
  h = fnv1a(h, d->cell_volume, N * sizeof(double));
  h = fnv1a(h, d->cell_center, 3 * N * sizeof(double));
  return h;
}

std::mutex g_tune_mu;
std::map<uint64_t, Tuned> g_tuned;

// ms per residual evaluation of handle m: 10 untimed evaluations, then the fastest of 3 timed runs of 10
// (a few milliseconds per layout; the first launches after the upload run cold).
int probe_residual(aero_hip_mesh* m, double* ms, char* err, size_t errlen) {
  const int N = m->N;
  const double g = 1.4, p0 = 101325.0, r0 = p0 / (287.05 * 288.15), a0 = std::sqrt(g * p0 / r0), V = 0.5 * a0;
  const double al = 2.0 * 3.14159265358979323846 / 180.0, u0 = V * std::cos(al), v0 = V * std::sin(al);
  std::vector<double> U(static_cast<size_t>(N) * 5);
  for (int c = 0; c < N; ++c) {
    const double x = m->h_ccx[c], y = m->h_ccy[c];
    const double s = std::sin(2.0 * 3.14159265358979323846 * x) * std::cos(2.0 * 3.14159265358979323846 * y);
    const double r = r0 * (1.0 + 0.02 * s), u = u0 + 0.02 * V * s, v = v0 - 0.02 * V * s, p = p0 * (1.0 + 0.02 * s);
    double* q = &U[5 * static_cast<size_t>(c)];
    q[0] = r;
    q[1] = r * u;
    q[2] = r * v;
    q[3] = 0.0;
    q[4] = p / (g - 1.0) + 0.5 * r * (u * u + v * v);
  }
  if (const int rc = aero_hip_set_state(m, U.data(), err, errlen)) return rc;
  aero_residual_config c{};
  c.farfield[0] = r0;
  c.farfield[1] = r0 * u0;
  c.farfield[2] = r0 * v0;
  c.farfield[4] = p0 / (g - 1.0) + 0.5 * r0 * V * V;
  c.gamma = g;
  c.flux_scheme = AERO_FLUX_HLLC;
  c.limiter = AERO_LIMITER_VENKAT;
  c.second_order = 1;
  c.low_mach_cutoff = 0.3;
  c.stab_mach_floor = 0.02;
  c.precond_mach_ref = 0.15;
  c.precond_mach_min = 1e-3;
  c.precond_beta_min = 0.2;
  c.precond_beta_max = 1.0;
  c.all_speed_cutoff = 0.25;
  c.all_speed_f_min = 0.05;
  c.all_speed_ramp = 1.0;
  for (int i = 0; i < 10; ++i)
    if (const int rc = aero_hip_residual_async(m, &c, err, errlen)) return rc;
  hipEvent_t e0, e1;
  AERO_HIP_TRY(hipEventCreate(&e0), "layout probe: hipEventCreate");
  AERO_HIP_TRY(hipEventCreate(&e1), "layout probe: hipEventCreate");
  int rc = AERO_OK;
  const int reps = 10;
  float best = -1.0f;
  for (int run = 0; run < 3 && !rc; ++run) {
    if (hipEventRecord(e0, m->stream) != hipSuccess) rc = AERO_EHIP;
    for (int i = 0; i < reps && !rc; ++i) rc = aero_hip_residual_async(m, &c, err, errlen);
    float t = 0.0f;
    if (!rc && (hipEventRecord(e1, m->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess)) {
      set_err(err, errlen, "layout probe: event timing failed");
      rc = AERO_EHIP;
    }
    if (!rc && (best < 0.0f || t < best)) best = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = static_cast<double>(best) / reps;
  return rc;
}

int tuned_layout(const aero_mesh_desc* d, const aero_hip_options* opts, int accum, Tuned* out, char* err,
                 size_t errlen) {
  const uint64_t key = mesh_fingerprint(d, opts, accum);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const auto it = g_tuned.find(key);
  if (it != g_tuned.end()) {
    *out = it->second;
    return AERO_OK;
  }
  Tuned t;
  double best = 1e300;
  const int bcs[3] = {128, 256, 512};
  for (int k = 0; k < 4; ++k) {
    const int mode = k < 3 ? AERO_ACCUM_GATHER : AERO_ACCUM_COLOUR;
    if (accum != AERO_ACCUM_AUTO && accum != mode) continue;
    aero_hip_options o = *opts;
    o.block_cells = k < 3 ? bcs[k] : 256;
    aero_hip_mesh* m = nullptr;
    char why[512] = {0};
    int rc = mesh_create_layout(d, &o, mode, &m, why, sizeof(why));
    if (rc == AERO_EINVAL && k < 3) continue;  // a block too large for LDS at this size: not a candidate
    if (rc) {
      set_err(err, errlen, std::string("layout probe: ") + why);
      return rc;
    }
    double ms = 0.0;
    rc = probe_residual(m, &ms, err, errlen);
    aero_hip_mesh_destroy(m);
    if (rc) return rc;
    t.ms[k] = ms;
    if (ms < best) {
      best = ms;
      t.BC = o.block_cells;
      t.accum = mode;
    }
  }
  if (best >= 1e300) {
    set_err(err, errlen, "layout probe: no candidate layout could be built");
    return AERO_EINVAL;
  }
  // Within 2 % of the fastest (the run-to-run spread of these timings) the gather layout with the largest
  // blocks is kept: it duplicates the fewest face slots, moves the fewest bytes and keeps the reference's
  // summation order, and the choice stays the same from run to run.
  for (int k = 2; k >= 0; --k) {
    if (t.ms[k] > 0.0 && t.ms[k] <= 1.02 * best) {
      t.BC = bcs[k];
      t.accum = AERO_ACCUM_GATHER;
      break;
    }
  }
  g_tuned[key] = t;
  *out = t;
  return AERO_OK;
}

}  // namespace

extern "C" {

int aero_hip_available(char* err, size_t errlen) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    set_err(err, errlen, std::string("hipGetDeviceCount: ") + (e != hipSuccess ? hipGetErrorString(e) : "no device"));
    return 0;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
    set_err(err, errlen, "hipGetDeviceProperties failed");
    return 0;
  }
  if (std::strstr(prop.gcnArchName, "gfx950") == nullptr) {
    set_err(err, errlen, std::string("device 0 is ") + prop.gcnArchName + ", this library is built for gfx950");
    return 0;
  }
  return 1;
}

}  // extern "C"

namespace {

// A mesh handle in one fixed layout: opts->block_cells is 128, 256 or 512; accum AERO_ACCUM_GATHER or
// AERO_ACCUM_COLOUR (the residual evaluations of the handle then use the face-coloured layout).
int mesh_create_layout(const aero_mesh_desc* d, const aero_hip_options* opts, int accum, aero_hip_mesh** out,
                              char* err, size_t errlen) {
  if (!d || !out) {
    set_err(err, errlen, "aero_hip_mesh_create: null argument");
    return AERO_EINVAL;
  }
  *out = nullptr;
  const int N = d->num_cells, F = d->num_faces;
  if (N <= 0 || F <= 0 || !d->face_owner || !d->face_neighbor || !d->face_normal || !d->face_area ||
      !d->cell_volume || !d->cell_center || !d->face_center) {
    set_err(err, errlen, "aero_hip_mesh_create: empty mesh or missing arrays");
    return AERO_EINVAL;
  }
  for (int f = 0; f < F; ++f) {
    const int o = d->face_owner[f], n = d->face_neighbor[f];
    if (o < 0 || o >= N || n >= N || n == o || n < -1) {
      set_err(err, errlen, "aero_hip_mesh_create: face " + std::to_string(f) + " has an invalid owner/neighbor");
      return AERO_EINVAL;
    }
  }
  if (d->wall_count < 0 || d->wall_start < 0 || d->wall_start + d->wall_count > F) {
    set_err(err, errlen, "aero_hip_mesh_create: wall patch range out of bounds");
    return AERO_EINVAL;
  }
  const int BC = opts && opts->block_cells > 0 ? opts->block_cells : 256;
  if (BC != 128 && BC != 256 && BC != 512) {
    set_err(err, errlen, "aero_hip_mesh_create: block_cells must be 0, 128, 256 or 512");
    return AERO_EINVAL;
  }
  const bool renumber = opts ? opts->renumber != 0 : true;
  if (opts && opts->gradient != AERO_GRAD_GREEN_GAUSS && opts->gradient != AERO_GRAD_LEAST_SQUARES) {
    set_err(err, errlen, "aero_hip_mesh_create: gradient must be AERO_GRAD_GREEN_GAUSS or AERO_GRAD_LEAST_SQUARES");
    return AERO_EINVAL;
  }
  char why[256] = {0};
  if (!aero_hip_available(why, sizeof(why))) {
    set_err(err, errlen, std::string("aero_hip_mesh_create: no gfx950 device: ") + why);
    return AERO_ENODEV;
  }

  aero_hip_mesh* m = new aero_hip_mesh();
  m->N = N;
  m->F = F;
  m->BC = BC;
  m->wall_start = d->wall_start;
  m->wall_count = d->wall_count;
  (void)hipGetDevice(&m->device);
  // Face normals re-normalised like the driver does before any use (mesh_geometry.cpp:49-80,
  // euler_solver.cpp:475-478): only normals deviating by more than 1e-5 from unit length change.
  m->h_nx.resize(F);
  m->h_ny.resize(F);
  for (int f = 0; f < F; ++f) {
    const double nx = d->face_normal[3 * f], ny = d->face_normal[3 * f + 1], nz = d->face_normal[3 * f + 2];
    const double nrm = std::sqrt(nx * nx + ny * ny + nz * nz);
    if (!std::isfinite(nrm) || nrm < 1.0e-12) {
      delete m;
      set_err(err, errlen, "Mesh contains invalid face normals.");
      return AERO_EINVAL;
    }
    if (std::abs(nrm - 1.0) > 1.0e-5) {
      const double inv = 1.0 / nrm;
      m->h_nx[f] = nx * inv;
      m->h_ny[f] = ny * inv;
    } else {
      m->h_nx[f] = nx;
      m->h_ny[f] = ny;
    }
  }
  m->h_owner.assign(d->face_owner, d->face_owner + F);
  m->h_neighbor.assign(d->face_neighbor, d->face_neighbor + F);
  m->h_bc.resize(F);
  m->h_fx.resize(F);
  m->h_fy.resize(F);
  for (int f = 0; f < F; ++f) {
    m->h_bc[f] = d->face_neighbor[f] >= 0 ? AERO_BC_INTERIOR
                                           : ((d->face_bc && d->face_bc[f] == AERO_BC_SLIP_WALL) ? AERO_BC_SLIP_WALL
                                                                                                  : AERO_BC_FARFIELD);
    m->h_fx[f] = d->face_center[3 * f];
    m->h_fy[f] = d->face_center[3 * f + 1];
  }
  m->h_vol.assign(d->cell_volume, d->cell_volume + N);
  m->h_ccx.resize(N);
  m->h_ccy.resize(N);
  for (int c = 0; c < N; ++c) {
    m->h_ccx[c] = d->cell_center[3 * c];
    m->h_ccy[c] = d->cell_center[3 * c + 1];
  }

  // Internal cell order.
  m->perm.resize(N);
  for (int i = 0; i < N; ++i) m->perm[i] = i;
  if (renumber) rcb(m->perm, 0, N, d->cell_center, BC);
  m->iperm.resize(N);
  for (int i = 0; i < N; ++i) m->iperm[m->perm[i]] = i;
  m->nblk = (N + BC - 1) / BC;

  // Per-cell face lists in increasing reference face index.
  std::vector<int> cf_off(N + 1, 0);
  for (int f = 0; f < F; ++f) {
    ++cf_off[d->face_owner[f] + 1];
    if (d->face_neighbor[f] >= 0) ++cf_off[d->face_neighbor[f] + 1];
  }
  for (int c = 0; c < N; ++c) cf_off[c + 1] += cf_off[c];
  std::vector<int> cf(cf_off[N]);
  {
    std::vector<int> fill(cf_off.begin(), cf_off.end() - 1);
    for (int f = 0; f < F; ++f) {
      cf[fill[d->face_owner[f]]++] = f;
      if (d->face_neighbor[f] >= 0) cf[fill[d->face_neighbor[f]]++] = f;
    }
  }
  // Per-block slot tables and per-cell gather entries.
  std::vector<int> slot_off(m->nblk + 1, 0), s_o, s_nm, s_face;
  std::vector<double> s_nx, s_ny, s_A, s_fx, s_fy;
  std::vector<int> c_off(N + 1, 0);
  std::vector<unsigned short> c_ent;  // a block's slots fit 16 bits with the side bit (LDS bound below)
  c_ent.reserve(cf.size());
  std::vector<int> stamp(F, -1), local(F, -1), bf;
  int smax = 0;
  for (int b = 0; b < m->nblk; ++b) {
    bf.clear();
    const int c0 = b * BC, c1 = std::min(N, c0 + BC);
    for (int c = c0; c < c1; ++c) {
      const int oc = m->perm[c];
      for (int e = cf_off[oc]; e < cf_off[oc + 1]; ++e)
        if (stamp[cf[e]] != b) {
          stamp[cf[e]] = b;
          bf.push_back(cf[e]);
        }
    }
    std::sort(bf.begin(), bf.end());
    slot_off[b + 1] = slot_off[b] + static_cast<int>(bf.size());
    smax = std::max(smax, static_cast<int>(bf.size()));
    for (int i = 0; i < static_cast<int>(bf.size()); ++i) {
      const int f = bf[i];
      local[f] = i;
      const int o = d->face_owner[f], n = d->face_neighbor[f];
      const int io = m->iperm[o];
      int bc = 0;
      if (n < 0) bc = (d->face_bc && d->face_bc[f] == AERO_BC_SLIP_WALL) ? 2 : 1;
      s_o.push_back(io);
      s_nm.push_back((n >= 0 ? m->iperm[n] : -1) * 8 + (bc | ((io / BC) == b ? aero_dev::kMetaPrimary : 0)));
      s_face.push_back(f);
      s_nx.push_back(m->h_nx[f]);
      s_ny.push_back(m->h_ny[f]);
      s_A.push_back(d->face_area[f]);
      s_fx.push_back(d->face_center[3 * f]);
      s_fy.push_back(d->face_center[3 * f + 1]);
    }
    for (int c = c0; c < c1; ++c) {
      const int oc = m->perm[c];
      for (int e = cf_off[oc]; e < cf_off[oc + 1]; ++e) {
        const int f = cf[e];
        c_ent.push_back(static_cast<unsigned short>((local[f] << 1) | (d->face_neighbor[f] == oc ? 1 : 0)));
      }
      c_off[c + 1] = static_cast<int>(c_ent.size());
    }
  }
  m->S = slot_off[m->nblk];
  // The kernels gather at 32-bit byte offsets of cell, slot and gather-entry indices (at most 8 bytes per
  // element): cells and slots stay below 2^28 (gather entries are at most twice the slots).
  if (N >= (1 << 28) || m->S >= (1 << 28)) {
    delete m;
    set_err(err, errlen, "aero_hip_mesh_create: the mesh exceeds 2^28 cells or face slots");
    return AERO_EINVAL;
  }
  const int grad = opts ? opts->gradient : AERO_GRAD_GREEN_GAUSS;
  // doubles staged per slot: the larger of the gradient kernel's (LSQ 8, Green-Gauss 4) and the flux kernel's (5)
  const size_t lds_per_slot = grad == AERO_GRAD_LEAST_SQUARES ? 8 : 5;
  if (static_cast<size_t>(smax) * lds_per_slot * sizeof(double) > 64 * 1024) {
    delete m;
    set_err(err, errlen, "aero_hip_mesh_create: a block touches " + std::to_string(smax) +
                           " faces; use renumber=1 or a smaller block_cells");
    return AERO_EINVAL;
  }
  std::vector<double> vol(N), ccx(N), ccy(N);
  for (int c = 0; c < N; ++c) {
    vol[c] = d->cell_volume[m->perm[c]];
    ccx[c] = d->cell_center[3 * m->perm[c]];
    ccy[c] = d->cell_center[3 * m->perm[c] + 1];
  }
  std::vector<int> w_o(d->wall_count);
  std::vector<double> w_nx(d->wall_count), w_ny(d->wall_count), w_A(d->wall_count), w_fx(d->wall_count),
    w_fy(d->wall_count);
  for (int i = 0; i < d->wall_count; ++i) {
    const int f = d->wall_start + i;
    w_o[i] = m->iperm[d->face_owner[f]];
    w_nx[i] = m->h_nx[f];
    w_ny[i] = m->h_ny[f];
    w_A[i] = d->face_area[f];
    w_fx[i] = d->face_center[3 * f];
    w_fy[i] = d->face_center[3 * f + 1];
  }
  m->h_area.assign(d->face_area, d->face_area + F);

#define AERO_SETUP(expr)                                                             \
  do {                                                                               \
    const hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                          \
      set_err(err, errlen, std::string("aero_hip_mesh_create: ") + hipGetErrorString(e_)); \
      free_all(m);                                                                   \
      delete m;                                                                      \
      return AERO_EHIP;                                                              \
    }                                                                                \
  } while (0)

  AERO_SETUP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  int *p_slot_off, *p_so, *p_snm, *p_face, *p_coff, *p_wo;
  unsigned short* p_cent;
  double *p_nx, *p_ny, *p_A, *p_fx, *p_fy, *p_vol, *p_ccx, *p_ccy, *p_wnx, *p_wny, *p_wA, *p_wfx, *p_wfy;
  AERO_SETUP(upload(m, &p_slot_off, slot_off));
  AERO_SETUP(upload(m, &p_so, s_o));
  AERO_SETUP(upload(m, &p_snm, s_nm));
  AERO_SETUP(upload(m, &p_face, s_face));
  AERO_SETUP(upload(m, &p_nx, s_nx));
  AERO_SETUP(upload(m, &p_ny, s_ny));
  AERO_SETUP(upload(m, &p_A, s_A));
  AERO_SETUP(upload(m, &p_fx, s_fx));
  AERO_SETUP(upload(m, &p_fy, s_fy));
  AERO_SETUP(upload(m, &p_coff, c_off));
  AERO_SETUP(upload(m, &p_cent, c_ent));
  AERO_SETUP(upload(m, &p_vol, vol));
  AERO_SETUP(upload(m, &p_ccx, ccx));
  AERO_SETUP(upload(m, &p_ccy, ccy));
  AERO_SETUP(upload(m, &p_wo, w_o));
  AERO_SETUP(upload(m, &p_wnx, w_nx));
  AERO_SETUP(upload(m, &p_wny, w_ny));
  AERO_SETUP(upload(m, &p_wA, w_A));
  AERO_SETUP(upload(m, &p_wfx, w_fx));
  AERO_SETUP(upload(m, &p_wfy, w_fy));
  AERO_SETUP(upload(m, &m->d_perm, m->perm));
  for (int k = 0; k < 3; ++k) AERO_SETUP(dev_alloc(m, &m->U[k], static_cast<size_t>(4) * N));
  AERO_SETUP(dev_alloc(m, &m->G, static_cast<size_t>(8) * N));
  AERO_SETUP(dev_alloc(m, &m->R, static_cast<size_t>(4) * N));
  AERO_SETUP(dev_alloc(m, &m->lam, static_cast<size_t>(N)));
  AERO_SETUP(dev_alloc(m, &m->diag, 8));  // counters [0..3], the iteration's first failing face [4]
  AERO_SETUP(dev_alloc(m, &m->partials, static_cast<size_t>(4) * m->nblk));
  AERO_SETUP(dev_alloc(m, &m->dt_scratch, 1));
  AERO_SETUP(dev_alloc(m, &m->iter_out, 16));
  AERO_SETUP(dev_alloc(m, &m->cp, std::max(1, d->wall_count)));
  AERO_SETUP(hipHostMalloc(reinterpret_cast<void**>(&m->h_iter), 64 * sizeof(double), hipHostMallocDefault));
  AERO_SETUP(hipMemsetAsync(m->U[0], 0, static_cast<size_t>(4) * N * sizeof(double), m->stream));
  AERO_SETUP(hipMemsetAsync(m->G, 0, static_cast<size_t>(8) * N * sizeof(double), m->stream));
  AERO_SETUP(hipStreamSynchronize(m->stream));
#undef AERO_SETUP

  DevMesh& dm = m->dm;
  dm.N = N;
  dm.nfaces = F;
  dm.ld = N;
  dm.nblk = m->nblk;
  dm.BC = BC;
  dm.smax = std::max(smax, 1);
  dm.xcd_remap = renumber ? 1 : 0;
  dm.grad = grad;
  dm.nslots = m->S;
  dm.grid = dm.xcd_remap ? ((m->nblk + 7) / 8) * 8 : m->nblk;
  dm.slot_off = p_slot_off;
  dm.s_o = p_so;
  dm.s_nm = p_snm;
  dm.s_face = p_face;
  dm.s_nx = p_nx;
  dm.s_ny = p_ny;
  dm.s_A = p_A;
  dm.s_fx = p_fx;
  dm.s_fy = p_fy;
  dm.c_off = p_coff;
  dm.c_ent = p_cent;
  dm.vol = p_vol;
  dm.ccx = p_ccx;
  dm.ccy = p_ccy;
  m->wall.n = d->wall_count;
  m->wall.owner = p_wo;
  m->wall.nx = p_wnx;
  m->wall.ny = p_wny;
  m->wall.A = p_wA;
  m->wall.fx = p_wfx;
  m->wall.fy = p_wfy;
  m->cur = m->base = 0;
  if (accum == AERO_ACCUM_COLOUR) {
    const int rc = build_colour_layout(m, d, err, errlen);
    if (rc) {
      free_all(m);
      delete m;
      return rc;
    }
  }
  *out = m;
  return AERO_OK;
}

}  // namespace

extern "C" {

int aero_hip_mesh_create(const aero_mesh_desc* d, const aero_hip_options* opts, aero_hip_mesh** out, char* err,
                         size_t errlen) {
  if (!d || !out) {
    set_err(err, errlen, "aero_hip_mesh_create: null argument");
    return AERO_EINVAL;
  }
  const int accum = opts ? opts->accumulation : AERO_ACCUM_AUTO;
  if (accum != AERO_ACCUM_AUTO && accum != AERO_ACCUM_GATHER && accum != AERO_ACCUM_COLOUR) {
    set_err(err, errlen, "aero_hip_mesh_create: accumulation must be AERO_ACCUM_AUTO, _GATHER or _COLOUR");
    return AERO_EINVAL;
  }
  if (!opts || opts->block_cells != 0)
    return mesh_create_layout(d, opts, accum == AERO_ACCUM_COLOUR ? AERO_ACCUM_COLOUR : AERO_ACCUM_GATHER, out, err,
                              errlen);
  // block_cells = 0: the layout probe picks block size and accumulation for this mesh (cached per mesh)
  Tuned t;
  int rc = tuned_layout(d, opts, accum, &t, err, errlen);
  if (rc) return rc;
  aero_hip_options o = *opts;
  o.block_cells = t.BC;
  rc = mesh_create_layout(d, &o, t.accum, out, err, errlen);
  if (rc) return rc;
  std::memcpy((*out)->probe_ms, t.ms, sizeof(t.ms));
  return AERO_OK;
}

int aero_hip_layout_probe(aero_hip_mesh* m, double* ms4) {
  if (!m || !ms4) return AERO_EINVAL;
  std::memcpy(ms4, m->probe_ms, sizeof(m->probe_ms));
  return AERO_OK;
}

void aero_hip_mesh_destroy(aero_hip_mesh* m) {
  if (!m) return;
  free_all(m);
  delete m;
}

void* aero_hip_stream(aero_hip_mesh* m) { return m ? static_cast<void*>(m->stream) : nullptr; }

int aero_hip_set_state(aero_hip_mesh* m, const double* conserved, char* err, size_t errlen) {
  if (!m || !conserved) {
    set_err(err, errlen, "aero_hip_set_state: null argument");
    return AERO_EINVAL;
  }
  std::vector<double> soa;
  to_internal_state(m, conserved, &soa);
  m->cur = m->base = 0;
  AERO_HIP_TRY(hipMemcpyAsync(m->U[0], soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, m->stream),
               "aero_hip_set_state: hipMemcpyAsync");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_set_state: hipStreamSynchronize");
  return AERO_OK;
}

int aero_hip_residual_async(aero_hip_mesh* m, const aero_residual_config* cfg, char* err, size_t errlen) {
  if (!m || !cfg) {
    set_err(err, errlen, "aero_hip_residual_async: null argument");
    return AERO_EINVAL;
  }
  if (cfg->flux_scheme < 0 || cfg->flux_scheme > 1 || cfg->limiter < 0 || cfg->limiter > 1) {
    set_err(err, errlen, "aero_hip_residual_async: unknown flux scheme or limiter");
    return AERO_EINVAL;
  }
  FaceOpts o;
  if (const int rc = mesh_face_opts(m, cfg, &o, err, errlen)) return rc;
  if (!cfg->second_order) {  // at second order the gradient kernel resets the counters
    hipLaunchKernelGGL(k_init_diag, dim3(1), dim3(64), 0, m->stream, m->diag);
    AERO_HIP_TRY(hipGetLastError(), "aero_hip_residual_async: k_init_diag");
  }
  const double* U = m->U[m->cur];
  const bool timed = m->tcount < m->tcap;
  if (timed) AERO_HIP_TRY(hipEventRecord(m->tev[3 * m->tcount + 0], m->stream), "hipEventRecord");
  if (cfg->second_order)
    AERO_HIP_TRY(aero_host::launch_gradient(m->dm, U, m->G, o.g, m->stream, m->diag), "k_gradient");
  if (timed) AERO_HIP_TRY(hipEventRecord(m->tev[3 * m->tcount + 1], m->stream), "hipEventRecord");
  StageArgs st{};
  if (m->colour)
    AERO_HIP_TRY(aero_host::launch_flux_colour(m->dm, m->cfaces, m->colour_off.data(), m->ncolours, o,
                                               cfg->flux_scheme, cfg->limiter, cfg->second_order != 0, U, m->G, m->R,
                                               m->lam, m->diag, m->stream),
                 "k_flux_colour");
  else
    AERO_HIP_TRY(aero_host::launch_flux(m->dm, o, cfg->flux_scheme, cfg->limiter, cfg->second_order != 0, 0, U, m->G,
                                        m->R, m->lam, m->diag, st, m->stream),
                 "k_flux");
  if (timed) {
    AERO_HIP_TRY(hipEventRecord(m->tev[3 * m->tcount + 2], m->stream), "hipEventRecord");
    ++m->tcount;
  }
  return AERO_OK;
}

int aero_hip_timing_begin(aero_hip_mesh* m, int32_t max_evals, char* err, size_t errlen) {
  if (!m || max_evals < 0) {
    set_err(err, errlen, "aero_hip_timing_begin: bad argument");
    return AERO_EINVAL;
  }
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_timing_begin: hipStreamSynchronize");
  free_events(m);
  m->tev.resize(static_cast<size_t>(3) * max_evals);
  for (auto& e : m->tev) AERO_HIP_TRY(hipEventCreate(&e), "hipEventCreate");
  m->tcap = max_evals;
  m->tcount = 0;
  return AERO_OK;
}

int aero_hip_timing_end(aero_hip_mesh* m, double* gradient_ms, double* flux_ms, double* span_ms, int32_t* n_evals,
                        char* err, size_t errlen) {
  if (!m) {
    set_err(err, errlen, "aero_hip_timing_end: null mesh");
    return AERO_EINVAL;
  }
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_timing_end: hipStreamSynchronize");
  double g = 0.0, f = 0.0, span = 0.0;
  for (int i = 0; i < m->tcount; ++i) {
    float a = 0.f, b = 0.f;
    AERO_HIP_TRY(hipEventElapsedTime(&a, m->tev[3 * i], m->tev[3 * i + 1]), "hipEventElapsedTime");
    AERO_HIP_TRY(hipEventElapsedTime(&b, m->tev[3 * i + 1], m->tev[3 * i + 2]), "hipEventElapsedTime");
    g += a;
    f += b;
  }
  if (m->tcount > 0) {
    float a = 0.f;
    AERO_HIP_TRY(hipEventElapsedTime(&a, m->tev[0], m->tev[3 * (m->tcount - 1) + 2]), "hipEventElapsedTime");
    span = a;
  }
  if (gradient_ms) *gradient_ms = g;
  if (flux_ms) *flux_ms = f;
  if (span_ms) *span_ms = span;
  if (n_evals) *n_evals = m->tcount;
  free_events(m);
  return AERO_OK;
}

int aero_hip_layout(aero_hip_mesh* m, int64_t* out8) {
  if (!m || !out8) return AERO_EINVAL;
  out8[0] = m->N;
  out8[1] = m->F;
  out8[2] = m->nblk;
  out8[3] = m->BC;
  out8[4] = m->S;
  out8[5] = m->dm.smax;
  out8[6] = m->colour ? AERO_ACCUM_COLOUR : AERO_ACCUM_GATHER;
  out8[7] = m->dm.grid;
  return AERO_OK;
}

int aero_hip_get_residual(aero_hip_mesh* m, double* residual, double* spectral_radius, aero_residual_diag* diag,
                          char* err, size_t errlen) {
  if (!m) {
    set_err(err, errlen, "aero_hip_get_residual: null mesh");
    return AERO_EINVAL;
  }
  std::vector<double> r(static_cast<size_t>(4) * m->N), l(m->N);
  int dg[4];
  AERO_HIP_TRY(hipMemcpyAsync(r.data(), m->R, r.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream), "R copy");
  AERO_HIP_TRY(hipMemcpyAsync(l.data(), m->lam, l.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream),
               "lambda copy");
  AERO_HIP_TRY(hipMemcpyAsync(dg, m->diag, sizeof(dg), hipMemcpyDeviceToHost, m->stream), "diag copy");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_get_residual: hipStreamSynchronize");
  if (residual) from_internal4(m, r, residual);
  if (spectral_radius)
    for (int c = 0; c < m->N; ++c) spectral_radius[m->perm[c]] = l[c];
  if (diag) {
    diag->first_order_fallback = dg[0];
    diag->diffusive_fallback = dg[1];
    diag->first_failure_face = dg[2] == INT_MAX ? -1 : dg[2];
  }
  return AERO_OK;
}

int aero_hip_get_state(aero_hip_mesh* m, double* conserved, char* err, size_t errlen) {
  if (!m || !conserved) {
    set_err(err, errlen, "aero_hip_get_state: null argument");
    return AERO_EINVAL;
  }
  std::vector<double> soa(4 * static_cast<size_t>(m->N));
  AERO_HIP_TRY(hipMemcpyAsync(soa.data(), m->U[m->cur], soa.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream),
               "aero_hip_get_state: copy");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_get_state: sync");
  from_internal4(m, soa, conserved);
  return AERO_OK;
}

int aero_hip_forces(aero_hip_mesh* m, const double* conserved, const aero_freestream_ref* ref, int32_t gauge,
                    double gamma, double* forces, double* wall_cp, char* err, size_t errlen) {
  if (!m || !ref || !forces) {
    set_err(err, errlen, "aero_hip_forces: null argument");
    return AERO_EINVAL;
  }
  if (!(gamma > 1.0)) {
    set_err(err, errlen, "aero_hip_forces: gamma must be > 1");
    return AERO_EINVAL;
  }
  std::vector<double> own;
  const double* U = conserved;
  if (!U) {
    own.resize(5 * static_cast<size_t>(m->N));
    const int rc = aero_hip_get_state(m, own.data(), err, errlen);
    if (rc) return rc;
    U = own.data();
  }
  // owner-cell pressure of conservative_to_primitive (euler_flux.cpp:115-127)
  auto pressure = [&](int cell) {
    const double* u = U + 5 * static_cast<size_t>(cell);
    const double rho = u[0] < 1.0e-8 ? 1.0e-8 : u[0];
    const double inv = 1.0 / rho;
    const double vx = u[1] * inv, vy = u[2] * inv;
    const double ke = 0.5 * rho * (vx * vx + vy * vy);
    const double p = (gamma - 1.0) * (u[4] - ke);
    return p < 1.0e-8 ? 1.0e-8 : p;
  };
  // integrate_pressure_forces (forces.cpp:250-294) over the wall patch in face order; nA = -n A
  const double q_inf = 0.5 * ref->rho_inf * ref->speed_inf * ref->speed_inf + 1.0e-12;
  double fx = 0.0, fy = 0.0, mom = 0.0;
  for (int i = 0; i < m->wall_count; ++i) {
    const int f = m->wall_start + i;
    const double p = pressure(m->h_owner[f]);
    const double pp = gauge ? p - ref->p_inf : p;
    const double nax = -m->h_nx[f] * m->h_area[f], nay = -m->h_ny[f] * m->h_area[f];
    fx += -pp * nax;
    fy += -pp * nay;
    const double dfx = -pp * nax, dfy = -pp * nay;
    mom += (m->h_fx[f] - ref->x_ref) * dfy - (m->h_fy[f] - ref->y_ref) * dfx;
    if (wall_cp) wall_cp[i] = (p - ref->p_inf) / q_inf;  // extract_wall_cp (forces.cpp:177-206)
  }
  const double pi = 3.14159265358979323846;
  const double a = ref->aoa_deg * (pi / 180.0);
  const double ca = std::cos(a), sa = std::sin(a);
  const double chord = std::max(ref->chord, 1.0e-8);
  const double s_ref = chord * 1.0;
  const double drag = fx * ca + fy * sa, lift = -fx * sa + fy * ca;
  forces[0] = fx;
  forces[1] = fy;
  forces[2] = lift;
  forces[3] = drag;
  forces[4] = mom;
  forces[5] = lift / (q_inf * s_ref);
  forces[6] = drag / (q_inf * s_ref);
  forces[7] = mom / (q_inf * s_ref * chord);
  return AERO_OK;
}

static_assert(aero_dev::kFsFields == AERO_FS_FIELDS && aero_dev::kFsWallMassFlux == AERO_FS_WALL_MASS_FLUX &&
                  aero_dev::kFsAcoustic == AERO_FS_ACOUSTIC,
              "face-scale columns of the kernel and the public header differ");

int aero_hip_face_scales(aero_hip_mesh* m, const aero_residual_config* cfg, const double* conserved,
                         double* face_scales, double* cell_scales, char* err, size_t errlen) {
  if (!m || !cfg || !face_scales) {
    set_err(err, errlen, "aero_hip_face_scales: null argument");
    return AERO_EINVAL;
  }
  if (!(cfg->gamma > 1.0)) {
    set_err(err, errlen, "aero_hip_face_scales: gamma must be > 1");
    return AERO_EINVAL;
  }
  const size_t F = m->F, N = m->N;
  const size_t nf = AERO_FS_FIELDS * F, nc = 2 * N, nu = conserved ? 4 * N : 0;
  void* q = nullptr;
  AERO_HIP_TRY(hipMalloc(&q, (nf + nc + nu) * sizeof(double)), "aero_hip_face_scales: hipMalloc");
  struct Free {
    void* p;
    ~Free() { (void)hipFree(p); }
  } guard{q};
  double* d_faces = static_cast<double*>(q);
  double* d_cells = d_faces + nf;
  const double* U = m->U[m->cur];
  if (conserved) {  // evaluated on a copy: the handle's state stays as it is
    std::vector<double> soa;
    to_internal_state(m, conserved, &soa);
    double* d_u = d_cells + nc;
    AERO_HIP_TRY(hipMemcpyAsync(d_u, soa.data(), nu * sizeof(double), hipMemcpyHostToDevice, m->stream),
                 "aero_hip_face_scales: state copy");
    U = d_u;
  }
  const FaceOpts o = face_opts(cfg);
  AERO_HIP_TRY(aero_host::launch_face_scales(m->dm, o, U, d_faces, d_cells, m->stream), "k_face_scales");
  AERO_HIP_TRY(hipMemcpyAsync(face_scales, d_faces, nf * sizeof(double), hipMemcpyDeviceToHost, m->stream),
               "aero_hip_face_scales: copy");
  std::vector<double> cells;
  if (cell_scales) {
    cells.resize(nc);
    AERO_HIP_TRY(hipMemcpyAsync(cells.data(), d_cells, nc * sizeof(double), hipMemcpyDeviceToHost, m->stream),
                 "aero_hip_face_scales: cell copy");
  }
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_face_scales: hipStreamSynchronize");
  for (size_t c = 0; c < (cell_scales ? N : 0); ++c) {
    cell_scales[m->perm[c]] = cells[c];
    cell_scales[N + m->perm[c]] = cells[N + c];
  }
  return AERO_OK;
}

int aero_hip_get_gradients(aero_hip_mesh* m, double* gradients, char* err, size_t errlen) {
  if (!m || !gradients) {
    set_err(err, errlen, "aero_hip_get_gradients: null argument");
    return AERO_EINVAL;
  }
  std::vector<double> g(static_cast<size_t>(8) * m->N);
  AERO_HIP_TRY(hipMemcpyAsync(g.data(), m->G, g.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream), "G copy");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_get_gradients: hipStreamSynchronize");
  const size_t N = m->N;
  for (int c = 0; c < m->N; ++c) {
    double* o = gradients + 10 * static_cast<size_t>(m->perm[c]);
    o[0] = g[0 * N + c];
    o[1] = g[1 * N + c];
    o[2] = g[2 * N + c];
    o[3] = g[3 * N + c];
    o[4] = g[4 * N + c];
    o[5] = g[5 * N + c];
    o[6] = 0.0;
    o[7] = 0.0;
    o[8] = g[6 * N + c];
    o[9] = g[7 * N + c];
  }
  return AERO_OK;
}

int aero_hip_residual_device(aero_hip_mesh* m, const aero_residual_config* cfg, const double* d_conserved,
                             double* d_residual, double* d_spectral_radius, aero_residual_diag* diag, char* err,
                             size_t errlen) {
  if (!m || !cfg || !d_conserved || !d_residual) {
    set_err(err, errlen, "aero_hip_residual_device: null argument");
    return AERO_EINVAL;
  }
  m->cur = m->base = 0;
  AERO_HIP_TRY(aero_host::launch_aos_to_internal(m->N, m->d_perm, d_conserved, m->U[0], m->stream),
               "aero_hip_residual_device: state");
  if (const int rc = aero_hip_residual_async(m, cfg, err, errlen)) return rc;
  AERO_HIP_TRY(aero_host::launch_internal_to_aos(m->N, m->d_perm, m->R, d_residual, m->lam, d_spectral_radius,
                                                 m->stream),
               "aero_hip_residual_device: residual");
  int dg[4];
  AERO_HIP_TRY(hipMemcpyAsync(dg, m->diag, sizeof(dg), hipMemcpyDeviceToHost, m->stream), "diag copy");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_residual_device: hipStreamSynchronize");
  if (diag) {
    diag->first_order_fallback = dg[0];
    diag->diffusive_fallback = dg[1];
    diag->first_failure_face = dg[2] == INT_MAX ? -1 : dg[2];
  }
  return AERO_OK;
}

int aero_hip_residual(aero_hip_mesh* m, const aero_residual_config* cfg, const double* conserved, double* residual,
                      double* spectral_radius, aero_residual_diag* diag, char* err, size_t errlen) {
  int rc = aero_hip_set_state(m, conserved, err, errlen);
  if (rc) return rc;
  rc = aero_hip_residual_async(m, cfg, err, errlen);
  if (rc) return rc;
  return aero_hip_get_residual(m, residual, spectral_radius, diag, err, errlen);
}

int aero_hip_freeze_limiter(aero_hip_mesh* m, const aero_residual_config* cfg, const double* conserved,
                            double* ratios, char* err, size_t errlen) {
  if (!m || !cfg) {
    set_err(err, errlen, "aero_hip_freeze_limiter: null argument");
    return AERO_EINVAL;
  }
  if (conserved) {
    const int rc = aero_hip_set_state(m, conserved, err, errlen);
    if (rc) return rc;
  }
  const size_t n = 8 * static_cast<size_t>(m->F);
  if (const int rc = ensure_ratio_buffers(m, "aero_hip_freeze_limiter", err, errlen)) return rc;
  const double* U = m->U[m->cur];
  AERO_HIP_TRY(hipMemsetAsync(m->lim_ratio, 0, n * sizeof(double), m->stream), "aero_hip_freeze_limiter");
  AERO_HIP_TRY(aero_host::launch_gradient(m->dm, U, m->G, cfg->gamma, m->stream), "k_gradient");
  AERO_HIP_TRY(aero_host::launch_limiter_capture(m->dm, cfg->limiter, U, m->G, cfg->gamma, m->lim_ratio, m->stream),
               "k_limiter_capture");
  AERO_HIP_TRY(aero_host::launch_ratio_slots(m->dm, m->lim_ratio, m->lim_slot, m->stream), "k_ratio_slots");
  if (ratios)
    AERO_HIP_TRY(hipMemcpyAsync(ratios, m->lim_ratio, n * sizeof(double), hipMemcpyDeviceToHost, m->stream),
                 "aero_hip_freeze_limiter: copy");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_freeze_limiter: sync");
  m->lim_valid = true;
  return AERO_OK;
}

int aero_hip_set_limiter_ratios(aero_hip_mesh* m, const double* ratios, char* err, size_t errlen) {
  if (!m || !ratios) {
    set_err(err, errlen, "aero_hip_set_limiter_ratios: null argument");
    return AERO_EINVAL;
  }
  const size_t n = 8 * static_cast<size_t>(m->F);
  if (const int rc = ensure_ratio_buffers(m, "aero_hip_set_limiter_ratios", err, errlen)) return rc;
  AERO_HIP_TRY(hipMemcpyAsync(m->lim_ratio, ratios, n * sizeof(double), hipMemcpyHostToDevice, m->stream),
               "aero_hip_set_limiter_ratios: copy");
  AERO_HIP_TRY(aero_host::launch_ratio_slots(m->dm, m->lim_ratio, m->lim_slot, m->stream), "k_ratio_slots");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_set_limiter_ratios: sync");
  m->lim_valid = true;
  return AERO_OK;
}

int aero_hip_iterate(aero_hip_mesh* m, const aero_residual_config* cfg, int32_t n_iters, int32_t rk_stages, double cfl,
                     double* rho_l2, char* err, size_t errlen) {
  if (!m || !cfg || n_iters < 0 || (rk_stages != 1 && rk_stages != 3)) {
    set_err(err, errlen, "aero_hip_iterate: bad argument (rk_stages must be 1 or 3)");
    return AERO_EINVAL;
  }
  std::vector<double> sums;
  if (rho_l2) sums.resize(static_cast<size_t>(n_iters));
  double* d_sums = nullptr;
  if (rho_l2 && n_iters > 0) {
    void* q = nullptr;
    AERO_HIP_TRY(hipMalloc(&q, n_iters * sizeof(double)), "aero_hip_iterate: hipMalloc");
    d_sums = static_cast<double*>(q);
  }
  struct Free {
    double* p;
    ~Free() {
      if (p) (void)hipFree(p);
    }
  } guard{d_sums};
  WallArgs wa = m->wall;
  for (int it = 0; it < n_iters; ++it) {
    for (int st = 0; st < rk_stages; ++st) {
      const int rc = aero_hip_stage(m, cfg, st, st == rk_stages - 1, cfl, 1, 0, err, errlen);
      if (rc) return rc;
    }
    if (d_sums) {
      AERO_HIP_TRY(aero_host::launch_iter_reduce(m->dm, m->partials, wa, m->U[m->cur], cfg->gamma, m->iter_out,
                                                 nullptr, m->stream),
                   "k_iter_reduce");
      AERO_HIP_TRY(hipMemcpyAsync(d_sums + it, m->iter_out + 3, sizeof(double), hipMemcpyDeviceToDevice, m->stream),
                   "aero_hip_iterate: norm");
    }
  }
  if (d_sums) {
    AERO_HIP_TRY(hipMemcpyAsync(sums.data(), d_sums, n_iters * sizeof(double), hipMemcpyDeviceToHost, m->stream),
                 "aero_hip_iterate: history");
  }
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_iterate: sync");
  for (int it = 0; it < n_iters && rho_l2; ++it)
    rho_l2[it] = std::sqrt(sums[it] / static_cast<double>(std::max(1, m->N)));
  return AERO_OK;
}

}  // extern "C"

// ---- allocation-free variant of aero_hip_iterate -----------------------------------------------------
// The per-iteration density-residual sums go to a file-scope device array instead of a hipMalloc'd
// buffer.  In host code `scratch` names a host-side stand-in, never the device memory: the device
// address comes from hipGetSymbolAddress, taken once per device with that device current and kept here.
namespace {
constexpr int kNoAllocMaxIters = AERO_NOALLOC_MAX_ITERS;
__device__ __align__(256) double scratch[kNoAllocMaxIters];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (taken on first use on that device).
int scratch_on(int device, double** out, char* err, size_t errlen) {
  if (device < 0 || device >= kMaxDevices) {
    set_err(err, errlen, "aero_hip_iterate_noalloc: device index out of range");
    return AERO_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int prev = 0;
    AERO_HIP_TRY(hipGetDevice(&prev), "aero_hip_iterate_noalloc: hipGetDevice");
    AERO_HIP_TRY(hipSetDevice(device), "aero_hip_iterate_noalloc: hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    AERO_HIP_TRY(e, "aero_hip_iterate_noalloc: hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return AERO_OK;
}
}  // namespace

extern "C" {

int aero_hip_iterate_noalloc(aero_hip_mesh* m, const aero_residual_config* cfg, int32_t n_iters, int32_t rk_stages,
                             double cfl, double* rho_l2, char* err, size_t errlen) {
  if (!m || !cfg || n_iters < 0 || (rk_stages != 1 && rk_stages != 3)) {
    set_err(err, errlen, "aero_hip_iterate_noalloc: bad argument (rk_stages must be 1 or 3)");
    return AERO_EINVAL;
  }
  if (rho_l2 && n_iters > kNoAllocMaxIters) {
    set_err(err, errlen, "aero_hip_iterate_noalloc: n_iters " + std::to_string(n_iters) +
                           " exceeds the history scratch of " + std::to_string(kNoAllocMaxIters) +
                           " iterations (AERO_NOALLOC_MAX_ITERS)");
    return AERO_EINVAL;
  }
  double* d_sums = nullptr;
  if (rho_l2 && n_iters > 0) {
    if (const int rc = scratch_on(m->device, &d_sums, err, errlen)) return rc;
  }
  const WallArgs wa = m->wall;
  for (int it = 0; it < n_iters; ++it) {
    for (int st = 0; st < rk_stages; ++st) {
      const int rc = aero_hip_stage(m, cfg, st, st == rk_stages - 1, cfl, 1, 0, err, errlen);
      if (rc) return rc;
    }
    if (d_sums) {
      AERO_HIP_TRY(aero_host::launch_iter_reduce(m->dm, m->partials, wa, m->U[m->cur], cfg->gamma, m->iter_out,
                                                 nullptr, m->stream),
                   "k_iter_reduce");
      AERO_HIP_TRY(hipMemcpyAsync(d_sums + it, m->iter_out + 3, sizeof(double), hipMemcpyDeviceToDevice, m->stream),
                   "aero_hip_iterate_noalloc: norm");
    }
  }
  if (d_sums)  // every entry read back was written by this call
    AERO_HIP_TRY(hipMemcpyAsync(rho_l2, d_sums, n_iters * sizeof(double), hipMemcpyDeviceToHost, m->stream),
                 "aero_hip_iterate_noalloc: history");
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "aero_hip_iterate_noalloc: sync");
  for (int it = 0; it < n_iters && rho_l2; ++it)
    rho_l2[it] = std::sqrt(rho_l2[it] / static_cast<double>(std::max(1, m->N)));
  return AERO_OK;
}

int aero_hip_stage(aero_hip_mesh* m, const aero_residual_config* cfg, int32_t stage, int32_t last_stage, double cfl,
                   int32_t lts, int32_t write_residual, char* err, size_t errlen) {
  if (!m || !cfg) {
    set_err(err, errlen, "aero_hip_stage: null argument");
    return AERO_EINVAL;
  }
  if (stage < 0 || stage > 2) {
    set_err(err, errlen, "aero_hip_stage: stage must be 0, 1 or 2");
    return AERO_EINVAL;
  }
  if (stage == 0) m->base = m->cur;
  int out = 0;
  while (out == m->cur || out == m->base) ++out;
  FaceOpts o;
  if (const int rc = mesh_face_opts(m, cfg, &o, err, errlen)) return rc;
  const double* U = m->U[m->cur];
  StageArgs st{};
  st.stage = stage;
  st.last_stage = last_stage;
  st.write_res = write_residual || !lts;
  st.cfl = cfl;
  st.Un = m->U[m->base];
  st.Unext = m->U[out];
  st.partials = m->partials;
  if (cfg->second_order) AERO_HIP_TRY(aero_host::launch_gradient(m->dm, U, m->G, o.g, m->stream), "k_gradient");
  if (lts) {
    AERO_HIP_TRY(aero_host::launch_flux(m->dm, o, cfg->flux_scheme, cfg->limiter, cfg->second_order != 0, 1, U, m->G,
                                        m->R, m->lam, m->diag, st, m->stream),
                 "k_flux(stage)");
  } else {
    AERO_HIP_TRY(aero_host::launch_flux(m->dm, o, cfg->flux_scheme, cfg->limiter, cfg->second_order != 0, 0, U, m->G,
                                        m->R, m->lam, m->diag, st, m->stream),
                 "k_flux");
    AERO_HIP_TRY(aero_host::launch_global_update(m->dm, o, U, m->R, m->lam, cfl, m->dt_scratch, m->diag, st, m->stream),
                 "global update");
    if (last_stage) AERO_HIP_TRY(aero_host::launch_norm_partials(m->dm, m->R, m->partials, m->stream), "norms");
  }
  m->cur = out;
  return AERO_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// The driver: run_euler_airfoil_case (euler_solver.cpp:463-1208) on the device-resident state.
// ------------------------------------------------------------------------------------------------
namespace {

struct Coef { double fx, fy, lift, drag, moment, cl, cd, cm; };

class Monitor {  // convergence_monitor.cpp:7-127
 public:
  Monitor(double target, double tol, int window, int min_it, double drift)
      : target_(target), tol_(tol), window_(window), min_it_(min_it), drift_(drift > 0.0 ? drift : tol) {}
  double update_residual(double l2) {
    const double b = std::max(l2, 1.0e-20);
    if (init_ < 0.0) init_ = b;
    return b / std::max(init_, 1.0e-20);
  }
  void update_forces(const Coef& f) {
    const int req = std::max(window_, 1);
    win_.push_back(f);
    while (static_cast<int>(win_.size()) > 2 * req) win_.pop_front();
  }
  bool stable() const {
    const int req = std::max(window_, 1);
    if (static_cast<int>(win_.size()) < req) return false;
    const int b = static_cast<int>(win_.size()) - req;
    double l0 = win_[b].cl, l1 = l0, d0 = win_[b].cd, d1 = d0, m0 = win_[b].cm, m1 = m0;
    for (int i = b; i < static_cast<int>(win_.size()); ++i) {
      l0 = std::min(l0, win_[i].cl);
      l1 = std::max(l1, win_[i].cl);
      d0 = std::min(d0, win_[i].cd);
      d1 = std::max(d1, win_[i].cd);
      m0 = std::min(m0, win_[i].cm);
      m1 = std::max(m1, win_[i].cm);
    }
    return (l1 - l0) <= tol_ && (d1 - d0) <= tol_ && (m1 - m0) <= tol_;
  }
  bool drift_stable() const {
    const int req = std::max(window_, 1);
    if (static_cast<int>(win_.size()) < 2 * req) return false;
    const int pb = static_cast<int>(win_.size()) - 2 * req, cb = static_cast<int>(win_.size()) - req;
    double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
    for (int i = pb; i < cb; ++i) {
      a0 += win_[i].cl;
      a1 += win_[i].cd;
      a2 += win_[i].cm;
    }
    for (int i = cb; i < static_cast<int>(win_.size()); ++i) {
      b0 += win_[i].cl;
      b1 += win_[i].cd;
      b2 += win_[i].cm;
    }
    const double inv = 1.0 / static_cast<double>(req);
    return std::abs(b0 * inv - a0 * inv) <= drift_ && std::abs(b1 * inv - a1 * inv) <= drift_ &&
           std::abs(b2 * inv - a2 * inv) <= drift_;
  }
  bool converged(int iter, double ratio) const {
    if (iter + 1 < min_it_) return false;
    if (ratio > target_) return false;
    return stable() && drift_stable();
  }

 private:
  double target_, tol_;
  int window_, min_it_;
  double drift_;
  double init_ = -1.0;
  std::deque<Coef> win_;
};

double clamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

}  // namespace

extern "C" int aero_hip_run_case(aero_hip_mesh* m, const aero_case_config* c, aero_run_output* out, char* err,
                                 size_t errlen) {
  if (!m || !c || !out) {
    set_err(err, errlen, "aero_hip_run_case: null argument");
    return AERO_EINVAL;
  }
  if (c->iterations <= 0) {
    set_err(err, errlen, "Euler iterations must be positive.");
    return AERO_EINVAL;
  }
  if (c->rk_stages != 1 && c->rk_stages != 3) {
    set_err(err, errlen, "Euler pseudo-time rk_stages must be 1 or 3.");
    return AERO_EINVAL;
  }
  if (m->wall_count <= 0) {
    set_err(err, errlen, "Wall patch face set is empty.");
    return AERO_EINVAL;
  }
  const double g = c->gamma;
  const bool precond = c->precond_on && c->precond_experimental_enabled;
  // build_precondition_config (euler_solver.cpp:341-355)
  const double minf = std::max(c->mach, 1.0e-4);
  const double auto_bmin = std::max(1.0e-4, std::min(1.0e-3, 0.1 * minf));
  const double mach_ref = std::max(c->precond_mach_ref > 0.0 ? c->precond_mach_ref : minf, 1.0e-4);
  const double mach_min = std::max(c->precond_mach_min > 0.0 ? c->precond_mach_min : 1.0e-6, 1.0e-6);
  const double beta_min = std::max(c->precond_beta_min > 0.0 ? c->precond_beta_min : auto_bmin, 1.0e-4);
  const double beta_max = std::max(c->precond_beta_max, beta_min);
  // resolve_stabilization_floor_schedule (:141-157)
  const double k0 = c->stab_floor_k_start >= 0.0 ? c->stab_floor_k_start : (c->stab_floor_start / minf);
  const double k1 = c->stab_floor_k_target >= 0.0 ? c->stab_floor_k_target : (c->stab_floor_target / minf);
  const double floor_start = clamp(std::max(k0, 0.0) * minf, 0.0, 1.0);
  const double floor_target = clamp(std::max(k1, 0.0) * minf, 0.0, 1.0);
  const int stab_ramp = std::max(c->stab_ramp_iters, 0);
  // all-speed schedule and staged controller (:492-528)
  const double as_fmin = clamp(c->all_speed_f_min, 1.0e-6, 1.0);
  const int as_on = std::max(c->all_speed_ramp_start_iter, 0);
  const int as_len = std::max(c->all_speed_ramp_iters, 0);
  const bool staged = c->all_speed_staged_controller != 0;
  const int a_min = std::max(c->all_speed_stage_a_min_iters, 0);
  const int a_max = std::max(c->all_speed_stage_a_max_iters, 0);
  const double f_high = clamp(c->all_speed_stage_f_min_high, 1.0e-6, 1.0);
  const double f_mid = clamp(std::min(c->all_speed_stage_f_min_mid, f_high), 1.0e-6, 1.0);
  const double f_low = clamp(std::min(c->all_speed_stage_f_min_low, f_mid), 1.0e-6, 1.0);
  const double auto_pjump = clamp(1.6e-2 * minf * minf * minf * minf, 1.0e-9, 2.0e-6);
  const double pjump_target = c->all_speed_pjump_wall_target <= 0.0 ? auto_pjump : c->all_speed_pjump_wall_target;
  const double spike_factor = std::max(c->all_speed_pjump_spike_factor, 1.0);
  const int hold_iters = std::max(c->all_speed_pjump_hold_iters, 1);
  const int freeze_iters = std::max(c->all_speed_freeze_iters, 0);
  const double drop = clamp(c->all_speed_cfl_drop_factor, 0.05, 1.0);
  const double restore = clamp(c->all_speed_cfl_restore_factor, 1.0, 1.2);
  const double cfl_min_scale = clamp(c->all_speed_cfl_min_scale, 0.01, 1.0);
  // freestream (:543-561)
  double rho_inf = c->rho_inf;
  if (rho_inf <= 0.0) rho_inf = c->p_inf / (c->gas_constant * c->t_inf);
  if (rho_inf <= 0.0) {
    set_err(err, errlen, "Invalid freestream density. Set rho_inf or p_inf/T_inf.");
    return AERO_EINVAL;
  }
  const double pi = 3.14159265358979323846;
  const double alpha = c->aoa_deg * (pi / 180.0);
  const double a_inf = std::sqrt(g * c->p_inf / rho_inf);
  const double v_inf = std::max(c->mach, 0.0) * a_inf;
  const double u_inf = v_inf * std::cos(alpha), vv_inf = v_inf * std::sin(alpha);
  double ff[5];
  {  // primitive_to_conservative (euler_flux.cpp:101-113)
    const double r = std::max(rho_inf, 1.0e-8);
    ff[0] = r;
    ff[1] = r * u_inf;
    ff[2] = r * vv_inf;
    ff[3] = r * 0.0;
    const double ke = 0.5 * r * (u_inf * u_inf + vv_inf * vv_inf + 0.0 * 0.0);
    ff[4] = std::max(c->p_inf, 1.0e-8) / (g - 1.0) + ke;
  }
  {
    std::vector<double> u0(static_cast<size_t>(m->N) * 5);
    for (int i = 0; i < m->N; ++i) std::memcpy(&u0[5 * static_cast<size_t>(i)], ff, sizeof(ff));
    const int rc = aero_hip_set_state(m, u0.data(), err, errlen);
    if (rc) return rc;
  }
  const double q_inf = 0.5 * rho_inf * v_inf * v_inf + 1.0e-12;
  const double pressure_ref = std::max(rho_inf * a_inf * a_inf, 1.0e-12);
  WallArgs wa = m->wall;
  wa.p_inf = c->p_inf;
  wa.q_inf = q_inf;
  wa.x_ref = c->x_ref;
  wa.y_ref = c->y_ref;
  const double ca = std::cos(alpha), sa = std::sin(alpha);
  const double chord = std::max(c->chord, 1.0e-8);
  const double sref = chord * 1.0;
  auto coefs = [&](const double* sums) {  // forces.cpp:263-293 (gauge pressure)
    Coef k{};
    k.fx = -sums[6];
    k.fy = -sums[7];
    k.moment = sums[8];
    k.drag = k.fx * ca + k.fy * sa;
    k.lift = -k.fx * sa + k.fy * ca;
    k.cd = k.drag / (q_inf * sref);
    k.cl = k.lift / (q_inf * sref);
    k.cm = k.moment / (q_inf * sref * chord);
    return k;
  };
  // Faces of the wall pressure-jump energy (staged controller): interior faces touching a wall-adjacent
  // cell, in increasing face index, as internal (owner, neighbour) ids; the sum runs on the device.
  int n_pj = 0;
  int2* d_pj_faces = nullptr;
  double* d_pj_terms = nullptr;
  struct PjFree {
    int2* f;
    double* t;
    ~PjFree() {
      if (f) (void)hipFree(f);
      if (t) (void)hipFree(t);
    }
  } pj_guard{nullptr, nullptr};
  if (staged) {
    std::vector<unsigned char> wall_adj(m->N, 0u);
    for (int i = 0; i < m->wall_count; ++i) wall_adj[m->h_owner[m->wall_start + i]] = 1u;
    std::vector<int2> pf;
    for (int f = 0; f < m->F; ++f) {
      const int o = m->h_owner[f], nb = m->h_neighbor[f];
      if (nb >= 0 && (wall_adj[o] || wall_adj[nb])) pf.push_back(make_int2(m->iperm[o], m->iperm[nb]));
    }
    n_pj = static_cast<int>(pf.size());
    if (n_pj > 0) {
      void* q = nullptr;
      AERO_HIP_TRY(hipMalloc(&q, pf.size() * sizeof(int2)), "aero_hip_run_case: hipMalloc");
      d_pj_faces = pj_guard.f = static_cast<int2*>(q);
      AERO_HIP_TRY(hipMalloc(&q, pf.size() * sizeof(double)), "aero_hip_run_case: hipMalloc");
      d_pj_terms = pj_guard.t = static_cast<double*>(q);
      AERO_HIP_TRY(hipMemcpy(d_pj_faces, pf.data(), pf.size() * sizeof(int2), hipMemcpyHostToDevice),
                   "aero_hip_run_case: p-jump faces");
    }
  }

  // Per-iteration diagnostics (out->diagnostics): the interior faces, wall-adjacent cells and farfield
  // owners of the reductions, in internal numbering.
  DiagArgs da{};
  int n_parts = 0, n_wall_int = 0;
  double* d_diag_parts = nullptr;
  double* d_diag_out = nullptr;
  std::vector<void*> diag_allocs;
  struct DiagFree {
    std::vector<void*>* v;
    ~DiagFree() {
      for (void* p : *v) (void)hipFree(p);
    }
  } diag_guard{&diag_allocs};
  if (out->diagnostics) {
    std::vector<unsigned char> adj(m->N, 0u);
    for (int i = 0; i < m->wall_count; ++i) adj[m->iperm[m->h_owner[m->wall_start + i]]] = 1u;
    std::vector<int2> faces;
    std::vector<int> ffo;
    for (int f = 0; f < m->F; ++f) {
      const int o = m->iperm[m->h_owner[f]];
      if (m->h_neighbor[f] >= 0) {
        const int n = m->iperm[m->h_neighbor[f]];
        faces.push_back(make_int2(o, n));
        if (adj[o] || adj[n]) ++n_wall_int;
      } else if (m->h_bc[f] == AERO_BC_FARFIELD) {
        ffo.push_back(o);
      }
    }
    auto dalloc = [&](void** p, size_t bytes) -> hipError_t {
      const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
      if (e == hipSuccess) diag_allocs.push_back(*p);
      return e;
    };
    void *pf = nullptr, *pa = nullptr, *po = nullptr, *pp = nullptr, *pd = nullptr;
    da.N = m->N;
    da.ld = m->dm.ld;
    da.nint = static_cast<int>(faces.size());
    da.nff = static_cast<int>(ffo.size());
    n_parts = std::min(256, std::max(1, (std::max(m->N, da.nint) + 255) / 256));
    AERO_HIP_TRY(dalloc(&pf, faces.size() * sizeof(int2)), "diagnostics: hipMalloc");
    AERO_HIP_TRY(dalloc(&pa, adj.size()), "diagnostics: hipMalloc");
    AERO_HIP_TRY(dalloc(&po, ffo.size() * sizeof(int)), "diagnostics: hipMalloc");
    AERO_HIP_TRY(dalloc(&pp, static_cast<size_t>(n_parts) * kDgSlots * sizeof(double)), "diagnostics: hipMalloc");
    AERO_HIP_TRY(dalloc(&pd, kDgSlots * sizeof(double)), "diagnostics: hipMalloc");
    if (!faces.empty())
      AERO_HIP_TRY(hipMemcpy(pf, faces.data(), faces.size() * sizeof(int2), hipMemcpyHostToDevice), "diagnostics");
    AERO_HIP_TRY(hipMemcpy(pa, adj.data(), adj.size(), hipMemcpyHostToDevice), "diagnostics");
    if (!ffo.empty())
      AERO_HIP_TRY(hipMemcpy(po, ffo.data(), ffo.size() * sizeof(int), hipMemcpyHostToDevice), "diagnostics");
    da.faces = static_cast<const int2*>(pf);
    da.wall_adj = static_cast<const unsigned char*>(pa);
    da.ff_owner = static_cast<const int*>(po);
    da.wall = wa;
    da.g = g;
    da.rho_inf = ff[0];
    da.u_inf = u_inf;
    da.v_inf = vv_inf;
    d_diag_parts = static_cast<double*>(pp);
    d_diag_out = static_cast<double*>(pd);
  }

  Monitor mon(c->residual_reduction_target, c->force_stability_tol, c->force_stability_window, c->min_iterations,
              c->force_mean_drift_tol);
  aero_residual_config rc{};
  std::memcpy(rc.farfield, ff, sizeof(ff));
  rc.gamma = g;
  rc.flux_scheme = c->flux_scheme;
  rc.limiter = c->limiter;
  rc.low_mach_fix = c->low_mach_fix;
  rc.low_mach_cutoff = c->mach_cutoff;
  rc.precond_farfield_bc = c->precond_farfield_bc;
  rc.precond_enabled = precond ? 1 : 0;
  rc.precond_mach_ref = mach_ref;
  rc.precond_mach_min = mach_min;
  rc.precond_beta_min = beta_min;
  rc.precond_beta_max = beta_max;
  rc.all_speed_cutoff = c->all_speed_mach_cutoff;

  double cfl_scale = 1.0, prev_l2 = -1.0;
  int stage_ab = staged ? 1 : 0;
  double b_fmin = f_high;
  int b_progress = 0, below = 0, freeze = 0;
  out->n_iter = 0;
  out->nonfinite = 0;
  Coef last{};
  const double* h = m->h_iter;
  const auto t0 = std::chrono::steady_clock::now();
  for (int iter = 0; iter < c->iterations; ++iter) {
    const bool second = iter >= std::max(c->startup_first_order_iters, 0);
    double base;
    if (c->cfl_ramp_iters <= 0) {
      base = c->cfl_max;
    } else {  // compute_cfl (:69-79)
      const double t = std::min(1.0, static_cast<double>(iter + 1) / static_cast<double>(std::max(c->cfl_ramp_iters, 1)));
      base = c->cfl_start + t * (c->cfl_max - c->cfl_start);
    }
    const double cfl = base * cfl_scale;
    double floor_cur;  // compute_stabilization_mach_floor (:95-106)
    {
      const double s0 = clamp(floor_start, 0.0, 1.0), s1 = clamp(floor_target, 0.0, 1.0);
      if (stab_ramp <= 0) {
        floor_cur = s1;
      } else {
        const double x = clamp(static_cast<double>(std::max(iter, 0)) / static_cast<double>(std::max(stab_ramp, 1)), 0.0, 1.0);
        floor_cur = s0 + x * (s1 - s0);
      }
    }
    if (!staged) {  // (:776-797)
      rc.all_speed_fix = c->all_speed_flux_fix;
      rc.all_speed_f_min = as_fmin;
      if (as_len <= 0)
        rc.all_speed_ramp = iter >= as_on ? 1.0 : 0.0;
      else
        rc.all_speed_ramp = clamp(static_cast<double>(iter - as_on) / static_cast<double>(std::max(as_len, 1)), 0.0, 1.0);
    } else if (stage_ab == 1) {
      rc.all_speed_fix = 0;
      rc.all_speed_f_min = f_high;
      rc.all_speed_ramp = 0.0;
    } else {
      rc.all_speed_fix = c->all_speed_flux_fix;
      rc.all_speed_f_min = clamp(b_fmin, f_low, f_high);
      rc.all_speed_ramp =
        as_len <= 0 ? 1.0 : clamp(static_cast<double>(b_progress) / static_cast<double>(std::max(as_len, 1)), 0.0, 1.0);
    }
    rc.second_order = second ? 1 : 0;
    rc.stab_mach_floor = floor_cur;
    if (second && c->limiter_freeze_iter >= 0 && iter == c->limiter_freeze_iter) {  // beyond the reference
      const int fr = aero_hip_freeze_limiter(m, &rc, nullptr, nullptr, err, errlen);
      if (fr) return fr;
      rc.limiter_frozen = 1;
    }
    hipLaunchKernelGGL(k_init_diag, dim3(1), dim3(64), 0, m->stream, m->diag);
    AERO_HIP_TRY(hipGetLastError(), "k_init_diag");
    for (int st = 0; st < c->rk_stages; ++st) {
      const int rcode = aero_hip_stage(m, &rc, st, st == c->rk_stages - 1, cfl, c->local_time_stepping, 1, err, errlen);
      if (rcode) return rcode;
      hipLaunchKernelGGL(k_latch_first_failure, dim3(1), dim3(64), 0, m->stream, m->diag);
      AERO_HIP_TRY(hipGetLastError(), "k_latch_first_failure");
    }
    if (c->aoa0_symmetry_enforce && c->aoa0_symmetry_enforce_interval > 0 && !(std::abs(c->aoa_deg) > 1.0e-12) &&
        ((iter + 1) % c->aoa0_symmetry_enforce_interval) == 0) {
      hipLaunchKernelGGL(k_aoa0_symmetry, dim3((m->N + 255) / 256), dim3(256), 0, m->stream, m->N, m->dm.ld,
                         m->U[m->cur], g);
      AERO_HIP_TRY(hipGetLastError(), "k_aoa0_symmetry");
    }
    AERO_HIP_TRY(aero_host::launch_iter_reduce(m->dm, m->partials, wa, m->U[m->cur], g, m->iter_out, nullptr, m->stream),
                 "k_iter_reduce");
    if (n_pj > 0) {  // iter_out[9]: the wall pressure-jump sum of the new state
      hipLaunchKernelGGL(k_pjump_sum, dim3(1), dim3(256), 0, m->stream, n_pj, d_pj_faces, m->U[m->cur], m->dm.ld, g,
                         d_pj_terms, m->iter_out + 9);
      AERO_HIP_TRY(hipGetLastError(), "k_pjump_sum");
    }
    AERO_HIP_TRY(hipMemcpyAsync(m->h_iter, m->iter_out, (n_pj > 0 ? 10 : 9) * sizeof(double), hipMemcpyDeviceToHost,
                                m->stream),
                 "iteration scalars");
    AERO_HIP_TRY(hipMemcpyAsync(m->h_iter + 10, m->diag, 5 * sizeof(int), hipMemcpyDeviceToHost, m->stream),
                 "iteration counters");
    if (d_diag_out) {  // h_iter[16 ..]: the diagnostics reductions of the new state
      hipLaunchKernelGGL(k_diag_partials, dim3(n_parts), dim3(256), 0, m->stream, da, m->U[m->cur], d_diag_parts);
      hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, m->stream, n_parts, d_diag_parts, d_diag_out);
      AERO_HIP_TRY(hipGetLastError(), "diagnostics reduction");
      AERO_HIP_TRY(hipMemcpyAsync(m->h_iter + 16, d_diag_out, kDgSlots * sizeof(double), hipMemcpyDeviceToHost,
                                  m->stream),
                   "diagnostics copy");
    }
    AERO_HIP_TRY(hipStreamSynchronize(m->stream), "iteration sync");
    const int* hd = reinterpret_cast<const int*>(m->h_iter + 10);
    if (hd[3]) {  // non-finite trap (:908-946)
      std::fprintf(stderr, "non_finite_trap iter=%d location=conserved\n", iter);
      out->nonfinite = 1;
      break;
    }
    const double N5 = static_cast<double>(std::max(1, m->N * 5));
    const double l1 = h[0], l2 = std::sqrt(h[1] / N5), linf = h[2];
    const double rho_l2 = std::sqrt(h[3] / static_cast<double>(std::max(1, m->N)));
    const double ratio = mon.update_residual(l2);
    if (prev_l2 > 0.0) {  // (:951-958)
      if (l2 > prev_l2 * 1.05)
        cfl_scale = std::max(cfl_scale * 0.65, 0.05);
      else if (l2 < prev_l2 * 0.90)
        cfl_scale = std::min(cfl_scale * 1.03, 1.0);
    }
    prev_l2 = l2;
    const Coef fg = coefs(h);
    last = fg;
    mon.update_forces(fg);
    const bool window_stable = mon.stable();
    int action = 0;  // the controller's action code this iteration (euler_solver.cpp:1070-1139)
    if (staged) {  // p-jump diagnostic (device sum, h[9]) + staged all-speed controller (:1026-1139)
      const double pj = n_pj > 0 ? h[9] : 0.0;
      const int pn = n_pj;
      const double pref2 = pressure_ref * pressure_ref + 1.0e-24;
      const double pjw = pn > 0 ? (pj / static_cast<double>(pn)) / pref2 : 0.0;
      if (stage_ab == 1) {
        const bool ready = (iter + 1 >= a_min) && window_stable;
        const bool timeout = (a_max > 0) && (iter + 1 >= a_max);
        if ((ready || timeout) && c->all_speed_flux_fix) {
          stage_ab = 2;
          b_fmin = f_high;
          b_progress = below = freeze = 0;
          action = ready ? 10 : 11;
        }
      } else if (stage_ab == 2) {
        const double thr = pjump_target > 0.0 ? pjump_target * spike_factor : std::numeric_limits<double>::max();
        const bool spike = pjw > thr;
        const bool good = (pjump_target <= 0.0) || (pjw <= pjump_target);
        if (spike) {
          freeze = freeze_iters;
          below = 0;
          b_fmin = std::min(f_high, b_fmin + 0.05);
          cfl_scale = std::max(cfl_scale * drop, cfl_min_scale);
          action = 20;
        } else {
          if (freeze > 0) --freeze;
          below = good ? below + 1 : 0;
          if (freeze == 0) {
            if (b_fmin > f_mid + 1.0e-6 && below >= hold_iters) {
              b_fmin = f_mid;
              below = 0;
              action = 31;
            } else if (b_fmin > f_low + 1.0e-6 && below >= hold_iters) {
              b_fmin = f_low;
              below = 0;
              action = 32;
            } else if (b_fmin <= f_low + 1.0e-6 && good) {
              ++b_progress;
              if (action == 0) action = 33;
            }
            if (good) {
              cfl_scale = std::min(cfl_scale * restore, 1.0);
            } else if (pjump_target > 0.0 && pjw > pjump_target) {
              cfl_scale = std::max(cfl_scale * drop, cfl_min_scale);
              if (action == 0) action = 21;
            }
          }
        }
      }
    }
    if (out->diagnostics && out->n_iter < out->hist_capacity) {  // euler_solver.cpp:921-1060, :1161-1203
      double* dr = out->diagnostics + static_cast<size_t>(out->n_iter) * AERO_DIAG_FIELDS;
      const double* dg = h + 16;
      const double pref = pressure_ref, pref2 = pressure_ref * pressure_ref + 1.0e-24;
      dr[AERO_D_MAX_WALL_UN] = dg[kDgMaxUn];
      dr[AERO_D_MAX_FF_DP] = dg[kDgFfDp];
      dr[AERO_D_MAX_FF_DRHO] = dg[kDgFfDrho];
      dr[AERO_D_MAX_FF_DU] = dg[kDgFfDu];
      dr[AERO_D_MAX_FF_DV] = dg[kDgFfDv];
      dr[AERO_D_WALL_CP_MIN] = m->wall_count > 0 ? dg[kDgCpMin] : 0.0;
      dr[AERO_D_WALL_CP_MAX] = m->wall_count > 0 ? dg[kDgCpMax] : 0.0;
      dr[AERO_D_WALL_CP_NONFINITE] = dg[kDgCpNonfinite];
      dr[AERO_D_FX_GAUGE] = fg.fx;
      dr[AERO_D_FY_GAUGE] = fg.fy;
      const double fxa = -h[4], fya = -h[5];
      dr[AERO_D_FX_ABS] = fxa;
      dr[AERO_D_FY_ABS] = fya;
      dr[AERO_D_CL_ABS] = (-fxa * sa + fya * ca) / (q_inf * sref);
      dr[AERO_D_CD_ABS] = (fxa * ca + fya * sa) / (q_inf * sref);
      dr[AERO_D_CM_ABS] = dg[kDgMomentAbs] / (q_inf * sref * chord);
      dr[AERO_D_LIFT] = fg.lift;
      dr[AERO_D_DRAG] = fg.drag;
      dr[AERO_D_PJUMP_ND] = da.nint > 0 ? (dg[kDgPjAll] / static_cast<double>(da.nint)) / pref2 : 0.0;
      dr[AERO_D_PJUMP_WALL_ND] = n_wall_int > 0 ? (dg[kDgPjWall] / static_cast<double>(n_wall_int)) / pref2 : 0.0;
      dr[AERO_D_PFLUCT_MAX] = dg[kDgPfMax] / pref;
      dr[AERO_D_PFLUCT_WALL] = dg[kDgPfWall] / pref;
      dr[AERO_D_PFLUCT_FARFIELD] = dg[kDgFfDp] / pref;
      dr[AERO_D_AS_STAGE] = stage_ab;
      dr[AERO_D_AS_FIX] = rc.all_speed_fix;
      dr[AERO_D_AS_FMIN] = rc.all_speed_f_min;
      dr[AERO_D_AS_RAMP] = rc.all_speed_ramp;
      dr[AERO_D_AS_FREEZE] = freeze;
      dr[AERO_D_AS_ACTION] = action;
      dr[AERO_D_CFL_SCALE] = cfl_scale;
      dr[AERO_D_FIRST_FAILURE_FACE] = hd[4] == INT_MAX ? -1 : hd[4];
    }
    if (out->history && out->n_iter < out->hist_capacity) {
      double* row = out->history + static_cast<size_t>(out->n_iter) * AERO_HIST_FIELDS;
      row[AERO_H_ITER] = iter;
      row[AERO_H_L1] = l1;
      row[AERO_H_L2] = l2;
      row[AERO_H_LINF] = linf;
      row[AERO_H_RATIO] = ratio;
      row[AERO_H_CL] = fg.cl;
      row[AERO_H_CD] = fg.cd;
      row[AERO_H_CM] = fg.cm;
      row[AERO_H_RHO_L2] = rho_l2;
      row[AERO_H_CFL] = cfl;
      row[AERO_H_FO_FALLBACK] = hd[0];
      row[AERO_H_DIFF_FALLBACK] = hd[1];
      ++out->n_iter;
    }
    if (mon.converged(iter, ratio)) break;
  }
  out->device_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // Final outputs in the reference numbering.
  if (out->wall_cp) {
    AERO_HIP_TRY(aero_host::launch_iter_reduce(m->dm, m->partials, wa, m->U[m->cur], g, m->iter_out, m->cp, m->stream),
                 "final reduce");
    AERO_HIP_TRY(hipMemcpyAsync(out->wall_cp, m->cp, m->wall_count * sizeof(double), hipMemcpyDeviceToHost, m->stream),
                 "cp copy");
  }
  if (out->conserved) {
    std::vector<double> u(static_cast<size_t>(4) * m->N);
    AERO_HIP_TRY(hipMemcpyAsync(u.data(), m->U[m->cur], u.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream),
                 "state copy");
    AERO_HIP_TRY(hipStreamSynchronize(m->stream), "final sync");
    from_internal4(m, u, out->conserved);
  }
  if (out->residual || out->spectral_radius) {
    const int rcode = aero_hip_get_residual(m, out->residual, out->spectral_radius, nullptr, err, errlen);
    if (rcode) return rcode;
  }
  AERO_HIP_TRY(hipStreamSynchronize(m->stream), "final sync");
  if (out->forces) {
    const double f8[8] = {last.fx, last.fy, last.lift, last.drag, last.moment, last.cl, last.cd, last.cm};
    std::memcpy(out->forces, f8, sizeof(f8));
  }
  if (out->final_config) *out->final_config = rc;
  return AERO_OK;
}
