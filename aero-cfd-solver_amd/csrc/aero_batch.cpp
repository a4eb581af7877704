// aero_batch.cpp -- batched pseudo-time solves: many cases (an AoA / Mach sweep, or meshes of uneven
// size) advanced together, one launch per RK stage over all (case, block) work items, with the driver's
// per-iteration control logic on the device so that a whole iteration is one replayable hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "aero_hip.h"
#include "aero_host.hpp"
#include "aero_internal.hpp"

using aero_dev::BatchCase;
using aero_dev::BatchParams;
using aero_dev::BatchStage;
using aero_dev::CaseCtl;
using aero_dev::DevMesh;
using aero_dev::FaceOpts;

// Cases iterate in groups sized to the last-level cache: a group's hot per-case data (state ring, gradients,
// frozen ratios) stays in MI355X's 256 MB Infinity Cache while the group's kernels follow one another, where
// one pass over every case of a large batch would stream all of it from HBM at every launch.  Groups run
// one after another over the same iterations; a case's arithmetic does not depend on the other cases, so
// results are bitwise the same for any grouping.
struct CaseGroup {
  int w0 = 0, nw = 0;  // range of the device work list (padded to a multiple of 8 items)
  int c0 = 0, nc = 0;  // range of the device case list
  bool plain = false;  // every case of the group has plain face options (the flux kernels' PLAIN instantiation)
  hipGraphExec_t graph[2][2] = {};  // [second order][frozen limiter]
};

// Default budget of a group's hot bytes: the size of the Infinity Cache (measured on the 64-case 66k-cell
// sweep: 3 groups of about 270 MB ran its frozen iterations in 1.382 ms, 4 groups of about 200 MB in
// 1.394 ms, one group of 64 in 1.56 ms); and the default least number of work items (workgroups) per group,
// about 16 per CU, below which a launch no longer fills the 256 CUs (8 groups of 8 cases: 1.70 ms).
constexpr size_t kGroupBytesDefault = static_cast<size_t>(256) << 20;
constexpr int kGroupItemsDefault = 4096;

struct aero_hip_batch {
  int ncases = 0, ngeom = 0, BC = 256, grad = 0, smax = 1, nwork = 0, work_cap = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  std::vector<aero_hip_mesh*> geoms;
  std::vector<BatchCase> h_cases;
  std::vector<aero_batch_case> spec;
  DevMesh* d_geoms = nullptr;
  BatchCase* d_cases = nullptr;
  CaseCtl* d_ctl = nullptr;
  int2* d_work = nullptr;          // work list of the active cases, group-major (capacity work_cap)
  int* d_gcases = nullptr;         // case indices of the groups, group-major
  std::vector<CaseGroup> groups;   // the active cases' groups
  size_t group_bytes = kGroupBytesDefault;
  int group_items = kGroupItemsDefault;
  std::vector<char> listed;        // the cases the device work list holds
  double* d_hist = nullptr;
  int hist_cap = 0;
  aero_case_config cfg{};
  bool frozen = false;               // limiter ratios captured (limiter_freeze_iter reached)
  std::vector<double> alpha, q_inf, chord;
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

template <class T> hipError_t balloc(aero_hip_batch* b, T** p, size_t n) {
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess) {
    b->allocs.push_back(q);
    *p = static_cast<T*>(q);
  }
  return e;
}

void destroy_graphs(aero_hip_batch* b) {
  for (CaseGroup& gr : b->groups)
    for (auto& row : gr.graph)
      for (auto& gx : row)
        if (gx) {
          (void)hipGraphExecDestroy(gx);
          gx = nullptr;
        }
}

void free_batch(aero_hip_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  destroy_graphs(b);
  for (void* p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

double clampd(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// The ring buffer that holds a case's current state between iterations (aero_internal.hpp kRingParity).
int state_sel(const aero_case_config& c) { return c.rk_stages == 3 ? 0 : aero_dev::kRingParity; }

// One pseudo-time iteration of the active cases of group gr, enqueued on the batch stream.
int enqueue_iteration(aero_hip_batch* b, bool second, bool frozen, const CaseGroup& gr, char* err, size_t errlen) {
  const aero_case_config& c = b->cfg;
  const int* list = b->d_gcases + gr.c0;
  const int2* work = b->d_work + gr.w0;
  AERO_TRY(aero_host::launch_batch_init_diag(b->d_cases, list, gr.nc, b->stream), "k_batch_init_diag");
  // Ring: RK3 starts and ends in buffer 0 (stage 2 updates in place: each thread reads U^n of its own cell
  // before writing it); RK1 ping-pongs between buffers 0 and 1 by each case's own iteration count.
  const int start = state_sel(c);
  int cur = start;
  int final_buf = start;
  for (int st = 0; st < c.rk_stages; ++st) {
    BatchStage bs{};
    bs.stage = st;
    bs.last = st == c.rk_stages - 1;
    bs.cur = cur;
    bs.base = start;
    if (c.rk_stages == 3)
      bs.out = st == 0 ? 1 : (st == 1 ? 2 : 0);
    else
      bs.out = aero_dev::kRingParityNext;
    if (second)
      AERO_TRY(aero_host::launch_batch_gradient(b->BC, b->grad, b->d_geoms, b->d_cases, b->d_ctl, work, gr.nw,
                                                b->smax, c.gamma, cur, b->stream),
               "k_gradient_batch");
    AERO_TRY(aero_host::launch_batch_flux(b->BC, c.flux_scheme, c.limiter, second, frozen, gr.plain, b->d_geoms,
                                          b->d_cases, b->d_ctl, work, gr.nw, b->smax, bs, b->stream),
             "k_flux_batch");
    cur = bs.out;
    final_buf = bs.out;
  }
  AERO_TRY(aero_host::launch_batch_reduce(b->d_cases, b->d_ctl, list, gr.nc, c.gamma, final_buf, b->stream),
           "k_batch_reduce");
  return 0;
}

// Hot bytes of case i while it iterates: three state buffers and the gradients per cell, the frozen ratios'
// 16-byte record per face slot.
size_t case_hot_bytes(const aero_hip_batch* b, int i) {
  const aero_hip_mesh* m = b->geoms[b->spec[i].mesh_index];
  return static_cast<size_t>(m->N) * (3 * 32 + 64) + static_cast<size_t>(m->S) * 16;
}

// Work-list items of every case, the most a grouping can hold: each group pads to a multiple of 8.
int work_capacity(const aero_hip_batch* b) {
  size_t w = 0;
  for (int i = 0; i < b->ncases; ++i) w += static_cast<size_t>(b->geoms[b->spec[i].mesh_index]->nblk);
  return static_cast<int>(w + 8 * static_cast<size_t>(b->ncases) + 8);
}

// The device work list of the active cases.  The cases, in index order, form k groups of about equal hot
// bytes: k = ceil(hot bytes / group_bytes), at most one per group_items work items and one per case; a
// case goes to group floor(k (bytes before it + half its own) / hot bytes).  Within a group, items are per
// geometry
// block-major / case-minor (the cases of a block share its face data in L2), then spread so that each XCD
// (workgroup index mod 8) receives a contiguous run.  Kernels launched over a group reach only its cases,
// so cases that stopped cost nothing; the graphs, whose grids depend on the groups, are rebuilt.
int upload_work_list(aero_hip_batch* b, const std::vector<char>& active, char* err, size_t errlen) {
  destroy_graphs(b);
  b->groups.clear();
  std::vector<int> gcases;
  std::vector<int2> work;
  std::vector<int> members;
  bool group_plain = false;
  auto close_group = [&]() {
    if (members.empty()) return;
    CaseGroup gr;
    gr.c0 = static_cast<int>(gcases.size());
    gr.nc = static_cast<int>(members.size());
    gr.plain = group_plain;
    gcases.insert(gcases.end(), members.begin(), members.end());
    std::vector<int2> items;
    for (int gi = 0; gi < b->ngeom; ++gi)
      for (int blk = 0; blk < b->geoms[gi]->nblk; ++blk)
        for (const int i : members)
          if (b->spec[i].mesh_index == gi) items.push_back(make_int2(i, blk));
    const int W = static_cast<int>(items.size());
    const int per = (W + 7) / 8;
    gr.w0 = static_cast<int>(work.size());
    gr.nw = per * 8;
    for (int j = 0; j < gr.nw; ++j) {
      const int src = (j & 7) * per + (j >> 3);
      work.push_back(src < W ? items[src] : make_int2(-1, -1));
    }
    b->groups.push_back(gr);
    members.clear();
  };
  // Cases with plain face options and the others are grouped apart, so that the plain groups run the
  // flux kernels' PLAIN instantiation.
  for (const bool want_plain : {true, false}) {
    std::vector<int> part;
    double total = 0.0;
    size_t items = 0;
    for (int i = 0; i < b->ncases; ++i)
      if (active[i] && aero_dev::plain_options(b->h_cases[i].fo) == want_plain) {
        part.push_back(i);
        total += static_cast<double>(case_hot_bytes(b, i));
        items += static_cast<size_t>(b->geoms[b->spec[i].mesh_index]->nblk);
      }
    if (part.empty()) continue;
    const size_t np = part.size();
    size_t k = static_cast<size_t>(std::ceil(total / static_cast<double>(std::max<size_t>(b->group_bytes, 1))));
    k = std::min(k, std::max<size_t>(1, items / static_cast<size_t>(std::max(b->group_items, 1))));
    k = std::max<size_t>(1, std::min(k, np));
    group_plain = want_plain;
    double before = 0.0;
    size_t cur = 0;
    for (const int i : part) {
      const double cb = static_cast<double>(case_hot_bytes(b, i));
      size_t gi = k == np ? cur + (members.empty() ? 0 : 1)
                          : static_cast<size_t>(static_cast<double>(k) * (before + 0.5 * cb) / total);
      gi = std::min(gi, k - 1);
      if (gi != cur && !members.empty()) close_group();
      cur = gi;
      members.push_back(i);
      before += cb;
    }
    close_group();
  }
  if (work.empty()) work.push_back(make_int2(-1, -1));
  if (gcases.empty()) gcases.push_back(0);
  if (static_cast<int>(work.size()) > b->work_cap || static_cast<int>(gcases.size()) > b->ncases) {
    set_err(err, errlen, "upload_work_list: work list exceeds its capacity");
    return AERO_EINVAL;
  }
  AERO_TRY(hipMemcpyAsync(b->d_work, work.data(), work.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream),
           "work list upload");
  AERO_TRY(hipMemcpyAsync(b->d_gcases, gcases.data(), gcases.size() * sizeof(int), hipMemcpyHostToDevice, b->stream),
           "case list upload");
  AERO_TRY(hipStreamSynchronize(b->stream), "work list upload");
  b->nwork = static_cast<int>(work.size());
  b->listed = active;
  return AERO_OK;
}

// Captures one iteration (enqueue) on the batch stream into an executable graph.  The captured graph is
// destroyed on every path, including a failed enqueue, end of capture or instantiation.
template <class Enqueue>
int capture_iteration(aero_hip_batch* b, Enqueue&& enqueue, hipGraphExec_t* gx, char* err, size_t errlen) {
  AERO_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  const int rc = enqueue();
  hipGraph_t captured = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->stream, &captured);
  hipError_t ei = hipSuccess;
  if (!rc && ec == hipSuccess) ei = hipGraphInstantiate(gx, captured, nullptr, nullptr, 0);
  if (captured) (void)hipGraphDestroy(captured);
  if (rc) return rc;
  AERO_TRY(ec, "hipStreamEndCapture");
  AERO_TRY(ei, "hipGraphInstantiate");
  return AERO_OK;
}

}  // namespace

extern "C" {

int aero_hip_batch_create(aero_hip_mesh* const* meshes, int32_t n_meshes, const aero_batch_case* cases,
                          int32_t n_cases, const aero_case_config* c, int32_t hist_capacity, aero_hip_batch** out,
                          char* err, size_t errlen) {
  if (!meshes || n_meshes <= 0 || !cases || n_cases <= 0 || !c || !out) {
    set_err(err, errlen, "aero_hip_batch_create: bad argument");
    return AERO_EINVAL;
  }
  *out = nullptr;
  if (c->rk_stages != 1 && c->rk_stages != 3) {
    set_err(err, errlen, "Euler pseudo-time rk_stages must be 1 or 3.");
    return AERO_EINVAL;
  }
  if (!c->local_time_stepping) {
    set_err(err, errlen, "aero_hip_batch_create: batched solves use local time stepping");
    return AERO_EINVAL;
  }
  if (c->all_speed_staged_controller) {
    set_err(err, errlen, "aero_hip_batch_create: the staged all-speed controller is single-case only");
    return AERO_EINVAL;
  }
  const int BC = meshes[0]->BC;
  for (int g = 0; g < n_meshes; ++g)
    if (!meshes[g] || meshes[g]->BC != BC || meshes[g]->dm.grad != meshes[0]->dm.grad) {
      set_err(err, errlen, "aero_hip_batch_create: all meshes need the same block_cells and gradient");
      return AERO_EINVAL;
    }
  for (int i = 0; i < n_cases; ++i) {
    if (cases[i].mesh_index < 0 || cases[i].mesh_index >= n_meshes) {
      set_err(err, errlen, "aero_hip_batch_create: case " + std::to_string(i) + " has a bad mesh_index");
      return AERO_EINVAL;
    }
    if (cases[i].precond < -1 || cases[i].precond > 1 || cases[i].all_speed_flux_fix < -1 ||
        cases[i].all_speed_flux_fix > 1) {
      set_err(err, errlen, "aero_hip_batch_create: case " + std::to_string(i) + " has a switch outside -1..1");
      return AERO_EINVAL;
    }
    const aero_case_config* cc = cases[i].config;
    if (cc && (cc->flux_scheme != c->flux_scheme || cc->limiter != c->limiter || cc->rk_stages != c->rk_stages ||
               cc->startup_first_order_iters != c->startup_first_order_iters ||
               cc->limiter_freeze_iter != c->limiter_freeze_iter || !cc->local_time_stepping ||
               cc->gamma != c->gamma || cc->all_speed_staged_controller)) {
      set_err(err, errlen, "aero_hip_batch_create: case " + std::to_string(i) +
                             "'s configuration differs from the base in flux scheme, limiter, rk_stages, "
                             "startup iterations, freeze iteration, time stepping, gamma or the staged controller");
      return AERO_EINVAL;
    }
  }
  aero_hip_batch* b = new aero_hip_batch();
  b->ncases = n_cases;
  b->ngeom = n_meshes;
  b->BC = BC;
  b->grad = meshes[0]->dm.grad;
  b->cfg = *c;
  b->spec.assign(cases, cases + n_cases);
  b->geoms.assign(meshes, meshes + n_meshes);
  b->hist_cap = std::max(hist_capacity, 0);
  for (auto* m : b->geoms) b->smax = std::max(b->smax, m->dm.smax);

#define AERO_BSETUP(expr)                                                                      \
  do {                                                                                         \
    const hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                                    \
      set_err(err, errlen, std::string("aero_hip_batch_create: ") + hipGetErrorString(e_));   \
      free_batch(b);                                                                           \
      delete b;                                                                                \
      return AERO_EHIP;                                                                        \
    }                                                                                          \
  } while (0)

  AERO_BSETUP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  const double g = c->gamma;
  std::vector<CaseCtl> ctl(n_cases);
  b->h_cases.resize(n_cases);
  b->alpha.resize(n_cases);
  b->q_inf.resize(n_cases);
  b->chord.resize(n_cases);
  for (int i = 0; i < n_cases; ++i) {
    aero_hip_mesh* m = meshes[cases[i].mesh_index];
    const aero_case_config* cc = cases[i].config ? cases[i].config : c;  // this case's configuration
    BatchCase& cs = b->h_cases[i];
    cs.geom = cases[i].mesh_index;
    cs.nblk = m->nblk;
    cs.N = m->N;
    const size_t N = m->N;
    for (int k = 0; k < 3; ++k) AERO_BSETUP(balloc(b, &cs.U[k], 4 * N));
    AERO_BSETUP(balloc(b, &cs.G, 8 * N));
    cs.R = nullptr;
    cs.lam = nullptr;
    AERO_BSETUP(balloc(b, &cs.partials, 4 * static_cast<size_t>(m->nblk)));
    cs.lim_ratio = nullptr;
    cs.lim_slot = nullptr;
    if (c->limiter_freeze_iter >= 0) {  // frozen limiter (beyond the reference); boundary faces keep 0
      AERO_BSETUP(balloc(b, &cs.lim_ratio, 8 * static_cast<size_t>(m->F)));
      AERO_BSETUP(hipMemset(cs.lim_ratio, 0, 8 * static_cast<size_t>(m->F) * sizeof(double)));
      AERO_BSETUP(balloc(b, &cs.lim_slot, 8 * static_cast<size_t>(m->S)));
      AERO_BSETUP(hipMemset(cs.lim_slot, 0, 8 * static_cast<size_t>(m->S) * sizeof(unsigned short)));
    }
    AERO_BSETUP(balloc(b, &cs.diag, 4));
    // freestream of this case (euler_solver.cpp:543-561) and its schedules
    const double mach = cases[i].mach;
    double rho = cc->rho_inf;
    if (rho <= 0.0) rho = cc->p_inf / (cc->gas_constant * cc->t_inf);
    const double pi = 3.14159265358979323846;
    const double alpha = cases[i].aoa_deg * (pi / 180.0);
    const double a = std::sqrt(g * cc->p_inf / rho);
    const double v = std::max(mach, 0.0) * a;
    const double u = v * std::cos(alpha), w = v * std::sin(alpha);
    const double r = std::max(rho, 1.0e-8);
    const double ke = 0.5 * r * (u * u + w * w + 0.0 * 0.0);
    FaceOpts& fo = cs.fo;
    fo = FaceOpts{};
    fo.ff.r = r;
    fo.ff.ru = r * u;
    fo.ff.rv = r * w;
    fo.ff.E = std::max(cc->p_inf, 1.0e-8) / (g - 1.0) + ke;
    fo.g = g;
    fo.low_mach_fix = cc->low_mach_fix;
    fo.low_mach_cutoff = cc->mach_cutoff;
    fo.char_bc = cc->precond_farfield_bc;
    const int precond = (cc->precond_on && cc->precond_experimental_enabled) ? 1 : 0;
    fo.precond = cases[i].precond == 0 ? precond : (cases[i].precond > 0 ? 1 : 0);
    fo.as_fix = cases[i].all_speed_flux_fix == 0 ? (cc->all_speed_flux_fix ? 1 : 0)
                                                 : (cases[i].all_speed_flux_fix > 0 ? 1 : 0);
    fo.as_cut = cc->all_speed_mach_cutoff;
    fo.as_fmin = clampd(cc->all_speed_f_min, 1.0e-6, 1.0);
    const double minf = std::max(mach, 1.0e-4);
    const double k0 = cc->stab_floor_k_start >= 0.0 ? cc->stab_floor_k_start : (cc->stab_floor_start / minf);
    const double k1 = cc->stab_floor_k_target >= 0.0 ? cc->stab_floor_k_target : (cc->stab_floor_target / minf);
    cs.floor_start = clampd(std::max(k0, 0.0) * minf, 0.0, 1.0);
    cs.floor_target = clampd(std::max(k1, 0.0) * minf, 0.0, 1.0);
    const double auto_bmin = std::max(1.0e-4, std::min(1.0e-3, 0.1 * minf));
    fo.mach_ref = std::max(cc->precond_mach_ref > 0.0 ? cc->precond_mach_ref : minf, 1.0e-4);
    fo.mach_min = std::max(cc->precond_mach_min > 0.0 ? cc->precond_mach_min : 1.0e-6, 1.0e-6);
    fo.beta_min = std::max(cc->precond_beta_min > 0.0 ? cc->precond_beta_min : auto_bmin, 1.0e-4);
    fo.beta_max = std::max(cc->precond_beta_max, fo.beta_min);
    cs.cfl_start = cc->cfl_start;
    cs.cfl_max = cc->cfl_max;
    cs.cfl_ramp = cc->cfl_ramp_iters;
    cs.floor_ramp = std::max(cc->stab_ramp_iters, 0);
    cs.as_on = std::max(cc->all_speed_ramp_start_iter, 0);
    cs.as_len = std::max(cc->all_speed_ramp_iters, 0);
    cs.wall = m->wall;
    cs.wall.p_inf = cc->p_inf;
    cs.wall.q_inf = 0.5 * rho * v * v + 1.0e-12;
    cs.wall.x_ref = cc->x_ref;
    cs.wall.y_ref = cc->y_ref;
    b->alpha[i] = alpha;
    b->q_inf[i] = cs.wall.q_inf;
    b->chord[i] = cc->chord;
    std::vector<double> u0(4 * N);
    for (size_t j = 0; j < N; ++j) {
      u0[j] = fo.ff.r;
      u0[N + j] = fo.ff.ru;
      u0[2 * N + j] = fo.ff.rv;
      u0[3 * N + j] = fo.ff.E;
    }
    AERO_BSETUP(hipMemcpy(cs.U[0], u0.data(), u0.size() * sizeof(double), hipMemcpyHostToDevice));
    AERO_BSETUP(hipMemset(cs.G, 0, 8 * N * sizeof(double)));
    // iteration-0 parameters (euler_solver.cpp:772-797)
    CaseCtl& k = ctl[i];
    std::memset(&k, 0, sizeof(k));
    k.active = 1;
    k.cfl_scale = 1.0;
    k.prev_l2 = -1.0;
    k.l2_0 = -1.0;
    k.rho_l2_0 = -1.0;
    double base;
    if (cc->cfl_ramp_iters <= 0) {
      base = cc->cfl_max;
    } else {
      const double t = std::min(1.0, 1.0 / static_cast<double>(std::max(cc->cfl_ramp_iters, 1)));
      base = cc->cfl_start + t * (cc->cfl_max - cc->cfl_start);
    }
    k.cfl = base;
    k.floor_cur = cs.floor_ramp <= 0 ? cs.floor_target : cs.floor_start;
    const int as_on = cs.as_on, as_len = cs.as_len;
    k.as_w = as_len <= 0 ? (0 >= as_on ? 1.0 : 0.0) : clampd(static_cast<double>(0 - as_on) / std::max(as_len, 1), 0.0, 1.0);
  }
  std::vector<DevMesh> gtab(n_meshes);
  for (int gi = 0; gi < n_meshes; ++gi) gtab[gi] = meshes[gi]->dm;
  AERO_BSETUP(balloc(b, &b->d_geoms, gtab.size()));
  AERO_BSETUP(hipMemcpy(b->d_geoms, gtab.data(), gtab.size() * sizeof(DevMesh), hipMemcpyHostToDevice));
  AERO_BSETUP(balloc(b, &b->d_cases, b->h_cases.size()));
  AERO_BSETUP(hipMemcpy(b->d_cases, b->h_cases.data(), b->h_cases.size() * sizeof(BatchCase), hipMemcpyHostToDevice));
  AERO_BSETUP(balloc(b, &b->d_ctl, ctl.size()));
  AERO_BSETUP(hipMemcpy(b->d_ctl, ctl.data(), ctl.size() * sizeof(CaseCtl), hipMemcpyHostToDevice));
  b->work_cap = work_capacity(b);
  AERO_BSETUP(balloc(b, &b->d_work, static_cast<size_t>(b->work_cap)));
  AERO_BSETUP(balloc(b, &b->d_gcases, static_cast<size_t>(n_cases)));
  if (b->hist_cap > 0) AERO_BSETUP(balloc(b, &b->d_hist, static_cast<size_t>(n_cases) * b->hist_cap * 4));
#undef AERO_BSETUP
  {  // every case starts active
    const int rc = upload_work_list(b, std::vector<char>(n_cases, 1), err, errlen);
    if (rc) {
      free_batch(b);
      delete b;
      return rc;
    }
  }
  *out = b;
  return AERO_OK;
}

void aero_hip_batch_destroy(aero_hip_batch* b) {
  if (!b) return;
  free_batch(b);
  delete b;
}

int aero_hip_batch_run(aero_hip_batch* b, int32_t max_iters, double rho_drop, int32_t fixed_iters,
                       int32_t iters_per_sync, int32_t use_graph, aero_batch_result* results, double* seconds,
                       char* err, size_t errlen) {
  if (!b || max_iters <= 0) {
    set_err(err, errlen, "aero_hip_batch_run: bad argument");
    return AERO_EINVAL;
  }
  const aero_case_config& c = b->cfg;
  BatchParams p{};
  p.ncases = b->ncases;
  p.max_iters = max_iters;
  p.fixed_iters = fixed_iters ? 1 : 0;
  p.rho_drop = rho_drop;
  p.hist = b->d_hist;
  p.hist_cap = b->hist_cap;
  const int startup = std::max(c.startup_first_order_iters, 0);
  const int sync_every = std::max(iters_per_sync, 1);
  const bool compact = true;  // relaunch over the active cases only, whenever the active set changes
  auto enqueue = [&](bool second, const CaseGroup& gr) -> int {
    int rc = enqueue_iteration(b, second, b->frozen, gr, err, errlen);
    if (rc) return rc;
    AERO_TRY(aero_host::launch_batch_control(b->d_cases, b->d_ctl, b->d_gcases + gr.c0, gr.nc, p, b->stream),
             "k_batch_control");
    return 0;
  };
  // Graphs are specific to (order, frozen limiter) and to the run parameters; rebuild them per run.
  destroy_graphs(b);
  // Frozen limiter (beyond the reference): at the start of iteration limiter_freeze_iter (second order)
  // every active case captures its ratios; the iterations run in chunks that end there.
  const int freeze_at =
    (c.limiter_freeze_iter >= 0 && !b->frozen) ? std::max(c.limiter_freeze_iter, std::max(c.startup_first_order_iters, 0))
                                               : -1;
  // Iterations of earlier runs; cases an earlier run stopped at its cap (not converged, not failed) resume.
  int done_iters = 0;
  {
    std::vector<CaseCtl> h0(b->ncases);
    AERO_TRY(hipMemcpy(h0.data(), b->d_ctl, h0.size() * sizeof(CaseCtl), hipMemcpyDeviceToHost), "control readback");
    std::vector<char> act(b->ncases, 0);
    for (int i = 0; i < b->ncases; ++i) {
      CaseCtl& k = h0[i];
      done_iters = std::max(done_iters, k.iter);
      if (!k.converged && !k.failed) k.active = 1;
      act[i] = k.active ? 1 : 0;
    }
    AERO_TRY(hipMemcpy(b->d_ctl, h0.data(), h0.size() * sizeof(CaseCtl), hipMemcpyHostToDevice), "control resume");
    if (act != b->listed) {
      const int rc = upload_work_list(b, act, err, errlen);
      if (rc) return rc;
    }
  }
  p.max_iters = done_iters + max_iters;
  std::vector<CaseCtl> h(b->ncases);
  const auto t0 = std::chrono::steady_clock::now();
  const int start = state_sel(c);
  int it = 0;
  while (it < max_iters) {
    const int g_it = done_iters + it;  // the cases' iteration index
    const bool second = g_it >= startup;
    if (freeze_at >= 0 && !b->frozen && g_it == freeze_at) {
      AERO_TRY(aero_host::launch_batch_gradient(b->BC, b->grad, b->d_geoms, b->d_cases, b->d_ctl, b->d_work, b->nwork,
                                                b->smax, c.gamma, start, b->stream),
               "k_gradient_batch");
      AERO_TRY(aero_host::launch_batch_limiter_capture(b->BC, c.limiter, b->d_geoms, b->d_cases, b->d_ctl, b->d_work,
                                                       b->nwork, c.gamma, start, b->stream),
               "k_limiter_capture_batch");
      for (const BatchCase& cs : b->h_cases)  // the flux kernels' slot-order copy of each case's ratios
        if (cs.lim_ratio)
          AERO_TRY(aero_host::launch_ratio_slots(b->geoms[cs.geom]->dm, cs.lim_ratio, cs.lim_slot, b->stream),
                   "k_ratio_slots");
      b->frozen = true;
    }
    int phase_end = second ? max_iters : std::min(startup - done_iters, max_iters);
    if (freeze_at >= 0 && !b->frozen && g_it < freeze_at) phase_end = std::min(phase_end, freeze_at - done_iters);
    const int n = std::min(sync_every, phase_end - it);
    for (CaseGroup& gr : b->groups)  // each group runs the chunk's n iterations in turn
      for (int k = 0; k < n; ++k) {
        if (use_graph) {
          hipGraphExec_t& gx = gr.graph[second ? 1 : 0][b->frozen ? 1 : 0];
          if (!gx) {
            const int rc = capture_iteration(b, [&] { return enqueue(second, gr); }, &gx, err, errlen);
            if (rc) return rc;
          }
          AERO_TRY(hipGraphLaunch(gx, b->stream), "hipGraphLaunch");
        } else {
          const int rc = enqueue(second, gr);
          if (rc) return rc;
        }
      }
    it += n;
    AERO_TRY(hipMemcpyAsync(h.data(), b->d_ctl, h.size() * sizeof(CaseCtl), hipMemcpyDeviceToHost, b->stream),
             "control readback");
    AERO_TRY(hipStreamSynchronize(b->stream), "aero_hip_batch_run: hipStreamSynchronize");
    std::vector<char> act(b->ncases, 0);
    bool any = false;
    for (int i = 0; i < b->ncases; ++i) {
      act[i] = h[i].active ? 1 : 0;
      any = any || h[i].active;
    }
    if (!any) break;
    if (compact && act != b->listed) {  // drop the stopped cases from the launches
      const int rc = upload_work_list(b, act, err, errlen);
      if (rc) return rc;
    }
  }
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < b->ncases; ++i) {
    const CaseCtl& k = h[i];
    aero_batch_result& r = results[i];
    r.iterations = k.iter;
    r.converged = k.converged;
    r.failed = k.failed;
    r.l2 = k.l2;
    r.rho_l2 = k.rho_l2;
    r.rho_ratio = k.rho_ratio;
    // forces.cpp:263-293 from the gauge sums of the last iteration
    const double fx = -k.sums[6], fy = -k.sums[7], mo = k.sums[8];
    const double ca = std::cos(b->alpha[i]), sa = std::sin(b->alpha[i]);
    const double chord = std::max(b->chord[i], 1.0e-8), q = b->q_inf[i];
    const double drag = fx * ca + fy * sa, lift = -fx * sa + fy * ca;
    r.cd = drag / (q * chord);
    r.cl = lift / (q * chord);
    r.cm = mo / (q * chord * chord);
  }
  return AERO_OK;
}

int aero_hip_batch_get_state(aero_hip_batch* b, int32_t case_index, double* conserved, char* err, size_t errlen) {
  if (!b || case_index < 0 || case_index >= b->ncases || !conserved) {
    set_err(err, errlen, "aero_hip_batch_get_state: bad argument");
    return AERO_EINVAL;
  }
  const BatchCase& cs = b->h_cases[case_index];
  aero_hip_mesh* m = b->geoms[cs.geom];
  const size_t N = cs.N;
  std::vector<double> u(4 * N);
  int buf = 0;  // RK3: buffer 0; RK1: the parity of this case's own iteration count
  if (b->cfg.rk_stages == 1) {
    CaseCtl k{};
    AERO_TRY(hipMemcpy(&k, b->d_ctl + case_index, sizeof(CaseCtl), hipMemcpyDeviceToHost), "control readback");
    buf = k.iter & 1;
  }
  AERO_TRY(hipMemcpy(u.data(), cs.U[buf], u.size() * sizeof(double), hipMemcpyDeviceToHost), "state copy");
  for (size_t ic = 0; ic < N; ++ic) {
    double* o = conserved + 5 * static_cast<size_t>(m->perm[ic]);
    o[0] = u[ic];
    o[1] = u[N + ic];
    o[2] = u[2 * N + ic];
    o[3] = 0.0;
    o[4] = u[3 * N + ic];
  }
  return AERO_OK;
}

int aero_hip_batch_history(aero_hip_batch* b, int32_t case_index, double* hist, int32_t cap, char* err,
                           size_t errlen) {
  if (!b || case_index < 0 || case_index >= b->ncases || !hist || cap < 0) {
    set_err(err, errlen, "aero_hip_batch_history: bad argument");
    return AERO_EINVAL;
  }
  const int n = std::min(cap, b->hist_cap);
  if (n <= 0) return AERO_OK;
  AERO_TRY(hipMemcpy(hist, b->d_hist + static_cast<size_t>(case_index) * b->hist_cap * 4, n * 4 * sizeof(double),
                     hipMemcpyDeviceToHost),
           "history copy");
  return AERO_OK;
}

int aero_hip_batch_set_state(aero_hip_batch* b, int32_t case_index, const double* conserved, char* err,
                             size_t errlen) {
  if (!b || case_index < 0 || case_index >= b->ncases || !conserved) {
    set_err(err, errlen, "aero_hip_batch_set_state: bad argument");
    return AERO_EINVAL;
  }
  const BatchCase& cs = b->h_cases[case_index];
  aero_hip_mesh* m = b->geoms[cs.geom];
  std::vector<double> soa;
  aero_hostutil::to_internal_state(m, conserved, &soa);
  int buf = 0;  // the buffer the case's next iteration starts from (see aero_hip_batch_get_state)
  if (b->cfg.rk_stages == 1) {
    CaseCtl k{};
    AERO_TRY(hipMemcpy(&k, b->d_ctl + case_index, sizeof(CaseCtl), hipMemcpyDeviceToHost), "control readback");
    buf = k.iter & 1;
  }
  AERO_TRY(hipMemcpy(cs.U[buf], soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice), "state copy");
  return AERO_OK;
}

int aero_hip_batch_set_limiter_ratios(aero_hip_batch* b, int32_t case_index, const double* ratios, char* err,
                                      size_t errlen) {
  if (!b || case_index < 0 || case_index >= b->ncases || !ratios) {
    set_err(err, errlen, "aero_hip_batch_set_limiter_ratios: bad argument");
    return AERO_EINVAL;
  }
  const BatchCase& cs = b->h_cases[case_index];
  if (!cs.lim_ratio) {
    set_err(err, errlen, "aero_hip_batch_set_limiter_ratios: the batch has no freeze iteration");
    return AERO_EINVAL;
  }
  aero_hip_mesh* m = b->geoms[cs.geom];
  const size_t n = 8 * static_cast<size_t>(m->F);
  AERO_TRY(hipMemcpyAsync(cs.lim_ratio, ratios, n * sizeof(double), hipMemcpyHostToDevice, b->stream),
           "ratios copy");
  AERO_TRY(aero_host::launch_ratio_slots(m->dm, cs.lim_ratio, cs.lim_slot, b->stream), "k_ratio_slots");
  AERO_TRY(hipStreamSynchronize(b->stream), "aero_hip_batch_set_limiter_ratios: sync");
  b->frozen = true;  // no capture at the freeze iteration from here on
  return AERO_OK;
}

int aero_hip_batch_set_groups(aero_hip_batch* b, int64_t group_bytes, int32_t group_items, int32_t* n_groups,
                              char* err, size_t errlen) {
  if (!b || group_bytes < 0 || group_items < 0) {
    set_err(err, errlen, "aero_hip_batch_set_groups: bad argument");
    return AERO_EINVAL;
  }
  b->group_bytes = group_bytes == 0 ? kGroupBytesDefault : static_cast<size_t>(group_bytes);
  b->group_items = group_items == 0 ? kGroupItemsDefault : group_items;
  const int rc = upload_work_list(b, b->listed, err, errlen);
  if (rc) return rc;
  if (n_groups) *n_groups = static_cast<int32_t>(b->groups.size());
  return AERO_OK;
}

int aero_hip_batch_limiter_ratios(aero_hip_batch* b, int32_t case_index, double* ratios, char* err, size_t errlen) {
  if (!b || case_index < 0 || case_index >= b->ncases || !ratios) {
    set_err(err, errlen, "aero_hip_batch_limiter_ratios: bad argument");
    return AERO_EINVAL;
  }
  const BatchCase& cs = b->h_cases[case_index];
  if (!cs.lim_ratio) {
    set_err(err, errlen, "aero_hip_batch_limiter_ratios: the batch has no freeze iteration");
    return AERO_EINVAL;
  }
  const size_t n = 8 * static_cast<size_t>(b->geoms[cs.geom]->F);
  AERO_TRY(hipMemcpy(ratios, cs.lim_ratio, n * sizeof(double), hipMemcpyDeviceToHost), "ratios copy");
  return AERO_OK;
}

}  // extern "C"
