// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// extern "C" entry points over the templated restatement in aero_oracle.hpp / aero_oracle_driver.hpp,
// one per scalar type (_f32 reproduces the reference's FP32 arithmetic, _f64 is the FP64 parity oracle).
// Loaded by tests/ (ctypes) and by bench.py's cpu_baseline leg only.
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstring>

#include "aero_oracle.hpp"
#include "aero_oracle_driver.hpp"

using namespace aero_oracle;

extern "C" {

// Mesh arrays in the reference layout; the pointer type follows the function suffix.
typedef struct ao_mesh {
  int num_cells, num_faces;
  const int* owner;
  const int* neighbor;
  const int* bc;
  const void* normal;       // [3F]
  const void* face_center;  // [3F]
  const void* area;         // [F]
  const void* volume;       // [N]
  const void* cell_center;  // [3N]
  int wall_start, wall_count;
  int gradient;  // 0: Green-Gauss, 1: weighted least squares
} ao_mesh;

// EulerResidualAssemblyConfig scalars (residual_assembly.hpp:12-31), carried as double.
typedef struct ao_residual_cfg {
  double farfield[5];
  double gamma;
  int scheme, limiter, second_order;
  int low_mach_fix;
  double low_mach_cutoff, stab_mach_floor;
  int precond_farfield_bc;
  int precond_enabled;
  double mach_ref, mach_min, beta_min, beta_max;
  int as_fix;
  double as_cutoff, as_fmin, as_ramp;
  int limiter_frozen;  // the frozen-ratio entry points below take the ratios explicitly
} ao_residual_cfg;

// EulerAirfoilCaseConfig scalars (euler_solver.hpp:14-74), carried as double / int.
typedef struct ao_case_cfg {
  int iterations, min_iterations, cfl_ramp_iters;
  double gamma, gas_constant, mach, aoa_deg, p_inf, t_inf, rho_inf, cfl_start, cfl_max;
  double residual_reduction_target, force_stability_tol;
  int force_stability_window;
  double force_mean_drift_tol;
  int startup_first_order_iters, rk_stages, local_time_stepping;
  int precond_on;
  double precond_mach_ref, precond_mach_min, precond_beta_min, precond_beta_max;
  int precond_farfield_bc;
  double stab_floor_start, stab_floor_target, stab_floor_k_start, stab_floor_k_target;
  int stab_ramp_iters;
  double x_ref, y_ref, chord;
  int flux_scheme, limiter;
  int low_mach_fix;
  double mach_cutoff;
  int all_speed_flux_fix;
  double all_speed_mach_cutoff, all_speed_f_min;
  int all_speed_ramp_start_iter, all_speed_ramp_iters;
  int all_speed_staged_controller, all_speed_stage_a_min_iters, all_speed_stage_a_max_iters;
  double all_speed_stage_f_min_high, all_speed_stage_f_min_mid, all_speed_stage_f_min_low;
  double all_speed_pjump_wall_target, all_speed_pjump_spike_factor;
  int all_speed_pjump_hold_iters, all_speed_freeze_iters;
  double all_speed_cfl_drop_factor, all_speed_cfl_restore_factor, all_speed_cfl_min_scale;
  int aoa0_symmetry_enforce, aoa0_symmetry_enforce_interval;
  int precond_experimental_enabled;
  int limiter_freeze_iter;
} ao_case_cfg;

int ao_hist_fields(void) { return kHistFields; }
int ao_sizeof_case_cfg(void) { return static_cast<int>(sizeof(ao_case_cfg)); }
int ao_sizeof_residual_cfg(void) { return static_cast<int>(sizeof(ao_residual_cfg)); }
}

namespace {

template <class R> MeshView<R> view(const ao_mesh* m) {
  MeshView<R> v;
  v.num_cells = m->num_cells;
  v.num_faces = m->num_faces;
  v.owner = m->owner;
  v.neighbor = m->neighbor;
  v.bc = m->bc;
  v.normal = static_cast<const R*>(m->normal);
  v.face_center = static_cast<const R*>(m->face_center);
  v.area = static_cast<const R*>(m->area);
  v.volume = static_cast<const R*>(m->volume);
  v.cell_center = static_cast<const R*>(m->cell_center);
  v.wall_start = m->wall_start;
  v.wall_count = m->wall_count;
  v.gradient = m->gradient;
  return v;
}

template <class R> AssemblyCfg<R> acfg(const ao_residual_cfg* c) {
  AssemblyCfg<R> a;
  a.farfield = Cons<R>{R(c->farfield[0]), R(c->farfield[1]), R(c->farfield[2]), R(c->farfield[3]), R(c->farfield[4])};
  a.gamma = R(c->gamma);
  a.scheme = c->scheme;
  a.limiter = c->limiter;
  a.second_order = c->second_order != 0;
  a.low_mach_fix = c->low_mach_fix != 0;
  a.low_mach_cutoff = R(c->low_mach_cutoff);
  a.stab_mach_floor = R(c->stab_mach_floor);
  a.precond_farfield_bc = c->precond_farfield_bc != 0;
  a.precond.enabled = c->precond_enabled != 0;
  a.precond.mach_ref = R(c->mach_ref);
  a.precond.mach_min = R(c->mach_min);
  a.precond.beta_min = R(c->beta_min);
  a.precond.beta_max = R(c->beta_max);
  a.as_fix = c->as_fix != 0;
  a.as_cutoff = R(c->as_cutoff);
  a.as_fmin = R(c->as_fmin);
  a.as_ramp = R(c->as_ramp);
  return a;
}

template <class R> CaseCfg<R> ccfg(const ao_case_cfg* c) {
  CaseCfg<R> o;
  o.iterations = c->iterations;
  o.min_iterations = c->min_iterations;
  o.cfl_ramp_iters = c->cfl_ramp_iters;
  o.gamma = R(c->gamma);
  o.gas_constant = R(c->gas_constant);
  o.mach = R(c->mach);
  o.aoa_deg = R(c->aoa_deg);
  o.p_inf = R(c->p_inf);
  o.t_inf = R(c->t_inf);
  o.rho_inf = R(c->rho_inf);
  o.cfl_start = R(c->cfl_start);
  o.cfl_max = R(c->cfl_max);
  o.residual_reduction_target = R(c->residual_reduction_target);
  o.force_stability_tol = R(c->force_stability_tol);
  o.force_stability_window = c->force_stability_window;
  o.force_mean_drift_tol = R(c->force_mean_drift_tol);
  o.startup_first_order_iters = c->startup_first_order_iters;
  o.rk_stages = c->rk_stages;
  o.local_time_stepping = c->local_time_stepping != 0;
  o.precond_on = c->precond_on != 0;
  o.precond_mach_ref = R(c->precond_mach_ref);
  o.precond_mach_min = R(c->precond_mach_min);
  o.precond_beta_min = R(c->precond_beta_min);
  o.precond_beta_max = R(c->precond_beta_max);
  o.precond_farfield_bc = c->precond_farfield_bc != 0;
  o.stab_floor_start = R(c->stab_floor_start);
  o.stab_floor_target = R(c->stab_floor_target);
  o.stab_floor_k_start = R(c->stab_floor_k_start);
  o.stab_floor_k_target = R(c->stab_floor_k_target);
  o.stab_ramp_iters = c->stab_ramp_iters;
  o.x_ref = R(c->x_ref);
  o.y_ref = R(c->y_ref);
  o.chord = R(c->chord);
  o.flux_scheme = c->flux_scheme;
  o.limiter = c->limiter;
  o.low_mach_fix = c->low_mach_fix != 0;
  o.mach_cutoff = R(c->mach_cutoff);
  o.all_speed_flux_fix = c->all_speed_flux_fix != 0;
  o.all_speed_mach_cutoff = R(c->all_speed_mach_cutoff);
  o.all_speed_f_min = R(c->all_speed_f_min);
  o.all_speed_ramp_start_iter = c->all_speed_ramp_start_iter;
  o.all_speed_ramp_iters = c->all_speed_ramp_iters;
  o.all_speed_staged_controller = c->all_speed_staged_controller != 0;
  o.all_speed_stage_a_min_iters = c->all_speed_stage_a_min_iters;
  o.all_speed_stage_a_max_iters = c->all_speed_stage_a_max_iters;
  o.all_speed_stage_f_min_high = R(c->all_speed_stage_f_min_high);
  o.all_speed_stage_f_min_mid = R(c->all_speed_stage_f_min_mid);
  o.all_speed_stage_f_min_low = R(c->all_speed_stage_f_min_low);
  o.all_speed_pjump_wall_target = R(c->all_speed_pjump_wall_target);
  o.all_speed_pjump_spike_factor = R(c->all_speed_pjump_spike_factor);
  o.all_speed_pjump_hold_iters = c->all_speed_pjump_hold_iters;
  o.all_speed_freeze_iters = c->all_speed_freeze_iters;
  o.all_speed_cfl_drop_factor = R(c->all_speed_cfl_drop_factor);
  o.all_speed_cfl_restore_factor = R(c->all_speed_cfl_restore_factor);
  o.all_speed_cfl_min_scale = R(c->all_speed_cfl_min_scale);
  o.aoa0_symmetry_enforce = c->aoa0_symmetry_enforce != 0;
  o.aoa0_symmetry_enforce_interval = c->aoa0_symmetry_enforce_interval;
  o.precond_experimental_enabled = c->precond_experimental_enabled != 0;
  o.limiter_freeze_iter = c->limiter_freeze_iter;
  return o;
}

template <class R> Cons<R> cons_at(const R* u, int i) {
  return Cons<R>{u[5 * i + 0], u[5 * i + 1], u[5 * i + 2], u[5 * i + 3], u[5 * i + 4]};
}
template <class R> A3<R> n_at(const R* n, int i) { return {n[3 * i + 0], n[3 * i + 1], n[3 * i + 2]}; }

template <class R> void limiter_t(int lim, const R* a, const R* b, R* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = apply_limiter<R>(lim, a[i], b[i]);
}

// Pointwise interior-face flux through the reference's dispatcher (euler_flux.cpp:366-390).
template <class R>
void flux_t(int n, int scheme, const R* ul, const R* ur, const R* nrm, const R* scale, R g, int as_fix, R as_cut,
            R as_fmin, R as_w, R* F, R* lam, int* fb) {
  for (int i = 0; i < n; ++i) {
    bool f = false;
    R w = R(0);
    const A5<R> x = face_flux(cons_at(ul, i), cons_at(ur, i), n_at(nrm, i), g, scheme, &w, scale[i], as_fix != 0,
                              as_cut, as_fmin, as_w, &f);
    for (int k = 0; k < 5; ++k) F[5 * i + k] = x[k];
    lam[i] = w;
    fb[i] = f ? 1 : 0;
  }
}

template <class R>
void bflux_t(int n, int bc, int scheme, const R* ui, const R* ff, const R* nrm, const R* scale, R g, int char_bc,
             int as_fix, R as_cut, R as_fmin, R as_w, R* F, R* lam, int* fb) {
  const Cons<R> far = cons_at(ff, 0);
  for (int i = 0; i < n; ++i) {
    bool f = false;
    R w = R(0);
    const A5<R> x = boundary_flux(bc, cons_at(ui, i), far, n_at(nrm, i), scheme, g, scale[i], char_bc != 0,
                                  as_fix != 0, as_cut, as_fmin, as_w, &w, &f);
    for (int k = 0; k < 5; ++k) F[5 * i + k] = x[k];
    lam[i] = w;
    fb[i] = f ? 1 : 0;
  }
}

template <class R> std::vector<Prim<R>> prims(int n, const R* U, R g) {
  std::vector<Prim<R>> p(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) p[i] = cons_to_prim(cons_at(U, i), g);
  return p;
}

template <class R> void gradients_t(const ao_mesh* m, const R* U, R g, R* grad) {
  const MeshView<R> v = view<R>(m);
  std::vector<R> gr;
  cell_gradients(v, prims(v.num_cells, U, g), &gr);
  std::memcpy(grad, gr.data(), gr.size() * sizeof(R));
}

// One residual evaluation as the driver's stage body does it (euler_solver.cpp:803-869):
// primitive conversion, Green-Gauss (or zeros), residual assembly.
template <class R>
void residual_t(const ao_mesh* m, const ao_residual_cfg* c, const R* U, R* res, R* lam, int* diag) {
  const MeshView<R> v = view<R>(m);
  const AssemblyCfg<R> a = acfg<R>(c);
  const std::vector<Prim<R>> p = prims(v.num_cells, U, a.gamma);
  std::vector<R> gr;
  if (a.second_order)
    cell_gradients(v, p, &gr);
  else
    gr.assign(static_cast<std::size_t>(v.num_cells) * 10, R(0));
  std::vector<R> r, l;
  AssemblyDiag d;
  assemble_residual(v, a, p, gr, &r, &l, &d);
  std::memcpy(res, r.data(), r.size() * sizeof(R));
  std::memcpy(lam, l.data(), l.size() * sizeof(R));
  if (diag) {
    diag[0] = d.first_order_fallback;
    diag[1] = d.diffusive_fallback;
    diag[2] = d.first_failure_face;
  }
}

// Per cell and component, sum over its faces of |F_k A| (the scale of the scaled residual tolerance).
template <class R> void flux_abs_sum_t(const ao_mesh* m, const ao_residual_cfg* c, const R* U, R* out) {
  const MeshView<R> v = view<R>(m);
  const AssemblyCfg<R> a = acfg<R>(c);
  const std::vector<Prim<R>> p = prims(v.num_cells, U, a.gamma);
  std::vector<R> gr, r, l, s;
  if (a.second_order)
    cell_gradients(v, p, &gr);
  else
    gr.assign(static_cast<std::size_t>(v.num_cells) * 10, R(0));
  assemble_residual(v, a, p, gr, &r, &l, nullptr, &s);
  std::memcpy(out, s.data(), s.size() * sizeof(R));
}

// Stage tail (euler_solver.cpp:893-899): preconditioned residual copy, dt/V, RK stage update.
template <class R>
void stage_tail_t(int n, int stage, const R* un, const R* us, const R* res, const R* lam, R cfl, int lts,
                  const ao_residual_cfg* c, R* next) {
  const AssemblyCfg<R> a = acfg<R>(c);
  const std::vector<Prim<R>> p = prims(n, us, a.gamma);
  std::vector<R> rp(res, res + 5 * static_cast<std::size_t>(n));
  precond_residual(p, a.gamma, a.precond, &rp);
  const std::vector<R> l(lam, lam + n);
  const std::vector<R> dt = dt_over_volume(l, cfl, lts != 0);
  std::vector<R> vn(un, un + 5 * static_cast<std::size_t>(n)), vs(us, us + 5 * static_cast<std::size_t>(n)), nx;
  rk3_stage_update(stage, vn, vs, rp, dt, a.gamma, &nx);
  std::memcpy(next, nx.data(), nx.size() * sizeof(R));
}

template <class R> void norms_t(const ao_mesh* m, const R* res, R* out4) {
  const MeshView<R> v = view<R>(m);
  const std::vector<R> r(res, res + 5 * static_cast<std::size_t>(v.num_cells));
  const Norms<R> nr = residual_norms(v, r);
  out4[0] = nr.l1;
  out4[1] = nr.l2;
  out4[2] = nr.linf;
  out4[3] = nr.rho_l2;
}

template <class R>
void forces_t(const ao_mesh* m, const R* p, const R* ref7, int subtract, R* out8) {
  const MeshView<R> v = view<R>(m);
  FreeRef<R> f;
  f.rho_inf = ref7[0];
  f.p_inf = ref7[1];
  f.aoa_deg = ref7[2];
  f.speed_inf = ref7[3];
  f.chord = ref7[4];
  f.x_ref = ref7[5];
  f.y_ref = ref7[6];
  const std::vector<R> pv(p, p + v.num_cells);
  const Coeffs<R> c = integrate_forces(v, pv, f, subtract != 0);
  const R o[8] = {c.fx, c.fy, c.lift, c.drag, c.moment, c.cl, c.cd, c.cm};
  std::memcpy(out8, o, sizeof(o));
}

// Full run; returns the number of history rows written (<= hist_cap rows), -1 on error.
// threaded: the stage residuals on all OpenMP threads (the serial run's results bit for bit).
template <class R>
int run_case_t(const ao_mesh* m, const ao_case_cfg* c, int hist_cap, R* hist, R* U, R* res, R* lam, R* forces8,
               R* cp, int* nonfinite, char* err, int errlen, bool threaded = false, R* diag = nullptr) {
  try {
    const RunOut<R> o = run_case(view<R>(m), ccfg<R>(c), threaded, diag != nullptr);
    const int rows = std::min(o.n_iter, hist_cap);
    std::memcpy(hist, o.hist.data(), static_cast<std::size_t>(rows) * kHistFields * sizeof(R));
    if (diag) std::memcpy(diag, o.diag.data(), static_cast<std::size_t>(rows) * kDiagFields * sizeof(R));
    if (U) std::memcpy(U, o.U.data(), o.U.size() * sizeof(R));
    if (res && !o.res.empty()) std::memcpy(res, o.res.data(), o.res.size() * sizeof(R));
    if (lam && !o.lam.empty()) std::memcpy(lam, o.lam.data(), o.lam.size() * sizeof(R));
    if (forces8) {
      const R f[8] = {o.forces.fx, o.forces.fy, o.forces.lift, o.forces.drag,
                      o.forces.moment, o.forces.cl, o.forces.cd, o.forces.cm};
      std::memcpy(forces8, f, sizeof(f));
    }
    if (cp) std::memcpy(cp, o.cp.data(), o.cp.size() * sizeof(R));
    if (nonfinite) *nonfinite = o.nonfinite ? 1 : 0;
    return rows;
  } catch (const std::exception& e) {
    if (err && errlen > 0) {
      std::strncpy(err, e.what(), static_cast<std::size_t>(errlen) - 1);
      err[errlen - 1] = '\0';
    }
    return -1;
  }
}

// The residual configuration of a run's last stage (the ramped stabilisation floor, the all-speed
// weights and the effective preconditioner), the configuration the post-run diagnostics use.
template <class R> int run_case_cfg_t(const ao_mesh* m, const ao_case_cfg* c, ao_residual_cfg* out) {
  try {
    const AssemblyCfg<R> a = run_case(view<R>(m), ccfg<R>(c)).cfg;
    *out = ao_residual_cfg{};
    out->farfield[0] = a.farfield.rho;
    out->farfield[1] = a.farfield.rhou;
    out->farfield[2] = a.farfield.rhov;
    out->farfield[3] = a.farfield.rhow;
    out->farfield[4] = a.farfield.rhoE;
    out->gamma = a.gamma;
    out->scheme = a.scheme;
    out->limiter = a.limiter;
    out->second_order = a.second_order;
    out->low_mach_fix = a.low_mach_fix;
    out->low_mach_cutoff = a.low_mach_cutoff;
    out->stab_mach_floor = a.stab_mach_floor;
    out->precond_farfield_bc = a.precond_farfield_bc;
    out->precond_enabled = a.precond.enabled;
    out->mach_ref = a.precond.mach_ref;
    out->mach_min = a.precond.mach_min;
    out->beta_min = a.precond.beta_min;
    out->beta_max = a.precond.beta_max;
    out->as_fix = a.as_fix;
    out->as_cutoff = a.as_cutoff;
    out->as_fmin = a.as_fmin;
    out->as_ramp = a.as_ramp;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// CPU baseline timing: `n_evals` back-to-back residual evaluations (primitive + gradient + assembly),
// timed with steady_clock around the loop only.  Returns seconds.
template <class R> double time_residual_t(const ao_mesh* m, const ao_residual_cfg* c, const R* U, int n_evals) {
  const MeshView<R> v = view<R>(m);
  const AssemblyCfg<R> a = acfg<R>(c);
  std::vector<R> gr, r, l;
  AssemblyDiag d;
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < n_evals; ++it) {
    const std::vector<Prim<R>> p = prims(v.num_cells, U, a.gamma);
    if (a.second_order)
      cell_gradients(v, p, &gr);
    else
      gr.assign(static_cast<std::size_t>(v.num_cells) * 10, R(0));
    assemble_residual(v, a, p, gr, &r, &l, &d);
  }
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// ---- the same evaluation on all host cores (the CPU baseline's threaded line) ------------------------
// Per-face values in parallel, then per cell the sum over its faces in increasing face index (CellFaces,
// aero_oracle.hpp): the serial result bit for bit.
template <class R>
void residual_threaded(const MeshView<R>& v, const AssemblyCfg<R>& a, const CellFaces& cf, const R* U,
                       std::vector<Prim<R>>* p, std::vector<R>* qf, std::vector<R>* gr, std::vector<R>* fa,
                       std::vector<R>* r, std::vector<R>* l, AssemblyDiag* d) {
  const int N = v.num_cells;
  p->resize(N);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < N; ++c) (*p)[c] = cons_to_prim(cons_at(U, c), a.gamma);
  if (a.second_order)
    cell_gradients_threaded(v, cf, *p, qf, gr);
  else
    gr->assign(static_cast<std::size_t>(N) * 10, R(0));
  assemble_residual_threaded(v, a, cf, *p, *gr, fa, r, l, d);
}

template <class R>
void residual_threaded_t(const ao_mesh* m, const ao_residual_cfg* c, const R* U, R* res, R* lam, int* diag) {
  const MeshView<R> v = view<R>(m);
  const AssemblyCfg<R> a = acfg<R>(c);
  const CellFaces cf = cell_faces(v);
  std::vector<Prim<R>> p;
  std::vector<R> qf, gr, fa, r, l;
  AssemblyDiag d;
  residual_threaded(v, a, cf, U, &p, &qf, &gr, &fa, &r, &l, &d);
  std::memcpy(res, r.data(), r.size() * sizeof(R));
  std::memcpy(lam, l.data(), l.size() * sizeof(R));
  if (diag) {
    diag[0] = d.first_order_fallback;
    diag[1] = d.diffusive_fallback;
    diag[2] = d.first_failure_face;
  }
}

template <class R>
double time_residual_threaded_t(const ao_mesh* m, const ao_residual_cfg* c, const R* U, int n_evals) {
  const MeshView<R> v = view<R>(m);
  const AssemblyCfg<R> a = acfg<R>(c);
  const CellFaces cf = cell_faces(v);
  std::vector<Prim<R>> p;
  std::vector<R> qf, gr, fa, r, l;
  AssemblyDiag d;
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < n_evals; ++it) residual_threaded(v, a, cf, U, &p, &qf, &gr, &fa, &r, &l, &d);
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace

extern "C" int ao_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

#define AO_EXPORT_ALL(SUF, T)                                                                                    \
  extern "C" void ao_limiter_##SUF(int lim, const T* a, const T* b, T* out, int n) { limiter_t<T>(lim, a, b, out, n); } \
  extern "C" void ao_flux_##SUF(int n, int scheme, const T* ul, const T* ur, const T* nrm, const T* scale, double g,    \
                                int as_fix, double as_cut, double as_fmin, double as_w, T* F, T* lam, int* fb) {         \
    flux_t<T>(n, scheme, ul, ur, nrm, scale, T(g), as_fix, T(as_cut), T(as_fmin), T(as_w), F, lam, fb);                  \
  }                                                                                                                      \
  extern "C" void ao_boundary_flux_##SUF(int n, int bc, int scheme, const T* ui, const T* ff, const T* nrm,             \
                                         const T* scale, double g, int char_bc, int as_fix, double as_cut,                \
                                         double as_fmin, double as_w, T* F, T* lam, int* fb) {                           \
    bflux_t<T>(n, bc, scheme, ui, ff, nrm, scale, T(g), char_bc, as_fix, T(as_cut), T(as_fmin), T(as_w), F, lam, fb);   \
  }                                                                                                                      \
  extern "C" void ao_gradients_##SUF(const ao_mesh* m, const T* U, double g, T* grad) { gradients_t<T>(m, U, T(g), grad); } \
  extern "C" void ao_residual_##SUF(const ao_mesh* m, const ao_residual_cfg* c, const T* U, T* res, T* lam, int* diag) { \
    residual_t<T>(m, c, U, res, lam, diag);                                                                              \
  }                                                                                                                      \
  extern "C" void ao_flux_abs_sum_##SUF(const ao_mesh* m, const ao_residual_cfg* c, const T* U, T* out) {           \
    flux_abs_sum_t<T>(m, c, U, out);                                                                                     \
  }                                                                                                                      \
  extern "C" void ao_stage_tail_##SUF(int n, int stage, const T* un, const T* us, const T* res, const T* lam,            \
                                      double cfl, int lts, const ao_residual_cfg* c, T* next) {                          \
    stage_tail_t<T>(n, stage, un, us, res, lam, T(cfl), lts, c, next);                                                   \
  }                                                                                                                      \
  extern "C" void ao_norms_##SUF(const ao_mesh* m, const T* res, T* out4) { norms_t<T>(m, res, out4); }                 \
  extern "C" void ao_forces_##SUF(const ao_mesh* m, const T* p, const T* ref7, int subtract, T* out8) {                 \
    forces_t<T>(m, p, ref7, subtract, out8);                                                                             \
  }                                                                                                                      \
  extern "C" int ao_run_case_##SUF(const ao_mesh* m, const ao_case_cfg* c, int hist_cap, T* hist, T* U, T* res,        \
                                   T* lam, T* forces8, T* cp, int* nonfinite, char* err, int errlen) {                   \
    return run_case_t<T>(m, c, hist_cap, hist, U, res, lam, forces8, cp, nonfinite, err, errlen);                        \
  }                                                                                                                      \
  extern "C" int ao_run_case_threaded_##SUF(const ao_mesh* m, const ao_case_cfg* c, int hist_cap, T* hist, T* U,      \
                                            T* res, T* lam, T* forces8, T* cp, int* nonfinite, char* err,                \
                                            int errlen) {                                                                \
    return run_case_t<T>(m, c, hist_cap, hist, U, res, lam, forces8, cp, nonfinite, err, errlen, true);                 \
  }                                                                                                                      \
  extern "C" int ao_run_case_diag_##SUF(const ao_mesh* m, const ao_case_cfg* c, int hist_cap, T* hist, T* U, T* res,   \
                                        T* lam, T* forces8, T* cp, int* nonfinite, T* diag, char* err, int errlen) {     \
    return run_case_t<T>(m, c, hist_cap, hist, U, res, lam, forces8, cp, nonfinite, err, errlen, true, diag);           \
  }                                                                                                                      \
  extern "C" int ao_run_case_cfg_##SUF(const ao_mesh* m, const ao_case_cfg* c, ao_residual_cfg* out) {              \
    return run_case_cfg_t<T>(m, c, out);                                                                                 \
  }                                                                                                                      \
  extern "C" double ao_time_residual_##SUF(const ao_mesh* m, const ao_residual_cfg* c, const T* U, int n_evals) {      \
    return time_residual_t<T>(m, c, U, n_evals);                                                                         \
  }                                                                                                                      \
  extern "C" void ao_residual_threaded_##SUF(const ao_mesh* m, const ao_residual_cfg* c, const T* U, T* res, T* lam,    \
                                             int* diag) {                                                                \
    residual_threaded_t<T>(m, c, U, res, lam, diag);                                                                     \
  }                                                                                                                      \
  extern "C" double ao_time_residual_threaded_##SUF(const ao_mesh* m, const ao_residual_cfg* c, const T* U,             \
                                                    int n_evals) {                                                       \
    return time_residual_threaded_t<T>(m, c, U, n_evals);                                                                \
  }

AO_EXPORT_ALL(f32, float)
AO_EXPORT_ALL(f64, double)

// ---- frozen limiter (FP64) ------------------------------------------------------------------------
// Limiter ratios [8][F] at the state U (gradients of U, second order), at the frozen limiter's 2^-15
// resolution; the _exact entry gives them unrounded (the pinning tests compare the two).
static void limiter_capture(const ao_mesh* m, const ao_residual_cfg* c, const double* U, double* ratios, bool q) {
  const MeshView<double> v = view<double>(m);
  const AssemblyCfg<double> a = acfg<double>(c);
  const std::vector<Prim<double>> p = prims(v.num_cells, U, a.gamma);
  std::vector<double> gr, out;
  cell_gradients(v, p, &gr);
  limiter_ratios(v, p, gr, a.limiter, &out, q);
  std::memcpy(ratios, out.data(), out.size() * sizeof(double));
}
extern "C" void ao_limiter_capture_f64(const ao_mesh* m, const ao_residual_cfg* c, const double* U, double* ratios) {
  limiter_capture(m, c, U, ratios, true);
}
extern "C" void ao_limiter_capture_exact_f64(const ao_mesh* m, const ao_residual_cfg* c, const double* U,
                                             double* ratios) {
  limiter_capture(m, c, U, ratios, false);
}

// Residual with frozen limiter ratios (second order).
extern "C" void ao_residual_frozen_f64(const ao_mesh* m, const ao_residual_cfg* c, const double* U,
                                       const double* ratios, double* res, double* lam, int* diag) {
  const MeshView<double> v = view<double>(m);
  AssemblyCfg<double> a = acfg<double>(c);
  a.frozen = ratios;
  const std::vector<Prim<double>> p = prims(v.num_cells, U, a.gamma);
  std::vector<double> gr, r, l;
  if (a.second_order)
    cell_gradients(v, p, &gr);
  else
    gr.assign(static_cast<std::size_t>(v.num_cells) * 10, 0.0);
  AssemblyDiag d;
  assemble_residual(v, a, p, gr, &r, &l, &d);
  std::memcpy(res, r.data(), r.size() * sizeof(double));
  std::memcpy(lam, l.data(), l.size() * sizeof(double));
  if (diag) {
    diag[0] = d.first_order_fallback;
    diag[1] = d.diffusive_fallback;
    diag[2] = d.first_failure_face;
  }
}

// ---- post-run scaling diagnostics (FP64) ----------------------------------------------------------
// The reference driver's per-face AcousticScalingInfo (euler_solver.cpp:392-416) with the exterior state
// of its post-processing loop (:1270-1279: interior at a slip wall, freestream at a farfield), the
// dissipation_debug.csv row's speeds and all-speed scale (:1412-1437) and the wall_flux.csv mass flux
// (:1312-1324), out[12][F] in the column order of the device's record; per-cell beta and local Mach
// (:1229-1246), cells[2][N].
extern "C" void ao_face_scales_f64(const ao_mesh* m, const ao_residual_cfg* c, const double* U, double* out,
                                   double* cells) {
  const MeshView<double> v = view<double>(m);
  const AssemblyCfg<double> a = acfg<double>(c);
  const double g = a.gamma;
  auto cons = [&](int i) {
    const double* u = U + 5 * static_cast<std::size_t>(i);
    Cons<double> s;
    s.rho = u[0];
    s.rhou = u[1];
    s.rhov = u[2];
    s.rhow = u[3];
    s.rhoE = u[4];
    return s;
  };
  const std::size_t F = static_cast<std::size_t>(v.num_faces);
  for (int f = 0; f < v.num_faces; ++f) {
    const Cons<double> l = cons(v.owner[f]);
    const int nb = v.neighbor[f];
    const bool wall = nb < 0 && v.bc[f] == kSlipWall;
    const Cons<double> r = nb >= 0 ? cons(nb) : (wall ? l : a.farfield);
    const A3<double> n = normal_of(v, f);
    const Prim<double> pl = cons_to_prim(l, g), pr = cons_to_prim(r, g);
    const double unl = pl.u * n[0] + pl.v * n[1] + pl.w * n[2];
    const double unr = pr.u * n[0] + pr.v * n[1] + pr.w * n[2];
    const double al = speed_of_sound(pl, g), ar = speed_of_sound(pr, g);
    // compute_acoustic_scaling_info, step by step (:399-415)
    const double face_m = 0.5 * (local_mach(pl, g) + local_mach(pr, g));
    const double beta = face_beta(l, r, g, a.precond);
    double base = a.precond.enabled ? sclamp(beta, 1.0e-4, 1.0) : 1.0;
    const double lms =
        a.low_mach_fix ? low_mach_dissipation_scale(l, r, g, true, a.low_mach_cutoff, a.precond.mach_ref) : 1.0;
    base = sclamp(smin(base, lms), 1.0e-4, 1.0);
    const double stab = sclamp(smooth_max(face_m, sclamp(a.stab_mach_floor, 0.0, 1.0)), 1.0e-4, 1.0);
    const double s = smax(base, stab);
    double* o = out + f;
    o[0 * F] = face_m;  // also the row's mach_local: each side's |v| / max(a, 1e-8)
    o[1 * F] = 0.5 * (ao_abs(unl) + ao_abs(unr));
    o[2 * F] = 0.5 * (al + ar);
    o[3 * F] = beta;
    o[4 * F] = lms;
    o[5 * F] = base;
    o[6 * F] = stab;
    o[7 * F] = s;
    o[8 * F] = all_speed_pressure_scale(l, r, n, g, a.as_fix, a.as_cutoff, a.as_fmin, a.as_ramp);
    o[9 * F] = smax(ao_abs(unl) + al, ao_abs(unr) + ar);
    o[10 * F] = smax(precond_spectral_radius(unl, al, s), precond_spectral_radius(unr, ar, s));
    double lam = 0.0;
    o[11 * F] = wall ? boundary_flux(static_cast<int>(kSlipWall), l, a.farfield, n, a.scheme, g, s, false, false, 0.0,
                                     0.0, 0.0, &lam, nullptr)[0]
                     : 0.0;
  }
  if (cells) {
    for (int i = 0; i < v.num_cells; ++i) {
      const Prim<double> p = cons_to_prim(cons(i), g);
      cells[i] = cell_beta(p, g, a.precond);
      cells[static_cast<std::size_t>(v.num_cells) + i] = local_mach(p, g);
    }
  }
}

// ---- tangent (dual-number) evaluation of the FP64 residual ---------------------------------------
#include "oracle_dual.hpp"

// A parameter direction: freestream-dependent options and geometry (reference layout; NULL = zero).
// The layout equals the C ABI's aero_tangent_seed.
typedef struct ao_tangent_seed {
  double d_farfield[5];
  double d_stab_floor, d_mach_ref, d_mach_min, d_beta_min, d_beta_max;
  const double* d_normal;
  const double* d_face_center;
  const double* d_area;
  const double* d_cell_center;
  const double* d_volume;
} ao_tangent_seed;

extern "C" int ao_sizeof_tangent_seed(void) { return static_cast<int>(sizeof(ao_tangent_seed)); }

// R and dR = R_U dU + R_p seed (dU and seed may be NULL) of the FP64 oracle residual.
extern "C" void ao_residual_jvp_f64(const ao_mesh* m, const ao_residual_cfg* c, const double* U, const double* dU,
                                    const ao_tangent_seed* seed, const double* frozen, double* res, double* dres) {
  using D = ODual;
  const int N = m->num_cells, F = m->num_faces;
  auto dual_vec = [](const void* val, const double* der, std::size_t n) {
    const double* v = static_cast<const double*>(val);
    std::vector<D> out(n);
    for (std::size_t i = 0; i < n; ++i) out[i] = D(v[i], der ? der[i] : 0.0);
    return out;
  };
  const std::vector<D> normal = dual_vec(m->normal, seed ? seed->d_normal : nullptr, 3 * static_cast<std::size_t>(F));
  const std::vector<D> fcen =
    dual_vec(m->face_center, seed ? seed->d_face_center : nullptr, 3 * static_cast<std::size_t>(F));
  const std::vector<D> area = dual_vec(m->area, seed ? seed->d_area : nullptr, static_cast<std::size_t>(F));
  const std::vector<D> vol = dual_vec(m->volume, seed ? seed->d_volume : nullptr, static_cast<std::size_t>(N));
  const std::vector<D> ccen =
    dual_vec(m->cell_center, seed ? seed->d_cell_center : nullptr, 3 * static_cast<std::size_t>(N));
  const std::vector<D> u = dual_vec(U, dU, 5 * static_cast<std::size_t>(N));
  MeshView<D> v;
  v.num_cells = N;
  v.num_faces = F;
  v.owner = m->owner;
  v.neighbor = m->neighbor;
  v.bc = m->bc;
  v.normal = normal.data();
  v.face_center = fcen.data();
  v.area = area.data();
  v.volume = vol.data();
  v.cell_center = ccen.data();
  v.wall_start = m->wall_start;
  v.wall_count = m->wall_count;
  v.gradient = m->gradient;
  const double z[5] = {0, 0, 0, 0, 0};
  const double* df = seed ? seed->d_farfield : z;
  AssemblyCfg<D> a;
  a.farfield = Cons<D>{D(c->farfield[0], df[0]), D(c->farfield[1], df[1]), D(c->farfield[2], df[2]),
                       D(c->farfield[3], df[3]), D(c->farfield[4], df[4])};
  a.gamma = D(c->gamma);
  a.scheme = c->scheme;
  a.limiter = c->limiter;
  a.second_order = c->second_order != 0;
  a.low_mach_fix = c->low_mach_fix != 0;
  a.low_mach_cutoff = D(c->low_mach_cutoff);
  a.stab_mach_floor = D(c->stab_mach_floor, seed ? seed->d_stab_floor : 0.0);
  a.precond_farfield_bc = c->precond_farfield_bc != 0;
  a.precond.enabled = c->precond_enabled != 0;
  a.precond.mach_ref = D(c->mach_ref, seed ? seed->d_mach_ref : 0.0);
  a.precond.mach_min = D(c->mach_min, seed ? seed->d_mach_min : 0.0);
  a.precond.beta_min = D(c->beta_min, seed ? seed->d_beta_min : 0.0);
  a.precond.beta_max = D(c->beta_max, seed ? seed->d_beta_max : 0.0);
  a.as_fix = c->as_fix != 0;
  a.as_cutoff = D(c->as_cutoff);
  a.as_fmin = D(c->as_fmin);
  a.as_ramp = D(c->as_ramp);
  a.frozen = frozen;
  const std::vector<Prim<D>> p = prims(N, u.data(), a.gamma);
  std::vector<D> gr;
  if (a.second_order)
    cell_gradients(v, p, &gr);
  else
    gr.assign(static_cast<std::size_t>(N) * 10, D(0.0));
  std::vector<D> r, l;
  AssemblyDiag d;
  assemble_residual(v, a, p, gr, &r, &l, &d);
  for (std::size_t i = 0; i < r.size(); ++i) {
    if (res) res[i] = r[i].v;
    dres[i] = r[i].d;
  }
}
