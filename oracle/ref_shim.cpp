// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" wrapper over the compiled FP32 reference (AstaWasTaken/Aero-CFD-Solver v0.1.1), linked
// with the reference's own sources where they lie (see oracle/Makefile, target `ref`), into
// oracle/_ref/libaero_ref.so.  It is used by tests/golden/make_golden.py (to write golden vectors), by
// tests/test_oracle_vs_ref.py (when the library exists) and by bench.py's CPU-baseline leg, which times
// the reference's residual beside the device kernels.  Nothing here is reference source: it only calls
// the reference's public C++ API.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "cfd_core/mesh.hpp"
#include "cfd_core/mesh/airfoil_mesh.hpp"
#include "cfd_core/numerics/euler_flux.hpp"
#include "cfd_core/numerics/reconstruction.hpp"
#include "cfd_core/post/forces.hpp"
#include "cfd_core/solvers/euler/boundary_conditions.hpp"
#include "cfd_core/solvers/euler/mesh_geometry.hpp"
#include "cfd_core/solvers/euler/preconditioning.hpp"
#include "cfd_core/solvers/euler/residual_assembly.hpp"
#include "cfd_core/solvers/euler/time_integrator.hpp"
#include "cfd_core/solvers/euler_solver.hpp"

using namespace cfd::core;

extern "C" {

typedef struct ref_mesh_cfg {
  int num_circumferential, num_radial;
  float farfield_radius, radial_stretch, chord;
  char naca_code[8];
} ref_mesh_cfg;

typedef struct ref_residual_cfg {
  float farfield[5];
  float gamma;
  int scheme, limiter, second_order;
  int low_mach_fix;
  float low_mach_cutoff, stab_mach_floor;
  int precond_farfield_bc;
  int precond_enabled;
  float mach_ref, mach_min, beta_min, beta_max;
  int as_fix;
  float as_cutoff, as_fmin, as_ramp;
} ref_residual_cfg;

typedef struct ref_case_cfg {
  int iterations, min_iterations, cfl_ramp_iters;
  float gamma, gas_constant, mach, aoa_deg, p_inf, t_inf, rho_inf, cfl_start, cfl_max;
  float residual_reduction_target, force_stability_tol;
  int force_stability_window;
  float force_mean_drift_tol;
  int startup_first_order_iters, rk_stages, local_time_stepping;
  int precond_on;
  float precond_mach_ref, precond_mach_min, precond_beta_min, precond_beta_max;
  int precond_farfield_bc;
  float stab_floor_start, stab_floor_target, stab_floor_k_start, stab_floor_k_target;
  int stab_ramp_iters;
  float x_ref, y_ref;
  int flux_scheme, limiter;
  int low_mach_fix;
  float mach_cutoff;
  int all_speed_flux_fix;
  float all_speed_mach_cutoff, all_speed_f_min;
  int all_speed_ramp_start_iter, all_speed_ramp_iters;
  int all_speed_staged_controller, all_speed_stage_a_min_iters, all_speed_stage_a_max_iters;
  float all_speed_stage_f_min_high, all_speed_stage_f_min_mid, all_speed_stage_f_min_low;
  float all_speed_pjump_wall_target, all_speed_pjump_spike_factor;
  int all_speed_pjump_hold_iters, all_speed_freeze_iters;
  float all_speed_cfl_drop_factor, all_speed_cfl_restore_factor, all_speed_cfl_min_scale;
  int aoa0_symmetry_enforce, aoa0_symmetry_enforce_interval;
} ref_case_cfg;
}

namespace {
AirfoilMeshConfig mesh_cfg(const ref_mesh_cfg* c) {
  AirfoilMeshConfig m;
  m.airfoil_source = "naca4";
  m.naca_code = std::string(c->naca_code);
  m.chord = c->chord;
  m.num_circumferential = c->num_circumferential;
  m.num_radial = c->num_radial;
  m.farfield_radius = c->farfield_radius;
  m.radial_stretch = c->radial_stretch;
  return m;
}

UnstructuredMesh make_mesh(const ref_mesh_cfg* c) {
  UnstructuredMesh m = make_airfoil_ogrid_mesh(mesh_cfg(c));
  normalize_face_normals(&m);
  return m;
}

ConservativeState cons_at(const float* u, int i) {
  return {u[5 * i + 0], u[5 * i + 1], u[5 * i + 2], u[5 * i + 3], u[5 * i + 4]};
}

EulerResidualAssemblyConfig assembly_cfg(const ref_residual_cfg* c, const UnstructuredMesh* mesh,
                                         const std::vector<PrimitiveState>* prim, const PrimitiveGradients* grad,
                                         const std::vector<EulerBoundaryConditionType>* bt) {
  EulerResidualAssemblyConfig a;
  a.mesh = mesh;
  a.primitive = prim;
  a.gradients = grad;
  a.boundary_type = bt;
  a.farfield_state = {c->farfield[0], c->farfield[1], c->farfield[2], c->farfield[3], c->farfield[4]};
  a.gamma = c->gamma;
  a.flux_scheme = static_cast<EulerFluxScheme>(c->scheme);
  a.limiter = static_cast<LimiterType>(c->limiter);
  a.use_second_order = c->second_order != 0;
  a.low_mach_dissipation_fix = c->low_mach_fix != 0;
  a.low_mach_dissipation_cutoff = c->low_mach_cutoff;
  a.stabilization_mach_floor = c->stab_mach_floor;
  a.preconditioned_farfield_characteristic = c->precond_farfield_bc != 0;
  a.precondition.enabled = c->precond_enabled != 0;
  a.precondition.mach_ref = c->mach_ref;
  a.precondition.mach_min = c->mach_min;
  a.precondition.beta_min = c->beta_min;
  a.precondition.beta_max = c->beta_max;
  a.all_speed_flux_fix = c->as_fix != 0;
  a.all_speed_mach_cutoff = c->as_cutoff;
  a.all_speed_f_min = c->as_fmin;
  a.all_speed_ramp_weight = c->as_ramp;
  return a;
}
}  // namespace

extern "C" {

void ref_limiter(int lim, const float* a, const float* b, float* out, int n) {
  for (int i = 0; i < n; ++i) out[i] = apply_limiter(static_cast<LimiterType>(lim), a[i], b[i]);
}

void ref_flux(int n, int scheme, const float* ul, const float* ur, const float* nrm, const float* scale, float g,
              int as_fix, float as_cut, float as_fmin, float as_w, float* F, float* lam, int* fb) {
  for (int i = 0; i < n; ++i) {
    FluxDebug dbg;
    float w = 0.0f;
    const std::array<float, 3> nn = {nrm[3 * i + 0], nrm[3 * i + 1], nrm[3 * i + 2]};
    const auto x = compute_euler_face_flux(cons_at(ul, i), cons_at(ur, i), nn, 1.0f, g,
                                           static_cast<EulerFluxScheme>(scheme), &w, scale[i], as_fix != 0, as_cut,
                                           as_fmin, as_w, &dbg);
    for (int k = 0; k < 5; ++k) F[5 * i + k] = x[k];
    lam[i] = w;
    fb[i] = dbg.used_diffusive_fallback ? 1 : 0;
  }
}

void ref_boundary_flux(int n, int bc, int scheme, const float* ui, const float* ff, const float* nrm,
                       const float* scale, float g, int char_bc, int as_fix, float as_cut, float as_fmin, float as_w,
                       float* F, float* lam, int* fb) {
  for (int i = 0; i < n; ++i) {
    FluxDebug dbg;
    const std::array<float, 3> nn = {nrm[3 * i + 0], nrm[3 * i + 1], nrm[3 * i + 2]};
    const BoundaryFluxResult r =
      compute_boundary_flux(static_cast<EulerBoundaryConditionType>(bc), cons_at(ui, i), cons_at(ff, 0), nn,
                            static_cast<EulerFluxScheme>(scheme), g, scale[i], char_bc != 0, as_fix != 0, as_cut,
                            as_fmin, as_w, &dbg);
    for (int k = 0; k < 5; ++k) F[5 * i + k] = r.flux[k];
    lam[i] = r.max_wave_speed;
    fb[i] = dbg.used_diffusive_fallback ? 1 : 0;
  }
}

// Mesh sizes, then arrays, from the reference generator (airfoil_mesh.cpp:459-544) after the
// driver's normal re-normalisation (euler_solver.cpp:475).
void ref_mesh_sizes(const ref_mesh_cfg* c, int* n_cells, int* n_faces, int* n_points, int* wall_start,
                    int* wall_count, int* far_start, int* far_count) {
  const UnstructuredMesh m = make_mesh(c);
  *n_cells = m.num_cells;
  *n_faces = m.num_faces;
  *n_points = static_cast<int>(m.points.size() / 3);
  *wall_start = *wall_count = *far_start = *far_count = 0;
  for (const auto& p : m.boundary_patches) {
    if (p.name == "wall") {
      *wall_start = p.start_face;
      *wall_count = p.face_count;
    } else if (p.name == "farfield") {
      *far_start = p.start_face;
      *far_count = p.face_count;
    }
  }
}

void ref_mesh_fill(const ref_mesh_cfg* c, int* owner, int* neighbor, int* bc, int* face_v, float* normal,
                   float* face_center, float* area, float* volume, float* cell_center, float* points) {
  const UnstructuredMesh m = make_mesh(c);
  const auto bt = build_euler_boundary_types(m);
  std::memcpy(owner, m.face_owner.data(), m.face_owner.size() * sizeof(int));
  std::memcpy(neighbor, m.face_neighbor.data(), m.face_neighbor.size() * sizeof(int));
  for (int f = 0; f < m.num_faces; ++f) bc[f] = static_cast<int>(bt[f]);
  std::memcpy(face_v, m.face_vertices.data(), m.face_vertices.size() * sizeof(int));
  std::memcpy(normal, m.face_normal.data(), m.face_normal.size() * sizeof(float));
  std::memcpy(face_center, m.face_center.data(), m.face_center.size() * sizeof(float));
  std::memcpy(area, m.face_area.data(), m.face_area.size() * sizeof(float));
  std::memcpy(volume, m.cell_volume.data(), m.cell_volume.size() * sizeof(float));
  std::memcpy(cell_center, m.cell_center.data(), m.cell_center.size() * sizeof(float));
  std::memcpy(points, m.points.data(), m.points.size() * sizeof(float));
}

void ref_gradients(const ref_mesh_cfg* c, const float* U, float g, float* grad) {
  const UnstructuredMesh m = make_mesh(c);
  std::vector<PrimitiveState> prim(static_cast<std::size_t>(m.num_cells));
  for (int i = 0; i < m.num_cells; ++i) prim[i] = conservative_to_primitive(cons_at(U, i), g);
  const PrimitiveGradients gr = compute_green_gauss_gradients(m, prim);
  std::memcpy(grad, gr.values.data(), gr.values.size() * sizeof(float));
}

}  // extern "C"

namespace {
// One residual evaluation exactly as the driver's stage body (euler_solver.cpp:803-869).
void residual_on(const UnstructuredMesh& m, const ref_residual_cfg* rc, const float* U, float* res, float* lam,
                 int* diag) {
  std::vector<PrimitiveState> prim(static_cast<std::size_t>(m.num_cells));
  for (int i = 0; i < m.num_cells; ++i) prim[i] = conservative_to_primitive(cons_at(U, i), rc->gamma);
  PrimitiveGradients gr;
  if (rc->second_order) {
    gr = compute_green_gauss_gradients(m, prim);
  } else {
    gr.num_components = 2;
    gr.values.assign(static_cast<std::size_t>(m.num_cells) * 10, 0.0f);
  }
  const auto bt = build_euler_boundary_types(m);
  std::vector<float> r(static_cast<std::size_t>(m.num_cells) * 5, 0.0f), l(static_cast<std::size_t>(m.num_cells), 0.0f);
  EulerResidualAssemblyDiagnostics d;
  assemble_euler_residual_cpu(assembly_cfg(rc, &m, &prim, &gr, &bt), &r, &l, &d);
  std::memcpy(res, r.data(), r.size() * sizeof(float));
  std::memcpy(lam, l.data(), l.size() * sizeof(float));
  diag[0] = d.num_faces_first_order_fallback;
  diag[1] = d.num_faces_diffusive_fallback;
  diag[2] = d.first_failure_face;
}
}  // namespace

extern "C" {

void ref_residual(const ref_mesh_cfg* c, const ref_residual_cfg* rc, const float* U, float* res, float* lam,
                  int* diag) {
  residual_on(make_mesh(c), rc, U, res, lam, diag);
}

// A reference mesh from caller arrays, for meshes the reference's generator does not build (the mixed,
// jittered meshes the HIP backend is timed on): unit normals and face / cell centres xyz-interleaved, the
// wall patch [wall_start, wall_start + wall_count), every other boundary face farfield.  Freed with
// ref_mesh_delete.
void* ref_mesh_new(int n_cells, int n_faces, const int* owner, const int* neighbor, const float* normal,
                   const float* face_center, const float* area, const float* volume, const float* cell_center,
                   int wall_start, int wall_count) {
  auto* m = new UnstructuredMesh();
  m->num_cells = n_cells;
  m->num_faces = n_faces;
  m->face_owner.assign(owner, owner + n_faces);
  m->face_neighbor.assign(neighbor, neighbor + n_faces);
  m->face_vertices.assign(2 * static_cast<std::size_t>(n_faces), 0);
  m->face_normal.assign(normal, normal + 3 * static_cast<std::size_t>(n_faces));
  m->face_center.assign(face_center, face_center + 3 * static_cast<std::size_t>(n_faces));
  m->face_area.assign(area, area + n_faces);
  m->cell_volume.assign(volume, volume + n_cells);
  m->cell_center.assign(cell_center, cell_center + 3 * static_cast<std::size_t>(n_cells));
  BoundaryPatchRange wall;
  wall.name = "wall";
  wall.start_face = wall_start;
  wall.face_count = wall_count;
  m->boundary_patches.push_back(wall);
  return m;
}

void ref_mesh_delete(void* mesh) { delete static_cast<UnstructuredMesh*>(mesh); }

// ref_residual on a mesh from ref_mesh_new.
void ref_residual_mesh(const void* mesh, const ref_residual_cfg* rc, const float* U, float* res, float* lam,
                       int* diag) {
  residual_on(*static_cast<const UnstructuredMesh*>(mesh), rc, U, res, lam, diag);
}

void ref_stage_tail(int n, int stage, const float* un, const float* us, const float* res, const float* lam, float cfl,
                    int lts, const ref_residual_cfg* rc, float* next) {
  std::vector<PrimitiveState> prim(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) prim[i] = conservative_to_primitive(cons_at(us, i), rc->gamma);
  LowMachPreconditionConfig pc;
  pc.enabled = rc->precond_enabled != 0;
  pc.mach_ref = rc->mach_ref;
  pc.mach_min = rc->mach_min;
  pc.beta_min = rc->beta_min;
  pc.beta_max = rc->beta_max;
  std::vector<float> rp(res, res + 5 * static_cast<std::size_t>(n));
  apply_low_mach_preconditioned_residual(prim, rc->gamma, pc, &rp);
  const std::vector<float> l(lam, lam + n);
  const std::vector<float> dt = compute_pseudo_time_step_over_volume(l, cfl, lts != 0);
  const std::vector<float> vn(un, un + 5 * static_cast<std::size_t>(n)), vs(us, us + 5 * static_cast<std::size_t>(n));
  std::vector<float> nx;
  apply_rk3_stage_update(stage, vn, vs, rp, dt, rc->gamma, &nx);
  std::memcpy(next, nx.data(), nx.size() * sizeof(float));
}

void ref_forces(const ref_mesh_cfg* c, const float* p, const float* ref7, int subtract, float* out8) {
  const UnstructuredMesh m = make_mesh(c);
  const FreestreamReference r = {ref7[0], ref7[1], ref7[2], ref7[3], ref7[4], ref7[5], ref7[6]};
  const std::vector<float> pv(p, p + m.num_cells);
  const ForceCoefficients f = integrate_pressure_forces(m, pv, r, "wall", subtract != 0);
  const float o[8] = {f.fx, f.fy, f.lift, f.drag, f.moment, f.cl, f.cd, f.cm};
  std::memcpy(out8, o, sizeof(o));
}

constexpr int kRunPost = 23;
static float g_run_post[kRunPost] = {};

// Full reference run (run_euler_airfoil_case_cpu).  History rows: iter, l1, l2, linf, cl, cd, cm.
// The precond gate reads AERO_ENABLE_EXPERIMENTAL_PRECOND from the process environment.
int ref_run_case(const ref_mesh_cfg* mc, const ref_case_cfg* c, const char* out_dir, int hist_cap, float* hist,
                 float* U, float* res, float* lam, float* forces8, float* cp_patch, float* cp_ordered_xyc,
                 int* n_cp_ordered) {
  EulerAirfoilCaseConfig e;
  e.mesh = mesh_cfg(mc);
  e.iterations = c->iterations;
  e.min_iterations = c->min_iterations;
  e.cfl_ramp_iters = c->cfl_ramp_iters;
  e.gamma = c->gamma;
  e.gas_constant = c->gas_constant;
  e.mach = c->mach;
  e.aoa_deg = c->aoa_deg;
  e.p_inf = c->p_inf;
  e.t_inf = c->t_inf;
  e.rho_inf = c->rho_inf;
  e.cfl_start = c->cfl_start;
  e.cfl_max = c->cfl_max;
  e.residual_reduction_target = c->residual_reduction_target;
  e.force_stability_tol = c->force_stability_tol;
  e.force_stability_window = c->force_stability_window;
  e.force_mean_drift_tol = c->force_mean_drift_tol;
  e.startup_first_order_iters = c->startup_first_order_iters;
  e.rk_stages = c->rk_stages;
  e.local_time_stepping = c->local_time_stepping != 0;
  e.precond_on = c->precond_on != 0;
  e.precond_mach_ref = c->precond_mach_ref;
  e.precond_mach_min = c->precond_mach_min;
  e.precond_beta_min = c->precond_beta_min;
  e.precond_beta_max = c->precond_beta_max;
  e.precond_farfield_bc = c->precond_farfield_bc != 0;
  e.stabilization_mach_floor_start = c->stab_floor_start;
  e.stabilization_mach_floor_target = c->stab_floor_target;
  e.stabilization_mach_floor_k_start = c->stab_floor_k_start;
  e.stabilization_mach_floor_k_target = c->stab_floor_k_target;
  e.stabilization_ramp_iters = c->stab_ramp_iters;
  e.x_ref = c->x_ref;
  e.y_ref = c->y_ref;
  e.flux_scheme = static_cast<EulerFluxScheme>(c->flux_scheme);
  e.limiter = static_cast<LimiterType>(c->limiter);
  e.low_mach_fix = c->low_mach_fix != 0;
  e.mach_cutoff = c->mach_cutoff;
  e.all_speed_flux_fix = c->all_speed_flux_fix != 0;
  e.all_speed_mach_cutoff = c->all_speed_mach_cutoff;
  e.all_speed_f_min = c->all_speed_f_min;
  e.all_speed_ramp_start_iter = c->all_speed_ramp_start_iter;
  e.all_speed_ramp_iters = c->all_speed_ramp_iters;
  e.all_speed_staged_controller = c->all_speed_staged_controller != 0;
  e.all_speed_stage_a_min_iters = c->all_speed_stage_a_min_iters;
  e.all_speed_stage_a_max_iters = c->all_speed_stage_a_max_iters;
  e.all_speed_stage_f_min_high = c->all_speed_stage_f_min_high;
  e.all_speed_stage_f_min_mid = c->all_speed_stage_f_min_mid;
  e.all_speed_stage_f_min_low = c->all_speed_stage_f_min_low;
  e.all_speed_pjump_wall_target = c->all_speed_pjump_wall_target;
  e.all_speed_pjump_spike_factor = c->all_speed_pjump_spike_factor;
  e.all_speed_pjump_hold_iters = c->all_speed_pjump_hold_iters;
  e.all_speed_freeze_iters = c->all_speed_freeze_iters;
  e.all_speed_cfl_drop_factor = c->all_speed_cfl_drop_factor;
  e.all_speed_cfl_restore_factor = c->all_speed_cfl_restore_factor;
  e.all_speed_cfl_min_scale = c->all_speed_cfl_min_scale;
  e.aoa0_symmetry_enforce = c->aoa0_symmetry_enforce != 0;
  e.aoa0_symmetry_enforce_interval = c->aoa0_symmetry_enforce_interval;
  e.output_dir = std::filesystem::path(out_dir);
  const EulerRunResult r = run_euler_airfoil_case_cpu(e);
  const int rows = std::min(static_cast<int>(r.history.size()), hist_cap);
  for (int i = 0; i < rows; ++i) {
    const auto& h = r.history[i];
    hist[7 * i + 0] = static_cast<float>(h.iter);
    hist[7 * i + 1] = h.residual_l1;
    hist[7 * i + 2] = h.residual_l2;
    hist[7 * i + 3] = h.residual_linf;
    hist[7 * i + 4] = h.cl;
    hist[7 * i + 5] = h.cd;
    hist[7 * i + 6] = h.cm;
  }
  std::memcpy(U, r.conserved.data(), r.conserved.size() * sizeof(float));
  std::memcpy(res, r.last_residual.data(), r.last_residual.size() * sizeof(float));
  std::memcpy(lam, r.last_spectral_radius.data(), r.last_spectral_radius.size() * sizeof(float));
  const float f[8] = {r.forces.fx, r.forces.fy, r.forces.lift, r.forces.drag,
                      r.forces.moment, r.forces.cl, r.forces.cd, r.forces.cm};
  std::memcpy(forces8, f, sizeof(f));
  // Wall Cp in patch order from the reference's own final pressure field.
  float rho_inf = e.rho_inf;
  if (rho_inf <= 0.0f) rho_inf = e.p_inf / (e.gas_constant * e.t_inf);
  const float a_inf = std::sqrt(e.gamma * e.p_inf / rho_inf);
  const float v_inf = std::max(e.mach, 0.0f) * a_inf;
  const float q = 0.5f * rho_inf * v_inf * v_inf + 1.0e-12f;
  int k = 0;
  for (const auto& p : r.mesh.boundary_patches) {
    if (p.name != "wall") continue;
    for (int i = 0; i < p.face_count; ++i) cp_patch[k++] = (r.p[r.mesh.face_owner[p.start_face + i]] - e.p_inf) / q;
  }
  // The reference's own ordered Cp samples (forces.cpp:177-206).
  const FreestreamReference fr = {rho_inf, e.p_inf, e.aoa_deg, v_inf, e.mesh.chord, e.x_ref, e.y_ref};
  const auto samples = extract_wall_cp(r.mesh, r.p, fr, "wall");
  *n_cp_ordered = static_cast<int>(samples.size());
  for (std::size_t i = 0; i < samples.size(); ++i) {
    cp_ordered_xyc[4 * i + 0] = samples[i].s;
    cp_ordered_xyc[4 * i + 1] = samples[i].x;
    cp_ordered_xyc[4 * i + 2] = samples[i].y;
    cp_ordered_xyc[4 * i + 3] = samples[i].cp;
  }
  const float post[kRunPost] = {r.max_wall_mass_flux,
                                r.beta_min,
                                r.beta_max,
                                r.beta_mean,
                                r.mach_local_min,
                                r.mach_local_max,
                                r.mach_local_mean,
                                r.acoustic_scale_min,
                                r.acoustic_scale_max,
                                r.acoustic_scale_mean,
                                r.acoustic_scale_final_min,
                                r.acoustic_scale_final_max,
                                r.acoustic_scale_final_mean,
                                r.wall_adjacent_acoustic_scale_min,
                                r.wall_adjacent_acoustic_scale_p01,
                                r.wall_adjacent_acoustic_scale_p50,
                                r.wall_adjacent_stabilization_scale_min,
                                r.wall_adjacent_stabilization_scale_p01,
                                r.wall_adjacent_stabilization_scale_p50,
                                r.wall_adjacent_acoustic_scale_final_min,
                                r.wall_adjacent_acoustic_scale_final_p01,
                                r.wall_adjacent_acoustic_scale_final_p50,
                                r.stabilization_mach_floor_final};
  std::memcpy(g_run_post, post, sizeof(post));
  return rows;
}

// The post-processing scalars of the last ref_run_case (EulerRunResult, euler_solver.hpp:101-141):
// max_wall_mass_flux, beta min/max/mean, mach_local min/max/mean, acoustic_scale min/max/mean,
// acoustic_scale_final min/max/mean, the wall-adjacent acoustic / stabilization / final scales'
// min/p01/p50, stabilization_mach_floor_final.
void ref_last_run_post(float* out) { std::memcpy(out, g_run_post, sizeof(g_run_post)); }

}  // extern "C"
