// aero_oracle_driver.hpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Restatement of the reference Euler driver's pseudo-time loop, run_euler_airfoil_case
// (src/cfd_core/solvers/euler_solver.cpp:463-1208), templated on the scalar type.  The mesh comes in
// as arrays (the reference builds it from make_airfoil_ogrid_mesh at :474; the mesh generator is pinned
// separately against the reference's own mesh arrays).  Host observability (CSV / VTU / debug files)
// is not restated; the per-iteration history that those CSVs carry is returned instead.
#pragma once

#include "aero_oracle.hpp"

namespace aero_oracle {

template <class R> struct CaseCfg {  // euler_solver.hpp:14-74 (mesh config excluded)
  int iterations = 400, min_iterations = 40, cfl_ramp_iters = 120;
  R gamma = AO_K(1.4), gas_constant = AO_K(287.05), mach = AO_K(0.15), aoa_deg = AO_K(2.0),
    p_inf = AO_K(101325.0), t_inf = AO_K(288.15), rho_inf = R(0), cfl_start = AO_K(0.2), cfl_max = AO_K(1.2);
  R residual_reduction_target = AO_K(1.0e-3), force_stability_tol = AO_K(2.0e-5);
  int force_stability_window = 6;
  R force_mean_drift_tol = AO_K(2.0e-5);
  int startup_first_order_iters = 25, rk_stages = 3;
  bool local_time_stepping = true;
  bool precond_on = false;
  R precond_mach_ref = AO_K(0.15), precond_mach_min = AO_K(1.0e-3), precond_beta_min = AO_K(0.2),
    precond_beta_max = AO_K(1.0);
  bool precond_farfield_bc = false;
  R stab_floor_start = AO_K(0.1), stab_floor_target = AO_K(0.02), stab_floor_k_start = AO_K(-1.0),
    stab_floor_k_target = AO_K(-1.0);
  int stab_ramp_iters = 400;
  R x_ref = AO_K(0.25), y_ref = R(0), chord = AO_K(1.0);
  int flux_scheme = kRusanov, limiter = kMinmod;
  bool low_mach_fix = false;
  R mach_cutoff = AO_K(0.3);
  bool all_speed_flux_fix = false;
  R all_speed_mach_cutoff = AO_K(0.25), all_speed_f_min = AO_K(0.05);
  int all_speed_ramp_start_iter = 0, all_speed_ramp_iters = 0;
  bool all_speed_staged_controller = false;
  int all_speed_stage_a_min_iters = 600, all_speed_stage_a_max_iters = 0;
  R all_speed_stage_f_min_high = AO_K(0.2), all_speed_stage_f_min_mid = AO_K(0.1),
    all_speed_stage_f_min_low = AO_K(0.05);
  R all_speed_pjump_wall_target = R(0), all_speed_pjump_spike_factor = AO_K(1.5);
  int all_speed_pjump_hold_iters = 120, all_speed_freeze_iters = 120;
  R all_speed_cfl_drop_factor = AO_K(0.8), all_speed_cfl_restore_factor = AO_K(1.02),
    all_speed_cfl_min_scale = AO_K(0.1);
  bool aoa0_symmetry_enforce = false;
  int aoa0_symmetry_enforce_interval = 0;
  // AERO_ENABLE_EXPERIMENTAL_PRECOND, read by the reference at euler_solver.cpp:482-489.
  bool precond_experimental_enabled = false;
  // Frozen limiter (beyond the reference): at the start of this iteration (second order) the limiter
  // ratios are captured from the state and used from then on; < 0: never.
  int limiter_freeze_iter = -1;
};

// One history row per iteration (the residuals.csv / forces.csv / diagnostics columns that the
// parity contract uses).
enum HistField : int {
  kHIter = 0, kHL1, kHL2, kHLinf, kHRatio, kHCl, kHCd, kHCm, kHRhoL2, kHCfl, kHFoFallback, kHDiffFallback,
  kHistFields
};

// Per-iteration diagnostics row of euler_diagnostics.csv: the columns that depend on the iteration's state
// or on the controller (euler_solver.cpp:971-1139, printed at :1171-1203), in the order of the device
// library's AERO_D_* fields.
enum DiagField : int {
  kDMaxWallUn = 0, kDFfDp, kDFfDrho, kDFfDu, kDFfDv, kDCpMin, kDCpMax, kDCpNonfinite, kDFxGauge, kDFyGauge,
  kDFxAbs, kDFyAbs, kDClAbs, kDCdAbs, kDCmAbs, kDLift, kDDrag, kDPjump, kDPjumpWall, kDPfluctMax, kDPfluctWall,
  kDPfluctFarfield, kDAsStage, kDAsFix, kDAsFmin, kDAsRamp, kDAsFreeze, kDAsAction, kDCflScale, kDFirstFailure,
  kDiagFields
};

template <class R> struct RunOut {
  std::vector<R> hist;   // [n_iter][kHistFields]
  std::vector<R> diag;   // [n_iter][kDiagFields] when the run was asked for diagnostics
  int n_iter = 0;
  bool nonfinite = false;
  std::vector<R> U;      // final conserved AoS [N][5]
  std::vector<R> res;    // last stage residual AoS [N][5]
  std::vector<R> lam;    // last stage spectral radius [N]
  Coeffs<R> forces;      // gauge forces of the final state
  std::vector<R> cp;     // wall Cp in wall-patch order
  AssemblyCfg<R> cfg;    // residual configuration of the last stage (frozen ratios not kept)
};

template <class R> R compute_cfl(const CaseCfg<R>& c, int iter, R scale) {  // euler_solver.cpp:69-79
  R base;
  if (c.cfl_ramp_iters <= 0) {
    base = c.cfl_max;
  } else {
    const R t = smin(AO_K(1.0), static_cast<R>(iter + 1) / static_cast<R>(std::max(c.cfl_ramp_iters, 1)));
    base = c.cfl_start + t * (c.cfl_max - c.cfl_start);
  }
  return base * scale;
}

template <class R> R stab_floor_at(R start, R target, int ramp, int iter) {  // euler_solver.cpp:95-106
  const R s = sclamp(start, R(0), AO_K(1.0));
  const R t = sclamp(target, R(0), AO_K(1.0));
  if (ramp <= 0) return t;
  const R x = sclamp(static_cast<R>(std::max(iter, 0)) / static_cast<R>(std::max(ramp, 1)), R(0), AO_K(1.0));
  return s + x * (t - s);
}

template <class R> R all_speed_ramp_weight(int on, int len, int iter) {  // euler_solver.cpp:108-114
  if (len <= 0) return (iter >= on) ? AO_K(1.0) : R(0);
  const R x = static_cast<R>(iter - on) / static_cast<R>(std::max(len, 1));
  return sclamp(x, R(0), AO_K(1.0));
}

template <class R> PrecondCfg<R> build_precond(const CaseCfg<R>& c, bool enabled) {  // euler_solver.cpp:341-355
  PrecondCfg<R> p;
  p.enabled = enabled;
  const R minf = smax(c.mach, AO_K(1.0e-4));
  const R auto_bmin = smax(AO_K(1.0e-4), smin(AO_K(1.0e-3), AO_K(0.1) * minf));
  p.mach_ref = smax(c.precond_mach_ref > R(0) ? c.precond_mach_ref : minf, AO_K(1.0e-4));
  p.mach_min = smax(c.precond_mach_min > R(0) ? c.precond_mach_min : AO_K(1.0e-6), AO_K(1.0e-6));
  p.beta_min = smax(c.precond_beta_min > R(0) ? c.precond_beta_min : auto_bmin, AO_K(1.0e-4));
  p.beta_max = smax(c.precond_beta_max, p.beta_min);
  return p;
}

// The driver.  `m` supplies the mesh; its normals are re-normalised into a private copy first
// (euler_solver.cpp:475-478).  threaded: the stage gradients and residuals run on all OpenMP threads
// (cell_gradients_threaded / assemble_residual_threaded: the serial results bit for bit).
template <class R> RunOut<R> run_case(MeshView<R> m, const CaseCfg<R>& c, bool threaded = false,
                                      bool diagnostics = false) {
  if (c.iterations <= 0) throw std::invalid_argument("Euler iterations must be positive.");
  if (c.rk_stages != 1 && c.rk_stages != 3) throw std::invalid_argument("Euler pseudo-time rk_stages must be 1 or 3.");
  std::vector<R> normals(m.normal, m.normal + 3 * static_cast<std::size_t>(m.num_faces));
  if (normalize_face_normals<R>(m.num_faces, normals.data(), AO_K(1.0e-5)) > 0)
    throw std::runtime_error("Mesh contains invalid face normals.");
  m.normal = normals.data();

  const int N = m.num_cells;
  const R g = c.gamma;
  const bool precond_eff = c.precond_on && c.precond_experimental_enabled;  // :482-489
  const PrecondCfg<R> pc = build_precond(c, precond_eff);

  // Stabilisation floor schedule (:141-157), from the un-gated config.
  const R minf = smax(c.mach, AO_K(1.0e-4));
  const R k0 = c.stab_floor_k_start >= R(0) ? c.stab_floor_k_start : (c.stab_floor_start / minf);
  const R k1 = c.stab_floor_k_target >= R(0) ? c.stab_floor_k_target : (c.stab_floor_target / minf);
  const R floor_start = sclamp(smax(k0, R(0)) * minf, R(0), AO_K(1.0));
  const R floor_target = sclamp(smax(k1, R(0)) * minf, R(0), AO_K(1.0));
  const int stab_ramp = std::max(c.stab_ramp_iters, 0);

  // All-speed schedule and staged controller settings (:492-528).
  const R as_fmin = sclamp(c.all_speed_f_min, AO_K(1.0e-6), AO_K(1.0));
  const int as_on = std::max(c.all_speed_ramp_start_iter, 0);
  const int as_len = std::max(c.all_speed_ramp_iters, 0);
  const bool staged = c.all_speed_staged_controller;
  const int a_min = std::max(c.all_speed_stage_a_min_iters, 0);
  const int a_max = std::max(c.all_speed_stage_a_max_iters, 0);
  const R f_high = sclamp(c.all_speed_stage_f_min_high, AO_K(1.0e-6), AO_K(1.0));
  const R f_mid = sclamp(smin(c.all_speed_stage_f_min_mid, f_high), AO_K(1.0e-6), AO_K(1.0));
  const R f_low = sclamp(smin(c.all_speed_stage_f_min_low, f_mid), AO_K(1.0e-6), AO_K(1.0));
  const R auto_pjump = sclamp(AO_K(1.6e-2) * minf * minf * minf * minf, AO_K(1.0e-9), AO_K(2.0e-6));
  const R pjump_target = c.all_speed_pjump_wall_target <= R(0) ? auto_pjump : c.all_speed_pjump_wall_target;
  const R spike_factor = smax(c.all_speed_pjump_spike_factor, AO_K(1.0));
  const int hold_iters = std::max(c.all_speed_pjump_hold_iters, 1);
  const int freeze_iters = std::max(c.all_speed_freeze_iters, 0);
  const R drop = sclamp(c.all_speed_cfl_drop_factor, AO_K(0.05), AO_K(1.0));
  const R restore = sclamp(c.all_speed_cfl_restore_factor, AO_K(1.0), AO_K(1.2));
  const R cfl_min_scale = sclamp(c.all_speed_cfl_min_scale, AO_K(0.01), AO_K(1.0));

  // Freestream (:543-561).
  R rho_inf = c.rho_inf;
  if (rho_inf <= R(0)) rho_inf = c.p_inf / (c.gas_constant * c.t_inf);
  if (rho_inf <= R(0)) throw std::invalid_argument("Invalid freestream density. Set rho_inf or p_inf/T_inf.");
  const R alpha = c.aoa_deg * (pi_const<R>() / AO_K(180.0));
  const R a_inf = std::sqrt(g * c.p_inf / rho_inf);
  const R v_inf = smax(c.mach, R(0)) * a_inf;
  Prim<R> pinf;
  pinf.rho = rho_inf;
  pinf.u = v_inf * std::cos(alpha);
  pinf.v = v_inf * std::sin(alpha);
  pinf.w = R(0);
  pinf.p = c.p_inf;
  const Cons<R> uinf = prim_to_cons(pinf, g);

  std::vector<R> U(static_cast<std::size_t>(N) * 5);
  for (int i = 0; i < N; ++i) {
    U[5 * i + 0] = uinf.rho;
    U[5 * i + 1] = uinf.rhou;
    U[5 * i + 2] = uinf.rhov;
    U[5 * i + 3] = uinf.rhow;
    U[5 * i + 4] = uinf.rhoE;
  }
  if (m.wall_count <= 0) throw std::runtime_error("Wall patch face set is empty.");
  std::vector<unsigned char> wall_adj(static_cast<std::size_t>(N), 0u);
  for (int i = 0; i < m.wall_count; ++i) wall_adj[m.owner[m.wall_start + i]] = 1u;

  FreeRef<R> ref;
  ref.rho_inf = rho_inf;
  ref.p_inf = c.p_inf;
  ref.aoa_deg = c.aoa_deg;
  ref.speed_inf = v_inf;
  ref.chord = c.chord;
  ref.x_ref = c.x_ref;
  ref.y_ref = c.y_ref;
  const R pressure_ref = smax(ref.rho_inf * a_inf * a_inf, AO_K(1.0e-12));

  MonitorCfg<R> mc;
  mc.ratio_target = c.residual_reduction_target;
  mc.force_tol = c.force_stability_tol;
  mc.window = c.force_stability_window;
  mc.min_iterations = c.min_iterations;
  mc.drift_tol = c.force_mean_drift_tol;
  Monitor<R> monitor(mc);

  RunOut<R> out;
  R cfl_scale = AO_K(1.0);
  R prev_l2 = AO_K(-1.0);
  bool as_fix_eff = c.all_speed_flux_fix;
  R as_fmin_eff = as_fmin;
  R as_w = AO_K(1.0);
  int stage_ab = staged ? 1 : 0;  // 0 single, 1 stage A, 2 stage B
  R b_fmin = f_high;
  int b_progress = 0, below = 0, freeze = 0;

  std::vector<Prim<R>> prim(static_cast<std::size_t>(N));
  std::vector<R> grad, res, resp, lam, dt, un, us, nxt, scratch_q, scratch_f;
  CellFaces cf;
  if (threaded) cf = cell_faces(m);
  std::vector<R> cell_p(static_cast<std::size_t>(N), c.p_inf);
  std::vector<double> frozen;  // frozen limiter ratios once captured

  for (int iter = 0; iter < c.iterations; ++iter) {  // :771
    const bool second = iter >= std::max(c.startup_first_order_iters, 0);
    const R cfl = compute_cfl(c, iter, cfl_scale);
    const R floor_cur = stab_floor_at(floor_start, floor_target, stab_ramp, iter);
    if (!staged) {
      as_fix_eff = c.all_speed_flux_fix;
      as_fmin_eff = as_fmin;
      as_w = all_speed_ramp_weight<R>(as_on, as_len, iter);
    } else if (stage_ab == 1) {
      as_fix_eff = false;
      as_fmin_eff = f_high;
      as_w = R(0);
    } else {
      as_fix_eff = c.all_speed_flux_fix;
      as_fmin_eff = sclamp(b_fmin, f_low, f_high);
      as_w = as_len <= 0 ? AO_K(1.0)
                         : sclamp(static_cast<R>(b_progress) / static_cast<R>(std::max(as_len, 1)), R(0), AO_K(1.0));
    }
    un = U;
    us = un;
    if (second && c.limiter_freeze_iter >= 0 && iter == c.limiter_freeze_iter) {
      for (int i = 0; i < N; ++i) {
        Cons<R> s{us[5 * i + 0], us[5 * i + 1], us[5 * i + 2], us[5 * i + 3], us[5 * i + 4]};
        prim[i] = cons_to_prim(s, g);
      }
      cell_gradients(m, prim, &grad);
      limiter_ratios(m, prim, grad, c.limiter, &frozen);
    }
    AssemblyDiag iter_diag;
    for (int st = 0; st < c.rk_stages; ++st) {  // :802-900
      for (int i = 0; i < N; ++i) {
        Cons<R> s{us[5 * i + 0], us[5 * i + 1], us[5 * i + 2], us[5 * i + 3], us[5 * i + 4]};
        prim[i] = cons_to_prim(s, g);
      }
      if (second && threaded)
        cell_gradients_threaded(m, cf, prim, &scratch_q, &grad);
      else if (second)
        cell_gradients(m, prim, &grad);
      else
        grad.assign(static_cast<std::size_t>(N) * 10, R(0));
      AssemblyCfg<R> ac;
      ac.farfield = uinf;
      ac.gamma = g;
      ac.scheme = c.flux_scheme;
      ac.limiter = c.limiter;
      ac.second_order = second;
      ac.low_mach_fix = c.low_mach_fix;
      ac.low_mach_cutoff = c.mach_cutoff;
      ac.stab_mach_floor = floor_cur;
      ac.frozen = frozen.empty() ? nullptr : frozen.data();
      ac.precond_farfield_bc = c.precond_farfield_bc;
      ac.precond = pc;
      ac.as_fix = as_fix_eff;
      ac.as_cutoff = c.all_speed_mach_cutoff;
      ac.as_fmin = as_fmin_eff;
      ac.as_ramp = as_w;
      AssemblyDiag d;
      if (threaded)
        assemble_residual_threaded(m, ac, cf, prim, grad, &scratch_f, &res, &lam, &d);
      else
        assemble_residual(m, ac, prim, grad, &res, &lam, &d);
      out.cfg = ac;
      out.cfg.frozen = nullptr;
      iter_diag.first_order_fallback += d.first_order_fallback;
      iter_diag.diffusive_fallback += d.diffusive_fallback;
      if (iter_diag.first_failure_face < 0 && d.first_failure_face >= 0) iter_diag.first_failure_face = d.first_failure_face;
      resp = res;
      precond_residual(prim, g, pc, &resp);
      dt = dt_over_volume(lam, cfl, c.local_time_stepping);
      rk3_stage_update(st, un, us, resp, dt, g, &nxt);
      us.swap(nxt);
    }
    U = us;
    // aoa=0 symmetry enforcement (:380-410, called at :904-906)
    if (c.aoa0_symmetry_enforce && c.aoa0_symmetry_enforce_interval > 0 && !(std::abs(c.aoa_deg) > AO_K(1.0e-12)) &&
        ((iter + 1) % c.aoa0_symmetry_enforce_interval) == 0) {
      for (int i = 0; i < N; ++i) {
        Cons<R> s{U[5 * i + 0], U[5 * i + 1], U[5 * i + 2], U[5 * i + 3], U[5 * i + 4]};
        Prim<R> p = cons_to_prim(s, g);
        p.v = R(0);
        p.w = R(0);
        const Cons<R> y = prim_to_cons(p, g);
        U[5 * i + 0] = y.rho;
        U[5 * i + 1] = y.rhou;
        U[5 * i + 2] = y.rhov;
        U[5 * i + 3] = y.rhow;
        U[5 * i + 4] = y.rhoE;
      }
    }
    out.res = res;
    out.lam = lam;
    // Non-finite trap (:908-946).
    bool bad = false;
    for (R x : U)
      if (!std::isfinite(x)) {
        bad = true;
        break;
      }
    if (!bad) {
      for (int i = 0; i < N; ++i) {
        Cons<R> s{U[5 * i + 0], U[5 * i + 1], U[5 * i + 2], U[5 * i + 3], U[5 * i + 4]};
        const Prim<R> p = cons_to_prim(s, g);
        if (!std::isfinite(p.rho) || !std::isfinite(p.u) || !std::isfinite(p.v) || !std::isfinite(p.w) ||
            !std::isfinite(p.p)) {
          bad = true;
          break;
        }
        prim[i] = p;
        cell_p[i] = p.p;
      }
    }
    if (bad) {
      out.nonfinite = true;
      break;
    }
    const Norms<R> nr = residual_norms(m, res);  // :948-949
    const R ratio = monitor.update_residual(nr.l2);
    if (prev_l2 > R(0)) {  // :951-958
      if (nr.l2 > prev_l2 * AO_K(1.05))
        cfl_scale = smax(cfl_scale * AO_K(0.65), AO_K(0.05));
      else if (nr.l2 < prev_l2 * AO_K(0.90))
        cfl_scale = smin(cfl_scale * AO_K(1.03), AO_K(1.0));
    }
    prev_l2 = nr.l2;
    const Coeffs<R> fg = integrate_forces(m, cell_p, ref, true);  // :960-969
    out.forces = fg;
    monitor.update_forces(fg);
    const bool window_stable = monitor.forces_stable();
    const bool stage_row_fix = as_fix_eff;  // the effective settings of this iteration, printed after the
    const R stage_row_fmin = as_fmin_eff;   // controller has moved on (:1193-1196)
    const R stage_row_ramp = as_w;
    int action = 0;

    if (staged) {  // :1026-1139
      double pj_wall = 0.0;
      int pj_wall_n = 0;
      for (int f = 0; f < m.num_faces; ++f) {
        const int o = m.owner[f], nb = m.neighbor[f];
        if (nb < 0) continue;
        const R dp = prim[o].p - prim[nb].p;
        const double e = static_cast<double>(dp) * static_cast<double>(dp);
        if (wall_adj[o] != 0u || wall_adj[nb] != 0u) {
          pj_wall += e;
          ++pj_wall_n;
        }
      }
      const double pref2 = static_cast<double>(pressure_ref) * static_cast<double>(pressure_ref) + 1.0e-24;
      const R pj = pj_wall_n > 0 ? static_cast<R>((pj_wall / static_cast<double>(pj_wall_n)) / pref2) : R(0);
      if (stage_ab == 1) {
        const bool ready = (iter + 1 >= a_min) && window_stable;
        const bool timeout = (a_max > 0) && (iter + 1 >= a_max);
        if ((ready || timeout) && c.all_speed_flux_fix) {
          stage_ab = 2;
          b_fmin = f_high;
          b_progress = 0;
          below = 0;
          freeze = 0;
          action = ready ? 10 : 11;
        }
      } else if (stage_ab == 2) {
        const R spike_thr = pjump_target > R(0) ? pjump_target * spike_factor : std::numeric_limits<R>::max();
        const bool spike = pj > spike_thr;
        const bool good = (pjump_target <= R(0)) || (pj <= pjump_target);
        if (spike) {
          freeze = freeze_iters;
          below = 0;
          b_fmin = smin(f_high, b_fmin + AO_K(0.05));
          cfl_scale = smax(cfl_scale * drop, cfl_min_scale);
          action = 20;
        } else {
          if (freeze > 0) --freeze;
          if (good)
            ++below;
          else
            below = 0;
          if (freeze == 0) {
            if (b_fmin > f_mid + AO_K(1.0e-6) && below >= hold_iters) {
              b_fmin = f_mid;
              below = 0;
              action = 31;
            } else if (b_fmin > f_low + AO_K(1.0e-6) && below >= hold_iters) {
              b_fmin = f_low;
              below = 0;
              action = 32;
            } else if (b_fmin <= f_low + AO_K(1.0e-6) && good) {
              ++b_progress;
              action = 33;
            }
            if (good) {
              cfl_scale = smin(cfl_scale * restore, AO_K(1.0));
            } else if (pjump_target > R(0) && pj > pjump_target) {
              cfl_scale = smax(cfl_scale * drop, cfl_min_scale);
              if (action == 0) action = 21;
            }
          }
        }
      }
    }

    if (diagnostics) {  // :971-1066 over the wall, farfield and interior faces of the updated state
      R d[kDiagFields] = {};
      const R q = AO_K(0.5) * ref.rho_inf * ref.speed_inf * ref.speed_inf + AO_K(1.0e-12);
      R cp_min = std::numeric_limits<R>::max(), cp_max = std::numeric_limits<R>::lowest();
      int cp_bad = 0;
      for (int i = 0; i < m.wall_count; ++i) {
        const int f = m.wall_start + i;
        const Prim<R>& p = prim[m.owner[f]];
        d[kDMaxWallUn] = smax(d[kDMaxWallUn], std::abs(p.u * m.normal[3 * f] + p.v * m.normal[3 * f + 1]));
        const R cp = (p.p - c.p_inf) / q;
        if (std::isfinite(cp)) {
          cp_min = smin(cp_min, cp);
          cp_max = smax(cp_max, cp);
        } else {
          ++cp_bad;
        }
      }
      d[kDCpMin] = cp_min;
      d[kDCpMax] = cp_max;
      d[kDCpNonfinite] = static_cast<R>(cp_bad);
      for (int f = 0; f < m.num_faces; ++f) {
        if (m.bc[f] != kFarfield) continue;
        const Prim<R>& p = prim[m.owner[f]];
        d[kDFfDp] = smax(d[kDFfDp], std::abs(p.p - c.p_inf));
        d[kDFfDrho] = smax(d[kDFfDrho], std::abs(p.rho - rho_inf));
        d[kDFfDu] = smax(d[kDFfDu], std::abs(p.u - pinf.u));
        d[kDFfDv] = smax(d[kDFfDv], std::abs(p.v - pinf.v));
      }
      const Coeffs<R> fa = integrate_forces(m, cell_p, ref, false);
      d[kDFxGauge] = fg.fx;
      d[kDFyGauge] = fg.fy;
      d[kDFxAbs] = fa.fx;
      d[kDFyAbs] = fa.fy;
      d[kDClAbs] = fa.cl;
      d[kDCdAbs] = fa.cd;
      d[kDCmAbs] = fa.cm;
      d[kDLift] = fg.lift;
      d[kDDrag] = fg.drag;
      double pj_all = 0.0, pj_wall = 0.0;
      int n_all = 0, n_wall = 0;
      for (int f = 0; f < m.num_faces; ++f) {
        const int o = m.owner[f], nb = m.neighbor[f];
        if (nb < 0) continue;
        const R dp = prim[o].p - prim[nb].p;
        const double e = static_cast<double>(dp) * static_cast<double>(dp);
        pj_all += e;
        ++n_all;
        if (wall_adj[o] != 0u || wall_adj[nb] != 0u) {
          pj_wall += e;
          ++n_wall;
        }
      }
      const double pref2 = static_cast<double>(pressure_ref) * static_cast<double>(pressure_ref) + 1.0e-24;
      d[kDPjump] = n_all > 0 ? static_cast<R>((pj_all / n_all) / pref2) : R(0);
      d[kDPjumpWall] = n_wall > 0 ? static_cast<R>((pj_wall / n_wall) / pref2) : R(0);
      R dp_all = 0, dp_wall = 0;
      for (int i = 0; i < N; ++i) {
        const R dp = std::abs(prim[i].p - c.p_inf);
        dp_all = smax(dp_all, dp);
        if (wall_adj[i] != 0u) dp_wall = smax(dp_wall, dp);
      }
      d[kDPfluctMax] = dp_all / pressure_ref;
      d[kDPfluctWall] = dp_wall / pressure_ref;
      d[kDPfluctFarfield] = d[kDFfDp] / pressure_ref;
      d[kDAsStage] = static_cast<R>(stage_ab);
      d[kDAsFix] = stage_row_fix ? R(1) : R(0);
      d[kDAsFmin] = stage_row_fmin;
      d[kDAsRamp] = stage_row_ramp;
      d[kDAsFreeze] = static_cast<R>(freeze);
      d[kDAsAction] = static_cast<R>(action);
      d[kDCflScale] = cfl_scale;
      d[kDFirstFailure] = static_cast<R>(iter_diag.first_failure_face);
      out.diag.insert(out.diag.end(), d, d + kDiagFields);
    }

    R row[kHistFields];
    row[kHIter] = static_cast<R>(iter);
    row[kHL1] = nr.l1;
    row[kHL2] = nr.l2;
    row[kHLinf] = nr.linf;
    row[kHRatio] = ratio;
    row[kHCl] = fg.cl;
    row[kHCd] = fg.cd;
    row[kHCm] = fg.cm;
    row[kHRhoL2] = nr.rho_l2;
    row[kHCfl] = cfl;
    row[kHFoFallback] = static_cast<R>(iter_diag.first_order_fallback);
    row[kHDiffFallback] = static_cast<R>(iter_diag.diffusive_fallback);
    out.hist.insert(out.hist.end(), row, row + kHistFields);
    ++out.n_iter;
    if (monitor.converged(iter, ratio)) break;
  }
  out.U = U;
  // Wall Cp in patch order from the final state (forces.cpp:177-206, ordering aside).
  const R q = AO_K(0.5) * ref.rho_inf * ref.speed_inf * ref.speed_inf + AO_K(1.0e-12);
  out.cp.resize(static_cast<std::size_t>(m.wall_count));
  for (int i = 0; i < m.wall_count; ++i) {
    const int o = m.owner[m.wall_start + i];
    Cons<R> s{U[5 * o + 0], U[5 * o + 1], U[5 * o + 2], U[5 * o + 3], U[5 * o + 4]};
    out.cp[i] = (cons_to_prim(s, g).p - ref.p_inf) / q;
  }
  return out;
}

}  // namespace aero_oracle
