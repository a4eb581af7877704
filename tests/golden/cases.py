"""Case definitions shared by the golden-vector generator and the tests.

The four driver runs mirror the reference's own CTest cases (tests/cpp/test_euler_regression.cpp:8-73,
test_euler_aoa0_regression.cpp:10-92, test_euler_wall_sanity.cpp:12-80,
test_euler_low_mach_asymptotic.cpp:135-197 / :280-329); two more cover paths those leave out
(venkat + low-Mach fix with the gated preconditioner, the preconditioned characteristic farfield, and
the staged all-speed controller).  Values are the reference's float literals; the FP32 runs use them
rounded to float32, the FP64 runs use the decimal values.
"""
from __future__ import annotations

MESHES = {
    "40x8": dict(num_circumferential=40, num_radial=8, farfield_radius=10.0, radial_stretch=1.25),
    "48x10": dict(num_circumferential=48, num_radial=10, farfield_radius=10.0, radial_stretch=1.25),
    "64x14": dict(num_circumferential=64, num_radial=14, farfield_radius=12.0, radial_stretch=1.35),
    "72x18": dict(num_circumferential=72, num_radial=18, farfield_radius=12.0, radial_stretch=1.35),
    "160x42": dict(num_circumferential=160, num_radial=42, farfield_radius=15.0, radial_stretch=1.45),
    "256x64": dict(num_circumferential=256, num_radial=64, farfield_radius=20.0, radial_stretch=445.79),
}

# EulerAirfoilCaseConfig overrides per run (defaults: euler_solver.hpp:14-74).
_LOW_MACH = dict(cfl_start=0.10, cfl_max=0.6, cfl_ramp_iters=180, residual_reduction_target=1.0e-3,
                 force_stability_tol=3.0e-3, force_stability_window=6, force_mean_drift_tol=3.0e-3,
                 startup_first_order_iters=30, rk_stages=3, local_time_stepping=1, precond_on=0,
                 precond_mach_ref=0.01, precond_mach_min=1.0e-3, precond_beta_min=1.0e-4, precond_beta_max=1.0,
                 precond_farfield_bc=0, stab_floor_k_start=5.0, stab_floor_k_target=1.0, stab_ramp_iters=700,
                 flux_scheme=1, limiter=0, low_mach_fix=0, mach_cutoff=0.3, all_speed_mach_cutoff=0.25,
                 all_speed_f_min=0.05, all_speed_ramp_start_iter=100, all_speed_ramp_iters=200,
                 all_speed_staged_controller=0, all_speed_stage_a_min_iters=0, all_speed_stage_a_max_iters=0,
                 all_speed_stage_f_min_high=0.2, all_speed_stage_f_min_mid=0.1, all_speed_stage_f_min_low=0.05,
                 all_speed_pjump_wall_target=0.0, all_speed_pjump_spike_factor=1.5, all_speed_pjump_hold_iters=160,
                 all_speed_freeze_iters=160, all_speed_cfl_drop_factor=0.7, all_speed_cfl_restore_factor=1.02,
                 all_speed_cfl_min_scale=0.05, mach=0.01, aoa_deg=0.0, iterations=1200, min_iterations=300)

RUNS = {
    "regression": dict(mesh="72x18", env_precond=False, cfg=dict(
        iterations=80, min_iterations=20, cfl_start=0.15, cfl_max=0.7, cfl_ramp_iters=80,
        residual_reduction_target=1.0e-2, force_stability_tol=5.0e-4, force_stability_window=6,
        flux_scheme=0, limiter=0, precond_on=0, precond_mach_ref=0.15, precond_mach_min=1.0e-3,
        startup_first_order_iters=25, mach=0.15, aoa_deg=2.0)),
    "aoa0": dict(mesh="64x14", env_precond=False, cfg=dict(
        iterations=12, min_iterations=12, cfl_start=0.15, cfl_max=0.7, cfl_ramp_iters=12,
        residual_reduction_target=1.0e-2, force_stability_tol=5.0e-4, force_stability_window=6,
        flux_scheme=0, limiter=0, precond_on=0, startup_first_order_iters=25, mach=0.15, aoa_deg=0.0)),
    "wall_sanity": dict(mesh="40x8", env_precond=True, cfg=dict(
        iterations=1, min_iterations=1, cfl_start=0.0, cfl_max=0.0, cfl_ramp_iters=1,
        residual_reduction_target=0.0, force_stability_tol=0.0, flux_scheme=1, limiter=1, precond_on=1,
        precond_mach_ref=0.15, precond_mach_min=1.0e-3, precond_beta_min=0.2, precond_beta_max=1.0,
        mach=0.2, aoa_deg=0.0)),
    "low_mach_off": dict(mesh="72x18", env_precond=False, cfg=dict(_LOW_MACH, all_speed_flux_fix=0)),
    "low_mach_on": dict(mesh="72x18", env_precond=False, cfg=dict(_LOW_MACH, all_speed_flux_fix=1)),
    # The naca0012_euler_2d.yaml numerics block (HLLC, venkat, low_mach_fix on, precond requested but
    # gated off) at BASELINE's flow point, on a small mesh.
    "case_file_numerics": dict(mesh="64x14", env_precond=False, cfg=dict(
        iterations=150, min_iterations=35, flux_scheme=1, limiter=1, precond_on=1, precond_mach_ref=0.15,
        precond_mach_min=1.0e-3, precond_beta_min=0.2, precond_beta_max=1.0, low_mach_fix=1, mach_cutoff=0.3,
        cfl_start=0.2, cfl_max=1.1, cfl_ramp_iters=120, startup_first_order_iters=25, mach=0.5, aoa_deg=1.25,
        residual_reduction_target=1.0e-3, force_stability_tol=2.0e-5, force_stability_window=6,
        force_mean_drift_tol=2.0e-5)),
    # The reference's CPU / CUDA backend parity case (tests/cpp/test_euler_parity.cpp:15-34): RK1, zero CFL,
    # first order throughout, gas constant 1, p_inf 0.05.
    "parity_rk1": dict(mesh="48x10", env_precond=False, cfg=dict(
        gamma=1.4, gas_constant=1.0, p_inf=0.05, t_inf=1.0, rho_inf=1.0, iterations=26, min_iterations=26,
        cfl_start=0.0, cfl_max=0.0, cfl_ramp_iters=1, residual_reduction_target=0.0, force_stability_tol=0.0,
        precond_on=0, rk_stages=1, startup_first_order_iters=1000)),
    # The global pseudo-time step (compute_pseudo_time_step_over_volume's non-local branch,
    # time_integrator.cpp:49-58) under RK1 and under SSP-RK3, through the first-order startup into second
    # order, with the CFL ramp and adaptation active.
    "rk1_global": dict(mesh="40x8", env_precond=False, cfg=dict(
        iterations=60, min_iterations=60, rk_stages=1, local_time_stepping=0, cfl_start=0.3, cfl_max=0.6,
        cfl_ramp_iters=20, startup_first_order_iters=20, residual_reduction_target=0.0, force_stability_tol=0.0,
        flux_scheme=1, limiter=1, mach=0.3, aoa_deg=2.0)),
    "rk3_global": dict(mesh="64x14", env_precond=False, cfg=dict(
        iterations=50, min_iterations=50, rk_stages=3, local_time_stepping=0, cfl_start=0.2, cfl_max=1.2,
        cfl_ramp_iters=30, startup_first_order_iters=10, residual_reduction_target=0.0, force_stability_tol=0.0,
        flux_scheme=0, limiter=0, mach=0.4, aoa_deg=1.0)),
    # Preconditioning actually on (env gate set), characteristic farfield, staged all-speed controller.
    "precond_staged": dict(mesh="64x14", env_precond=True, cfg=dict(_LOW_MACH, mach=0.05, precond_on=1,
                                                                    precond_mach_ref=0.05, precond_farfield_bc=1,
                                                                    all_speed_flux_fix=1,
                                                                    all_speed_staged_controller=1,
                                                                    all_speed_stage_a_min_iters=40,
                                                                    all_speed_pjump_hold_iters=20,
                                                                    all_speed_freeze_iters=10, iterations=160,
                                                                    min_iterations=100)),
    # The staged controller's other branches: stage A ends on its iteration cap, and a wall p-jump target
    # near the run's own p-jump level gives spikes (freeze, f_min raise, CFL drop) and CFL drops.
    "staged_spike": dict(mesh="64x14", env_precond=True, cfg=dict(_LOW_MACH, mach=0.05, precond_on=1,
                                                                  precond_mach_ref=0.05, precond_farfield_bc=1,
                                                                  all_speed_flux_fix=1,
                                                                  all_speed_staged_controller=1,
                                                                  all_speed_stage_a_min_iters=1000,
                                                                  all_speed_stage_a_max_iters=30,
                                                                  all_speed_pjump_wall_target=5.0e-11,
                                                                  all_speed_pjump_spike_factor=1.1,
                                                                  all_speed_pjump_hold_iters=20,
                                                                  all_speed_freeze_iters=10, iterations=160,
                                                                  min_iterations=100)),
    # A run past its stability limit: reconstruction failures (first-order fallbacks, the first failing
    # face) and then the driver's non-finite trap (euler_solver.cpp:908-946).
    "trap_rk1": dict(mesh="40x8", env_precond=False, cfg=dict(
        iterations=200, min_iterations=200, rk_stages=1, cfl_start=20.0, cfl_max=20.0, cfl_ramp_iters=0,
        residual_reduction_target=0.0, force_stability_tol=0.0, flux_scheme=1, limiter=1,
        startup_first_order_iters=5, mach=0.5, aoa_deg=2.0)),
}

# Runs that end in the driver's non-finite trap: their final state is not a flow field, so the post-run
# records of that state are not compared.
TRAPPED = {"trap_rk1"}

# Residual-evaluation option combinations for the single-evaluation golden vectors:
# (scheme, limiter, second_order, low_mach_fix, precond, all_speed_fix, char_farfield_bc)
RESIDUAL_COMBOS = [
    (s, l, so, lm, pc, asf, cbc)
    for s in (0, 1) for l in (0, 1) for so in (1,) for lm in (0, 1) for pc in (0, 1) for asf in (0, 1) for cbc in (0, 1)
    if not (s == 0 and asf == 1)  # the all-speed fix only enters HLLC
] + [(0, 0, 0, 0, 0, 0, 0), (1, 0, 0, 1, 1, 0, 0)]


def residual_cfg_values(combo, farfield):
    s, l, so, lm, pc, asf, cbc = combo
    return dict(farfield=list(farfield), gamma=1.4, flux_scheme=s, limiter=l, second_order=so, low_mach_fix=lm,
                low_mach_cutoff=0.3, stab_mach_floor=0.02, precond_farfield_bc=cbc, precond_enabled=pc,
                precond_mach_ref=0.15, precond_mach_min=1.0e-3, precond_beta_min=0.2, precond_beta_max=1.0,
                all_speed_fix=asf, all_speed_cutoff=0.25, all_speed_f_min=0.05, all_speed_ramp=0.7)
