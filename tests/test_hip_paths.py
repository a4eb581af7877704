"""HIP driver paths and failure handling against the FP64 oracle, through the C ABI.

* The config-1 workload as a run: the case file's numerics (cases/naca0012_euler_2d.yaml:22-45: HLLC,
  venkat, low-Mach fix, SSP-RK3 with local time stepping, first-order startup, CFL ramp 0.2 -> 1.1) on the
  256 x 64 quad O-grid at M 0.5, AoA 1.25 deg.  The reference's limiter cycle keeps this run from meeting
  its convergence test within thousands of iterations (the FP64 oracle: a residual ratio of 3.7e-3 after
  4000), so the run is compared over a fixed 2000 iterations: the final state and the L-infinity history
  bit for bit, the other residual norms, the force history, CL / CD / CM and the wall Cp at 1e-9 relative.
* The low-Mach preconditioned all-speed residual at M 0.01 on a 256 x 64 mixed mesh.
* Forced failures, set up from the host: a cell with negative pressure and a cell with a NaN density
  (first-order fallback counts, diffusive fallback counts and the first failing face), and runs driven
  past their stability limit until the driver's non-finite trap stops them (euler_solver.cpp:908-946).
"""
import sys

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol

pytestmark = pytest.mark.gpu
COMP = [0, 1, 2, 4]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available(), "no gfx950 device visible to the HIP backend"
    return aero_hip


def compare_runs(out, ref, label, rtol=1e-9):
    hr, ho = ref["history"], out["history"]
    assert ho.shape == hr.shape, (label, ho.shape, hr.shape)
    for col in (1, 2, 3, 5, 6, 7, 8, 9):
        a, b = ho[:, col], hr[:, col]
        both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
        atol = 1e-12 if col in (5, 6, 7) else 0.0
        with np.errstate(invalid="ignore"):  # inf - inf where both blew up
            rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
            ok = both_inf | (rel <= rtol) | (np.abs(a - b) <= atol)
        assert np.all(ok), (label, col, float(np.nanmax(np.where(both_inf, 0.0, rel))))
    assert np.array_equal(ho[:, 10:12], hr[:, 10:12]), f"{label}: fallback counters differ"


def test_config1_quad256x64_run_parity(hip):
    run = dict(cases.RUNS["case_file_numerics"]["cfg"])
    iters = 2000
    run.update(iterations=iters, mach=0.5, aoa_deg=1.25)
    m = helpers.mesh("256x64", cells="quad")
    assert m.num_cells == 16384 and m.num_faces == 33024
    om = ol.OracleMesh(m, np.float64)
    say(f"[config1] oracle driver (threaded), {iters} iterations on N={m.num_cells}")
    ref = ol.run_case(om, helpers.case_cfg(run, False), iters, threaded=True)
    from aero_hip import CaseConfig
    case = CaseConfig(**run)
    case.precond_experimental_enabled = 0
    say("[config1] HIP driver")
    with hip.HipMesh(m) as hm:
        out = hm.run_case(case)
    compare_runs(out, ref, "config1")
    assert np.array_equal(out["history"][:, 3], ref["history"][:, 3]), "L-infinity residual history differs"
    assert np.array_equal(out["conserved"], ref["conserved"]), "final state differs from the oracle driver's"
    for k in (5, 6, 7):
        assert abs(out["forces"][k] - ref["forces"][k]) <= 1e-9 * abs(ref["forces"][k]) + 1e-12
    assert np.allclose(out["wall_cp"], ref["cp"], rtol=1e-9, atol=1e-12)
    say(f"[config1] bitwise state: {np.array_equal(out['conserved'], ref['conserved'])}, "
        f"final L2 ratio {ref['history'][-1, 4]:.3e}")


def lowmach_cfg(ff, flux, limiter):
    """The bench's low-Mach line: the reference low-Mach test settings (test_euler_low_mach_asymptotic.cpp:
    135-191) with the preconditioner and the all-speed HLLC fix on."""
    return dict(farfield=list(ff), gamma=1.4, flux_scheme=flux, limiter=limiter, second_order=1, low_mach_fix=0,
                low_mach_cutoff=0.3, stab_mach_floor=0.0, precond_farfield_bc=0, precond_enabled=1,
                precond_mach_ref=0.01, precond_mach_min=1e-3, precond_beta_min=1e-4, precond_beta_max=1.0,
                all_speed_fix=1, all_speed_cutoff=0.25, all_speed_f_min=0.05, all_speed_ramp=1.0)


@pytest.mark.parametrize("flux,limiter,cbc", [(1, 1, 0), (1, 0, 1), (0, 0, 0)])
def test_lowmach_precond_residual_256x64_mixed(hip, flux, limiter, cbc):
    from aero_hip import freestream
    m = helpers.mesh("256x64", cells="mixed", quad_layers=12, jitter=0.1, seed=12345, rcm=True)
    om = ol.OracleMesh(m, np.float64)
    ff = freestream(0.01, 2.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.02, 20260101)
    vals = lowmach_cfg(ff, flux, limiter)
    vals["precond_farfield_bc"] = cbc
    cfg = helpers.res_cfg(vals)
    say(f"[lowmach] M 0.01 precond all-speed flux={flux} limiter={limiter} char_bc={cbc} N={m.num_cells}")
    with hip.HipMesh(m) as hm:
        Rh, lh, dh = hm.residual(U, cfg)
    Ro, lo, do = ol.residual(om, cfg, U)
    scale = helpers.flux_scale(om, cfg, U)
    err = helpers.scaled_err(Rh[:, COMP], Ro[:, COMP], scale[:, COMP])
    assert err.max() <= 1e-12, err.max()
    assert (np.abs(lh - lo) / np.abs(lo)).max() <= 1e-12
    assert (dh["first_order_fallback"], dh["diffusive_fallback"], dh["first_failure_face"]) == tuple(do)


def forced_state(m, ff, kind):
    """A smooth state with one cell made unphysical from the host: negative pressure (energy below the
    kinetic energy) or a NaN density."""
    U = helpers.perturbed_state(m.cell_center, ff, 0.05, 31)
    c = int(m.owner[m.wall_start + m.wall_count // 3])  # a wall cell, where the slip-wall flux reads it too
    d = int(m.neighbor[np.nonzero(m.neighbor >= 0)[0][len(m.neighbor) // 2]])  # an interior cell
    for cell in (c, d):
        if kind == "negative_pressure":
            ke = 0.5 * (U[cell, 1] ** 2 + U[cell, 2] ** 2) / U[cell, 0]
            U[cell, 4] = 0.5 * ke
        else:
            U[cell, 0] = np.nan
    return U


@pytest.mark.parametrize("kind", ["negative_pressure", "nan_density"])
@pytest.mark.parametrize("combo", [(1, 1, 1, 1, 0, 0, 0), (0, 0, 1, 0, 0, 0, 0), (1, 1, 1, 1, 1, 1, 1)])
def test_forced_failure_residual(hip, kind, combo):
    from aero_hip import freestream
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=5, rcm=True)
    om = ol.OracleMesh(m, np.float64)
    ff = freestream(0.5, 2.0)[1]
    U = forced_state(m, ff, kind)
    cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
    say(f"[failure] {kind} combo={combo}")
    with hip.HipMesh(m) as hm:
        Rh, lh, dh = hm.residual(U, cfg)
    Ro, lo, do = ol.residual(om, cfg, U)
    assert (dh["first_order_fallback"], dh["diffusive_fallback"], dh["first_failure_face"]) == tuple(do), (dh, do)
    assert dh["first_order_fallback"] > 0 and dh["first_failure_face"] >= 0
    fin_h, fin_o = np.isfinite(Rh[:, COMP]), np.isfinite(Ro[:, COMP])
    assert np.array_equal(fin_h, fin_o), "non-finite residual entries differ"
    assert np.array_equal(np.isfinite(lh), np.isfinite(lo))
    if kind == "nan_density":
        assert not fin_o.all()
    scale = helpers.flux_scale(om, cfg, np.nan_to_num(U, nan=ff[0]))
    ok = fin_o.all(axis=1)
    err = helpers.scaled_err(Rh[ok][:, COMP], Ro[ok][:, COMP], scale[ok][:, COMP])
    assert err.max() <= 1e-12, err.max()


@pytest.mark.parametrize("rk,cfl", [(1, 20.0), (3, 100.0)])
def test_nonfinite_trap_run(hip, rk, cfl):
    """A run driven far past its stability limit blows up; the driver's non-finite trap stops it at the
    same iteration as the oracle's, after the same history."""
    m = helpers.mesh("40x8")
    om = ol.OracleMesh(m, np.float64)
    run = dict(iterations=200, min_iterations=200, rk_stages=rk, cfl_start=cfl, cfl_max=cfl, cfl_ramp_iters=0,
               residual_reduction_target=0.0, force_stability_tol=0.0, flux_scheme=1, limiter=1,
               startup_first_order_iters=5, mach=0.5, aoa_deg=2.0)
    ref = ol.run_case(om, helpers.case_cfg(run, False), 200)
    assert ref["nonfinite"] and ref["n_iter"] < 200
    from aero_hip import CaseConfig
    case = CaseConfig(**run)
    case.precond_experimental_enabled = 0
    say(f"[trap] rk={rk} cfl={cfl}: oracle trapped after {ref['n_iter']} iterations")
    with hip.HipMesh(m) as hm:
        out = hm.run_case(case)
    assert out["nonfinite"], "the HIP driver did not trap the non-finite state"
    compare_runs(out, ref, f"trap rk={rk}")
