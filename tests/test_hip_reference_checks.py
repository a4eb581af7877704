"""The reference's own acceptance checks, applied to HIP-backend runs of the same configurations.

Each test runs a reference CTest case (the configurations in tests/golden/cases.py, which the golden
driver runs pin against the compiled reference) through the device driver and asserts that test's
windows:

* tests/cpp/test_euler_regression.cpp:38-73: the residual stays bounded (final L2 <= 1.2 initial),
  0.16 <= CL <= 0.30, |CD| <= 0.25, finite coefficients, and the run's output files exist;
* tests/cpp/test_euler_aoa0_regression.cpp:60-88: |CL| < 0.05 and |CD| < 0.06 at zero incidence;
* tests/cpp/test_euler_low_mach_asymptotic.cpp:286-320 with tests/data/baselines/low_mach_euler/
  expected.yaml (M 0.01, 1200 iterations, f_min 0.05, ramp from 100 over 200): the all-speed fix lowers
  |CD * M| below 0.95 of the run without it;
* tests/cpp/test_euler_uniform_invariance.cpp:13-61: a uniform state (rho 1.05, u 0.32, v -0.08,
  p 0.9) with preconditioning on leaves |R| <= 1e-6; here on the 40x8 O-grid with the farfield at that
  state, in every cell off the wall (a wall face adds its pressure term to its cell).

The slip-wall mass flux the regression and wall-sanity tests bound is zero by construction here: the
wall flux's mass and energy components are set to exactly 0 (boundary_conditions.cpp:97-111).
"""
import sys

import numpy as np
import pytest

import cases
import helpers

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available()
    return aero_hip


def _run(hip, name):
    run = cases.RUNS[name]
    cfg = dict(run["cfg"])
    case = hip.CaseConfig(**cfg)
    case.precond_experimental_enabled = 1 if run["env_precond"] else 0
    m = helpers.mesh(run["mesh"], precision=4).astype(np.float64)
    with hip.HipMesh(m) as hm:
        return case, hm.run_case(case)


def test_regression_window(hip, tmp_path):
    case, out = _run(hip, "regression")
    h = out["history"]
    assert len(h) >= 2
    l2_0, l2_n = h[0, 2], h[-1, 2]
    assert np.isfinite(l2_n) and 0.0 < l2_n <= 1.2 * l2_0
    assert 0.16 <= out["cl"] <= 0.30, out["cl"]
    assert abs(out["cd"]) <= 0.25
    assert np.isfinite([out["cl"], out["cd"], out["cm"]]).all()
    # the same case through the file runner writes the reference's outputs
    c = case
    text = "\n".join([
        "case_type=euler_airfoil_2d", "num_circumferential=72", "num_radial=18", "farfield_radius=12",
        "radial_stretch=1.35", f"iterations={c.iterations}", f"min_iterations={c.min_iterations}",
        f"cfl_start={c.cfl_start}", f"cfl_max={c.cfl_max}", f"cfl_ramp_iters={c.cfl_ramp_iters}",
        f"residual_reduction_target={c.residual_reduction_target}", f"force_stability_tol={c.force_stability_tol}",
        f"force_stability_window={c.force_stability_window}", "flux=rusanov", "limiter=minmod", "precond=off",
        "low_mach_fix=off", f"startup_first_order_iters={c.startup_first_order_iters}", f"mach={c.mach}",
        f"aoa_deg={c.aoa_deg}"]) + "\n"
    (tmp_path / "native_case.cfg").write_text(text)
    s = hip.run_case_file(tmp_path / "native_case.cfg", tmp_path / "out")
    assert s["cl"] == out["cl"] and s["iterations"] == len(h)
    for f in ("residuals.csv", "forces.csv", "cp_wall.csv", "field_0000.vtu"):
        assert (tmp_path / "out" / f).exists()


def test_aoa0_window(hip):
    _, out = _run(hip, "aoa0")
    assert np.isfinite([out["cl"], out["cd"], out["cm"]]).all()
    assert abs(out["cl"]) < 0.05, out["cl"]
    assert abs(out["cd"]) < 0.06, out["cd"]


def test_low_mach_all_speed_mechanism(hip):
    say("[reference checks] low-Mach fix off / on, 1200 iterations each")
    case_off, off = _run(hip, "low_mach_off")
    case_on, on = _run(hip, "low_mach_on")
    assert case_off.mach == 0.01 and case_on.all_speed_flux_fix == 1 and case_off.all_speed_flux_fix == 0
    cdm_off, cdm_on = off["cd"] * case_off.mach, on["cd"] * case_on.mach
    assert np.isfinite([cdm_off, cdm_on]).all()
    assert abs(cdm_on) < 0.95 * abs(cdm_off), (cdm_on, cdm_off)


def test_uniform_invariance(hip):
    m = helpers.mesh("40x8")
    g = 1.4
    prim = (1.05, 0.32, -0.08, 0.9)
    r, u, v, p = prim
    cons = np.array([r, r * u, r * v, 0.0, p / (g - 1.0) + 0.5 * r * (u * u + v * v)])
    U = np.tile(cons, (m.num_cells, 1))
    cfg = hip.residual_config(cons, gamma=g, flux="hllc", limiter="venkat", second_order=True, precond=True,
                              mach_ref=0.15, mach_min=1e-3, beta_min=0.2, beta_max=1.0)
    with hip.HipMesh(m) as hm:
        R, lam, _ = hm.residual(U, cfg)
    scale = np.abs(cons).max() * np.abs(m.area).max()
    wall_owners = m.owner[m.wall_start:m.wall_start + m.wall_count]
    interior = np.ones(m.num_cells, bool)
    interior[wall_owners] = False
    assert np.abs(R[interior]).max() <= 1e-6 * scale
    assert np.isfinite(R).all() and np.isfinite(lam).all()
