"""The driver's per-iteration diagnostics (aero_run_output.diagnostics, the state and controller columns of
euler_diagnostics.csv: euler_solver.cpp:971-1139, :1161-1203) against the FP64 oracle driver's rows.

The oracle's rows are pinned by test_oracle_golden: in FP32 they reproduce the reference's
euler_diagnostics.csv character for character on every golden run, which between them take every staged
controller action (10, 11, 20, 21, 31, 32, 33), reconstruction failures with a first failing face, and the
non-finite trap.  Here the device rows match the FP64 oracle's: the integer columns (controller stage,
effective fix, freeze counter, action, non-finite Cp count, first failing face) exactly, the others to
1e-9 of each column's scale over the run (the parity bar of the run histories), with an absolute floor
for columns that stay near zero: 1e-9 of the freestream quantity the column is measured in (a state that
agrees to 1e-9 of its scale moves such a column by that much).
"""
import sys

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available(), "no gfx950 device visible to the HIP backend"
    return aero_hip


@pytest.mark.parametrize("name", list(cases.RUNS))
def test_run_diagnostics_parity(hip, name):
    from aero_hip import CaseConfig, _lib
    run = cases.RUNS[name]
    m = helpers.mesh(run["mesh"])
    om = ol.OracleMesh(m, np.float64)
    say(f"[diagnostics] {name} mesh={run['mesh']} iterations={run['cfg']['iterations']}")
    ref = ol.run_case(om, helpers.case_cfg(run["cfg"], run["env_precond"]), run["cfg"]["iterations"],
                      diagnostics=True)
    case = CaseConfig(**run["cfg"])
    case.precond_experimental_enabled = 1 if run["env_precond"] else 0
    with hip.HipMesh(m) as hm:
        out = hm.run_case(case, diagnostics=True)
    do, dr = out["diagnostics"], ref["diagnostics"]
    assert do.shape == dr.shape == (ref["n_iter"], _lib.DIAG_FIELDS), (do.shape, dr.shape)
    assert out["nonfinite"] == ref["nonfinite"]
    from aero_hip.outputs import rho_inf_of, speed_inf_of
    rho, V = rho_inf_of(case), speed_inf_of(case)
    a = np.sqrt(case.gamma * case.p_inf / rho)
    q, pref, p, L = 0.5 * rho * V * V, rho * a * a, case.p_inf, max(case.chord, 1e-8)
    e_pj = 1e-9 * p / pref
    floor = dict(max_wall_un=V + a, max_farfield_abs_dp=p, max_farfield_abs_drho=rho, max_farfield_abs_du=V + a,
                 max_farfield_abs_dv=V + a, wall_cp_min=p / q, wall_cp_max=p / q, Fx_gauge=p * L, Fy_gauge=p * L,
                 Fx_abs=p * L, Fy_abs=p * L, L=p * L, D=p * L, cl_abs=p / q, cd_abs=p / q, cm_abs=p / q,
                 p_fluct_nd_max=p / pref, p_fluct_nd_wall=p / pref, p_fluct_nd_farfield=p / pref)
    exact = {"wall_cp_nonfinite", "all_speed_stage", "all_speed_flux_fix_effective", "all_speed_freeze_counter",
             "all_speed_controller_action", "first_failure_face"}
    for k, col in enumerate(_lib.DIAG_COLUMNS):
        x, y = do[:, k], dr[:, k]
        if col in exact:
            assert np.array_equal(x, y), (name, col, np.flatnonzero(x != y)[:5])
        else:
            same = x == y  # equal infinities (a sum that overflowed on the way to the trap) included
            assert np.array_equal(np.isfinite(x), np.isfinite(y)) and np.all(same | np.isfinite(y)), (name, col)
            x, y = x[~same], y[~same]
            scale = float(np.abs(dr[np.isfinite(dr[:, k]), k]).max(initial=0.0))
            if col.startswith("p_jump"):  # a mean of squared pressure jumps
                atol = 2.0 * np.sqrt(scale) * e_pj + e_pj * e_pj
            else:
                atol = 1e-9 * floor.get(col, 0.0)
            err = float(np.abs(x - y).max(initial=0.0))
            assert err <= 1e-9 * scale + atol + 1e-300, (name, col, err, scale)
