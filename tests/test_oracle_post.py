"""The oracle's post-run scaling diagnostics against the compiled reference's own outputs.

For each golden run, the FP64 oracle's per-face scaling record (ao_face_scales_f64) and per-cell beta and
local Mach, evaluated at the reference's final FP32 state under the residual configuration of the oracle
driver's last stage (the driver is pinned bit for bit by test_oracle_golden), reproduce the reference's
dissipation_debug.csv rows, its wall_flux.csv and the EulerRunResult statistics that run.log prints
(euler_solver.cpp:1217-1465).  The reference computes in FP32 and prints six significant digits, so the
comparison is relative 3e-5 of each column's scale.  The face selection and the statistics are the
runner's host code (aero_hip.outputs), so this also pins that code.
"""
import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol
from aero_hip import _lib as hl
from aero_hip import outputs
from aero_hip.solver import freestream

POST_KEYS = ("max_wall_mass_flux", "beta_min", "beta_max", "beta_mean", "mach_local_min", "mach_local_max",
             "mach_local_mean", "acoustic_scale_min", "acoustic_scale_max", "acoustic_scale_mean",
             "acoustic_scale_final_min", "acoustic_scale_final_max", "acoustic_scale_final_mean",
             "wall_adjacent_acoustic_scale_min", "wall_adjacent_acoustic_scale_p01",
             "wall_adjacent_acoustic_scale_p50", "wall_adjacent_stabilization_scale_min",
             "wall_adjacent_stabilization_scale_p01", "wall_adjacent_stabilization_scale_p50",
             "wall_adjacent_acoustic_scale_final_min", "wall_adjacent_acoustic_scale_final_p01",
             "wall_adjacent_acoustic_scale_final_p50")
RTOL = 3.0e-5


def _close(ours, ref, what):
    ours, ref = np.asarray(ours, float), np.asarray(ref, float)
    scale = max(float(np.abs(ref).max(initial=0.0)), 1.0e-30)
    err = float(np.abs(ours - ref).max(initial=0.0)) / scale
    assert err <= RTOL, f"{what}: max error {err:.2e} of scale {scale:.3e}"


def _setup(golden, name):
    g = golden("runs")
    run = cases.RUNS[name]
    m = helpers.mesh(run["mesh"], precision=4)
    cfg = ol.run_case_cfg(ol.OracleMesh(m, np.float32), helpers.case_cfg(run["cfg"], run["env_precond"],
                                                                          np.float32))
    U = g[f"run_{name}_U"].astype(np.float64)
    fs, cs = ol.face_scales(ol.OracleMesh(m, np.float64), cfg, U)
    return g, m, cfg, fs, cs


@pytest.mark.parametrize("name", [n for n in cases.RUNS if n not in cases.TRAPPED])
def test_dissipation_rows(golden, oracle_lib, name):
    g, m, cfg, fs, _ = _setup(golden, name)
    ref = g[f"run_{name}_dissipation"]
    sel = outputs.dissipation_faces(m)
    assert [f for f, _ in sel] == ref[:, 0].astype(int).tolist()
    assert [outputs.DISSIPATION_REGIONS.index(r) for _, r in sel] == ref[:, 1].astype(int).tolist()
    faces = ref[:, 0].astype(int)
    assert float(cfg.stab_mach_floor) == pytest.approx(float(g[f"run_{name}_post"][22]), rel=1e-6)
    cols = [hl.FS_MACH, hl.FS_UN_ABS, hl.FS_SOUND, hl.FS_BETA, hl.FS_LOW_MACH, hl.FS_ACOUSTIC_BASE, hl.FS_STAB,
            hl.FS_ACOUSTIC, None, hl.FS_ALL_SPEED, hl.FS_SMAX_PHYSICAL, hl.FS_SMAX_DISSIPATION]
    for k, c in enumerate(cols):
        if c is None:  # stabilization_m0: the configuration's floor
            _close(np.full(len(faces), min(max(cfg.stab_mach_floor, 0.0), 1.0)), ref[:, 5 + k], "stabilization_m0")
        else:
            _close(fs[c, faces], ref[:, 5 + k], f"{name} column {5 + k}")


@pytest.mark.parametrize("name", [n for n in cases.RUNS if n not in cases.TRAPPED])
def test_statistics_and_wall_flux(golden, oracle_lib, name):
    g, m, _, fs, cs = _setup(golden, name)
    st = outputs.scaling_statistics(m, fs, cs)
    post = g[f"run_{name}_post"]
    for k, key in enumerate(POST_KEYS):
        assert st[key] == pytest.approx(float(post[k]), rel=RTOL, abs=1e-7), key
    wf = g[f"run_{name}_wall_flux"]
    wall = np.flatnonzero(m.bc == hl.BC_SLIP_WALL)
    assert wf[:, 0].astype(int).tolist() == wall.tolist()
    assert np.array_equal(fs[hl.FS_WALL_MASS_FLUX, wall], wf[:, 3])  # both exactly zero


def test_statistics_reduce_to_closed_forms(oracle_lib):
    """Preconditioner and low-Mach fix off, zero floor: beta = 1, base scale = 1, and at a uniform state the
    stabilisation scale is clamp(smooth_max(M, 0)) of the freestream Mach."""
    m = helpers.mesh("40x8", precision=4)
    om = ol.OracleMesh(m, np.float64)
    ff = np.asarray(freestream(0.3, 2.0)[1], float)
    cfg = helpers.res_cfg(dict(farfield=ff, gamma=1.4, flux_scheme=1, limiter=1, second_order=1))
    U = np.tile(ff, (m.num_cells, 1))
    fs, cs = ol.face_scales(om, cfg, U)
    assert np.all(fs[hl.FS_BETA] == 1.0) and np.all(cs[0] == 1.0)
    assert np.all(fs[hl.FS_ACOUSTIC_BASE] == 1.0) and np.all(fs[hl.FS_ACOUSTIC] == 1.0)
    rho, u, v = ff[0], ff[1] / ff[0], ff[2] / ff[0]
    p = 0.4 * (ff[4] - 0.5 * rho * (u * u + v * v))
    mach = np.hypot(u, v) / np.sqrt(1.4 * p / rho)
    assert np.allclose(cs[1], mach, rtol=1e-14)
    stab = 0.5 * (mach + np.sqrt(mach * mach + 1e-12))
    assert np.allclose(fs[hl.FS_STAB], stab, rtol=1e-14)


@pytest.mark.parametrize("name", ["regression", "case_file_numerics", "precond_staged"])
def test_writer_reproduces_reference_files(golden, oracle_lib, name, tmp_path):
    """The runner's writer (aero_hip.outputs.write_post_files), fed the oracle's record, writes the
    reference's wall_flux.csv and dissipation_debug.csv: the same faces, regions and geometry columns and
    the same values to the printed digits."""
    g, m, cfg, fs, _ = _setup(golden, name)
    paths = outputs.write_post_files(tmp_path, m, cfg, fs)
    lines = paths["dissipation"].read_text().splitlines()
    assert lines[0] == ("face,region,x,y,area,mach_local,un_abs,a,beta,low_mach_scale,acoustic_scale,s_stab,"
                        "acoustic_scale_final,stabilization_m0,all_speed_pressure_scale,smax_physical,"
                        "smax_dissipation")
    ref = g[f"run_{name}_dissipation"]
    got = np.array([[float(f[0]), float(outputs.DISSIPATION_REGIONS.index(f[1]))] + [float(x) for x in f[2:]]
                    for f in (line.split(",") for line in lines[1:])])
    assert got.shape == ref.shape
    assert np.array_equal(got[:, :2], ref[:, :2])
    for k in range(2, ref.shape[1]):
        _close(got[:, k], ref[:, k], f"{name} dissipation column {k}")
    wf = np.loadtxt(paths["wall_flux"], delimiter=",", skiprows=1, ndmin=2)
    ref_wf = g[f"run_{name}_wall_flux"]
    assert np.array_equal(wf[:, [0, 3]], ref_wf[:, [0, 3]])
    _close(wf[:, 1:3], ref_wf[:, 1:3], f"{name} wall_flux x, y")
