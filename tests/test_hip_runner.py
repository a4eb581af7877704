"""The case-file runner and sweep of the HIP backend (python -m aero_hip run / sweep): outputs in the
reference runner's layouts, equal to the driver's own results; every sweep run gets its own AoA."""
import sys

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu

CASE = """case_type=euler_airfoil_2d
airfoil_source=naca4
naca_code=0012
num_circumferential=40
num_radial=8
farfield_radius=10
radial_stretch=1.25
mach=0.5
aoa_deg=1.25
iterations=60
min_iterations=1
flux=hllc
limiter=venkat
cfl_max=1.1
residual_reduction_target=1e-12
"""


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available()
    return aero_hip


def _csv(path):
    lines = path.read_text().strip().splitlines()
    return lines[0].split(","), np.array([[float(x) for x in l.split(",")] for l in lines[1:]])


def test_run_case_file_outputs(hip, tmp_path, golden):
    from aero_hip.case_file import euler_case_from_kv, read_native_case
    cfg = tmp_path / "native_case.cfg"
    cfg.write_text(CASE)
    say("[runner] run_case_file")
    s = hip.run_case_file(cfg, tmp_path / "out")
    out = tmp_path / "out"
    for name in ("residuals.csv", "forces.csv", "cp_wall.csv", "field_0000.vtu", "run.log"):
        assert (out / name).exists(), name
    head, res = _csv(out / "residuals.csv")
    assert head == ["iter", "residual_l1", "residual_l2", "residual_linf", "residual_ratio"]
    assert len(res) == 60 and s["iterations"] == 60 and np.array_equal(res[:, 0], np.arange(60))
    fh, frc = _csv(out / "forces.csv")
    assert fh[:4] == ["iter", "cl", "cd", "cm"] and len(fh) == 16
    # the same run through the driver directly, on the reference generator's mesh
    case, _ = euler_case_from_kv(read_native_case(cfg))
    mesh = helpers.mesh("40x8", precision=4).astype(np.float64)
    with hip.HipMesh(mesh) as hm:
        ref = hm.run_case(case)
    assert s["cl"] == ref["cl"] and s["cd"] == ref["cd"] and s["cm"] == ref["cm"]
    assert np.allclose(frc[:, 1], ref["history"][:, 5], rtol=1e-5, atol=1e-7)   # six significant digits
    assert np.allclose(res[:, 2], ref["history"][:, 2], rtol=1e-5)
    ch, cp = _csv(out / "cp_wall.csv")
    assert ch == ["s", "x", "y", "Cp"] and len(cp) == mesh.wall_count
    assert np.all(np.diff(cp[:, 0]) > 0)
    from aero_hip.outputs import order_wall_faces
    order = order_wall_faces(mesh)
    assert np.allclose(cp[:, 3], ref["wall_cp"][order - mesh.wall_start], rtol=1e-5, atol=1e-6)
    assert np.allclose(cp[:, 1], mesh.face_center[order, 0], rtol=1e-5, atol=1e-6)
    assert "backend=hip" in (out / "run.log").read_text()
    dlines = (out / "euler_diagnostics.csv").read_text().splitlines()
    golden_header = str(golden("runs")["run_regression_diagnostics"][0])  # the reference's header
    assert dlines[0] == golden_header and len(dlines) == 61
    dh, dg = dlines[0].split(","), [l.split(",") for l in dlines[1:]]
    assert [int(r[0]) for r in dg] == list(range(60))
    assert all(r[dh.index("all_speed_stage")] == "single" for r in dg)
    assert np.allclose([float(r[dh.index("Cl")]) for r in dg], ref["history"][:, 5], rtol=1e-5, atol=1e-7)


def test_run_case_on_a_mesh_file(hip, tmp_path):
    """The runner on a binary mesh file of the case's O-grid (`--mesh`) reproduces the generated-mesh run:
    the same summary and the same output files, byte for byte."""
    from aero_hip.__main__ import main
    cfg = tmp_path / "native_case.cfg"
    cfg.write_text(CASE)
    mfile = tmp_path / "case.amsh"
    assert main(["mesh", str(cfg), str(mfile)]) == 0
    say("[runner] generated mesh vs mesh file")
    a = hip.run_case_file(cfg, tmp_path / "gen")
    b = hip.run_case_file(cfg, tmp_path / "file", mesh_path=mfile)
    assert (a["cl"], a["cd"], a["cm"], a["iterations"]) == (b["cl"], b["cd"], b["cm"], b["iterations"])
    for name in ("residuals.csv", "forces.csv", "cp_wall.csv", "field_0000.vtu"):
        assert (tmp_path / "gen" / name).read_bytes() == (tmp_path / "file" / name).read_bytes(), name


def test_sweep_runs_each_aoa(hip, tmp_path):
    (tmp_path / "base.cfg").write_text(CASE)
    (tmp_path / "sweep.yaml").write_text("base_case: base.cfg\nsweep:\n  aoa_deg: [0.0, 2.0]\n")
    say("[runner] serial sweep")
    r = hip.run_sweep_file(tmp_path / "sweep.yaml", tmp_path / "serial")
    cls = [run["cl"] for run in r["runs"]]
    assert [run["aoa_deg"] for run in r["runs"]] == [0.0, 2.0]
    assert abs(cls[0]) < 0.05 * abs(cls[1])   # the AoA reaches each run
    assert (tmp_path / "serial" / "run_001" / "forces.csv").exists()
    say("[runner] batched sweep")
    rb = hip.run_sweep_file(tmp_path / "sweep.yaml", tmp_path / "batched", batched=True)
    assert len(rb["runs"]) == 2 and rb["batched"]
    assert all(run["iterations"] == 60 for run in rb["runs"])
    assert (tmp_path / "batched" / "run_000" / "cp_wall.csv").exists()
    # same arithmetic as the serial runs until the serial driver's own stopping test: 60 iterations both
    assert rb["runs"][1]["cl"] == pytest.approx(cls[1], rel=1e-12)


def test_forces_match_oracle(hip):
    """aero_hip_forces (gauge and absolute) against the FP64 oracle's integrate_pressure_forces on the owner
    pressures of the same state, and equal to the driver's final coefficients."""
    import ctypes as C
    import oracle_lib as ol
    m = helpers.mesh("40x8")
    om = ol.OracleMesh(m, np.float64)
    case = hip.CaseConfig(mach=0.5, aoa_deg=1.25, flux_scheme=1, limiter=1, iterations=40, min_iterations=1,
                          residual_reduction_target=0.0)
    with hip.HipMesh(m) as hm:
        out = hm.run_case(case)
        U = out["conserved"]
        rho_inf = case.p_inf / (case.gas_constant * case.t_inf)
        V = case.mach * np.sqrt(case.gamma * case.p_inf / rho_inf)
        for gauge in (True, False):
            f = hm.forces(rho_inf, case.p_inf, case.aoa_deg, V, gauge=gauge)
            rho = np.maximum(U[:, 0], 1e-8)
            inv = 1.0 / rho
            u, v = U[:, 1] * inv, U[:, 2] * inv
            p = np.maximum((1.4 - 1.0) * (U[:, 4] - 0.5 * rho * (u * u + v * v)), 1e-8)
            ref7 = np.array([rho_inf, case.p_inf, case.aoa_deg, V, 1.0, 0.25, 0.0])
            o = np.zeros(8)
            ol.oracle().ao_forces_f64(C.byref(om.c), ol.ptr(p), ol.ptr(ref7), int(gauge), ol.ptr(o))
            got = np.array([f[k] for k in ("fx", "fy", "lift", "drag", "moment", "cl", "cd", "cm")])
            assert np.allclose(got, o, rtol=1e-13, atol=1e-9 * np.abs(o).max()), (gauge, got, o)
            if gauge:
                assert f["cl"] == pytest.approx(out["cl"], rel=1e-12)
                assert np.allclose(f["wall_cp"], out["wall_cp"], rtol=1e-12, atol=1e-14)
