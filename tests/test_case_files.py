"""Case files and output writers of the HIP runner (no GPU): the native key=value semantics of
residual_stub.cpp:160-189 and :402-559, the YAML mapping, the convergence-monitor columns of forces.csv,
the wall-face order of cp_wall.csv and the VTU layout."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import helpers
from aero_hip.case_file import euler_case_from_kv, load_case, native_kv_from_yaml, read_native_case
from aero_hip.outputs import force_monitor, order_wall_faces, wall_cp_from_state, write_vtu
from aero_hip.solver import CaseConfig


def test_native_parsing_rules(tmp_path):
    p = tmp_path / "native_case.cfg"
    p.write_text("# comment\n\ncase_type=euler_airfoil_2d\n  mach = 0.5  \naoa_deg:1.25\niterations=12.7\n"
                 "cfl_max=1.5abc\nflux=HLLC\nlimiter=Venkatakrishnan\nnum_radial=64\nbogus line\n=orphan\n")
    kv = read_native_case(p)
    assert kv["case_type"] == "euler_airfoil_2d" and kv["mach"] == "0.5" and kv["aoa_deg"] == "1.25"
    assert "" not in kv and "bogus line" not in kv
    c, mesh = euler_case_from_kv(kv)
    assert c.mach == 0.5 and c.aoa_deg == 1.25
    assert c.iterations == 12          # std::stoi reads the integer prefix
    assert c.cfl_max == 1.5            # std::stof reads the numeric prefix
    assert c.flux_scheme == 1 and c.limiter == 1
    assert mesh["num_radial"] == 64 and mesh["num_circumferential"] == 160  # AirfoilMeshConfig default


def test_defaults_are_the_native_ones():
    c, mesh = euler_case_from_kv({})
    ref = CaseConfig()
    assert c.iterations == 400 and c.min_iterations == ref.min_iterations
    assert c.flux_scheme == 0 and c.limiter == 0 and c.precond_on == 0 and c.low_mach_fix == 0
    assert mesh == dict(airfoil_source="naca4", naca="0012", airfoil_file="", chord=1.0, num_circumferential=160,
                        num_radial=48, farfield_radius=15.0, radial_stretch=1.4)


def test_precond_low_mach_coupling_aliases_and_clamps():
    # no precond key: precond follows the low_mach_fix token; low_mach_fix given explicitly
    c, _ = euler_case_from_kv({"low_mach_fix": "on"})
    assert c.precond_on == 1 and c.low_mach_fix == 1
    # precond given, low_mach_fix absent: low_mach_fix follows precond
    c, _ = euler_case_from_kv({"precond": "on"})
    assert c.precond_on == 1 and c.low_mach_fix == 1
    c, _ = euler_case_from_kv({"precond_on": "false", "low_mach_fix": "yes"})
    assert c.precond_on == 0 and c.low_mach_fix == 1
    # unknown tokens keep the fallback
    c, _ = euler_case_from_kv({"precond": "maybe"})
    assert c.precond_on == 0
    c, _ = euler_case_from_kv({"mach_ref": "0.05", "all_speed_theta_min": "-1", "all_speed_mach_cutoff": "0",
                               "stabilization_mach_floor": "0.07", "all_speed_ramp_iters": "-5",
                               "all_speed_pjump_hold_iters": "0", "force_mean_drift_tol": "0",
                               "force_stability_tol": "3e-4"})
    assert c.precond_mach_ref == 0.05
    assert c.all_speed_f_min == 0.05 and c.all_speed_mach_cutoff == 0.25
    assert c.stab_floor_start == 0.07 and c.stab_floor_target == 0.07
    assert c.all_speed_ramp_iters == 0 and c.all_speed_pjump_hold_iters == 1
    assert c.force_mean_drift_tol == 3e-4
    c, _ = euler_case_from_kv({"all_speed_f_min": "7"})
    assert c.all_speed_f_min == 1.0


def test_yaml_mapping(tmp_path):
    y = tmp_path / "case.yaml"
    y.write_text("case_type: euler_airfoil_2d\ngeometry: {type: airfoil2d, profile: NACA2412, chord: 1.0}\n"
                 "mesh: {generator: structured_o, circumferential: 96, radial: 24, farfield_radius: 12.0,\n"
                 "       radial_stretch: 1.3}\nflow: {mach: 0.4, aoa_deg: 3.0}\n"
                 "numerics: {flux: hllc, limiter: venkat, precond: off, low_mach_fix: off, cfl_max: 1.1}\n"
                 "post: {x_ref: 0.3}\n")
    kv = load_case(y)
    assert kv["naca_code"] == "2412" and kv["num_circumferential"] == "96" and kv["num_radial"] == "24"
    assert kv["iterations"] == "300"  # the YAML layer's default
    c, mesh = euler_case_from_kv(kv)
    assert (c.mach, c.aoa_deg, c.flux_scheme, c.limiter, c.cfl_max, c.x_ref) == (0.4, 3.0, 1, 1, 1.1, 0.3)
    assert c.precond_on == 0 and c.low_mach_fix == 0 and c.iterations == 300
    assert mesh["naca"] == "2412" and mesh["farfield_radius"] == 12.0 and mesh["radial_stretch"] == 1.3
    assert native_kv_from_yaml({"geometry": {"profile": "flat"}})["naca_code"] == "0012"


def test_force_monitor_columns():
    n, w, tol = 20, 4, 1e-3
    cl = 0.5 + 1e-4 * np.arange(n)           # linear drift of 1e-4 per iteration
    cd = np.full(n, 0.01)
    cm = np.zeros(n)
    m = force_monitor(cl, cd, cm, w, tol, 2e-4)
    assert np.allclose(m["d"][1:, 0], 1e-4) and m["d"][0, 0] == 0.0
    assert not m["stable"][: w - 1].any() and m["stable"][w - 1:].all()   # range 3e-4 <= 1e-3
    assert not m["mean_stable"][: 2 * w - 1].any()
    assert np.allclose(m["mean_d"][2 * w - 1:, 0], w * 1e-4)               # mean of the window moves by w * slope
    assert not m["mean_stable"][2 * w - 1:].any()                          # 4e-4 > 2e-4
    m2 = force_monitor(cl, cd, cm, w, tol, 0.0)                            # drift tol falls back to tol
    assert m2["converged"][2 * w - 1:].all()


def test_wall_face_order_and_cp():
    mesh = helpers.mesh("40x8")
    order = order_wall_faces(mesh)
    assert sorted(order.tolist()) == list(range(mesh.wall_start, mesh.wall_start + mesh.wall_count))
    fv = mesh.face_vertices
    for a, b in zip(order[:-1], order[1:]):          # a connected chain
        assert set(fv[a]) & set(fv[b])
    pts = mesh.points
    te = np.argmax(pts[np.unique(fv[order]), 0])
    first = mesh.face_center[order[0]]
    assert first[1] > 0.0 and first[0] > 0.9        # leaves the trailing edge over the upper surface
    assert np.isclose(mesh.face_center[order[len(order) // 2], 0], pts[:, 0][np.unique(fv[order])].min(), atol=0.05)
    # Cp of a uniform state at freestream pressure is 0
    case = CaseConfig(mach=0.5, aoa_deg=1.0)
    from aero_hip import freestream
    U = np.tile(freestream(0.5, 1.0)[1], (mesh.num_cells, 1))
    assert np.allclose(wall_cp_from_state(mesh, case, U), 0.0, atol=1e-9)
    del te


def test_vtu_layout(tmp_path):
    mesh = helpers.mesh("40x8", cells="mixed", quad_layers=3)
    U = np.tile([1.2, 10.0, 1.0, 0.0, 2.5e5], (mesh.num_cells, 1))
    R = np.zeros((mesh.num_cells, 5))
    R[:, 0] = mesh.volume
    p = tmp_path / "field_0000.vtu"
    write_vtu(p, mesh, U, R, 1.4)
    root = ET.parse(p).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfCells")) == mesh.num_cells and int(piece.get("NumberOfPoints")) == len(mesh.points)
    arrays = {a.get("Name"): a.text.split() for a in piece.iter("DataArray") if a.get("Name")}
    assert set(arrays) >= {"connectivity", "offsets", "types", "rho", "u", "v", "p", "Mach", "residual_magnitude"}
    types = np.array(arrays["types"], int)
    assert set(types.tolist()) == {5, 9} and len(types) == mesh.num_cells
    assert np.allclose(np.array(arrays["residual_magnitude"], float), 1.0, rtol=1e-5)
    assert np.allclose(np.array(arrays["rho"], float), 1.2)
