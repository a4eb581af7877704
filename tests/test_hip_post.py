"""The device's post-run scaling diagnostics (aero_hip_face_scales) against the FP64 oracle, and the
runner's wall_flux.csv, dissipation_debug.csv and run.log statistics built from them.

The oracle function is pinned to the reference's own dissipation_debug.csv, wall_flux.csv and run
statistics in test_oracle_post.  The device record performs the same FP64 operations in the same order
and equals the oracle's bit for bit."""
import sys

import numpy as np
import pytest

import helpers
import oracle_lib as ol
from aero_hip import _lib as hl
from aero_hip.solver import freestream

pytestmark = pytest.mark.gpu

CONFIGS = {
    "hllc_venkat_plain": dict(flux_scheme=1, limiter=1),
    "precond_lowmach_allspeed": dict(flux_scheme=1, limiter=0, precond_enabled=1, precond_mach_ref=0.15,
                                     precond_mach_min=0.001, precond_beta_min=0.2, precond_beta_max=1.0,
                                     low_mach_fix=1, low_mach_cutoff=0.3, stab_mach_floor=0.05, all_speed_fix=1,
                                     all_speed_cutoff=0.25, all_speed_f_min=0.05, all_speed_ramp=0.6),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available()
    return aero_hip


MESHES = {"40x8_tri": ("40x8", {}), "64x14_mixed": ("64x14", dict(cells="mixed", quad_layers=4, jitter=0.1, seed=3))}


@pytest.mark.parametrize("mesh_name", list(MESHES))
@pytest.mark.parametrize("cname", list(CONFIGS))
def test_face_scales_match_oracle(hip, mesh_name, cname):
    name, kw = MESHES[mesh_name]
    m = helpers.mesh(name, **kw)
    _, ff = freestream(0.3, 2.0)
    U = helpers.perturbed_state(m.cell_center, ff, 0.05, 7)
    cfg = helpers.res_cfg(dict(farfield=ff, gamma=1.4, second_order=1, **CONFIGS[cname]))
    ofs, ocs = ol.face_scales(ol.OracleMesh(m, np.float64), cfg, U)
    say(f"[post] face scales {mesh_name} {cname} N={m.num_cells} F={m.num_faces}")
    with hip.HipMesh(m) as hm:
        hm.set_state(np.tile(ff, (m.num_cells, 1)))
        S0 = hm.state()
        fs, cs = hm.face_scales(cfg, U)
        assert np.array_equal(hm.state(), S0)  # evaluated on a copy: the handle's state is unchanged
    for c in range(hl.FS_FIELDS):
        assert np.array_equal(fs[c], ofs[c]), (c, float(np.abs(fs[c] - ofs[c]).max()))
    for c in range(2):
        assert np.array_equal(cs[c], ocs[c]), c
    say(f"[post] bitwise faces={np.array_equal(fs, ofs)} cells={np.array_equal(cs, ocs)}")


def test_runner_post_files(hip, tmp_path):
    """run_case_file writes wall_flux.csv and dissipation_debug.csv in the reference's layouts and the
    run.log statistics; the rows equal the oracle's record at the run's final state and configuration."""
    import test_hip_runner as tr
    from aero_hip import outputs
    cfgf = tmp_path / "case.cfg"
    cfgf.write_text(tr.CASE)
    say("[post] runner")
    hip.run_case_file(cfgf, tmp_path / "out")
    out = tmp_path / "out"
    case, m = outputs.case_and_mesh(hip.load_case(cfgf))
    with hip.HipMesh(m) as hm:
        r = hm.run_case(case)
    U = r["conserved"]
    ofs, ocs = ol.face_scales(ol.OracleMesh(m, np.float64), r["final_config"], U)
    wf = np.loadtxt(out / "wall_flux.csv", delimiter=",", skiprows=1, ndmin=2)
    wall = np.flatnonzero(m.bc == hl.BC_SLIP_WALL)
    assert wf[:, 0].astype(int).tolist() == wall.tolist() and np.all(wf[:, 3] == 0.0)
    lines = (out / "dissipation_debug.csv").read_text().splitlines()
    assert lines[0].startswith("face,region,x,y,area,mach_local,un_abs,a,beta,")
    sel = outputs.dissipation_faces(m)
    assert len(lines) - 1 == len(sel) == 20
    for line, (face, region) in zip(lines[1:], sel):
        f = line.split(",")
        assert int(f[0]) == face and f[1] == region
        got = np.array([float(x) for x in f[5:13]])
        want = ofs[[hl.FS_MACH, hl.FS_UN_ABS, hl.FS_SOUND, hl.FS_BETA, hl.FS_LOW_MACH, hl.FS_ACOUSTIC_BASE,
                    hl.FS_STAB, hl.FS_ACOUSTIC], face]
        assert np.allclose(got, want, rtol=1e-5, atol=1e-9), (face, got, want)
    log = dict(line.split("=", 1) for line in (out / "run.log").read_text().splitlines() if "=" in line)
    st = outputs.scaling_statistics(m, ofs, ocs)
    for key in ("beta_min", "mach_local_max", "acoustic_scale_final_mean", "wall_adjacent_acoustic_scale_p01"):
        assert float(log[key]) == pytest.approx(st[key], rel=1e-5), key
    fa = None
    with hip.HipMesh(m) as hm:
        fa = hm.forces(outputs.rho_inf_of(case), case.p_inf, case.aoa_deg, outputs.speed_inf_of(case), case.chord,
                       case.x_ref, case.y_ref, gauge=False, gamma=case.gamma, U=U)
    assert float(log["cl_abs"]) == pytest.approx(fa["cl"], rel=1e-5)
    assert float(log["Fx_abs"]) == pytest.approx(fa["fx"], rel=1e-5)
    wall_nA = -(m.normal[wall, :2] * m.area[wall, None]).sum(axis=0)
    assert float(log["sum_nA_mag"]) == pytest.approx(float(np.hypot(*wall_nA)), rel=1e-5, abs=1e-12)
    assert log["force_physical_converged_final"] in ("0", "1") and log["rk_stages"] == "3"
