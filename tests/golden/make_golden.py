"""Write the golden vectors under tests/golden/ from the compiled FP32 reference.

Run in the container that has the reference tree (the library is built by `make -C oracle ref` from
the reference's own sources into oracle/_ref/):

    python tests/golden/make_golden.py

The fixtures are data only (inputs and the reference's outputs), little-endian .npz without pickles.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(ROOT / "tests"))
import cases  # noqa: E402

REF_LIB = ROOT / "oracle" / "_ref" / "libaero_ref.so"
f32p = lambda a: a.ctypes.data_as(C.c_void_p)


class RefMeshCfg(C.Structure):
    _fields_ = [("num_circumferential", C.c_int), ("num_radial", C.c_int), ("farfield_radius", C.c_float),
                ("radial_stretch", C.c_float), ("chord", C.c_float), ("naca_code", C.c_char * 8)]


class RefResCfg(C.Structure):
    _fields_ = [("farfield", C.c_float * 5), ("gamma", C.c_float), ("scheme", C.c_int), ("limiter", C.c_int),
                ("second_order", C.c_int), ("low_mach_fix", C.c_int), ("low_mach_cutoff", C.c_float),
                ("stab_mach_floor", C.c_float), ("precond_farfield_bc", C.c_int), ("precond_enabled", C.c_int),
                ("mach_ref", C.c_float), ("mach_min", C.c_float), ("beta_min", C.c_float),
                ("beta_max", C.c_float), ("as_fix", C.c_int), ("as_cutoff", C.c_float), ("as_fmin", C.c_float),
                ("as_ramp", C.c_float)]


_CASE_NAMES = [
    ("iterations", C.c_int), ("min_iterations", C.c_int), ("cfl_ramp_iters", C.c_int), ("gamma", C.c_float),
    ("gas_constant", C.c_float), ("mach", C.c_float), ("aoa_deg", C.c_float), ("p_inf", C.c_float),
    ("t_inf", C.c_float), ("rho_inf", C.c_float), ("cfl_start", C.c_float), ("cfl_max", C.c_float),
    ("residual_reduction_target", C.c_float), ("force_stability_tol", C.c_float),
    ("force_stability_window", C.c_int), ("force_mean_drift_tol", C.c_float),
    ("startup_first_order_iters", C.c_int), ("rk_stages", C.c_int), ("local_time_stepping", C.c_int),
    ("precond_on", C.c_int), ("precond_mach_ref", C.c_float), ("precond_mach_min", C.c_float),
    ("precond_beta_min", C.c_float), ("precond_beta_max", C.c_float), ("precond_farfield_bc", C.c_int),
    ("stab_floor_start", C.c_float), ("stab_floor_target", C.c_float), ("stab_floor_k_start", C.c_float),
    ("stab_floor_k_target", C.c_float), ("stab_ramp_iters", C.c_int), ("x_ref", C.c_float), ("y_ref", C.c_float),
    ("flux_scheme", C.c_int), ("limiter", C.c_int), ("low_mach_fix", C.c_int), ("mach_cutoff", C.c_float),
    ("all_speed_flux_fix", C.c_int), ("all_speed_mach_cutoff", C.c_float), ("all_speed_f_min", C.c_float),
    ("all_speed_ramp_start_iter", C.c_int), ("all_speed_ramp_iters", C.c_int),
    ("all_speed_staged_controller", C.c_int), ("all_speed_stage_a_min_iters", C.c_int),
    ("all_speed_stage_a_max_iters", C.c_int), ("all_speed_stage_f_min_high", C.c_float),
    ("all_speed_stage_f_min_mid", C.c_float), ("all_speed_stage_f_min_low", C.c_float),
    ("all_speed_pjump_wall_target", C.c_float), ("all_speed_pjump_spike_factor", C.c_float),
    ("all_speed_pjump_hold_iters", C.c_int), ("all_speed_freeze_iters", C.c_int),
    ("all_speed_cfl_drop_factor", C.c_float), ("all_speed_cfl_restore_factor", C.c_float),
    ("all_speed_cfl_min_scale", C.c_float), ("aoa0_symmetry_enforce", C.c_int),
    ("aoa0_symmetry_enforce_interval", C.c_int)]


class RefCaseCfg(C.Structure):
    _fields_ = _CASE_NAMES


# EulerAirfoilCaseConfig defaults (euler_solver.hpp:14-74), for fields a run leaves unset.
CASE_DEFAULTS = dict(
    iterations=400, min_iterations=40, cfl_ramp_iters=120, gamma=1.4, gas_constant=287.05, mach=0.15, aoa_deg=2.0,
    p_inf=101325.0, t_inf=288.15, rho_inf=0.0, cfl_start=0.2, cfl_max=1.2, residual_reduction_target=1.0e-3,
    force_stability_tol=2.0e-5, force_stability_window=6, force_mean_drift_tol=2.0e-5, startup_first_order_iters=25,
    rk_stages=3, local_time_stepping=1, precond_on=0, precond_mach_ref=0.15, precond_mach_min=1.0e-3,
    precond_beta_min=0.2, precond_beta_max=1.0, precond_farfield_bc=0, stab_floor_start=0.1, stab_floor_target=0.02,
    stab_floor_k_start=-1.0, stab_floor_k_target=-1.0, stab_ramp_iters=400, x_ref=0.25, y_ref=0.0, flux_scheme=0,
    limiter=0, low_mach_fix=0, mach_cutoff=0.3, all_speed_flux_fix=0, all_speed_mach_cutoff=0.25,
    all_speed_f_min=0.05, all_speed_ramp_start_iter=0, all_speed_ramp_iters=0, all_speed_staged_controller=0,
    all_speed_stage_a_min_iters=600, all_speed_stage_a_max_iters=0, all_speed_stage_f_min_high=0.2,
    all_speed_stage_f_min_mid=0.1, all_speed_stage_f_min_low=0.05, all_speed_pjump_wall_target=0.0,
    all_speed_pjump_spike_factor=1.5, all_speed_pjump_hold_iters=120, all_speed_freeze_iters=120,
    all_speed_cfl_drop_factor=0.8, all_speed_cfl_restore_factor=1.02, all_speed_cfl_min_scale=0.1,
    aoa0_symmetry_enforce=0, aoa0_symmetry_enforce_interval=0)


def full_case(over: dict) -> dict:
    d = dict(CASE_DEFAULTS)
    d.update(over)
    return d


def mesh_cfg(name: str) -> RefMeshCfg:
    m = cases.MESHES[name]
    return RefMeshCfg(m["num_circumferential"], m["num_radial"], m["farfield_radius"], m["radial_stretch"], 1.0,
                      b"0012")


def ref_mesh(ref, name: str) -> dict:
    cfg = mesh_cfg(name)
    n = [C.c_int() for _ in range(7)]
    ref.ref_mesh_sizes(C.byref(cfg), *[C.byref(x) for x in n])
    N, F, P, ws, wc, fs, fc = [x.value for x in n]
    a = dict(owner=np.zeros(F, np.int32), neighbor=np.zeros(F, np.int32), bc=np.zeros(F, np.int32),
             face_vertices=np.zeros(2 * F, np.int32), normal=np.zeros(3 * F, np.float32),
             face_center=np.zeros(3 * F, np.float32), area=np.zeros(F, np.float32), volume=np.zeros(N, np.float32),
             cell_center=np.zeros(3 * N, np.float32), points=np.zeros(3 * P, np.float32))
    ref.ref_mesh_fill(C.byref(cfg), *[f32p(a[k]) for k in ("owner", "neighbor", "bc", "face_vertices", "normal",
                                                            "face_center", "area", "volume", "cell_center", "points")])
    a.update(num_cells=N, num_faces=F, wall_start=ws, wall_count=wc, farfield_start=fs, farfield_count=fc)
    return a


def mesh_digest(a: dict) -> str:
    h = hashlib.sha256()
    for k in ("owner", "neighbor", "bc", "face_vertices", "normal", "face_center", "area", "volume", "cell_center",
              "points"):
        h.update(np.ascontiguousarray(a[k]).tobytes())
    return h.hexdigest()


def freestream32(mach, aoa, p=101325.0, t=288.15, g=1.4, rgas=287.05):
    f = np.float32
    rho = f(p) / (f(rgas) * f(t))
    alpha = f(aoa) * (f(3.14159265358979323846) / f(180.0))
    a = np.sqrt(f(g) * f(p) / rho, dtype=np.float32)
    v = f(max(mach, 0.0)) * a
    u, w = v * np.cos(alpha, dtype=np.float32), v * np.sin(alpha, dtype=np.float32)
    ke = f(0.5) * rho * (u * u + w * w + f(0) * f(0))
    return np.array([rho, rho * u, rho * w, rho * f(0), f(p) / (f(g) - f(1)) + ke], np.float32)


def perturbed_state(cc: np.ndarray, uinf32: np.ndarray, amp: float, seed: int) -> np.ndarray:
    """Conservative FP32 state: freestream primitives times (1 + amp sin cos) with seeded phases."""
    rng = np.random.default_rng(seed)
    x, y = cc[:, 0].astype(np.float64), cc[:, 1].astype(np.float64)
    g = 1.4
    rho0 = float(uinf32[0])
    u0, v0 = float(uinf32[1]) / rho0, float(uinf32[2]) / rho0
    p0 = (g - 1.0) * (float(uinf32[4]) - 0.5 * rho0 * (u0 * u0 + v0 * v0))
    ph = rng.uniform(0.0, 2.0 * np.pi, (4, 2))
    s = lambda k: np.sin(2 * np.pi * x + ph[k, 0]) * np.cos(2 * np.pi * y + ph[k, 1])
    vref = max(abs(u0), abs(v0), 1.0)
    rho = rho0 * (1.0 + amp * s(0))
    u = u0 + amp * vref * s(1)
    v = v0 + amp * vref * s(2)
    p = p0 * (1.0 + amp * s(3))
    U = np.zeros((len(x), 5))
    U[:, 0] = rho
    U[:, 1] = rho * u
    U[:, 2] = rho * v
    U[:, 4] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    return U.astype(np.float32)


def ref_res_cfg(vals: dict) -> RefResCfg:
    return RefResCfg((C.c_float * 5)(*vals["farfield"]), vals["gamma"], vals["flux_scheme"], vals["limiter"],
                     vals["second_order"], vals["low_mach_fix"], vals["low_mach_cutoff"], vals["stab_mach_floor"],
                     vals["precond_farfield_bc"], vals["precond_enabled"], vals["precond_mach_ref"],
                     vals["precond_mach_min"], vals["precond_beta_min"], vals["precond_beta_max"],
                     vals["all_speed_fix"], vals["all_speed_cutoff"], vals["all_speed_f_min"], vals["all_speed_ramp"])


def pointwise(ref) -> dict:
    rng = np.random.default_rng(20260101)
    out = {}
    # Limiter grid: signs, zeros, tiny, huge and non-finite arguments.
    base = np.array([0.0, -0.0, 1e-30, -1e-30, 1e-7, -1e-7, 0.3, -0.3, 1.0, -1.0, 2.5, -2.5, 1e20, -1e20,
                     np.inf, -np.inf, np.nan], np.float32)
    a, b = np.meshgrid(base, base, indexing="ij")
    ra = np.concatenate([a.ravel(), rng.normal(0, 1, 4000).astype(np.float32)])
    rb = np.concatenate([b.ravel(), rng.normal(0, 1, 4000).astype(np.float32)])
    out["lim_a"], out["lim_b"] = ra, rb
    for lim in (0, 1):
        o = np.zeros_like(ra)
        ref.ref_limiter(lim, f32p(ra), f32p(rb), f32p(o), len(ra))
        out[f"lim_{lim}"] = o
    # Face flux inputs: primitive pairs over a Mach range, some near-vacuum / negative-pressure states.
    n = 3000
    g = np.float32(1.4)

    def states(k):
        rho = np.exp(rng.uniform(-3, 2, k))
        p = np.exp(rng.uniform(-2, 12, k))
        a = np.sqrt(1.4 * p / rho)
        mach = np.exp(rng.uniform(np.log(1e-3), np.log(3.0), k))
        th = rng.uniform(0, 2 * np.pi, k)
        u, v = mach * a * np.cos(th), mach * a * np.sin(th)
        U = np.stack([rho, rho * u, rho * v, np.zeros(k), p / 0.4 + 0.5 * rho * (u * u + v * v)], 1)
        bad = rng.random(k) < 0.03
        U[bad, 4] *= rng.uniform(0.0, 0.9, bad.sum())  # negative internal energy -> pressure floor
        tiny = rng.random(k) < 0.02
        U[tiny, 0] = 1e-10
        return U.astype(np.float32)

    UL, UR = states(n), states(n)
    same = rng.random(n) < 0.05
    UR[same] = UL[same]
    near = rng.random(n) < 0.2  # nearby states (smooth-flow faces)
    UR[near] = (UL[near] * (1.0 + 0.01 * rng.normal(size=(near.sum(), 5)))).astype(np.float32)
    UR[near, 3] = 0
    th = rng.uniform(0, 2 * np.pi, n)
    nrm = np.stack([np.cos(th), np.sin(th), np.zeros(n)], 1).astype(np.float32)
    scale = np.exp(rng.uniform(np.log(1e-5), np.log(1.2), n)).astype(np.float32)
    out.update(flux_UL=UL, flux_UR=UR, flux_n=nrm, flux_scale=scale)
    for scheme in (0, 1):
        for asf, w in ((0, 1.0), (1, 0.0), (1, 0.5), (1, 1.0)):
            if scheme == 0 and asf:
                continue
            F = np.zeros((n, 5), np.float32)
            lam = np.zeros(n, np.float32)
            fb = np.zeros(n, np.int32)
            ref.ref_flux(n, scheme, f32p(UL), f32p(UR), f32p(nrm), f32p(scale), C.c_float(g), asf, C.c_float(0.25),
                         C.c_float(0.05), C.c_float(w), f32p(F), f32p(lam), f32p(fb))
            tag = f"{scheme}_{asf}_{w}"
            out[f"flux_F_{tag}"], out[f"flux_lam_{tag}"], out[f"flux_fb_{tag}"] = F, lam, fb
    ff = freestream32(0.3, 2.0)
    out["bflux_ff"] = ff
    for bc, cbc in ((2, 0), (1, 0), (1, 1)):
        for scheme in (0, 1):
            F = np.zeros((n, 5), np.float32)
            lam = np.zeros(n, np.float32)
            fb = np.zeros(n, np.int32)
            ref.ref_boundary_flux(n, bc, scheme, f32p(UL), f32p(ff), f32p(nrm), f32p(scale), C.c_float(g), cbc, 0,
                                  C.c_float(0.25), C.c_float(0.05), C.c_float(1.0), f32p(F), f32p(lam), f32p(fb))
            tag = f"{bc}_{cbc}_{scheme}"
            out[f"bflux_F_{tag}"], out[f"bflux_lam_{tag}"], out[f"bflux_fb_{tag}"] = F, lam, fb
    return out


def residuals(ref) -> dict:
    out = {}
    for mname, combos, seeds in (("40x8", cases.RESIDUAL_COMBOS, (7, 8)), ("64x14", cases.RESIDUAL_COMBOS[::6], (9,))):
        mesh = ref_mesh(ref, mname)
        cc = mesh["cell_center"].reshape(-1, 3)
        N = mesh["num_cells"]
        ff = freestream32(0.3, 2.0)
        for si, seed in enumerate(seeds):
            amp = 0.05 if si == 0 else 0.9  # the second state is strong enough to trigger fallbacks
            U = perturbed_state(cc, ff, amp, seed)
            out[f"res_{mname}_U{si}"] = U
            G = np.zeros((N, 5, 2), np.float32)
            ref.ref_gradients(C.byref(mesh_cfg(mname)), f32p(U), C.c_float(1.4), f32p(G))
            out[f"grad_{mname}_U{si}"] = G
            for ci, combo in enumerate(combos if si == 0 else combos[::4]):
                vals = cases.residual_cfg_values(combo, ff)
                R = np.zeros((N, 5), np.float32)
                lam = np.zeros(N, np.float32)
                dg = np.zeros(3, np.int32)
                ref.ref_residual(C.byref(mesh_cfg(mname)), C.byref(ref_res_cfg(vals)), f32p(U), f32p(R), f32p(lam),
                                 f32p(dg))
                key = f"res_{mname}_U{si}_" + "".join(str(x) for x in combo)
                out[key + "_R"], out[key + "_lam"], out[key + "_diag"] = R, lam, dg
    # Stage tail: preconditioned residual, dt/V (local and global), RK3 stages.
    rng = np.random.default_rng(31)
    n = 500
    un = perturbed_state(rng.uniform(-1, 1, (n, 3)), freestream32(0.3, 2.0), 0.05, 3)
    us = perturbed_state(rng.uniform(-1, 1, (n, 3)), freestream32(0.3, 2.0), 0.05, 4)
    res = (rng.normal(0, 1, (n, 5)) * np.array([0.1, 30, 30, 0, 1e4])).astype(np.float32)
    lam = np.exp(rng.uniform(0, 8, n)).astype(np.float32)
    out.update(stage_un=un, stage_us=us, stage_res=res, stage_lam=lam)
    for pc in (0, 1):
        vals = cases.residual_cfg_values((1, 1, 1, 0, pc, 0, 0), freestream32(0.3, 2.0))
        for lts in (0, 1):
            for st in (0, 1, 2):
                nxt = np.zeros((n, 5), np.float32)
                ref.ref_stage_tail(n, st, f32p(un), f32p(us), f32p(res), f32p(lam), C.c_float(0.7), lts,
                                   C.byref(ref_res_cfg(vals)), f32p(nxt))
                out[f"stage_{pc}_{lts}_{st}"] = nxt
    # Forces over the 40x8 wall on a perturbed pressure field.
    mesh = ref_mesh(ref, "40x8")
    N = mesh["num_cells"]
    p = (101325.0 * (1.0 + 0.02 * np.sin(3.0 * mesh["cell_center"].reshape(-1, 3)[:, 0]))).astype(np.float32)
    ref7 = np.array([1.225, 101325.0, 2.0, 100.0, 1.0, 0.25, 0.0], np.float32)
    out["forces_p"], out["forces_ref7"] = p, ref7
    for sub in (0, 1):
        o = np.zeros(8, np.float32)
        ref.ref_forces(C.byref(mesh_cfg("40x8")), f32p(p), f32p(ref7), sub, f32p(o))
        out[f"forces_{sub}"] = o
    return out


DISSIPATION_REGIONS = ("wall_adjacent_interior", "farfield_boundary")


def post_csvs(out_dir: Path) -> tuple[np.ndarray, np.ndarray]:
    """The reference run's dissipation_debug.csv rows as [face, region (index in DISSIPATION_REGIONS), x, y,
    area, mach_local, un_abs, a, beta, low_mach_scale, acoustic_scale, s_stab, acoustic_scale_final,
    stabilization_m0, all_speed_pressure_scale, smax_physical, smax_dissipation] and its wall_flux.csv rows
    as [face, x, y, mass_flux_n] (euler_solver.cpp:1304-1465), as printed (six significant digits)."""
    rows = []
    lines = (out_dir / "dissipation_debug.csv").read_text().splitlines()
    for line in lines[1:]:
        f = line.split(",")
        rows.append([float(f[0]), float(DISSIPATION_REGIONS.index(f[1]))] + [float(x) for x in f[2:]])
    dissipation = np.array(rows, np.float64).reshape(-1, 17)
    lines = (out_dir / "wall_flux.csv").read_text().splitlines()
    wall_flux = np.array([[float(x) for x in line.split(",")] for line in lines[1:]], np.float64).reshape(-1, 4)
    return dissipation, wall_flux


def runs(ref) -> dict:
    out = {}
    for name, run in cases.RUNS.items():
        cfg = full_case(run["cfg"])
        rc = RefCaseCfg(*[cfg[k] for k, _ in _CASE_NAMES])
        mesh = ref_mesh(ref, run["mesh"])
        N = mesh["num_cells"]
        cap = cfg["iterations"]
        hist = np.zeros((cap, 7), np.float32)
        U = np.zeros((N, 5), np.float32)
        R = np.zeros((N, 5), np.float32)
        lam = np.zeros(N, np.float32)
        forces = np.zeros(8, np.float32)
        cp = np.zeros(mesh["wall_count"], np.float32)
        cpo = np.zeros((mesh["wall_count"], 4), np.float32)
        ncpo = C.c_int()
        if run["env_precond"]:
            os.environ["AERO_ENABLE_EXPERIMENTAL_PRECOND"] = "1"
        else:
            os.environ.pop("AERO_ENABLE_EXPERIMENTAL_PRECOND", None)
        with tempfile.TemporaryDirectory() as td:
            rows = ref.ref_run_case(C.byref(mesh_cfg(run["mesh"])), C.byref(rc), td.encode(), cap, f32p(hist),
                                    f32p(U), f32p(R), f32p(lam), f32p(forces), f32p(cp), f32p(cpo), C.byref(ncpo))
            post = np.zeros(23, np.float32)
            ref.ref_last_run_post(f32p(post))
            dissipation, wall_flux = post_csvs(Path(td))
            diagnostics = np.array((Path(td) / "euler_diagnostics.csv").read_text().splitlines())
        os.environ.pop("AERO_ENABLE_EXPERIMENTAL_PRECOND", None)
        out[f"run_{name}_hist"] = hist[:rows]
        out[f"run_{name}_U"] = U
        out[f"run_{name}_R"] = R
        out[f"run_{name}_lam"] = lam
        out[f"run_{name}_forces"] = forces
        out[f"run_{name}_cp"] = cp
        out[f"run_{name}_cp_ordered"] = cpo[: ncpo.value]
        out[f"run_{name}_post"] = post
        out[f"run_{name}_dissipation"] = dissipation
        out[f"run_{name}_wall_flux"] = wall_flux
        out[f"run_{name}_diagnostics"] = diagnostics  # euler_diagnostics.csv as written: header, then one line a row
        print(f"  run {name}: {rows} iterations, cl={forces[5]:.6f} cd={forces[6]:.6f}")
    return out


def main() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    ref = C.CDLL(str(REF_LIB))
    ref.ref_run_case.restype = C.c_int
    meshes = {}
    for name in cases.MESHES:
        a = ref_mesh(ref, name)
        meshes[f"digest_{name}"] = np.array(mesh_digest(a))
        meshes[f"sizes_{name}"] = np.array([a["num_cells"], a["num_faces"], a["wall_start"], a["wall_count"],
                                            a["farfield_start"], a["farfield_count"]], np.int64)
        if name in ("40x8",):
            for k in ("owner", "neighbor", "bc", "face_vertices", "normal", "face_center", "area", "volume",
                      "cell_center", "points"):
                meshes[f"{name}_{k}"] = a[k]
    np.savez_compressed(HERE / "golden_mesh.npz", **meshes)
    print("pointwise ...")
    np.savez_compressed(HERE / "golden_pointwise.npz", **pointwise(ref))
    print("residuals ...")
    np.savez_compressed(HERE / "golden_residual.npz", **residuals(ref))
    print("runs ...")
    np.savez_compressed(HERE / "golden_runs.npz", **runs(ref))
    for p in sorted(HERE.glob("golden_*.npz")):
        print(f"{p.name}: {p.stat().st_size / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
