"""Run a case file on the HIP backend and write the reference runner's outputs.

`run_case_file(case_path, out_dir)` plays the role of the native `run_case(case_path, out_dir, backend)`
behind the reference's pybind module (bindings/pybind_module.cpp:18-38, numerics/residual_stub.cpp:395-785)
for euler_airfoil_2d cases: it builds the case's O-grid with the reference's generator (FP32 point set,
so the mesh is the reference's mesh), runs the driver on the device and writes, in the reference's
column layouts,

  residuals.csv   iter,residual_l1,residual_l2,residual_linf,residual_ratio  (euler_solver.cpp:722, :1161)
  forces.csv      iter,cl,cd,cm,lift,drag,moment,dcl,dcd,dcm,mean_dcl,mean_dcd,mean_dcm,
                  force_window_stable,force_mean_drift_stable,force_physical_converged  (:723-724, :1163-1169)
  cp_wall.csv     s,x,y,Cp over the wall faces ordered from the trailing edge over the upper surface
                  (post/forces.cpp:39-129, :177-206; euler_solver.cpp:1467-1476)
  field_0000.vtu  cell rho, u, v, p, Mach, residual_magnitude (io/io_vtk_stub.cpp:95-166)
  wall_flux.csv   face,x,y,mass_flux_n over the slip-wall faces (euler_solver.cpp:1304-1331)
  dissipation_debug.csv  the scaling record of ten wall-adjacent interior and ten farfield faces
                  (:1333-1465)
  run.log         the run summary lines (residual_stub.cpp:600-762, the subset this backend knows, with
                  the beta / local-Mach / acoustic-scale statistics of the final state)

and returns the same summary dict as the pybind run_case.  Numbers are printed like a C++ ostream's
default float format (six significant digits).
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np

from . import _lib
from .case_file import euler_case_from_kv, load_case
from .mesh import Mesh, generate_mesh, load_mesh
from .solver import CaseConfig, HipMesh

VERSION = "0.1.1-hip"


def _g(v) -> str:
    """A C++ ostream's default formatting of a float (precision 6, %g-style)."""
    v = float(v)
    if math.isnan(v):
        return "nan" if math.copysign(1.0, v) > 0 else "-nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return f"{v:.6g}"


def force_monitor(cl, cd, cm, window: int, tol: float, drift_tol: float) -> dict:
    """Per-iteration force columns of ConvergenceMonitor (solvers/euler/convergence_monitor.cpp:21-117):
    change from the previous iteration, stability of the last `window` values (max - min <= tol each),
    mean drift between the last two windows and its test (<= drift_tol; tol when drift_tol <= 0).  The
    arithmetic is the monitor's, in the coefficients' own precision: window sums accumulated in order,
    means as sum * (1 / window)."""
    F = np.stack([np.asarray(cl), np.asarray(cd), np.asarray(cm)], axis=1)
    if F.dtype not in (np.float32, np.float64):
        F = F.astype(np.float64)
    t = F.dtype.type
    n = len(F)
    w = max(int(window), 1)
    tol_t = t(tol)
    dtol = t(drift_tol) if drift_tol > 0.0 else tol_t
    d = np.zeros((n, 3), F.dtype)
    if n > 1:
        d[1:] = np.abs(F[1:] - F[:-1])
    md = np.zeros((n, 3), F.dtype)
    stable = np.zeros(n, bool)
    mstable = np.zeros(n, bool)
    if n >= w:
        win = np.lib.stride_tricks.sliding_window_view(F, w, axis=0)  # [n - w + 1, 3, w]: F[e+1-w .. e]
        stable[w - 1:] = np.all(win.max(axis=-1) - win.min(axis=-1) <= tol_t, axis=1)
    if n >= 2 * w:
        ends = np.arange(2 * w - 1, n)
        prev = np.zeros((len(ends), 3), F.dtype)
        curr = np.zeros((len(ends), 3), F.dtype)
        for k in range(w):  # in order, oldest first
            prev = prev + F[ends + 1 - 2 * w + k]
            curr = curr + F[ends + 1 - w + k]
        inv = t(1.0) / t(w)
        md[ends] = np.abs(curr * inv - prev * inv)
        mstable[ends] = np.all(md[ends] <= dtol, axis=1)
    return dict(d=d, mean_d=md, stable=stable, mean_stable=mstable, converged=stable & mstable)


def order_wall_faces(mesh: Mesh) -> np.ndarray:
    """Wall faces in the order of order_wall_faces (post/forces.cpp:39-129): from the wall vertex of
    largest x (ties within 1e-8: largest y) towards its neighbour of larger y, then along the chain;
    by descending atan2(y, x) of the face centres when the chain does not close."""
    faces = np.arange(mesh.wall_start, mesh.wall_start + mesh.wall_count)
    if len(faces) <= 2:
        return faces
    fv = mesh.face_vertices
    pts = mesh.points
    edge = {}
    nbrs: dict[int, list[int]] = {}
    for f in faces:
        a, b = int(fv[f, 0]), int(fv[f, 1])
        edge[(min(a, b), max(a, b))] = int(f)
        nbrs.setdefault(a, []).append(b)
        nbrs.setdefault(b, []).append(a)
    start, bx, by = -1, -1.0e30, -1.0e30
    for v in sorted(nbrs):
        x, y = float(pts[v, 0]), float(pts[v, 1])
        if x > bx or (abs(x - bx) < 1.0e-8 and y > by):
            start, bx, by = v, x, y
    first = nbrs.get(start, [])
    if start < 0 or len(first) < 2:
        return faces
    nxt = first[0]
    if pts[first[1], 1] - pts[start, 1] > pts[first[0], 1] - pts[start, 1]:
        nxt = first[1]
    out = []
    prev, cur = start, nxt
    for _ in range(len(faces)):
        f = edge.get((min(prev, cur), max(prev, cur)))
        if f is None:
            out = []
            break
        out.append(f)
        cand = [c for c in nbrs[cur] if c != prev]
        if not cand:
            out = []
            break
        prev, cur = cur, cand[0]
        if prev == start:
            break
    if len(out) != len(faces):
        ang = np.arctan2(mesh.face_center[faces, 1], mesh.face_center[faces, 0])
        return faces[np.argsort(-ang, kind="stable")]
    return np.asarray(out)


def _primitives(U: np.ndarray, gamma: float):
    rho = np.maximum(U[:, 0], 1.0e-8)
    u, v = U[:, 1] / rho, U[:, 2] / rho
    p = np.maximum((gamma - 1.0) * (U[:, 4] - 0.5 * rho * (u * u + v * v)), 1.0e-8)
    a = np.sqrt(gamma * np.maximum(p, 1.0e-8) / rho)
    mach = np.sqrt(u * u + v * v) / np.maximum(a, 1.0e-8)
    return rho, u, v, p, mach


def write_vtu(path: Path, mesh: Mesh, U: np.ndarray, R: np.ndarray, gamma: float) -> None:
    rho, u, v, p, mach = _primitives(U, gamma)
    inv = 1.0 / np.maximum(mesh.volume.astype(float), 1.0e-12)
    rmag = np.sqrt(np.sum((R * inv[:, None]) ** 2, axis=1))
    offs = np.asarray(mesh.cell_offsets)
    sizes = np.diff(np.concatenate([[0], offs]))
    types = np.where(sizes == 3, 5, np.where(sizes == 4, 9, 7))  # triangle, quad, polygon
    arr = lambda a: " ".join(_g(x) for x in np.ravel(a))
    iarr = lambda a: " ".join(str(int(x)) for x in np.ravel(a))
    with open(path, "w", encoding="utf-8") as o:
        o.write('<?xml version="1.0"?>\n')
        o.write('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
        o.write("  <UnstructuredGrid>\n")
        o.write(f'    <Piece NumberOfPoints="{len(mesh.points)}" NumberOfCells="{mesh.num_cells}">\n')
        o.write("      <Points>\n")
        o.write(f'        <DataArray type="Float32" NumberOfComponents="3" format="ascii">{arr(mesh.points)}</DataArray>\n')
        o.write("      </Points>\n      <Cells>\n")
        o.write(f'        <DataArray type="Int32" Name="connectivity" format="ascii">{iarr(mesh.cell_connectivity)}</DataArray>\n')
        o.write(f'        <DataArray type="Int32" Name="offsets" format="ascii">{iarr(offs)}</DataArray>\n')
        o.write(f'        <DataArray type="UInt8" Name="types" format="ascii">{iarr(types)}</DataArray>\n')
        o.write("      </Cells>\n      <PointData/>\n")
        o.write('      <CellData Scalars="rho">\n')
        for name, a in (("rho", rho), ("u", u), ("v", v), ("p", p), ("Mach", mach), ("residual_magnitude", rmag)):
            o.write(f'        <DataArray type="Float32" Name="{name}" format="ascii">{arr(a)}</DataArray>\n')
        o.write("      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n")


def write_outputs(out_dir: Path, mesh: Mesh, case: CaseConfig, out: dict) -> dict:
    """Write residuals.csv, forces.csv, cp_wall.csv and field_0000.vtu of one run; returns the paths."""
    out_dir.mkdir(parents=True, exist_ok=True)
    H = out["history"]
    rho_inf = case.rho_inf if case.rho_inf > 0.0 else case.p_inf / (case.gas_constant * case.t_inf)
    V = max(case.mach, 0.0) * math.sqrt(case.gamma * case.p_inf / max(rho_inf, 1.0e-12))
    q = 0.5 * rho_inf * V * V + 1.0e-12
    chord = max(case.chord, 1.0e-8)
    s_ref = chord * 1.0
    paths = dict(residuals=out_dir / "residuals.csv", forces=out_dir / "forces.csv", cp=out_dir / "cp_wall.csv",
                 vtu=out_dir / "field_0000.vtu")
    with open(paths["residuals"], "w", encoding="utf-8") as f:
        f.write("iter,residual_l1,residual_l2,residual_linf,residual_ratio\n")
        for r in H:
            f.write(f"{int(r[_lib.H_ITER])},{_g(r[_lib.H_L1])},{_g(r[_lib.H_L2])},{_g(r[_lib.H_LINF])},"
                    f"{_g(r[_lib.H_RATIO])}\n")
    cl, cd, cm = H[:, _lib.H_CL], H[:, _lib.H_CD], H[:, _lib.H_CM]
    mon = force_monitor(cl, cd, cm, case.force_stability_window, case.force_stability_tol, case.force_mean_drift_tol)
    with open(paths["forces"], "w", encoding="utf-8") as f:
        f.write("iter,cl,cd,cm,lift,drag,moment,dcl,dcd,dcm,mean_dcl,mean_dcd,mean_dcm,"
                "force_window_stable,force_mean_drift_stable,force_physical_converged\n")
        for i, r in enumerate(H):
            vals = [cl[i], cd[i], cm[i], cl[i] * q * s_ref, cd[i] * q * s_ref, cm[i] * q * s_ref * chord,
                    *mon["d"][i], *mon["mean_d"][i]]
            f.write(f"{int(r[_lib.H_ITER])}," + ",".join(_g(x) for x in vals) +
                    f",{int(mon['stable'][i])},{int(mon['mean_stable'][i])},{int(mon['converged'][i])}\n")
    order = order_wall_faces(mesh)
    cp = np.asarray(out["wall_cp"])
    with open(paths["cp"], "w", encoding="utf-8") as f:
        f.write("s,x,y,Cp\n")
        s = 0.0
        for face in order:
            f.write(f"{_g(s)},{_g(mesh.face_center[face, 0])},{_g(mesh.face_center[face, 1])},"
                    f"{_g(cp[face - mesh.wall_start])}\n")
            s += float(mesh.area[face])
    write_vtu(paths["vtu"], mesh, out["conserved"], out["residual"], case.gamma)
    return paths


DIAGNOSTICS_HEADER = (
    "iter,residual_ratio,force_window_stable,integrated_wall_faces,wall_patch_faces,wall_patch_bc_mismatch,"
    "sum_wall_ds,sum_nA_x,sum_nA_y,sum_nA_mag,max_wall_un,max_wall_mass_flux_n,max_farfield_abs_dp,"
    "max_farfield_abs_drho,max_farfield_abs_du,max_farfield_abs_dv,max_farfield_abs_dw,wall_cp_min,wall_cp_max,"
    "wall_cp_nonfinite,q_inf,rho_inf,V_inf,S_ref,chord,Fx_gauge,Fy_gauge,Fx_abs,Fy_abs,L,D,Cl,Cd,Cm,cl_abs,cd_abs,"
    "cm_abs,dcl,dcd,dcm,force_mean_drift_stable,force_physical_converged,mean_dcl,mean_dcd,mean_dcm,cd_sign_flip,"
    "p_jump_energy_nd,p_jump_energy_wall_nd,all_speed_stage,all_speed_flux_fix_effective,all_speed_f_min_effective,"
    "all_speed_ramp_weight_effective,all_speed_freeze_counter,all_speed_controller_action,cfl_scale_effective,"
    "cd_times_mach,p_fluct_nd_max,p_fluct_nd_wall,p_fluct_nd_farfield,num_faces_first_order_fallback,"
    "num_faces_diffusive_fallback,first_failure_face,first_failure_owner,first_failure_neighbor")
STAGE_NAMES = ("single", "A", "B")


def diagnostics_rows(mesh: Mesh, case: CaseConfig, out: dict) -> list[list]:
    """The rows of euler_diagnostics.csv (euler_solver.cpp:1161-1203) from a run's history, its per-iteration
    diagnostics (aero_run_output.diagnostics) and the mesh: one list of values per iteration, in the
    header's order (all_speed_stage as the reference's stage name).  The wall-patch sums run in patch
    order in the mesh's precision (summarize_face_geometry, compute_pressure_force_diagnostics:
    euler_solver.cpp:197-228, post/forces.cpp:208-246)."""
    H, D = out["history"], out["diagnostics"]
    rho_inf = rho_inf_of(case)
    V = speed_inf_of(case)
    q = 0.5 * rho_inf * V * V + 1.0e-12
    chord = max(case.chord, 1.0e-8)
    wall_patch = np.arange(mesh.wall_start, mesh.wall_start + mesh.wall_count)
    wall_bc = np.flatnonzero(mesh.bc == _lib.BC_SLIP_WALL)
    mismatch = len(np.setxor1d(wall_patch, wall_bc))
    area = mesh.area[wall_patch]
    sum_ds = float(np.cumsum(area)[-1]) if len(area) else 0.0
    nA = -mesh.normal[wall_patch, :2] * area[:, None]  # face_nA: the body-outward normal times the area
    nax = float(np.cumsum(nA[:, 0])[-1]) if len(area) else 0.0
    nay = float(np.cumsum(nA[:, 1])[-1]) if len(area) else 0.0
    nam = float(np.sqrt(np.asarray(nax * nax + nay * nay, area.dtype)))
    cl, cd, cm = H[:, _lib.H_CL], H[:, _lib.H_CD], H[:, _lib.H_CM]
    mon = force_monitor(cl, cd, cm, case.force_stability_window, case.force_stability_tol, case.force_mean_drift_tol)
    col = {n: i for i, n in enumerate(_lib.DIAG_COLUMNS)}
    mach = H.dtype.type(max(case.mach, 0.0))
    rows = []
    for i, (h, d) in enumerate(zip(H, D)):
        ff = int(d[col["first_failure_face"]])
        flip = i > 0 and ((cd[i - 1] >= 0.0) != (cd[i] >= 0.0))
        rows.append([
            int(h[_lib.H_ITER]), h[_lib.H_RATIO], int(mon["stable"][i]), mesh.wall_count, mesh.wall_count, mismatch,
            sum_ds, nax, nay, nam, d[col["max_wall_un"]], 0.0,  # the slip-wall flux carries no mass
            d[col["max_farfield_abs_dp"]], d[col["max_farfield_abs_drho"]], d[col["max_farfield_abs_du"]],
            d[col["max_farfield_abs_dv"]], 0.0, d[col["wall_cp_min"]], d[col["wall_cp_max"]],
            int(d[col["wall_cp_nonfinite"]]), q, rho_inf, V, chord, chord, d[col["Fx_gauge"]], d[col["Fy_gauge"]],
            d[col["Fx_abs"]], d[col["Fy_abs"]], d[col["L"]], d[col["D"]], cl[i], cd[i], cm[i], d[col["cl_abs"]],
            d[col["cd_abs"]], d[col["cm_abs"]], *mon["d"][i], int(mon["mean_stable"][i]), int(mon["converged"][i]),
            *mon["mean_d"][i], int(flip), d[col["p_jump_energy_nd"]], d[col["p_jump_energy_wall_nd"]],
            STAGE_NAMES[int(d[col["all_speed_stage"]])], int(d[col["all_speed_flux_fix_effective"]]),
            d[col["all_speed_f_min_effective"]], d[col["all_speed_ramp_weight_effective"]],
            int(d[col["all_speed_freeze_counter"]]), int(d[col["all_speed_controller_action"]]),
            d[col["cfl_scale_effective"]], cd[i] * mach, d[col["p_fluct_nd_max"]],
            d[col["p_fluct_nd_wall"]], d[col["p_fluct_nd_farfield"]], int(h[_lib.H_FO]),
            int(h[_lib.H_DIFF]), ff, int(mesh.owner[ff]) if ff >= 0 else -1,
            int(mesh.neighbor[ff]) if ff >= 0 else -1])
    return rows


def diagnostics_lines(mesh: Mesh, case: CaseConfig, out: dict) -> list[str]:
    """euler_diagnostics.csv as lines (euler_solver.cpp:717-738 header, :1161-1203 rows)."""
    fmt = lambda v: v if isinstance(v, str) else (str(v) if isinstance(v, (int, np.integer)) else _g(v))
    return [DIAGNOSTICS_HEADER] + [",".join(fmt(v) for v in r) for r in diagnostics_rows(mesh, case, out)]


def write_diagnostics(path: Path, mesh: Mesh, case: CaseConfig, out: dict) -> None:
    """Write euler_diagnostics.csv of a run."""
    with open(path, "w", encoding="utf-8") as f:
        f.write("\n".join(diagnostics_lines(mesh, case, out)) + "\n")


DISSIPATION_REGIONS = ("wall_adjacent_interior", "farfield_boundary")


def _quantile_sorted(v: np.ndarray, q: float) -> float:
    """linear_quantile_from_sorted (euler_solver.cpp:380-390)."""
    if v.size == 0:
        return 1.0
    idx = min(max(q, 0.0), 1.0) * (v.size - 1)
    i0 = int(math.floor(idx))
    i1 = min(i0 + 1, v.size - 1)
    t = idx - i0
    return float((1.0 - t) * v[i0] + t * v[i1])


def _wall_adjacent_faces(mesh: Mesh) -> np.ndarray:
    """Interior faces with a cell that owns a slip-wall face, in face order (euler_solver.cpp:1334-1355)."""
    wall = np.flatnonzero(mesh.bc == _lib.BC_SLIP_WALL)
    mask = np.zeros(mesh.num_cells, bool)
    mask[mesh.owner[wall]] = True
    nb = mesh.neighbor
    interior = nb >= 0
    touch = np.zeros(mesh.num_faces, bool)
    touch[interior] = mask[mesh.owner[interior]] | mask[nb[interior]]
    return np.flatnonzero(touch)


def scaling_statistics(mesh: Mesh, fs: np.ndarray, cs: np.ndarray) -> dict:
    """run.log's beta, local-Mach and acoustic-scale statistics of a state (EulerRunResult fields,
    euler_solver.cpp:1217-1402) from aero_hip_face_scales records: fs [FS_FIELDS][F], cs [2][N]."""
    st = {}
    beta, mach = cs[0], cs[1]
    if mesh.num_cells > 0:
        st.update(beta_min=beta.min(), beta_max=beta.max(), beta_mean=beta.mean(), mach_local_min=mach.min(),
                  mach_local_max=mach.max(), mach_local_mean=mach.mean())
    else:
        st.update(beta_min=1.0, beta_max=1.0, beta_mean=1.0, mach_local_min=0.0, mach_local_max=0.0,
                  mach_local_mean=0.0)
    base, final = fs[_lib.FS_ACOUSTIC_BASE], fs[_lib.FS_ACOUSTIC]
    for key, v in (("acoustic_scale", base), ("acoustic_scale_final", final)):
        st.update({f"{key}_min": v.min() if v.size else 1.0, f"{key}_max": v.max() if v.size else 1.0,
                   f"{key}_mean": v.mean() if v.size else 1.0})
    wa = _wall_adjacent_faces(mesh)
    for key, col in (("wall_adjacent_acoustic_scale", _lib.FS_ACOUSTIC_BASE),
                     ("wall_adjacent_stabilization_scale", _lib.FS_STAB),
                     ("wall_adjacent_acoustic_scale_final", _lib.FS_ACOUSTIC)):
        v = np.sort(fs[col, wa])
        st.update({f"{key}_min": float(v[0]) if v.size else 1.0, f"{key}_p01": _quantile_sorted(v, 0.01),
                   f"{key}_p50": _quantile_sorted(v, 0.50)})
    wall = mesh.bc == _lib.BC_SLIP_WALL
    st["max_wall_mass_flux"] = float(np.abs(fs[_lib.FS_WALL_MASS_FLUX, wall]).max()) if wall.any() else 0.0
    return {k: float(v) for k, v in st.items()}


def dissipation_faces(mesh: Mesh) -> list[tuple[int, str]]:
    """The faces of dissipation_debug.csv: the first ten wall-adjacent interior faces, then the first ten
    farfield faces, in face order (euler_solver.cpp:1445-1463)."""
    far = np.flatnonzero(mesh.bc == _lib.BC_FARFIELD)
    return ([(int(f), DISSIPATION_REGIONS[0]) for f in _wall_adjacent_faces(mesh)[:10]] +
            [(int(f), DISSIPATION_REGIONS[1]) for f in far[:10]])


def write_post_files(out_dir: Path, mesh: Mesh, cfg: _lib.ResidualConfig, fs: np.ndarray) -> dict:
    """wall_flux.csv (euler_solver.cpp:1304-1331) and dissipation_debug.csv (:1406-1463) from the
    aero_hip_face_scales records fs of the final state; returns the paths."""
    paths = dict(wall_flux=out_dir / "wall_flux.csv", dissipation=out_dir / "dissipation_debug.csv")
    with open(paths["wall_flux"], "w", encoding="utf-8") as f:
        f.write("face,x,y,mass_flux_n\n")
        for face in np.flatnonzero(mesh.bc == _lib.BC_SLIP_WALL):
            f.write(f"{face},{_g(mesh.face_center[face, 0])},{_g(mesh.face_center[face, 1])},"
                    f"{_g(fs[_lib.FS_WALL_MASS_FLUX, face])}\n")
    m0 = min(max(cfg.stab_mach_floor, 0.0), 1.0)
    cols = (_lib.FS_MACH, _lib.FS_UN_ABS, _lib.FS_SOUND, _lib.FS_BETA, _lib.FS_LOW_MACH, _lib.FS_ACOUSTIC_BASE,
            _lib.FS_STAB, _lib.FS_ACOUSTIC)
    with open(paths["dissipation"], "w", encoding="utf-8") as f:
        f.write("face,region,x,y,area,mach_local,un_abs,a,beta,low_mach_scale,acoustic_scale,s_stab,"
                "acoustic_scale_final,stabilization_m0,all_speed_pressure_scale,smax_physical,smax_dissipation\n")
        for face, region in dissipation_faces(mesh):
            vals = [mesh.face_center[face, 0], mesh.face_center[face, 1], mesh.area[face],
                    *(fs[c, face] for c in cols), m0, fs[_lib.FS_ALL_SPEED, face], fs[_lib.FS_SMAX_PHYSICAL, face],
                    fs[_lib.FS_SMAX_DISSIPATION, face]]
            f.write(f"{face},{region}," + ",".join(_g(v) for v in vals) + "\n")
    return paths


def case_and_mesh(kv: dict, mesh_path=None) -> tuple[CaseConfig, Mesh]:
    """CaseConfig and the reference generator's O-grid (FP32 point set, FP64 arrays) of native keys, or
    the mesh of a binary mesh file (save_mesh) when mesh_path is given."""
    case_type = kv.get("case_type", "scalar_advect_demo")
    if case_type != "euler_airfoil_2d":
        raise ValueError(f"the HIP backend runs euler_airfoil_2d cases (case_type={case_type})")
    case, ms = euler_case_from_kv(kv)
    if mesh_path is not None:
        return case, load_mesh(mesh_path).astype(np.float64)
    if ms["airfoil_source"].lower() != "naca4":
        raise ValueError("the HIP backend's mesh generator builds NACA 4-digit airfoils (airfoil_source=naca4)")
    mesh = generate_mesh(naca=ms["naca"], num_circumferential=ms["num_circumferential"], num_radial=ms["num_radial"],
                         farfield_radius=ms["farfield_radius"], radial_stretch=ms["radial_stretch"],
                         chord=ms["chord"], cells="tri", precision=4).astype(np.float64)
    return case, mesh


def run_case_file(case_path: str | Path, out_dir: str | Path, block_cells: int = 256,
                  overrides: dict | None = None, mesh_path=None) -> dict:
    """Run an euler_airfoil_2d case file (native key=value or YAML) on the device and write the reference
    runner's outputs into out_dir; returns the pybind run_case summary dict.  overrides: native keys
    replacing the file's (e.g. {"aoa_deg": "4"}); mesh_path: a binary mesh file to run on instead of the
    generated O-grid."""
    kv = load_case(case_path)
    kv.update({k: str(v) for k, v in (overrides or {}).items()})
    case_type = kv.get("case_type", "scalar_advect_demo")
    case, mesh = case_and_mesh(kv, mesh_path)
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    with HipMesh(mesh, block_cells=block_cells) as hm:
        out = hm.run_case(case, diagnostics=True)
        if out["nonfinite"]:
            raise RuntimeError("Euler pseudo-time iteration produced a non-finite state.")
        fs, cs = hm.face_scales(out["final_config"], out["conserved"])
        fa = hm.forces(rho_inf_of(case), case.p_inf, case.aoa_deg, speed_inf_of(case), case.chord, case.x_ref,
                       case.y_ref, gauge=False, gamma=case.gamma, U=out["conserved"])
    write_outputs(out_dir, mesh, case, out)
    write_diagnostics(out_dir / "euler_diagnostics.csv", mesh, case, out)
    post = write_post_files(out_dir, mesh, out["final_config"], fs)
    stats = scaling_statistics(mesh, fs, cs)
    H = out["history"]
    last = H[-1]
    summary = dict(status="ok", backend="hip", case_type=case_type, run_log=str(out_dir / "run.log"),
                   iterations=int(len(H)), residual_l1=float(last[_lib.H_L1]), residual_l2=float(last[_lib.H_L2]),
                   residual_linf=float(last[_lib.H_LINF]), cl=float(out["cl"]), cd=float(out["cd"]),
                   cm=float(out["cm"]))
    rho_inf = case.rho_inf if case.rho_inf > 0.0 else case.p_inf / (case.gas_constant * case.t_inf)
    V = max(case.mach, 0.0) * math.sqrt(case.gamma * case.p_inf / max(rho_inf, 1.0e-12))
    a = math.radians(case.aoa_deg)
    f = out["forces"]
    onoff = lambda v: "on" if v else "off"
    c = case
    wall = np.flatnonzero(mesh.bc == _lib.BC_SLIP_WALL)
    nax = float(np.sum(-mesh.normal[wall, 0] * mesh.area[wall]))  # face_nA = -n A (forces.cpp:233-238)
    nay = float(np.sum(-mesh.normal[wall, 1] * mesh.area[wall]))
    mon = force_monitor(H[:, _lib.H_CL], H[:, _lib.H_CD], H[:, _lib.H_CM], c.force_stability_window,
                        c.force_stability_tol, c.force_mean_drift_tol)
    # residual_stub.cpp:609-762, in its order; the keys this backend has values for
    lines = ["AeroCFD Euler airfoil 2D", f"version={VERSION}", f"case_path={case_path}", "backend=hip",
             f"case_type={case_type}", f"num_cells={mesh.num_cells}", f"num_faces={mesh.num_faces}",
             f"iterations={len(H)}", f"residual_l1={_g(summary['residual_l1'])}",
             f"residual_l2={_g(summary['residual_l2'])}", f"residual_linf={_g(summary['residual_linf'])}",
             f"cl={_g(f[5])}", f"cd={_g(f[6])}", f"cm={_g(f[7])}", f"aoa_deg={_g(c.aoa_deg)}",
             f"aoa_rad={_g(a)}", f"u_inf={_g(V * math.cos(a))}", f"v_inf={_g(V * math.sin(a))}",
             f"wall_faces_integrated={len(wall)}", "normal_is_unit=1", f"sum_nA_x={_g(nax)}", f"sum_nA_y={_g(nay)}",
             f"sum_nA_mag={_g(math.hypot(nax, nay))}", f"Fx_abs={_g(fa['fx'])}", f"Fy_abs={_g(fa['fy'])}",
             f"Fx_gauge={_g(f[0])}", f"Fy_gauge={_g(f[1])}", f"q_inf={_g(0.5 * rho_inf * V * V + 1e-12)}",
             f"V_inf={_g(V)}", f"rho_inf={_g(rho_inf)}", f"S_ref={_g(max(c.chord, 1e-8))}",
             f"chord={_g(max(c.chord, 1e-8))}", f"Fx={_g(f[0])}", f"Fy={_g(f[1])}", f"L={_g(f[2])}",
             f"D={_g(f[3])}", f"cl_abs={_g(fa['cl'])}", f"cd_abs={_g(fa['cd'])}", f"cm_abs={_g(fa['cm'])}",
             f"flux={'hllc' if c.flux_scheme else 'rusanov'}", f"limiter={'venkat' if c.limiter else 'minmod'}",
             f"low_mach_fix={onoff(c.low_mach_fix)}", f"mach_cutoff={_g(c.mach_cutoff)}",
             f"all_speed_flux_fix={onoff(c.all_speed_flux_fix)}", f"all_speed_mach_cutoff={_g(c.all_speed_mach_cutoff)}",
             f"all_speed_f_min={_g(c.all_speed_f_min)}", f"all_speed_ramp_start_iter={c.all_speed_ramp_start_iter}",
             f"all_speed_ramp_iters={c.all_speed_ramp_iters}",
             f"all_speed_staged_controller={onoff(c.all_speed_staged_controller)}",
             f"all_speed_stage_a_min_iters={c.all_speed_stage_a_min_iters}",
             f"all_speed_stage_a_max_iters={c.all_speed_stage_a_max_iters}",
             f"all_speed_stage_f_min_high={_g(c.all_speed_stage_f_min_high)}",
             f"all_speed_stage_f_min_mid={_g(c.all_speed_stage_f_min_mid)}",
             f"all_speed_stage_f_min_low={_g(c.all_speed_stage_f_min_low)}",
             f"all_speed_pjump_wall_target={_g(c.all_speed_pjump_wall_target)}",
             f"all_speed_pjump_spike_factor={_g(c.all_speed_pjump_spike_factor)}",
             f"all_speed_pjump_hold_iters={c.all_speed_pjump_hold_iters}",
             f"all_speed_freeze_iters={c.all_speed_freeze_iters}",
             f"all_speed_cfl_drop_factor={_g(c.all_speed_cfl_drop_factor)}",
             f"all_speed_cfl_restore_factor={_g(c.all_speed_cfl_restore_factor)}",
             f"all_speed_cfl_min_scale={_g(c.all_speed_cfl_min_scale)}",
             f"aoa0_symmetry_enforce={onoff(c.aoa0_symmetry_enforce)}",
             f"aoa0_symmetry_enforce_interval={c.aoa0_symmetry_enforce_interval}",
             f"precond={onoff(c.precond_on)}", f"precond_mach_ref={_g(c.precond_mach_ref)}",
             f"precond_mach_min={_g(c.precond_mach_min)}", f"precond_beta_min={_g(c.precond_beta_min)}",
             f"precond_beta_max={_g(c.precond_beta_max)}", f"precond_farfield_bc={onoff(c.precond_farfield_bc)}",
             f"stabilization_mach_floor_start={_g(c.stab_floor_start)}",
             f"stabilization_mach_floor_target={_g(c.stab_floor_target)}",
             f"stabilization_mach_floor_k_start={_g(c.stab_floor_k_start)}",
             f"stabilization_mach_floor_k_target={_g(c.stab_floor_k_target)}",
             f"stabilization_ramp_iters={c.stab_ramp_iters}",
             f"stabilization_mach_floor_final={_g(out['final_config'].stab_mach_floor)}",
             *(f"{k}={_g(v)}" for k, v in stats.items() if k != "max_wall_mass_flux"),
             f"dissipation_debug_csv={post['dissipation']}",
             f"startup_first_order_iters={c.startup_first_order_iters}", f"rk_stages={c.rk_stages}",
             f"local_time_stepping={onoff(c.local_time_stepping)}",
             f"force_stability_window={c.force_stability_window}",
             f"force_mean_drift_tol={_g(c.force_mean_drift_tol)}",
             f"force_window_stable_final={int(mon['stable'][-1])}",
             f"force_mean_drift_stable_final={int(mon['mean_stable'][-1])}",
             f"force_physical_converged_final={int(mon['converged'][-1])}",
             f"force_mean_dcl_final={_g(mon['mean_d'][-1][0])}", f"force_mean_dcd_final={_g(mon['mean_d'][-1][1])}",
             f"force_mean_dcm_final={_g(mon['mean_d'][-1][2])}",
             f"device_seconds={out['seconds']:.6g}"]
    (out_dir / "run.log").write_text("\n".join(lines) + "\n", encoding="utf-8")
    return summary


def rho_inf_of(case: CaseConfig) -> float:
    """Freestream density: rho_inf, or p_inf / (R T_inf) when it is not set (euler_solver.cpp:543-549)."""
    return case.rho_inf if case.rho_inf > 0.0 else case.p_inf / (case.gas_constant * case.t_inf)


def speed_inf_of(case: CaseConfig) -> float:
    """Freestream speed M a_inf (euler_solver.cpp:551-553)."""
    return max(case.mach, 0.0) * math.sqrt(case.gamma * case.p_inf / max(rho_inf_of(case), 1.0e-12))


def wall_cp_from_state(mesh: Mesh, case: CaseConfig, U: np.ndarray) -> np.ndarray:
    """Cp of the wall faces (wall-patch order) from the owner cells' pressure (forces.cpp:177-206)."""
    rho_inf = case.rho_inf if case.rho_inf > 0.0 else case.p_inf / (case.gas_constant * case.t_inf)
    V = max(case.mach, 0.0) * math.sqrt(case.gamma * case.p_inf / max(rho_inf, 1.0e-12))
    q = 0.5 * rho_inf * V * V + 1.0e-12
    _, _, _, p, _ = _primitives(U, case.gamma)
    owners = mesh.owner[mesh.wall_start:mesh.wall_start + mesh.wall_count]
    return (p[owners] - case.p_inf) / q


def run_sweep_file(sweep_path: str | Path, out_dir: str | Path, batched: bool = False, block_cells: int = 256,
                   jobs: int = 1, progress=None, mesh_path=None) -> dict:
    """A sweep YAML (base_case: a case file relative to the sweep file; sweep: aoa_deg and optionally
    mach lists, crossed) as python/cfd/sweep.py:14-59 runs it, except that every run gets its own AoA
    and Mach.  Serial: one device run per case with the reference's outputs in run_000, run_001, ...
    Batched: all cases in one aero_hip_batch over the shared mesh, each stopping at a density-residual drop
    of residual_reduction_target or at `iterations`; per run cp_wall.csv and field_0000.vtu."""
    import yaml
    sp = Path(sweep_path)
    if not sp.exists():
        raise FileNotFoundError(f"Sweep file not found: {sp}")
    cfg = yaml.safe_load(sp.read_text(encoding="utf-8")) or {}
    if not isinstance(cfg, dict):
        raise ValueError("Sweep YAML must contain a top-level mapping.")
    if not cfg.get("base_case"):
        raise ValueError("Sweep YAML must define 'base_case'.")
    base_path = (sp.parent / str(cfg["base_case"])).resolve()
    kv = load_case(base_path)
    case0, mesh = case_and_mesh(kv, mesh_path)
    sweep = cfg.get("sweep", {}) or {}
    aoas = [float(a) for a in sweep.get("aoa_deg", [case0.aoa_deg])]
    machs = [float(m) for m in sweep.get("mach", [case0.mach])]
    points = [(m, a) for m in machs for a in aoas]
    root = Path(out_dir)
    root.mkdir(parents=True, exist_ok=True)
    runs = []
    if not batched:
        for i, (m, a) in enumerate(points):
            s = run_case_file(base_path, root / f"run_{i:03d}", block_cells=block_cells,
                              overrides={"aoa_deg": repr(a), "mach": repr(m)}, mesh_path=mesh_path)
            s.update(aoa_deg=a, mach=m)
            runs.append(s)
            if progress:
                progress(i + 1, len(points))
    else:
        from dataclasses import replace
        from .solver import HipBatch, case_linearization
        with HipMesh(mesh, block_cells=block_cells) as hm:
            with HipBatch([hm], [(0, m, a) for m, a in points], case0) as b:
                r = b.run(case0.iterations, case0.residual_reduction_target, iters_per_sync=200)
                states = [b.state(i) for i in range(len(points))]
            for i, ((m, a), row) in enumerate(zip(points, r["cases"])):
                d = root / f"run_{i:03d}"
                d.mkdir(parents=True, exist_ok=True)
                ci = replace(case0, mach=m, aoa_deg=a, iterations=max(row["iterations"], 1))
                rc, _, _, _ = case_linearization(ci)  # the residual configuration of the last iteration
                R, _, _ = hm.residual(states[i], rc)
                cp = wall_cp_from_state(mesh, ci, states[i])
                order = order_wall_faces(mesh)
                with open(d / "cp_wall.csv", "w", encoding="utf-8") as f:
                    f.write("s,x,y,Cp\n")
                    s_acc = 0.0
                    for face in order:
                        f.write(f"{_g(s_acc)},{_g(mesh.face_center[face, 0])},{_g(mesh.face_center[face, 1])},"
                                f"{_g(cp[face - mesh.wall_start])}\n")
                        s_acc += float(mesh.area[face])
                write_vtu(d / "field_0000.vtu", mesh, states[i], R, ci.gamma)
                runs.append(dict(status="ok", backend="hip", case_type="euler_airfoil_2d", aoa_deg=a, mach=m,
                                 iterations=row["iterations"], converged=row["converged"],
                                 rho_residual_ratio=row["rho_ratio"], residual_l2=row["l2"], cl=row["cl"],
                                 cd=row["cd"], cm=row["cm"]))
    return dict(status="ok", jobs_requested=int(jobs), output_root=str(root), batched=bool(batched), runs=runs)

