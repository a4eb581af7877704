"""Case-level sensitivities through the discrete adjoint (no reference counterpart).

For an objective J = w_cl CL + w_cd CD + w_cm CM of the converged state U (R(U, p) = 0):

    R_U^T psi = dJ/dU,      dJ/dp = dJ/dp|_U - psi . dR/dp|_U

for the angle of attack, the Mach number and any geometry direction (Hicks-Henne amplitudes through
``deform_mesh`` tangents).  One adjoint solve serves every parameter; ``tangent=True`` also solves the
forward problem R_U v = -dR/dp per parameter, whose dJ/dp = dJ/dp|_U + dJ/dU . v must agree.
``gmres_restart`` > 0 accelerates both solves with restarted GMRES on the pseudo-time fixed point, which
converges at limit-cycling second-order states where the plain iteration diverges.
"""
from __future__ import annotations

import time

import numpy as np

from dataclasses import replace

from .mesh import Mesh, deform_mesh
from .solver import CaseConfig, HipBatch, HipMesh, Seed, adjoint_solve_batch, case_linearization


def geometry_seed(tangents: dict, i: int) -> Seed:
    """The Seed of shape parameter i from ``deform_mesh`` tangents."""
    return Seed(d_face_normal=tangents["normal"][i], d_face_center=tangents["face_center"][i],
                d_face_area=tangents["area"][i], d_cell_center=tangents["cell_center"][i],
                d_cell_volume=tangents["volume"][i])


def sensitivities(hm: HipMesh, case: CaseConfig, U: np.ndarray, weights=(1.0, 0.0, 0.0), shape_tangents=None,
                  shape_params=None, cfl: float | None = None, max_iters: int = 20000, tol: float = 1e-10,
                  tangent: bool = False, gmres_restart: int = 0, gmres_sweeps: int = 1) -> dict:
    """dJ/d(aoa [per degree], mach, shape_i) at the converged state U of `case` on `hm`'s mesh.

    The residual configuration is the case's last iteration (aero_case_linearization)."""
    rc, seed_a, obj, _ = case_linearization(case, param="aoa")
    _, seed_m, _, dq_m = case_linearization(case, param="mach")
    obj.w_cl, obj.w_cd, obj.w_cm = weights
    cfl = case.cfl_max if cfl is None else cfl
    ob = hm.objective(obj, U)
    t0 = time.perf_counter()
    psi, info = hm.adjoint_solve(rc, ob["dJdU"], U, cfl=cfl, rk_stages=case.rk_stages, max_iters=max_iters, tol=tol,
                                 gmres_restart=gmres_restart, gmres_sweeps=gmres_sweeps)
    t_adj = time.perf_counter() - t0
    params = [("aoa", seed_a, ob["dJ_daoa"]), ("mach", seed_m, ob["dJ_dq"] * dq_m)]
    if shape_tangents is not None:
        idx = range(shape_tangents["area"].shape[0]) if shape_params is None else shape_params
        for i in idx:
            s = geometry_seed(shape_tangents, i)
            params.append((f"shape{i}", s, hm.objective(obj, U, seed=s, gradient=False)["dJ_seed"]))
    out = dict(J=ob["J"], adjoint=info, adjoint_seconds=t_adj, d={}, d_tangent={}, tangent_info={})
    for name, seed, j_p in params:
        _, r_p = hm.residual_jvp(rc, U, None, seed)
        out["d"][name] = j_p - float(np.sum(psi * r_p))
        if tangent:
            v, ti = hm.tangent_solve(rc, seed, U, cfl=cfl, rk_stages=case.rk_stages, max_iters=max_iters, tol=tol,
                                     gmres_restart=gmres_restart, gmres_sweeps=gmres_sweeps)
            out["d_tangent"][name] = j_p + float(np.sum(ob["dJdU"] * v))
            out["tangent_info"][name] = ti
    return out


def sensitivities_batch(items: list, objectives=(("cl", (1.0, 0.0, 0.0)), ("cd", (0.0, 1.0, 0.0))),
                        cfl: float | None = None, max_iters: int = 20000, tol: float = 1e-10,
                        iters_per_sync: int = 200, progress=None) -> dict:
    """dJ/d(aoa [per degree], mach, shape_i) for every (case, objective) pair through one batched adjoint
    solve (aero_hip_adjoint_solve_batch).

    items: dicts with hm (HipMesh, holding the case's frozen limiter ratios when the case froze them),
    case (CaseConfig whose `iterations` is the converged iteration count), U (the converged state) and
    optionally shape_tangents (deform_mesh tangents).  Each case iterates at `cfl` (default: its own
    cfl_max); all share rk_stages; progress(iterations, active) is called at every host check of the
    solve.  Returns per-item results and the wall times of the solve and of the gradient assembly."""
    if not items:
        raise ValueError("sensitivities_batch: no cases")
    stages = {it["case"].rk_stages for it in items}
    if len(stages) != 1:
        raise ValueError("sensitivities_batch: cases need one rk_stages")
    solves, prep = [], []
    for it in items:
        hm, case, U = it["hm"], it["case"], it["U"]
        rc, seed_a, obj, _ = case_linearization(case, param="aoa")
        _, seed_m, _, dq_m = case_linearization(case, param="mach")
        hm.set_state(U)
        obs = {}
        for name, w in objectives:
            obj.w_cl, obj.w_cd, obj.w_cm = w
            obs[name] = hm.objective(obj, None)
            solves.append((hm, rc, obs[name]["dJdU"], None, case.cfl_max if cfl is None else cfl))
        prep.append((rc, seed_a, seed_m, dq_m, obj, obs))
    psis, infos, t_solve = adjoint_solve_batch(solves, rk_stages=stages.pop(), max_iters=max_iters,
                                               tol=tol, iters_per_sync=iters_per_sync, progress=progress)
    t0 = time.perf_counter()
    out = []
    k = 0
    for it, (rc, seed_a, seed_m, dq_m, obj, obs) in zip(items, prep):
        hm, U, tan = it["hm"], it["U"], it.get("shape_tangents")
        hm.set_state(U)
        params = [("aoa", seed_a, {n: o["dJ_daoa"] for n, o in obs.items()}),
                  ("mach", seed_m, {n: o["dJ_dq"] * dq_m for n, o in obs.items()})]
        if tan is not None:
            for i in range(tan["area"].shape[0]):
                s = geometry_seed(tan, i)
                jp = {}
                for name, w in objectives:
                    obj.w_cl, obj.w_cd, obj.w_cm = w
                    jp[name] = hm.objective(obj, None, seed=s, gradient=False)["dJ_seed"]
                params.append((f"shape{i}", s, jp))
        psi = {name: psis[k + j] for j, (name, _) in enumerate(objectives)}
        info = {name: infos[k + j] for j, (name, _) in enumerate(objectives)}
        k += len(objectives)
        d = {name: {} for name, _ in objectives}
        for pname, seed, jp in params:
            _, r_p = hm.residual_jvp(rc, None, None, seed)
            for name, _ in objectives:
                d[name][pname] = jp[name] - float(np.sum(psi[name] * r_p))
        out.append(dict(J={n: o["J"] for n, o in obs.items()}, d=d, adjoint=info, psi=psi))
    return dict(cases=out, solve_seconds=t_solve, assembly_seconds=time.perf_counter() - t0)



# Central-difference steps: AoA in degrees, Mach, Hicks-Henne amplitude (in chords; the bump geometry is
# nonlinear enough in the amplitude that 1e-5 leaves a 2e-6 relative truncation error on the 16k workload,
# 2.5e-6 about 1e-7).
FD_STEPS = {"aoa": 1e-3, "mach": 1e-4, "shape": 2.5e-6}


def fd_sensitivities_batch(base_mesh: Mesh, amplitudes, case: CaseConfig, U: np.ndarray, ratios: np.ndarray,
                           params=None, steps=None, cfl: float | None = None, max_iters: int = 200000,
                           chunk: int = 2000, j_tol: float = 1e-14, block_cells: int = 256,
                           shape_kw: dict | None = None) -> dict:
    """Central differences of the converged (CL, CD) of the frozen-limiter discretisation, every perturbed
    solution continued from the converged state U in one batch (aero_hip_batch_set_state /
    aero_hip_batch_set_limiter_ratios): for p in params ("aoa" [per degree], "mach", "shape<i>"), the cases
    p + h and p - h with h from steps (FD_STEPS by default), all with the frozen ratios of U's run.

    base_mesh / amplitudes: the undeformed mesh and the Hicks-Henne amplitudes of the converged case (None
    amplitudes: the case's mesh is base_mesh itself, no shape parameters).  The continuation keeps the
    converged run's final discretisation (no first-order startup, floor and all-speed ramps at their end
    values, the same frozen ratios) and ramps the CFL again from the case's cfl_start to `cfl` (default
    its cfl_max), which changes the path, not the converged solution.  The batch runs in chunks of `chunk`
    iterations until no case's CL or CD moved by more than j_tol over a chunk (or max_iters).  Returns
    {"d": {param: {"cl", "cd", "converged", "iterations", "h"}}, "seconds"}."""
    shape_kw = shape_kw or {}
    amp = None if amplitudes is None else np.asarray(amplitudes, dtype=np.float64)
    n_shape = 0 if amp is None else len(amp)
    if params is None:
        params = ["aoa", "mach"] + [f"shape{i}" for i in range(n_shape)]
    steps = dict(FD_STEPS, **(steps or {}))
    mesh0 = base_mesh if amp is None else deform_mesh(base_mesh, amp, bumps_per_surface=n_shape // 2,
                                                      tangents=False, **shape_kw)[0]
    handles = [HipMesh(mesh0, block_cells=block_cells)]
    cases, hs = [], {}
    for p in params:
        kind = "shape" if p.startswith("shape") else p
        h = steps[kind]
        hs[p] = h
        for sgn in (1.0, -1.0):
            if kind == "aoa":
                cases.append((0, case.mach, case.aoa_deg + sgn * h))
            elif kind == "mach":
                cases.append((0, case.mach + sgn * h, case.aoa_deg))
            else:
                e = np.zeros(n_shape)
                e[int(p[5:])] = sgn * h
                mi = deform_mesh(base_mesh, amp + e, bumps_per_surface=n_shape // 2, tangents=False, **shape_kw)[0]
                handles.append(HipMesh(mi, block_cells=block_cells))
                cases.append((len(handles) - 1, case.mach, case.aoa_deg))
    cont = replace(case, startup_first_order_iters=0, cfl_max=case.cfl_max if cfl is None else cfl,
                   stab_ramp_iters=0, all_speed_ramp_iters=0, all_speed_ramp_start_iter=0,
                   limiter_freeze_iter=max(case.limiter_freeze_iter, 0))
    seconds, done = 0.0, 0
    moved = np.full(len(cases), np.inf)
    try:
        with HipBatch(handles, cases, cont) as b:
            for i in range(len(cases)):
                b.set_state(i, U)
                b.set_limiter_ratios(i, ratios)
            prev = None
            while done < max_iters:
                r = b.run(min(chunk, max_iters - done), fixed_iters=True, iters_per_sync=chunk)
                seconds += r["seconds"]
                done += min(chunk, max_iters - done)
                now = np.array([[c["cl"], c["cd"]] for c in r["cases"]])
                if prev is not None:
                    moved = np.abs(now - prev).max(axis=1)
                    if moved.max() <= j_tol:
                        break
                prev = now
    finally:
        for hm in handles:
            hm.close()
    out = {}
    for k, p in enumerate(params):
        a, m = r["cases"][2 * k], r["cases"][2 * k + 1]
        h = hs[p]
        out[p] = dict(cl=(a["cl"] - m["cl"]) / (2 * h), cd=(a["cd"] - m["cd"]) / (2 * h), h=h,
                      converged=[bool(moved[2 * k] <= j_tol), bool(moved[2 * k + 1] <= j_tol)],
                      iterations=[a["iterations"], m["iterations"]],
                      rho_l2=[a["rho_l2"], m["rho_l2"]])
    return dict(d=out, seconds=seconds, iterations=done)
