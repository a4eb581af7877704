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

from .solver import CaseConfig, HipMesh, Seed, case_linearization


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
