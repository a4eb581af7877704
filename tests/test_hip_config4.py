"""The discrete adjoint on a reduced copy of the Hicks-Henne workload: a 16k-cell jittered mixed mesh
deformed by 20 bumps (10 per surface, peaks x/c 0.05..0.95, sin^3, amplitudes U(+-0.002c) from seed 4242,
IDW), M 0.6, AoA 2 deg, HLLC + venkat with the limiter frozen at iteration 1000.  For all 22 parameters
(AoA, Mach, the 20 amplitudes) and both CL and CD, the batched adjoint gradient must agree to 1e-6 with:

* central differences of converged solutions (aero_hip.adjoint.fd_sensitivities_batch: every perturbed
  case continued from the converged state in one batch, with the same frozen ratios, until no CL or CD
  moves by more than 1e-14 over 2000 iterations; steps 1e-3 deg, 1e-4 and 2.5e-6 chords);
* the CPU tangent evaluation: the FP64 oracle's dual-number residual derivative R_p along each
  parameter's direction and its force integration (central differences of the wall forces, exact for
  these bilinear sums up to round-off), assembled with the device adjoint state as J_p - psi . R_p.
The adjoint state itself is checked against the oracle's linearisation: <psi, R_U v> (oracle tangent)
equals J_U v (oracle forces) for random v.  The tolerance is relative to each derivative, with a floor of
1e-3 of the largest derivative of the same objective for the few near-zero ones.
"""
import ctypes as C
import sys
from dataclasses import replace

import numpy as np
import pytest

import helpers
import oracle_lib as ol

pytestmark = pytest.mark.gpu
TOL = 1e-6


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def workload():
    import aero_hip
    from aero_hip import CaseConfig, HipBatch, HipMesh, deform_mesh, generate_mesh
    from aero_hip.adjoint import sensitivities_batch
    assert aero_hip.hip_available()
    base = generate_mesh(precision=8, num_circumferential=128, num_radial=78, farfield_radius=20.0,
                         radial_stretch=200.0, cells="mixed", quad_layers=30, jitter=0.1, seed=12345, rcm=True)
    amp = np.random.default_rng(4242).uniform(-0.002, 0.002, 20)
    mesh, tan = deform_mesh(base, amp)
    cap = 100000
    case = CaseConfig(mach=0.6, aoa_deg=2.0, flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=2.0,
                      cfl_ramp_iters=120, startup_first_order_iters=25, stab_floor_start=0.1, stab_floor_target=0.02,
                      stab_ramp_iters=400, precond_experimental_enabled=0, limiter_freeze_iter=1000,
                      residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, iterations=cap)
    hm = HipMesh(mesh)
    say(f"[config4] primal on {mesh.num_cells} cells")
    with HipBatch([hm], [(0, case.mach, case.aoa_deg)], case) as b:
        r = b.run(cap, 1e-10, iters_per_sync=1000)
        U = b.state(0)
        ratios = b.limiter_ratios(0)
    row = r["cases"][0]
    assert row["converged"], row
    hm.set_limiter_ratios(ratios)
    done = replace(case, iterations=row["iterations"])
    say(f"[config4] primal {row['iterations']} iterations; adjoints of CL and CD")
    sb = sensitivities_batch([dict(hm=hm, case=done, U=U, shape_tangents=tan)], cfl=2.4, max_iters=4 * cap,
                             tol=1e-11, iters_per_sync=1000)
    c = sb["cases"][0]
    for name, info in c["adjoint"].items():
        assert info["converged"], (name, info)
    yield dict(base=base, amp=amp, mesh=mesh, tan=tan, hm=hm, case=done, U=U, ratios=ratios, d=c["d"], psi=c["psi"])
    hm.close()


def check(adj: dict, ref: dict, label: str):
    worst = 0.0
    for obj in ("cl", "cd"):
        floor = 1e-3 * max(abs(v) for v in ref[obj].values())
        for p, v in ref[obj].items():
            err = abs(adj[obj][p] - v) / max(abs(v), floor)
            worst = max(worst, err)
            assert err <= TOL, f"{label} d{obj}/d{p}: adjoint {adj[obj][p]:.12e} vs {v:.12e} ({err:.2e})"
    say(f"[config4] adjoint vs {label}: worst relative difference {worst:.2e} over 22 parameters x (CL, CD)")


def test_adjoint_matches_central_differences(workload):
    from aero_hip.adjoint import fd_sensitivities_batch
    w = workload
    say("[config4] 44 perturbed solutions in one batch")
    fd = fd_sensitivities_batch(w["base"], w["amp"], w["case"], w["U"], w["ratios"], cfl=2.0, max_iters=200000)
    for p, v in fd["d"].items():
        assert all(v["converged"]), (p, v)
    ref = {obj: {p: v[obj] for p, v in fd["d"].items()} for obj in ("cl", "cd")}
    check(w["d"], ref, "central differences")


def cpu_forces(om, U, rho_inf, p_inf, aoa, speed, gamma=1.4):
    """(CL, CD) of the gauge wall pressure forces of U on the oracle's mesh (forces.cpp:250-294)."""
    from aero_hip.outputs import _primitives
    p = np.ascontiguousarray(_primitives(U, gamma)[3], np.float64)
    ref7 = np.array([rho_inf, p_inf, aoa, speed, 1.0, 0.25, 0.0])
    o = np.zeros(8)
    ol.oracle().ao_forces_f64(C.byref(om.c), ol.ptr(p), ol.ptr(ref7), 1, ol.ptr(o))
    return np.array([o[5], o[6]])


def test_adjoint_matches_cpu_tangent(workload):
    from aero_hip import case_linearization
    from aero_hip.adjoint import geometry_seed
    from aero_hip.outputs import rho_inf_of, speed_inf_of
    w = workload
    mesh, U, case, tan = w["mesh"], w["U"], w["case"], w["tan"]
    om = ol.OracleMesh(mesh, np.float64)
    rc, seed_a, _, _ = case_linearization(case, param="aoa")
    _, seed_m, _, _ = case_linearization(case, param="mach")
    rho, V = rho_inf_of(case), speed_inf_of(case)
    psi = w["psi"]

    def J(om_, U_, aoa=case.aoa_deg, mach=case.mach):
        c2 = replace(case, mach=mach, aoa_deg=aoa)
        return cpu_forces(om_, U_, rho_inf_of(c2), c2.p_inf, aoa, speed_inf_of(c2))

    cpu = {"cl": {}, "cd": {}}
    for p in ["aoa", "mach"] + [f"shape{i}" for i in range(20)]:
        if p == "aoa":
            seed, e = seed_a, 1e-4
            jp = (J(om, U, aoa=case.aoa_deg + e) - J(om, U, aoa=case.aoa_deg - e)) / (2 * e)
        elif p == "mach":
            seed, e = seed_m, 1e-6
            jp = (J(om, U, mach=case.mach + e) - J(om, U, mach=case.mach - e)) / (2 * e)
        else:
            i = int(p[5:])
            seed = geometry_seed(tan, i)
            e = 1e-3
            js = []
            for sgn in (1, -1):  # the forces are bilinear in the wall geometry: central differences are exact
                mp = mesh.astype(np.float64)
                mp.normal = mesh.normal + sgn * e * tan["normal"][i]
                mp.area = mesh.area + sgn * e * tan["area"][i]
                mp.face_center = mesh.face_center + sgn * e * tan["face_center"][i]
                js.append(J(ol.OracleMesh(mp, np.float64), U))
            jp = (js[0] - js[1]) / (2 * e)
        _, r_p = ol.residual_jvp(om, rc, U, None, seed, frozen=w["ratios"])
        for k, obj in enumerate(("cl", "cd")):
            cpu[obj][p] = float(jp[k]) - float(np.sum(psi[obj] * r_p))
    check(w["d"], cpu, "CPU tangent evaluation")
    # the adjoint state against the oracle's linearisation: <psi, R_U v> = J_U v
    rng = np.random.default_rng(7)
    for _ in range(2):
        v = rng.standard_normal(U.shape) * np.abs(U).max(axis=0) * 1e-3
        v[:, 3] = 0.0
        _, ruv = ol.residual_jvp(om, rc, U, v, None, frozen=w["ratios"])
        h = 1e-4
        juv = (J(om, U + h * v) - J(om, U - h * v)) / (2 * h)
        for k, obj in enumerate(("cl", "cd")):
            lhs = float(np.sum(psi[obj] * ruv))
            assert abs(lhs - juv[k]) <= 1e-6 * abs(juv[k]), (obj, lhs, juv[k])
