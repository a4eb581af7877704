"""Frozen limiter (beyond the reference) on the device against the oracle: captured ratios, frozen
residual, a driver run with a freeze iteration, the frozen tangent, and second-order sensitivities
(adjoint = tangent = central differences) at a converged frozen-limiter state."""
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


def _mesh():
    return helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)


def test_capture_and_frozen_residual_parity(hip):
    m = _mesh()
    om = ol.OracleMesh(m, np.float64)
    from aero_hip import freestream
    ff = freestream(0.5, 2.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.05, 17)
    comp = [0, 1, 2, 4]
    with hip.HipMesh(m) as hm:
        for combo in [(1, 1, 1, 0, 0, 0, 0), (0, 0, 1, 1, 0, 0, 0), (1, 1, 1, 1, 1, 1, 1)]:
            vals = cases.residual_cfg_values(combo, ff)
            cfg = helpers.res_cfg(vals)
            say(f"[frozen] capture + residual combo={combo}")
            rh = hm.freeze_limiter(cfg, U)
            ro = ol.limiter_capture(om, cfg, U)
            assert np.array_equal(rh, ro), np.abs(rh - ro).max()
            cfz = helpers.res_cfg(dict(vals, limiter_frozen=1))
            Rh, lh, _ = hm.residual(U, cfz)
            Ro, lo, _ = ol.residual_frozen(om, cfg, U, ro)
            scale = helpers.flux_scale(om, cfg, U)
            assert (np.abs(Rh - Ro)[:, comp] / (scale[:, comp] + 1e-300)).max() <= 1e-12
            assert (np.abs(lh - lo) / lo).max() <= 1e-12
            assert np.array_equal(Rh[:, comp], Ro[:, comp]) and np.array_equal(lh, lo), combo
            # frozen tangent against the oracle's
            rng = np.random.default_rng(2)
            dU = rng.standard_normal(U.shape) * np.abs(U).max(axis=0) * 1e-3
            dU[:, 3] = 0.0
            _, dRh = hm.residual_jvp(cfz, U, dU)
            _, dRo = ol.residual_jvp(om, cfg, U, dU, frozen=ro)
            sc = np.abs(dRo[:, comp]).max(axis=0, keepdims=True)
            assert (np.abs(dRh[:, comp] - dRo[:, comp]) / sc).max() <= 1e-11
            assert np.array_equal(dRh[:, comp], dRo[:, comp]), combo
            # frozen adjoint: dot-product identity
            psi = rng.standard_normal(U.shape)
            psi[:, 3] = 0.0
            a = float(np.sum(psi * dRh))
            b = float(np.sum(hm.residual_vjp(cfz, psi, U) * dU))
            assert abs(a - b) <= 1e-12 * float(np.sum(np.abs(psi * dRh)))


def test_frozen_run_case_parity(hip):
    m = _mesh()
    om = ol.OracleMesh(m, np.float64)
    over = dict(flux_scheme=1, limiter=1, cfl_start=0.5, cfl_max=1.5, cfl_ramp_iters=50, startup_first_order_iters=25,
                residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, iterations=600, mach=0.5,
                aoa_deg=2.0, limiter_freeze_iter=300)
    ref = ol.run_case(om, helpers.case_cfg(over, False), 600)
    case = hip.CaseConfig(**over)
    case.precond_experimental_enabled = 0
    say("[frozen] run_case with freeze at 300")
    with hip.HipMesh(m) as hm:
        out = hm.run_case(case)
    hr, ho = ref["history"], out["history"]
    assert ho.shape == hr.shape
    for col in (1, 2, 3, 5, 6, 7, 8):
        rel = np.abs(ho[:, col] - hr[:, col]) / np.maximum(np.abs(hr[:, col]), 1e-300)
        assert rel.max() <= 1e-9, (col, rel.max())
    assert np.array_equal(ho[:, 3], hr[:, 3]), "L-infinity residual history differs"
    assert np.array_equal(out["conserved"], ref["conserved"]), "final state differs from the oracle driver's"


def _frozen_converge(hm, rc, U0, iters, cfl):
    """Pseudo-time iterations from U0 with the handle's frozen ratios (aero_hip_iterate); returns the
    density-residual L2 after the first and the last iteration."""
    hm.set_state(U0)
    h = hm.iterate(rc, iters, cfl)
    return h[0], h[-1]


def test_frozen_second_order_sensitivities(hip):
    """HLLC + venkat, limiter frozen at iteration 300, converged: adjoint dCL/d(aoa, mach) equals the
    tangent solve and central differences of converged CL with the same frozen ratios."""
    from aero_hip.adjoint import sensitivities
    m = _mesh()
    over = dict(flux_scheme=1, limiter=1, cfl_start=0.5, cfl_max=1.5, cfl_ramp_iters=50, startup_first_order_iters=25,
                residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, iterations=24000,
                mach=0.5, aoa_deg=2.0, limiter_freeze_iter=300)
    case = hip.CaseConfig(**over)
    with hip.HipMesh(m) as hm:
        say("[frozen sens] primal")
        out = hm.run_case(case)
        drop = out["history"][-1, 8] / out["history"][0, 8]
        assert drop < 1e-11, drop
        U = out["conserved"]
        say("[frozen sens] adjoint + tangent")
        s = sensitivities(hm, case, U, cfl=1.5, tol=1e-12, max_iters=40000, tangent=True)
        assert s["adjoint"]["res"] <= 1e-11 * s["adjoint"]["res0"], s["adjoint"]
        for name in ("aoa", "mach"):
            assert abs(s["d"][name] - s["d_tangent"][name]) <= 1e-8 * abs(s["d"][name]), name
        h = {"aoa": 1e-3, "mach": 1e-4}
        for name in ("aoa", "mach"):
            cls = []
            for sgn in (1, -1):
                kw = dict(over)
                kw["aoa_deg" if name == "aoa" else "mach"] += sgn * h[name]
                cp = hip.CaseConfig(**kw)
                rc, _, obj, _ = hip.case_linearization(cp)
                assert rc.limiter_frozen == 1
                say(f"[frozen sens] finite difference {name} {sgn:+d}")
                r_first, r_last = _frozen_converge(hm, rc, U, 12000, 1.5)
                assert r_last <= 1e-7 * r_first, (r_first, r_last)
                obj.w_cl = 1.0
                cls.append(hm.objective(obj, None, gradient=False)["J"])
            fd = (cls[0] - cls[1]) / (2 * h[name])
            say(f"[frozen sens] {name}: rel {abs(s['d'][name] - fd) / abs(fd):.2e}")
            assert abs(s["d"][name] - fd) <= 1e-6 * abs(fd), (name, s["d"][name], fd)
