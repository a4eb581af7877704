"""Frozen limiter (an option beyond the reference) in the oracle: the captured ratios are the limiter's
phi(a, b) / a held at 2^-15 resolution (the nearest multiple, ties to even), the frozen residual with the
unrounded ratios equals the limited one at the capture state and with the held ratios differs from it
by no more than the rounding of the increments allows, its tangent matches central differences, and a
driver run with a freeze iteration follows the reference run up to it."""
import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol


def _setup():
    from aero_hip import freestream
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    om = ol.OracleMesh(m, np.float64)
    ff = freestream(0.5, 2.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.05, 17)
    return m, om, ff, U


@pytest.mark.parametrize("limiter", [0, 1])
def test_capture_and_frozen_residual(oracle_lib, limiter):
    m, om, ff, U = _setup()
    for scheme in (0, 1):
        cfg = helpers.res_cfg(cases.residual_cfg_values((scheme, limiter, 1, 1, 0, 0, 0), ff))
        r = ol.limiter_capture(om, cfg, U, exact=True)
        assert r.shape == (8, m.num_faces)
        assert np.all(r >= 0.0) and np.all(r <= 1.0 + 1e-15)
        assert np.all(r[:, m.neighbor < 0] == 0.0)
        assert np.count_nonzero(r) > 0.5 * r.size * (m.neighbor >= 0).mean()
        Rf, lf, _ = ol.residual_frozen(om, cfg, U, r)
        R0, l0, _ = ol.residual(om, cfg, U)
        scale = helpers.flux_scale(om, cfg, U)
        comp = [0, 1, 2, 4]
        assert (np.abs(Rf - R0)[:, comp] / (scale[:, comp] + 1e-300)).max() <= 1e-13
        assert np.array_equal(lf, l0)
        # the held ratios: the nearest multiple of 2^-15 of the exact ones (numpy's rint rounds ties to even)
        rq = ol.limiter_capture(om, cfg, U)
        assert np.array_equal(rq, np.rint(np.clip(r, 0.0, 1.0) * 32768.0) / 32768.0)
        assert np.all(np.abs(rq - r) <= 2.0 ** -16)
        assert np.array_equal(rq * 32768.0, np.floor(rq * 32768.0))
        # with the held ratios the residual moves by the rounding of the increments only: a few 1e-5 of the
        # face-flux scale, nowhere near the limited-vs-unlimited difference
        Rq, lq, _ = ol.residual_frozen(om, cfg, U, rq)
        dq = (np.abs(Rq - R0)[:, comp] / (scale[:, comp] + 1e-300)).max()
        assert 0.0 < dq <= 1e-4, dq


def test_frozen_tangent_matches_central_difference(oracle_lib):
    m, om, ff, U = _setup()
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 1, 1, 1, 1), ff))
    r = ol.limiter_capture(om, cfg, U)
    rng = np.random.default_rng(5)
    dU = rng.standard_normal(U.shape) * np.abs(U).max(axis=0) * 1e-3
    dU[:, 3] = 0.0
    R0, dR = ol.residual_jvp(om, cfg, U, dU, frozen=r)
    assert np.array_equal(R0, ol.residual_frozen(om, cfg, U, r)[0])
    h = 1e-5
    fd = (ol.residual_frozen(om, cfg, U + h * dU, r)[0] - ol.residual_frozen(om, cfg, U - h * dU, r)[0]) / (2 * h)
    assert np.linalg.norm(dR - fd) <= 1e-7 * np.linalg.norm(fd)


def test_driver_freeze_follows_reference_until_freeze(oracle_lib):
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    om = ol.OracleMesh(m, np.float64)
    base = dict(flux_scheme=1, limiter=1, cfl_start=0.5, cfl_max=1.5, cfl_ramp_iters=50, startup_first_order_iters=25,
                residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, iterations=400, mach=0.5,
                aoa_deg=2.0)
    h0 = ol.run_case(om, helpers.case_cfg(base, False), 400)["history"]
    h1 = ol.run_case(om, helpers.case_cfg(dict(base, limiter_freeze_iter=200), False), 400)["history"]
    assert np.array_equal(h0[:200], h1[:200])
    assert not np.array_equal(h0[200:], h1[200:])
    assert np.all(np.isfinite(h1))
