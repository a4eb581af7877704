"""The oracle's tangent (the FP64 restatement instantiated on dual numbers) against central finite
differences of the oracle residual: in the state, in the freestream-dependent options and in the
geometry.  This pins the tangent the device kernels are compared with.  Also pins the case
linearisation (aero_case_linearization, host code) against finite differences of the freestream."""
import ctypes as C

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol

COMBOS = [(0, 0, 0, 0, 0, 0, 0), (1, 1, 0, 0, 0, 0, 0), (0, 0, 1, 0, 0, 0, 0), (1, 1, 1, 0, 0, 0, 0),
          (1, 1, 1, 1, 1, 1, 1), (1, 0, 1, 1, 0, 1, 0)]


def _setup(gradient=0, seed=3):
    from aero_hip import freestream
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=seed, rcm=True)
    om = ol.OracleMesh(m, np.float64, gradient=gradient)
    ff = freestream(0.5, 2.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.05, 17)
    return m, om, ff, U


def _fd(fun, h):
    return (fun(h) - fun(-h)) / (2.0 * h)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("combo", COMBOS)
@pytest.mark.parametrize("gradient", [0, 1])
def test_state_tangent_matches_central_difference(oracle_lib, combo, gradient):
    m, om, ff, U = _setup(gradient)
    cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
    rng = np.random.default_rng(5)
    dU = rng.standard_normal(U.shape) * np.abs(U).max(axis=0) * 1e-3
    dU[:, 3] = 0.0
    R0, dR = ol.residual_jvp(om, cfg, U, dU)
    Rv, _, _ = ol.residual(om, cfg, U)
    assert np.array_equal(R0, Rv), "the tangent's value part is the residual"
    fd = _fd(lambda h: ol.residual(om, cfg, U + h * dU)[0], 1e-5)
    assert _rel(dR, fd) < 1e-6, _rel(dR, fd)


def test_parameter_tangent_matches_central_difference(oracle_lib):
    from aero_hip import Seed
    m, om, ff, U = _setup()
    vals = cases.residual_cfg_values((1, 1, 1, 1, 1, 1, 1), ff)
    d_ff = np.array([1e-3, 2.0, -3.0, 0.0, 400.0])
    seed = Seed(d_farfield=d_ff, d_stab_floor=0.3, d_mach_ref=0.2, d_beta_min=0.1)
    _, dR = ol.residual_jvp(om, helpers.res_cfg(vals), U, None, seed)

    def R(h):
        v = dict(vals)
        v["farfield"] = list(np.asarray(vals["farfield"]) + h * d_ff)
        v["stab_mach_floor"] = vals["stab_mach_floor"] + 0.3 * h
        v["precond_mach_ref"] = vals["precond_mach_ref"] + 0.2 * h
        v["precond_beta_min"] = vals["precond_beta_min"] + 0.1 * h
        return ol.residual(om, helpers.res_cfg(v), U)[0]
    fd = _fd(R, 1e-6)
    assert np.linalg.norm(dR) > 0
    assert _rel(dR, fd) < 1e-6, _rel(dR, fd)


@pytest.mark.parametrize("gradient", [0, 1])
def test_geometry_tangent_matches_central_difference(oracle_lib, gradient):
    from aero_hip import Seed
    m, om, ff, U = _setup(gradient)
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 0, 0, 0, 0), ff))
    rng = np.random.default_rng(9)
    dn = rng.standard_normal(m.normal.shape) * 1e-2
    dn[:, 2] = 0.0
    dfc = rng.standard_normal(m.face_center.shape) * 1e-3
    dfc[:, 2] = 0.0
    dA = rng.standard_normal(m.area.shape) * 1e-3
    dcc = rng.standard_normal(m.cell_center.shape) * 1e-3
    dcc[:, 2] = 0.0
    dV = rng.standard_normal(m.volume.shape) * 1e-4
    seed = Seed(d_face_normal=dn, d_face_center=dfc, d_face_area=dA, d_cell_center=dcc, d_cell_volume=dV)
    _, dR = ol.residual_jvp(om, cfg, U, None, seed)

    def R(h):
        import copy
        mm = copy.copy(m)
        mm.normal = m.normal + h * dn
        mm.face_center = m.face_center + h * dfc
        mm.area = m.area + h * dA
        mm.cell_center = m.cell_center + h * dcc
        mm.volume = m.volume + h * dV
        return ol.residual(ol.OracleMesh(mm, np.float64, gradient=gradient), cfg, U)[0]
    fd = _fd(R, 1e-6)
    assert _rel(dR, fd) < 1e-6, _rel(dR, fd)


@pytest.mark.parametrize("param", ["aoa", "mach"])
def test_case_linearization_matches_central_difference(param):
    """d(freestream state, floor, preconditioner bounds, q_inf)/d(param) from aero_case_linearization
    against differences of the same function at param +- h."""
    from aero_hip import CaseConfig, case_linearization
    base = dict(mach=0.3, aoa_deg=2.0, precond_on=1, precond_mach_ref=0.0, precond_beta_min=0.0,
                stab_floor_k_start=0.4, stab_floor_k_target=0.2, stab_ramp_iters=100)
    case = CaseConfig(**base)
    case.precond_experimental_enabled = 1
    rc, seed, obj, dq = case_linearization(case, iteration=50, param=param)
    h = 1e-6 if param == "mach" else 1e-4

    def at(d):
        c = CaseConfig(**dict(base, **{("mach" if param == "mach" else "aoa_deg"): base[
            "mach" if param == "mach" else "aoa_deg"] + d}))
        c.precond_experimental_enabled = 1
        r, _, o, _ = case_linearization(c, iteration=50, param="none")
        return (np.array(list(r.farfield) + [r.stab_mach_floor, r.precond_mach_ref, r.precond_beta_min,
                                             r.precond_beta_max]), o.q_inf)
    (vp, qp), (vm, qm) = at(h), at(-h)
    fd = (vp - vm) / (2 * h)
    got = np.array(list(seed.c.d_farfield) + [seed.c.d_stab_floor, seed.c.d_mach_ref, seed.c.d_beta_min,
                                               seed.c.d_beta_max])
    scale = np.abs(fd).max()
    assert np.abs(got - fd).max() <= 1e-6 * scale, (got, fd)
    assert abs(dq - (qp - qm) / (2 * h)) <= 1e-6 * abs(dq) + 1e-9
