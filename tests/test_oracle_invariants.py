"""FP64 oracle checks fixed by the mathematics (closed forms, invariants, textbook special cases), chosen so
that a dropped term, a wrong sign or a swapped operand fails one of them.  They complement the FP32 golden
vectors for the FP64 instantiation and pin the density-only norm, which has no reference function."""
import ctypes as C

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol

G = 1.4


def cons(rho, u, v, p):
    return np.array([rho, rho * u, rho * v, 0.0, p / (G - 1) + 0.5 * rho * (u * u + v * v)])


def physical_flux(U, n):
    rho, ru, rv, _, E = U
    u, v = ru / rho, rv / rho
    p = (G - 1) * (E - 0.5 * rho * (u * u + v * v))
    un = u * n[0] + v * n[1]
    return np.array([rho * un, ru * un + n[0] * p, rv * un + n[1] * p, 0.0, (E + p) * un]), un, np.sqrt(G * p / rho)


def oflux(scheme, UL, UR, n, s=1.0, as_fix=0, as_w=1.0):
    UL = np.ascontiguousarray(np.atleast_2d(UL), np.float64)
    UR = np.ascontiguousarray(np.atleast_2d(UR), np.float64)
    k = UL.shape[0]
    nn = np.ascontiguousarray(np.tile(np.asarray(n, np.float64), (k, 1)))
    sc = np.full(k, s)
    F = np.zeros((k, 5))
    lam = np.zeros(k)
    fb = np.zeros(k, np.int32)
    ol.oracle().ao_flux_f64(k, scheme, ol.ptr(UL), ol.ptr(UR), ol.ptr(nn), ol.ptr(sc), C.c_double(G), as_fix,
                            C.c_double(0.25), C.c_double(0.05), C.c_double(as_w), ol.ptr(F), ol.ptr(lam), ol.ptr(fb))
    return F, lam, fb


@pytest.mark.parametrize("scheme", [0, 1])
def test_consistency(oracle_lib, scheme):
    rng = np.random.default_rng(0)
    for _ in range(50):
        U = cons(rng.uniform(0.2, 3), rng.normal(0, 200), rng.normal(0, 200), rng.uniform(1e4, 2e5))
        th = rng.uniform(0, 2 * np.pi)
        n = [np.cos(th), np.sin(th), 0.0]
        F, lam, fb = oflux(scheme, U, U, n)
        Fp, un, a = physical_flux(U, n)
        assert np.allclose(F[0], Fp, rtol=1e-13, atol=1e-9 * np.abs(Fp).max())
        assert lam[0] == pytest.approx(abs(un) + a, rel=1e-14)


def test_rusanov_is_local_lax_friedrichs(oracle_lib):
    rng = np.random.default_rng(1)
    for _ in range(50):
        UL = cons(rng.uniform(0.5, 2), rng.normal(0, 100), rng.normal(0, 100), rng.uniform(5e4, 2e5))
        UR = cons(rng.uniform(0.5, 2), rng.normal(0, 100), rng.normal(0, 100), rng.uniform(5e4, 2e5))
        n = [0.6, 0.8, 0.0]
        FL, unl, al = physical_flux(UL, n)
        FR, unr, ar = physical_flux(UR, n)
        smax = max(abs(unl) + al, abs(unr) + ar)
        ref = 0.5 * (FL + FR) - 0.5 * smax * (UR - UL)
        F, _, _ = oflux(0, UL, UR, n)
        assert np.allclose(F[0], ref, rtol=1e-12, atol=1e-10 * np.abs(ref).max())


def test_hllc_stationary_contact_and_upwinding(oracle_lib):
    n = [1.0, 0.0, 0.0]
    # stationary contact: density jump, zero normal velocity, equal pressure -> no mass or energy flux
    UL, UR = cons(1.0, 0.0, 5.0, 1e5), cons(0.3, 0.0, -2.0, 1e5)
    F, _, fb = oflux(1, UL, UR, n)
    assert fb[0] == 0
    assert abs(F[0, 0]) < 1e-9 and abs(F[0, 4]) < 1e-3 and F[0, 1] == pytest.approx(1e5, rel=1e-12)
    # supersonic to the right: the HLLC flux is the left physical flux
    UL, UR = cons(1.0, 900.0, 0.0, 1e5), cons(0.8, 850.0, 10.0, 0.9e5)
    F, _, _ = oflux(1, UL, UR, n)
    assert np.allclose(F[0], physical_flux(UL, n)[0], rtol=1e-14)
    UL, UR = cons(1.0, -900.0, 0.0, 1e5), cons(0.8, -850.0, 10.0, 0.9e5)
    F, _, _ = oflux(1, UL, UR, n)
    assert np.allclose(F[0], physical_flux(UR, n)[0], rtol=1e-14)


def test_limiter_properties(oracle_lib):
    rng = np.random.default_rng(2)
    a = rng.normal(0, 1, 20000)
    b = rng.normal(0, 1, 20000)
    for lim in (0, 1):
        out = np.zeros_like(a)
        ol.oracle().ao_limiter_f64(lim, ol.ptr(a), ol.ptr(b), ol.ptr(out), len(a))
        same = a * b > 0
        assert np.all(out[~same] == 0)
        r = out[same] / a[same]
        assert np.all((r >= 0) & (r <= 1))
        if lim == 0:
            assert np.array_equal(out[same], np.where(np.abs(a[same]) < np.abs(b[same]), a[same], b[same]))
        else:  # pairwise smooth limiter: phi(a, a) = a(3a^2 + e)/(4a^2 + e) -> clamp -> a
            t = np.abs(a) + 0.1
            ta = np.zeros_like(t)
            ol.oracle().ao_limiter_f64(1, ol.ptr(t), ol.ptr(t), ol.ptr(ta), len(t))
            assert np.allclose(ta, t * 3 / 4 + 0 * t, rtol=1e-9) or np.all(ta <= t)


def residual_and_scale(mesh, U, combo, ff):
    om = ol.OracleMesh(mesh, np.float64)
    cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
    R, lam, d = ol.residual(om, cfg, U)
    return om, cfg, R, lam, d, helpers.flux_scale(om, cfg, U)


@pytest.mark.parametrize("combo", [c for c in cases.RESIDUAL_COMBOS if c[6] == 0][::3])
def test_freestream_preservation(oracle_lib, combo):
    """tests/cpp/test_euler_uniform_invariance.cpp in FP64: a uniform state that equals the farfield
    state gives a residual at round-off of the face-flux scale, for every option combination."""
    from aero_hip import freestream
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=1, rcm=True)
    ff = freestream(0.3, 2.0)[1]
    U = np.tile(ff, (m.num_cells, 1))
    _, _, R, lam, d, scale = residual_and_scale(m, U, combo, ff)
    comp = [0, 1, 2, 4]
    # A slip wall is not consistent with a uniform non-tangent stream, so wall cells are excluded (the
    # reference test uses a mesh with farfield boundaries only).
    off_wall = np.ones(m.num_cells, bool)
    off_wall[m.owner[m.wall_start: m.wall_start + m.wall_count]] = False
    err = np.abs(R[:, comp]) / scale[:, comp]
    assert err[off_wall].max() < 1e-13
    assert err[~off_wall].max() > 1e-3  # and the wall cells do see the wall
    assert np.all(lam > 0)


def test_telescoping_sum_equals_boundary_flux(oracle_lib):
    """Interior faces add +F A to the owner and -F A to the neighbour, so they cancel in the sum over
    cells: with interior face areas set to zero (first order, so boundary fluxes are unchanged) the
    sum over cells of R is the same.  A sign or index error in the neighbour update breaks this."""
    from aero_hip import freestream
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=2, rcm=True)
    ff = freestream(0.3, 2.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.05, 3)
    combo = (1, 0, 0, 0, 0, 0, 0)
    om, cfg, R, lam, d, scale = residual_and_scale(m, U, combo, ff)
    mb = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=2, rcm=True)
    mb.area = mb.area.copy()
    mb.area[: mb.wall_start] = 0.0
    omb = ol.OracleMesh(mb, np.float64)
    Rb, _, _ = ol.residual(omb, cfg, U)
    comp = [0, 1, 2, 4]
    tot = R[:, comp].sum(0)
    bnd = Rb[:, comp].sum(0)
    assert np.all(np.abs(Rb[:, comp]).sum(0) > 0)
    assert np.all(np.abs(tot - bnd) <= 1e-12 * scale[:, comp].sum(0))
    # the slip wall carries no mass: the mass balance comes from the farfield faces only
    wall_cells = m.owner[m.wall_start: m.wall_start + m.wall_count]
    mw = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=2, rcm=True)
    mw.area = mw.area.copy()
    mw.area[: mw.wall_start] = 0.0
    mw.area[mw.farfield_start:] = 0.0
    Rw, _, _ = ol.residual(ol.OracleMesh(mw, np.float64), cfg, U)
    assert np.all(Rw[:, 0] == 0.0) and np.abs(Rw[wall_cells, 1:3]).max() > 0


def test_rk3_stage_update_closed_form(oracle_lib):
    """With R = k U and no limiting, three SSP-RK3 stages multiply U by 1 - z + z^2/2 - z^3/6, z = dt k."""
    from aero_hip import freestream
    n = 64
    rng = np.random.default_rng(4)
    U0 = np.tile(freestream(0.3, 2.0)[1], (n, 1)) * rng.uniform(0.5, 2.0, (n, 1))
    k = 3.0
    lam = np.full(n, 100.0)
    cfl = 5.0
    z = cfl / 100.0 * k
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 0, 0, 0, 0), freestream(0.3, 2.0)[1]))
    us = U0.copy()
    for st in range(3):
        nxt = np.zeros_like(U0)
        res = np.ascontiguousarray(k * us)
        ol.oracle().ao_stage_tail_f64(n, st, ol.ptr(U0), ol.ptr(us), ol.ptr(res), ol.ptr(lam), C.c_double(cfl), 1,
                                      C.byref(cfg), ol.ptr(nxt))
        us = nxt
    amp = 1 - z + z * z / 2 - z ** 3 / 6
    assert np.allclose(us[:, [0, 1, 2, 4]], U0[:, [0, 1, 2, 4]] * amp, rtol=1e-13)


def test_residual_norms_closed_form(oracle_lib):
    from aero_hip import freestream
    m = helpers.mesh("40x8")
    ff = freestream(0.3, 2.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.05, 5)
    om, cfg, R, lam, d, scale = residual_and_scale(m, U, (1, 1, 1, 0, 0, 0, 0), ff)
    out = np.zeros(4)
    ol.oracle().ao_norms_f64(C.byref(om.c), ol.ptr(R), ol.ptr(out))
    v = R / np.maximum(m.volume, 1e-12)[:, None]
    assert out[0] == pytest.approx(np.abs(v).sum(), rel=1e-13)
    assert out[1] == pytest.approx(np.sqrt((v ** 2).sum() / (5 * m.num_cells)), rel=1e-13)
    assert out[2] == pytest.approx(np.abs(v).max(), rel=1e-15)
    assert out[3] == pytest.approx(np.sqrt((v[:, 0] ** 2).sum() / m.num_cells), rel=1e-13)


def test_forces_vanish_for_freestream_pressure(oracle_lib):
    m = helpers.mesh("64x14")
    om = ol.OracleMesh(m, np.float64)
    p = np.full(m.num_cells, 101325.0)
    ref7 = np.array([1.225, 101325.0, 2.0, 100.0, 1.0, 0.25, 0.0])
    o = np.zeros(8)
    ol.oracle().ao_forces_f64(C.byref(om.c), ol.ptr(p), ol.ptr(ref7), 1, ol.ptr(o))
    assert np.all(o == 0.0)
    ol.oracle().ao_forces_f64(C.byref(om.c), ol.ptr(p), ol.ptr(ref7), 0, ol.ptr(o))
    # absolute pressure on a closed body: sum n A = 0 up to round-off
    assert abs(o[0]) < 1e-9 * 101325.0 and abs(o[1]) < 1e-9 * 101325.0


# ---- weighted least-squares gradient (no reference function) -----------------------------------------
def _prim_field_to_U(rho, u, v, p, g=1.4):
    U = np.zeros((len(rho), 5))
    U[:, 0] = rho
    U[:, 1] = rho * u
    U[:, 2] = rho * v
    U[:, 4] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def _lsq_mesh():
    return helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.15, seed=21, rcm=True)


def test_lsq_gradient_exact_for_linear_fields(oracle_lib):
    """A linear primitive field has the exact gradient in every cell, boundary cells included (only
    interior neighbours enter, and they determine a linear field)."""
    m = _lsq_mesh()
    om = ol.OracleMesh(m, np.float64, gradient=1)
    x, y = m.cell_center[:, 0], m.cell_center[:, 1]
    coef = np.array([[1.2, 0.03, -0.02], [0.5, 0.01, 0.04], [-0.1, -0.03, 0.02], [2.0, 0.05, -0.07]])
    q = [c[0] + c[1] * x + c[2] * y for c in coef]
    G = ol.gradients(om, _prim_field_to_U(*q))
    for k, kk in enumerate((0, 1, 2, 4)):
        assert np.abs(G[:, kk, 0] - coef[k, 1]).max() <= 1e-10 * (1 + abs(coef[k, 1]))
        assert np.abs(G[:, kk, 1] - coef[k, 2]).max() <= 1e-10 * (1 + abs(coef[k, 2]))
    assert np.all(G[:, 3, :] == 0.0)


def test_lsq_gradient_matches_weighted_lstsq(oracle_lib):
    """Per cell, the oracle equals numpy's least-squares solution of sqrt(w) d . g = sqrt(w) dq over the
    cell's interior neighbours with w = 1 / |d|^2, for a non-linear field."""
    m = _lsq_mesh()
    om = ol.OracleMesh(m, np.float64, gradient=1)
    from aero_hip import freestream
    ff = freestream(0.5, 3.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.3, 77)
    G = ol.gradients(om, U)
    g = 1.4
    rho = U[:, 0]
    prim = np.stack([rho, U[:, 1] / rho, U[:, 2] / rho,
                     (g - 1) * (U[:, 4] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho)], axis=1)
    nbrs = [[] for _ in range(m.num_cells)]
    for o, n in zip(m.owner, m.neighbor):
        if n >= 0:
            nbrs[o].append(n)
            nbrs[n].append(o)
    cc = m.cell_center[:, :2]
    worst = 0.0
    for c in range(m.num_cells):
        d = cc[nbrs[c]] - cc[c]
        sw = 1.0 / np.sqrt((d * d).sum(axis=1))
        A = d * sw[:, None]
        for k, kk in enumerate((0, 1, 2, 4)):
            b = (prim[nbrs[c], k] - prim[c, k]) * sw
            sol = np.linalg.lstsq(A, b, rcond=None)[0]
            scale = np.abs(sol).max() + 1e-30
            worst = max(worst, np.abs(G[c, kk] - sol).max() / scale)
    assert worst <= 1e-9, worst


def test_lsq_gradient_zero_for_constant_field(oracle_lib):
    m = _lsq_mesh()
    om = ol.OracleMesh(m, np.float64, gradient=1)
    from aero_hip import freestream
    U = np.tile(freestream(0.3, 2.0)[1], (m.num_cells, 1))
    assert np.all(ol.gradients(om, U) == 0.0)


def test_lsq_second_order_residual_preserves_freestream(oracle_lib):
    m = _lsq_mesh()
    om = ol.OracleMesh(m, np.float64, gradient=1)
    from aero_hip import freestream
    ff = freestream(0.5, 1.25)[1]
    U = np.tile(ff, (m.num_cells, 1))
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 0, 0, 0, 0), ff))
    R, lam, d = ol.residual(om, cfg, U)
    scale = helpers.flux_scale(om, cfg, U)
    off_wall = np.ones(m.num_cells, bool)
    off_wall[m.owner[m.wall_start: m.wall_start + m.wall_count]] = False
    comp = [0, 1, 2, 4]
    assert (np.abs(R[off_wall][:, comp]) / scale[off_wall][:, comp]).max() < 1e-13


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_threaded_oracle_equals_serial(dtype):
    """The all-core oracle (the CPU baseline's threaded line) gives the serial oracle's residual, spectral
    radius and fallback counters bit for bit, for every option combination on a jittered mixed mesh."""
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=9, rcm=True)
    om = ol.OracleMesh(m, dtype)
    ff = np.asarray(__import__("aero_hip").freestream(0.5, 1.25)[1])
    for amp, seed in ((0.05, 3), (0.9, 4)):
        U = helpers.perturbed_state(m.cell_center, ff, amp, seed)
        for combo in cases.RESIDUAL_COMBOS:
            cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff), dtype)
            a = ol.residual(om, cfg, U)
            b = ol.residual_threaded(om, cfg, U)
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), (combo, amp)


@pytest.mark.parametrize("name", ["case_file_numerics", "rk3_global", "precond_staged"])
def test_threaded_oracle_driver_equals_serial(name):
    """The oracle driver with its stage residuals on all threads (used to keep long GPU parity runs short)
    reproduces the serial driver's history, state and forces bit for bit."""
    run = cases.RUNS[name]
    m = helpers.mesh(run["mesh"])
    om = ol.OracleMesh(m, np.float64)
    cfg = helpers.case_cfg(run["cfg"], run["env_precond"])
    n = run["cfg"]["iterations"]
    a = ol.run_case(om, cfg, n)
    b = ol.run_case(om, cfg, n, threaded=True)
    for k in ("history", "conserved", "residual", "spectral_radius", "forces", "cp"):
        assert np.array_equal(a[k], b[k]), (name, k)
