"""Device tangent and adjoint of the residual (no reference counterpart), through the C ABI.

* tangent: device dR = R_U dU + R_p . seed equal bit for bit to the oracle's dual-number tangent (which
  tests/test_oracle_tangent.py pins to central differences), in the state, the freestream options
  and the geometry;
* adjoint: R_U^T psi against the transpose of the oracle's dense Jacobian on a small mesh, and the
  dot-product identity <psi, R_U v> = <R_U^T psi, v> on a larger one;
* objective gradient against central differences of the objective;
* sensitivities: adjoint and tangent solves agree (dual consistency), and both agree with central
  differences of converged solutions.
"""
import sys

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol

pytestmark = pytest.mark.gpu

COMBOS = [(0, 0, 0, 0, 0, 0, 0), (1, 1, 0, 0, 0, 0, 0), (0, 0, 1, 0, 0, 0, 0), (1, 1, 1, 0, 0, 0, 0),
          (1, 1, 1, 1, 1, 1, 1), (1, 0, 1, 1, 0, 1, 0)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available(), "no gfx950 device visible to the HIP backend"
    return aero_hip


def _mesh64():
    return helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)


def _state(m, mach=0.5, aoa=2.0, amp=0.05, seed=17):
    from aero_hip import freestream
    ff = freestream(mach, aoa)[1]
    return ff, helpers.perturbed_state(m.cell_center, ff, amp, seed)


def _maxrel(a, b):
    scale = np.abs(b).max(axis=0, keepdims=True) + 1e-300
    return (np.abs(a - b) / scale).max()


@pytest.mark.parametrize("gradient", ["green_gauss", "least_squares"])
def test_tangent_parity(hip, gradient):
    m = _mesh64()
    gk = 1 if gradient == "least_squares" else 0
    om = ol.OracleMesh(m, np.float64, gradient=gk)
    ff, U = _state(m)
    rng = np.random.default_rng(4)
    dU = rng.standard_normal(U.shape) * np.abs(U).max(axis=0) * 1e-3
    dU[:, 3] = 0.0
    dn = rng.standard_normal(m.normal.shape) * 1e-2
    dn[:, 2] = 0.0
    dfc = rng.standard_normal(m.face_center.shape) * 1e-3
    dfc[:, 2] = 0.0
    dcc = rng.standard_normal(m.cell_center.shape) * 1e-3
    dcc[:, 2] = 0.0
    seed = hip.Seed(d_farfield=[1e-3, 2.0, -3.0, 0.0, 400.0], d_stab_floor=0.3, d_mach_ref=0.2, d_beta_min=0.1,
                    d_face_normal=dn, d_face_center=dfc, d_face_area=rng.standard_normal(m.num_faces) * 1e-3,
                    d_cell_center=dcc, d_cell_volume=rng.standard_normal(m.num_cells) * 1e-4)
    with hip.HipMesh(m, gradient=gradient) as hm:
        for combo in COMBOS:
            cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
            say(f"[tangent] {gradient} combo={combo}")
            for s in (None, seed):
                Rh, dRh = hm.residual_jvp(cfg, U, dU, s)
                Ro, dRo = ol.residual_jvp(om, cfg, U, dU, s)
                Rp, _, _ = hm.residual(U, cfg)
                comp = [0, 1, 2, 4]
                assert np.array_equal(Rh[:, comp], Rp[:, comp]), "tangent value part differs from the residual"
                err = _maxrel(dRh[:, comp], dRo[:, comp])
                assert np.array_equal(dRh[:, comp], dRo[:, comp]), (combo, s is not None, err)


def _dense_jacobian(om, cfg, U):
    """d R[:, (0,1,2,4)] / d U[:, (0,1,2,4)] of the oracle, column by column from its tangent."""
    N = om.N
    comp = [0, 1, 2, 4]
    J = np.zeros((4 * N, 4 * N))
    for c in range(N):
        for j, cj in enumerate(comp):
            dU = np.zeros((N, 5))
            dU[c, cj] = 1.0
            _, dR = ol.residual_jvp(om, cfg, U, dU)
            J[:, 4 * c + j] = dR[:, comp].ravel()
    return J


@pytest.mark.parametrize("gradient", ["green_gauss", "least_squares"])
def test_adjoint_matches_dense_oracle_transpose(hip, gradient):
    m = helpers.mesh("40x8")
    gk = 1 if gradient == "least_squares" else 0
    om = ol.OracleMesh(m, np.float64, gradient=gk)
    ff, U = _state(m, 0.4, 1.0, 0.05, 3)
    rng = np.random.default_rng(8)
    psi = rng.standard_normal((m.num_cells, 5))
    psi[:, 3] = 0.0
    comp = [0, 1, 2, 4]
    with hip.HipMesh(m, gradient=gradient) as hm:
        for combo in [(0, 0, 1, 0, 0, 0, 0), (1, 1, 1, 0, 0, 0, 0), (1, 1, 1, 1, 1, 1, 1)]:
            cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
            say(f"[adjoint dense] {gradient} combo={combo}")
            J = _dense_jacobian(om, cfg, U)
            ref = (J.T @ psi[:, comp].ravel()).reshape(m.num_cells, 4)
            got = hm.residual_vjp(cfg, psi, U)
            assert np.all(got[:, 3] == 0.0)
            err = _maxrel(got[:, comp], ref)
            assert err <= 1e-11, (combo, err)


@pytest.mark.parametrize("gradient", ["green_gauss", "least_squares"])
def test_adjoint_dot_product_identity(hip, gradient):
    m = helpers.mesh("256x64", cells="mixed", quad_layers=12, jitter=0.1, seed=12345, rcm=True)
    ff, U = _state(m, 0.5, 1.25, 0.02, 20260101)
    rng = np.random.default_rng(1)
    v = rng.standard_normal(U.shape) * np.abs(U).max(axis=0) * 1e-3
    v[:, 3] = 0.0
    psi = rng.standard_normal(U.shape)
    psi[:, 3] = 0.0
    with hip.HipMesh(m, gradient=gradient) as hm:
        for combo in COMBOS:
            cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
            say(f"[adjoint dot] {gradient} combo={combo}")
            _, Jv = hm.residual_jvp(cfg, U, v)
            JTpsi = hm.residual_vjp(cfg, psi, U)
            a = float(np.sum(psi * Jv))
            b = float(np.sum(JTpsi * v))
            scale = float(np.sum(np.abs(psi * Jv)))
            assert abs(a - b) <= 1e-12 * scale, (combo, a, b)


def test_objective_and_gradient(hip):
    """J against the oracle's forces on the same state; dJ/dU against central differences of J."""
    m = helpers.mesh("64x14")
    ff, U = _state(m, 0.5, 3.0, 0.05, 2)
    case = hip.CaseConfig(mach=0.5, aoa_deg=3.0)
    _, _, obj, _ = hip.case_linearization(case, param="none")
    obj.w_cl, obj.w_cd, obj.w_cm = 1.0, 0.7, -0.3
    with hip.HipMesh(m) as hm:
        out = hm.objective(obj, U)
        # forces of the oracle (forces.cpp:208-294) from the same pressures
        om = ol.OracleMesh(m, np.float64)
        g = 1.4
        inv = 1.0 / U[:, 0]
        u, v = U[:, 1] * inv, U[:, 2] * inv
        p = (g - 1) * (U[:, 4] - 0.5 * U[:, 0] * (u * u + v * v))
        import ctypes as C
        rho_inf, speed = ff[0], np.hypot(ff[1], ff[2]) / ff[0]
        ref7 = np.array([rho_inf, obj.p_inf, obj.aoa_deg, speed, obj.chord, obj.x_ref, obj.y_ref])
        f8 = np.zeros(8)
        ol.oracle().ao_forces_f64(C.byref(om.c), ol.ptr(np.ascontiguousarray(p)), ol.ptr(ref7), 1, ol.ptr(f8))
        Jo = 1.0 * f8[5] + 0.7 * f8[6] - 0.3 * f8[7]
        assert abs(out["J"] - Jo) <= 1e-9 * abs(Jo), (out["J"], Jo)
        g_dev = out["dJdU"]
        wall_cells = np.unique(m.owner[m.wall_start:m.wall_start + m.wall_count])
        rng = np.random.default_rng(3)
        for c in rng.choice(wall_cells, 6, replace=False):
            for k in (0, 1, 2, 4):
                h = 1e-6 * abs(U[c, k]) + 1e-9
                Up, Um = U.copy(), U.copy()
                Up[c, k] += h
                Um[c, k] -= h
                fd = (hm.objective(obj, Up, gradient=False)["J"] - hm.objective(obj, Um, gradient=False)["J"]) / (2 * h)
                assert abs(g_dev[c, k] - fd) <= 1e-6 * np.abs(g_dev[c]).max() + 1e-12, (c, k, g_dev[c, k], fd)
        interior = np.setdiff1d(np.arange(m.num_cells), wall_cells)
        assert np.all(g_dev[interior] == 0.0)
        # explicit partials: dJ/daoa at fixed state, dJ/dq
        h = 1e-4
        o2 = hip.case_linearization(case, param="none")[2]
        o2.w_cl, o2.w_cd, o2.w_cm = 1.0, 0.7, -0.3
        o2.aoa_deg = obj.aoa_deg + h
        Jp = hm.objective(o2, U, gradient=False)["J"]
        o2.aoa_deg = obj.aoa_deg - h
        Jm = hm.objective(o2, U, gradient=False)["J"]
        assert abs(out["dJ_daoa"] - (Jp - Jm) / (2 * h)) <= 1e-7 * abs(out["dJ_daoa"]) + 1e-12


def _converged_case(hip, m, aoa, mach, iters):
    case = hip.CaseConfig(mach=mach, aoa_deg=aoa, flux_scheme=0, limiter=0, cfl_start=0.5, cfl_max=1.5,
                          cfl_ramp_iters=50, startup_first_order_iters=10 ** 6, residual_reduction_target=0.0,
                          force_stability_tol=0.0, min_iterations=1, iterations=iters, low_mach_fix=0)
    return case


def test_sensitivities_adjoint_tangent_and_finite_differences(hip):
    """First-order Rusanov on the 40x8 reference mesh converges to round-off; there dCL/dalpha and
    dCL/dMach from the adjoint equal the tangent solve's (dual consistency) and central differences of
    converged CL."""
    m = helpers.mesh("40x8")
    iters = 9000
    with hip.HipMesh(m) as hm:
        case = _converged_case(hip, m, 2.0, 0.5, iters)
        say("[sens] primal")
        out = hm.run_case(case)
        assert out["history"][-1, 8] / out["history"][0, 8] < 1e-11
        U = out["conserved"]
        results = {}
        for param in ("aoa", "mach"):
            rc, seed, obj, dq = hip.case_linearization(case, param=param)
            obj.w_cl = 1.0
            ob = hm.objective(obj, U, seed=None)
            J_p = ob["dJ_daoa"] if param == "aoa" else ob["dJ_dq"] * dq
            _, R_p = hm.residual_jvp(rc, U, None, seed)
            say(f"[sens] adjoint {param}")
            psi, ai = hm.adjoint_solve(rc, ob["dJdU"], U, cfl=1.5, max_iters=20000, tol=1e-12)
            assert ai["res"] <= 1e-11 * ai["res0"], ai
            d_adj = J_p - float(np.sum(psi * R_p))
            say(f"[sens] tangent {param}")
            v, ti = hm.tangent_solve(rc, seed, U, cfl=1.5, max_iters=20000, tol=1e-12)
            assert ti["res"] <= 1e-11 * ti["res0"], ti
            d_tan = J_p + float(np.sum(ob["dJdU"] * v))
            assert abs(d_adj - d_tan) <= 1e-8 * abs(d_adj), (param, d_adj, d_tan)
            say(f"[sens] adjoint GMRES {param}")
            psi_g, gi = hm.adjoint_solve(rc, ob["dJdU"], U, cfl=1.5, max_iters=3000, tol=1e-12, gmres_restart=30)
            assert gi["res"] <= 1e-12 * gi["res0"] and gi["iterations"] < ai["iterations"], (gi, ai["iterations"])
            assert abs(J_p - float(np.sum(psi_g * R_p)) - d_adj) <= 1e-8 * abs(d_adj)
            results[param] = d_adj
        h = {"aoa": 1e-3, "mach": 1e-4}
        for param in ("aoa", "mach"):
            cls = []
            for s in (1, -1):
                kw = dict(aoa=2.0 + (s * h[param] if param == "aoa" else 0.0),
                          mach=0.5 + (s * h[param] if param == "mach" else 0.0))
                say(f"[sens] finite difference {param} {s:+d}")
                o = hm.run_case(_converged_case(hip, m, kw["aoa"], kw["mach"], iters))
                cls.append(o["cl"])
            fd = (cls[0] - cls[1]) / (2 * h[param])
            say(f"[sens] {param}: adjoint {results[param]:.12e} FD {fd:.12e} rel {abs(results[param] - fd) / abs(fd):.2e}")
            assert abs(results[param] - fd) <= 1e-6 * abs(fd), (param, results[param], fd)


def test_shape_sensitivities(hip):
    """Hicks-Henne amplitudes on a deformed 40x8 mesh (first-order Rusanov, converged to round-off):
    adjoint dCL/da_i equals the tangent solve and central differences of converged CL over
    re-deformed meshes."""
    from aero_hip.adjoint import sensitivities
    base = helpers.mesh("40x8")
    rng = np.random.default_rng(4242)
    amp = rng.uniform(-0.002, 0.002, 20)
    m, tan = hip.deform_mesh(base, amp)
    iters = 9000
    case = _converged_case(hip, m, 2.0, 0.5, iters)
    with hip.HipMesh(m) as hm:
        say("[shape] primal")
        out = hm.run_case(case)
        assert out["history"][-1, 8] / out["history"][0, 8] < 1e-11
        params = [2, 7, 13]
        say("[shape] adjoint + tangent")
        s = sensitivities(hm, case, out["conserved"], shape_tangents=tan, shape_params=params, cfl=1.5, tol=1e-12,
                          tangent=True)
        assert s["adjoint"]["res"] <= 1e-11 * s["adjoint"]["res0"]
        for name in s["d"]:
            assert abs(s["d"][name] - s["d_tangent"][name]) <= 1e-8 * max(abs(s["d"][name]), 1e-12), name
    h = 1e-5
    for i in params:
        cls = []
        for sgn in (1, -1):
            e = np.zeros(20)
            e[i] = sgn * h
            mi, _ = hip.deform_mesh(base, amp + e, tangents=False)
            say(f"[shape] finite difference a{i} {sgn:+d}")
            with hip.HipMesh(mi) as hmi:
                cls.append(hmi.run_case(case)["cl"])
        fd = (cls[0] - cls[1]) / (2 * h)
        say(f"[shape] a{i}: rel {abs(s['d'][f'shape{i}'] - fd) / abs(fd):.2e}")
        assert abs(s["d"][f"shape{i}"] - fd) <= 1e-6 * abs(fd), (i, s["d"][f"shape{i}"], fd)


def _first_order_state(hip, hm, mach, aoa, iters, **kw):
    case = _converged_case(hip, hm.mesh, aoa, mach, iters)
    for k, v in kw.items():
        setattr(case, k, v)
    return case, hm.run_case(case)["conserved"]


@pytest.mark.parametrize("rk_stages", [3, 1])
def test_adjoint_batch_matches_single_solves(hip, rk_stages):
    """Batched adjoint solves (CL and CD on two meshes of different size, one preconditioned; the two solves
    of a mesh share work items, a third solve with its own CFL runs alone) against single solves checked
    after every iteration: bitwise identical under RK3; under RK1, whose single solve checks every second
    iteration, equal to the solve tolerance."""
    m1 = helpers.mesh("40x8")
    m2 = _mesh64()
    with hip.HipMesh(m1) as h1, hip.HipMesh(m2) as h2:
        say("[adj-batch] primal states")
        c1, U1 = _first_order_state(hip, h1, 0.5, 2.0, 2000)
        # preconditioner settings under which the linearised iteration contracts quickly (at the defaults
        # beta_min 0.2, mach_ref 0.15 and M 0.2 its amplification reaches 0.9997 on this mesh)
        c2, U2 = _first_order_state(hip, h2, 0.4, 3.0, 2000, precond_on=1, precond_experimental_enabled=1,
                                    precond_mach_ref=0.3, precond_beta_min=0.5)
        solves = []
        for hm, case, U, cfl in ((h1, c1, U1, None), (h2, c2, U2, 1.2)):  # the second solve has its own CFL
            rc, _, obj, _ = hip.case_linearization(case)
            for w in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0)):
                obj.w_cl, obj.w_cd, obj.w_cm = w
                rhs = hm.objective(obj, U)["dJdU"]
                solves.append((hm, rc, rhs, U, cfl))
        solves.append(solves[0][:4] + (1.3,))  # a third solve on the first mesh, alone (its own CFL)
        kw = dict(rk_stages=rk_stages, max_iters=20000, tol=1e-10)
        say(f"[adj-batch] batched solve rk{rk_stages}")
        psis, infos, _ = hip.adjoint_solve_batch(solves, cfl=1.5, iters_per_sync=150, **kw)
        for i, (hm, rc, rhs, U, cfl) in enumerate(solves):
            say(f"[adj-batch] single solve {i}")
            psi, info = hm.adjoint_solve(rc, rhs, U, check_every=1, cfl=cfl or 1.5, **kw)
            assert infos[i]["converged"] and not infos[i]["failed"], infos[i]
            if rk_stages == 3:
                assert infos[i]["iterations"] == info["iterations"], (i, infos[i], info["iterations"])
                assert np.array_equal(psis[i], psi), f"solve {i} differs from its single solve"
            else:
                assert abs(infos[i]["iterations"] - info["iterations"]) <= 1
                assert _maxrel(psis[i], psi) <= 1e-8


def _mixed(nc, nr):
    return helpers.mesh("256x64", num_circumferential=nc, num_radial=nr, farfield_radius=20.0, radial_stretch=200.0,
                        cells="mixed", quad_layers=30, jitter=0.1, seed=12345, rcm=True)


@pytest.mark.parametrize("gradient", ["green_gauss", "least_squares"])
def test_adjoint_launch_shapes_agree(hip, gradient):
    """The adjoint kernels run four lanes per cell below 1024 work items and one thread per cell from 1024
    up, with the same per-cell sums.  Four solves with their own CFLs on a 66k-cell mesh (4 x 258 work
    items: one thread per cell) against the same solves alone (258 entry groups: four lanes per cell),
    bitwise after 300 RK3 iterations."""
    m = _mixed(256, 144)
    assert m.num_cells == 66048
    ff, U = _state(m, 0.5, 1.25, 0.02, 20260101)
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 0, 0, 0, 0), ff))
    rhs = np.random.default_rng(5).standard_normal(U.shape)
    rhs[:, 3] = 0.0
    kw = dict(rk_stages=3, max_iters=300, tol=1e-30)
    with hip.HipMesh(m, gradient=gradient) as hm:
        solves = [(hm, cfg, rhs, U, cfl) for cfl in (0.5, 0.55, 0.6, 0.65)]
        say(f"[adj-shapes] {gradient}: batched, one thread per cell")
        psis, infos, _ = hip.adjoint_solve_batch(solves, iters_per_sync=100, **kw)
        for i, (_, _, _, _, cfl) in enumerate(solves):
            say(f"[adj-shapes] {gradient}: single solve at CFL {cfl}, four lanes per cell")
            psi, info = hm.adjoint_solve(cfg, rhs, U, cfl=cfl, check_every=1, **kw)
            assert not infos[i]["failed"] and infos[i]["iterations"] == info["iterations"] == 300, (infos[i], info)
            assert np.all(np.isfinite(psi))
            assert np.array_equal(psis[i], psi), f"solve {i}: launch shapes differ"


@pytest.mark.parametrize("gradient", ["green_gauss", "least_squares"])
def test_adjoint_product_one_thread_per_cell(hip, gradient):
    """The adjoint product in its one-thread-per-cell shape (262k cells: 1024 entry groups) against the
    tangent product: psi . (R_U v) = (R_U^T psi) . v."""
    m = _mixed(512, 271)
    assert m.num_cells == 262144
    ff, U = _state(m, 0.5, 1.25, 0.02, 20260101)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(U.shape) * np.abs(U).max(axis=0) * 1e-3
    v[:, 3] = 0.0
    psi = rng.standard_normal(U.shape)
    psi[:, 3] = 0.0
    with hip.HipMesh(m, gradient=gradient) as hm:
        for combo in [(1, 1, 1, 0, 0, 0, 0), (0, 0, 1, 1, 1, 1, 1)]:
            cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
            say(f"[adj-262k] {gradient} combo={combo}")
            _, Jv = hm.residual_jvp(cfg, U, v)
            JTpsi = hm.residual_vjp(cfg, psi, U)
            a = float(np.sum(psi * Jv))
            b = float(np.sum(JTpsi * v))
            scale = float(np.sum(np.abs(psi * Jv)))
            assert abs(a - b) <= 1e-12 * scale, (combo, a, b)


def test_sensitivities_batch_matches_single_case_sensitivities(hip):
    """sensitivities_batch (one batched solve for every case and objective) against sensitivities per
    case and objective, with shape tangents on a deformed mesh."""
    from aero_hip.adjoint import sensitivities, sensitivities_batch
    base = helpers.mesh("40x8")
    amp = np.random.default_rng(7).uniform(-0.002, 0.002, 20)
    m, tan = hip.deform_mesh(base, amp)
    with hip.HipMesh(m) as h1, hip.HipMesh(base) as h2:
        say("[sens-batch] primal states")
        c1, U1 = _first_order_state(hip, h1, 0.5, 2.0, 3000)
        c2, U2 = _first_order_state(hip, h2, 0.4, 1.0, 3000)
        items = [dict(hm=h1, case=c1, U=U1, shape_tangents=tan), dict(hm=h2, case=c2, U=U2)]
        say("[sens-batch] batched")
        sb = sensitivities_batch(items, cfl=1.5, max_iters=20000, tol=1e-12)
        for it, res in zip(items, sb["cases"]):
            for name, w in (("cl", (1.0, 0.0, 0.0)), ("cd", (0.0, 1.0, 0.0))):
                say(f"[sens-batch] single {name}")
                s = sensitivities(it["hm"], it["case"], it["U"], weights=w, shape_tangents=it.get("shape_tangents"),
                                  cfl=1.5, max_iters=20000, tol=1e-12)
                assert res["J"][name] == pytest.approx(s["J"], rel=1e-13)
                assert set(res["d"][name]) == set(s["d"])
                for p, v in s["d"].items():
                    assert abs(res["d"][name][p] - v) <= 1e-8 * max(abs(v), 1e-10), (name, p, res["d"][name][p], v)

