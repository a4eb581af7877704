"""The torch.autograd.Function layer over the device residual and adjoint."""
import sys

import numpy as np
import pytest

import cases
import helpers

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available()
    return aero_hip


def test_residual_autograd(hip):
    import torch
    from aero_hip.torch_ops import residual
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    from aero_hip import freestream
    ff = freestream(0.5, 2.0)[1]
    U0 = helpers.perturbed_state(m.cell_center, ff, 0.05, 17)
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 1, 0, 0, 0), ff))
    rng = np.random.default_rng(3)
    w = torch.from_numpy(rng.standard_normal(U0.shape))
    with hip.HipMesh(m) as hm:
        U = torch.from_numpy(U0.copy()).requires_grad_(True)
        say("[torch] residual forward / backward")
        loss = (w * residual(U, hm, cfg)).sum()
        loss.backward()
        ref = hm.residual_vjp(cfg, w.numpy(), U0)
        assert np.array_equal(U.grad.numpy(), ref)
        d = rng.standard_normal(U0.shape) * np.abs(U0).max(axis=0) * 1e-3
        d[:, 3] = 0.0
        h = 1e-5
        lp = float((w * residual(torch.from_numpy(U0 + h * d), hm, cfg)).sum())
        lm = float((w * residual(torch.from_numpy(U0 - h * d), hm, cfg)).sum())
        fd = (lp - lm) / (2 * h)
        an = float((U.grad * torch.from_numpy(d)).sum())
        assert abs(an - fd) <= 1e-6 * abs(fd), (an, fd)


def test_residual_autograd_on_device_tensors(hip):
    """With the state on the GPU the Function keeps every array there (the device-pointer entry points) and
    gives the host path's residual and R_U^T w bit for bit."""
    import torch
    from aero_hip.torch_ops import residual
    m = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    from aero_hip import freestream
    ff = freestream(0.5, 2.0)[1]
    U0 = helpers.perturbed_state(m.cell_center, ff, 0.05, 17)
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 1, 0, 0, 0), ff))
    rng = np.random.default_rng(4)
    w0 = rng.standard_normal(U0.shape)
    with hip.HipMesh(m) as hm:
        say("[torch] residual forward / backward on cuda tensors")
        U = torch.from_numpy(U0.copy()).to("cuda").requires_grad_(True)
        w = torch.from_numpy(w0).to("cuda")
        R = residual(U, hm, cfg)
        assert R.is_cuda
        (w * R).sum().backward()
        assert U.grad.is_cuda
        Rh, _, _ = hm.residual(U0, cfg)
        ref = hm.residual_vjp(cfg, w0, U0)
        assert np.array_equal(R.detach().cpu().numpy(), Rh)
        assert np.array_equal(U.grad.cpu().numpy(), ref)


def test_steady_coefficients_gradient(hip):
    """d(CL + 0.5 CD)/d(aoa, mach, amplitudes) through the autograd Function against central differences of
    the converged coefficients (first-order Rusanov on the 40x8 mesh, converged to round-off)."""
    import torch
    from aero_hip.torch_ops import steady_coefficients
    base = helpers.mesh("40x8")
    case = hip.CaseConfig(flux_scheme=0, limiter=0, cfl_start=0.5, cfl_max=1.5, cfl_ramp_iters=50,
                          startup_first_order_iters=10 ** 6, residual_reduction_target=0.0, force_stability_tol=0.0,
                          min_iterations=1, iterations=9000, low_mach_fix=0)
    rng = np.random.default_rng(4242)
    p0 = np.concatenate([[2.0, 0.5], rng.uniform(-0.002, 0.002, 20)])
    solve_kw = dict(cfl=1.5, tol=1e-12, max_iters=20000)
    params = torch.tensor(p0, requires_grad=True)
    say("[torch] steady coefficients forward")
    c = steady_coefficients(params, base, case, solve_kw=solve_kw)
    say("[torch] backward (adjoint)")
    (c[0] + 0.5 * c[1]).backward()
    g = params.grad.numpy()
    for i, h in ((0, 1e-3), (1, 1e-4), (7, 1e-5), (14, 1e-5)):
        vals = []
        for sgn in (1, -1):
            p = p0.copy()
            p[i] += sgn * h
            say(f"[torch] finite difference parameter {i} {sgn:+d}")
            with torch.no_grad():
                cc = steady_coefficients(torch.tensor(p), base, case, solve_kw=solve_kw)
            vals.append(float(cc[0] + 0.5 * cc[1]))
        fd = (vals[0] - vals[1]) / (2 * h)
        say(f"[torch] parameter {i}: rel {abs(g[i] - fd) / abs(fd):.2e}")
        assert abs(g[i] - fd) <= 1e-6 * abs(fd), (i, g[i], fd)


def test_steady_coefficients_batch_primal_and_two_backwards(hip):
    """The batch primal of the Function (a one-case device batch to a density drop) against the driver
    primal, and two backward passes over one retained graph (dCL, then dCD)."""
    import torch
    from aero_hip.torch_ops import last_run, steady_coefficients
    base = helpers.mesh("40x8")
    case = hip.CaseConfig(flux_scheme=0, limiter=0, cfl_start=0.5, cfl_max=1.5, cfl_ramp_iters=50,
                          startup_first_order_iters=10 ** 6, residual_reduction_target=0.0, force_stability_tol=0.0,
                          min_iterations=1, iterations=9000, low_mach_fix=0)
    p0 = np.concatenate([[2.0, 0.5], np.random.default_rng(5).uniform(-0.002, 0.002, 4)])
    grads = {}
    for primal, jac in (("driver", False), ("batch", False), ("batch", True)):
        params = torch.tensor(p0, requires_grad=True)
        say(f"[torch] {primal} primal, jacobian={jac}")
        c = steady_coefficients(params, base, case, solve_kw=dict(primal=primal, rho_drop=1e-11, cfl=1.5,
                                                                   tol=1e-12, max_iters=20000, jacobian=jac))
        if primal == "batch":
            assert last_run["primal_converged"] and last_run["primal_rho_ratio"] <= 1e-11
        (g_cl,) = torch.autograd.grad(c[0], params, retain_graph=True)
        (g_cd,) = torch.autograd.grad(c[1], params)
        if jac:
            assert last_run["adjoint_solves"] == "cached"  # the second row came from the first solve
        grads[(primal, jac)] = (c.detach().numpy(), g_cl.numpy(), g_cd.numpy())
    (cd_, gl_d, gd_d), (cb, gl_b, gd_b) = grads[("driver", False)], grads[("batch", False)]
    assert np.allclose(cb, cd_, rtol=1e-9)
    assert np.allclose(gl_b, gl_d, rtol=1e-6, atol=1e-9) and np.allclose(gd_b, gd_d, rtol=1e-6, atol=1e-9)
    assert not np.allclose(gl_b, gd_b)
    _, gl_j, gd_j = grads[("batch", True)]
    assert np.allclose(gl_j, gl_b, rtol=1e-8, atol=1e-11) and np.allclose(gd_j, gd_b, rtol=1e-8, atol=1e-11)
