"""torch.autograd.Function wrappers over the device residual and its discrete adjoint.

* ``residual(U, hm, cfg)``: R(U) [N, 5] with backward R_U^T grad (aero_hip_residual_vjp), so the residual
  can sit inside any torch computation.
* ``steady_coefficients(params, base_mesh, case, ...)``: (CL, CD) of the converged solution for
  params = [aoa_deg, mach, a_1 .. a_P] (Hicks-Henne amplitudes deforming ``base_mesh``); the backward
  pass runs one adjoint solve for J = grad_CL * CL + grad_CD * CD and returns dJ/dparams.  The primal is
  the device driver (``primal="driver"``: aero_hip_run_case with the case's own stopping test) or a
  one-case device batch to a density-residual drop (``primal="batch"``, ``rho_drop``), which also carries
  a frozen limiter's ratios over to the adjoint.  With ``jacobian=True`` the first backward solves the CL
  and CD adjoints as one batched pair and keeps both gradient rows for further backward passes.

``residual`` keeps GPU tensors on the GPU: their FP64 storage goes to the C ABI's device-pointer entry
points (aero_hip_residual_device / aero_hip_residual_vjp_device) after the caller's stream is
synchronised; CPU tensors cross as FP64 host arrays.  ``steady_coefficients`` takes its few parameters on
any device.  Only the Python layer uses torch: the numerics are the HIP kernels behind ``aero_hip``.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import torch

from .adjoint import sensitivities, sensitivities_batch
from .mesh import deform_mesh
from .solver import CaseConfig, HipBatch, HipMesh

# Timings of the last SteadyCoefficients forward (primal) and backward (adjoint) on this process.
last_run: dict = {}


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).numpy()


def _dev64(t: torch.Tensor) -> torch.Tensor:
    """t as a contiguous FP64 tensor on its GPU, its producers finished (the C ABI's device entry points
    run on the mesh handle's own stream)."""
    t = t.detach().to(torch.float64).contiguous()
    torch.cuda.current_stream(t.device).synchronize()
    return t


class ResidualFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, hm, cfg):
        ctx.hm, ctx.cfg = hm, cfg
        ctx.save_for_backward(U)
        if U.is_cuda:
            u = _dev64(U)
            R = torch.empty_like(u)
            hm.residual_device(u.data_ptr(), cfg, R.data_ptr())
            return R.to(U.dtype)
        R, _, _ = hm.residual(_np(U), cfg)
        return torch.from_numpy(R).to(U.device)

    @staticmethod
    def backward(ctx, grad_R):
        (U,) = ctx.saved_tensors
        if U.is_cuda:
            u, g = _dev64(U), _dev64(grad_R)
            out = torch.empty_like(u)
            ctx.hm.residual_vjp_device(ctx.cfg, g.data_ptr(), u.data_ptr(), out.data_ptr())
            return out.to(grad_R.dtype), None, None
        out = ctx.hm.residual_vjp(ctx.cfg, _np(grad_R), _np(U))
        return torch.from_numpy(out).to(U.device, grad_R.dtype), None, None


def residual(U: torch.Tensor, hm: HipMesh, cfg) -> torch.Tensor:
    """Residual R(U) of ``hm``'s mesh under ``cfg`` as a differentiable torch op."""
    return ResidualFunction.apply(U, hm, cfg)


class SteadyCoefficients(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, base_mesh, case, shape_kw, solve_kw):
        p = _np(params)
        case_p = replace(case, aoa_deg=float(p[0]), mach=float(p[1]))
        n_shape = len(p) - 2
        tan = None
        mesh = base_mesh
        if n_shape > 0:
            mesh, tan = deform_mesh(base_mesh, p[2:], bumps_per_surface=n_shape // 2, **shape_kw)
        hm = HipMesh(mesh)
        kw = dict(solve_kw)
        primal = kw.pop("primal", "driver")
        rho_drop = kw.pop("rho_drop", 1e-8)
        if primal == "batch":
            with HipBatch([hm], [(0, case_p.mach, case_p.aoa_deg)], case_p) as b:
                res = b.run(case_p.iterations, rho_drop, iters_per_sync=500)
                row = res["cases"][0]
                last_run.update(primal_seconds=res["seconds"], primal_iterations=row["iterations"],
                                primal_rho_ratio=row["rho_ratio"], primal_converged=bool(row["converged"]))
                U = b.state(0)
                if case_p.limiter_freeze_iter >= 0:
                    hm.set_limiter_ratios(b.limiter_ratios(0))
            case_p = replace(case_p, iterations=max(row["iterations"], 1))
            cl, cd = row["cl"], row["cd"]
        elif primal == "driver":
            out = hm.run_case(case_p)
            U, cl, cd = out["conserved"], out["cl"], out["cd"]
            last_run.update(primal_seconds=out["seconds"], primal_iterations=out["n_iter"])
        else:
            raise ValueError("primal must be 'driver' or 'batch'")
        ctx.saved = (hm, case_p, U, tan, n_shape, kw)
        ctx.device = params.device
        return torch.tensor([cl, cd], dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(ctx, grad):
        hm, case_p, U, tan, n_shape, solve_kw = ctx.saved
        g = _np(grad)
        names = ["aoa", "mach"] + [f"shape{i}" for i in range(n_shape)]
        if solve_kw.get("jacobian", False):
            # Both rows of d(CL, CD)/dparams from one batched solve of the two adjoints (they share the
            # face Jacobians), kept on ctx: a later backward over the same graph needs no solve.
            if getattr(ctx, "jac", None) is None:
                kw = {k: v for k, v in solve_kw.items() if k in ("cfl", "max_iters", "tol")}
                sb = sensitivities_batch([dict(hm=hm, case=case_p, U=U, shape_tangents=tan)], **kw)
                c = sb["cases"][0]
                ctx.jac = np.array([[c["d"]["cl"][n] for n in names], [c["d"]["cd"][n] for n in names]])
                last_run.update(adjoint_seconds=sb["solve_seconds"],
                                adjoint_iterations=max(a["iterations"] for a in c["adjoint"].values()),
                                adjoint_res0=c["adjoint"]["cl"]["res0"], adjoint_res=c["adjoint"]["cl"]["res"],
                                adjoint_solves="cl+cd")
            else:
                last_run.update(adjoint_seconds=0.0, adjoint_solves="cached")
            d = (g[0] * ctx.jac[0] + g[1] * ctx.jac[1]).tolist()
            return torch.tensor(d, dtype=torch.float64, device=ctx.device), None, None, None, None
        kw = {k: v for k, v in solve_kw.items() if k != "jacobian"}
        s = sensitivities(hm, case_p, U, weights=(float(g[0]), float(g[1]), 0.0), shape_tangents=tan, **kw)
        last_run.update(adjoint_seconds=s["adjoint_seconds"], adjoint_iterations=s["adjoint"]["iterations"],
                        adjoint_res0=s["adjoint"]["res0"], adjoint_res=s["adjoint"]["res"], adjoint_solves="weighted")
        d = [s["d"][n] for n in names]
        # the device mesh stays with ctx (freed with the graph), so a retained graph can be differentiated
        # again, e.g. once for CL and once for CD
        return torch.tensor(d, dtype=torch.float64, device=ctx.device), None, None, None, None


def steady_coefficients(params: torch.Tensor, base_mesh, case: CaseConfig, shape_kw: dict | None = None,
                        solve_kw: dict | None = None) -> torch.Tensor:
    """(CL, CD) of the converged case at params = [aoa_deg, mach, amplitudes...], differentiable through
    the discrete adjoint.  shape_kw: deform_mesh options (h_min, h_max, idw_exponent); solve_kw:
    primal ("driver" or "batch"), rho_drop (batch primal), jacobian (True: the first backward solves the
    CL and CD adjoints together and later backwards reuse them) and the sensitivities options (cfl, tol,
    max_iters, gmres_restart, ...)."""
    return SteadyCoefficients.apply(params, base_mesh, case, shape_kw or {}, solve_kw or {})
