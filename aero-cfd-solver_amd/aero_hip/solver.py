"""Host-side mirror of the reference's Euler backend interface over the HIP C ABI.

`HipMesh` plays the role of cfd::cuda_backend::EulerDeviceBuffers + its RAII guard
(src/cfd_core/solvers/euler_solver.cpp:452-460): it owns the device mesh, state and stream.
`CaseConfig` mirrors EulerAirfoilCaseConfig (src/cfd_core/include/cfd_core/solvers/euler_solver.hpp:14-74)
with the reference's defaults.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import asdict, dataclass, fields

import numpy as np

from . import _lib
from .mesh import Mesh


@dataclass
class CaseConfig:
    iterations: int = 400
    min_iterations: int = 40
    cfl_ramp_iters: int = 120
    gamma: float = 1.4
    gas_constant: float = 287.05
    mach: float = 0.15
    aoa_deg: float = 2.0
    p_inf: float = 101325.0
    t_inf: float = 288.15
    rho_inf: float = 0.0
    cfl_start: float = 0.2
    cfl_max: float = 1.2
    residual_reduction_target: float = 1.0e-3
    force_stability_tol: float = 2.0e-5
    force_stability_window: int = 6
    force_mean_drift_tol: float = 2.0e-5
    startup_first_order_iters: int = 25
    rk_stages: int = 3
    local_time_stepping: int = 1
    precond_on: int = 0
    precond_mach_ref: float = 0.15
    precond_mach_min: float = 1.0e-3
    precond_beta_min: float = 0.2
    precond_beta_max: float = 1.0
    precond_farfield_bc: int = 0
    stab_floor_start: float = 0.1
    stab_floor_target: float = 0.02
    stab_floor_k_start: float = -1.0
    stab_floor_k_target: float = -1.0
    stab_ramp_iters: int = 400
    x_ref: float = 0.25
    y_ref: float = 0.0
    chord: float = 1.0
    flux_scheme: int = 0
    limiter: int = 0
    low_mach_fix: int = 0
    mach_cutoff: float = 0.3
    all_speed_flux_fix: int = 0
    all_speed_mach_cutoff: float = 0.25
    all_speed_f_min: float = 0.05
    all_speed_ramp_start_iter: int = 0
    all_speed_ramp_iters: int = 0
    all_speed_staged_controller: int = 0
    all_speed_stage_a_min_iters: int = 600
    all_speed_stage_a_max_iters: int = 0
    all_speed_stage_f_min_high: float = 0.2
    all_speed_stage_f_min_mid: float = 0.1
    all_speed_stage_f_min_low: float = 0.05
    all_speed_pjump_wall_target: float = 0.0
    all_speed_pjump_spike_factor: float = 1.5
    all_speed_pjump_hold_iters: int = 120
    all_speed_freeze_iters: int = 120
    all_speed_cfl_drop_factor: float = 0.8
    all_speed_cfl_restore_factor: float = 1.02
    all_speed_cfl_min_scale: float = 0.1
    aoa0_symmetry_enforce: int = 0
    aoa0_symmetry_enforce_interval: int = 0
    precond_experimental_enabled: int = -1  # -1: read AERO_ENABLE_EXPERIMENTAL_PRECOND like the reference
    limiter_freeze_iter: int = -1  # beyond the reference: freeze the limiter at this iteration (-1: never)

    def resolved(self) -> "CaseConfig":
        c = CaseConfig(**asdict(self))
        if c.precond_experimental_enabled < 0:  # euler_solver.cpp:38-45, :482-489
            v = os.environ.get("AERO_ENABLE_EXPERIMENTAL_PRECOND", "")
            c.precond_experimental_enabled = 1 if v in ("1", "true", "TRUE", "on", "ON") else 0
        return c

    def to_c(self) -> _lib.CaseConfigC:
        r = self.resolved()
        return _lib.CaseConfigC(*[getattr(r, name) for name, _ in _lib.CASE_FIELDS])


def freestream(mach: float, aoa_deg: float, p_inf: float = 101325.0, t_inf: float = 288.15, gamma: float = 1.4,
               gas_constant: float = 287.05, rho_inf: float = 0.0):
    """Freestream primitive and conservative states (euler_solver.cpp:543-561, euler_flux.cpp:101-113)."""
    rho = rho_inf if rho_inf > 0.0 else p_inf / (gas_constant * t_inf)
    alpha = aoa_deg * (math.pi / 180.0)
    a = math.sqrt(gamma * p_inf / rho)
    v = max(mach, 0.0) * a
    u, w = v * math.cos(alpha), v * math.sin(alpha)
    r = max(rho, 1e-8)
    ke = 0.5 * r * (u * u + w * w + 0.0 * 0.0)
    cons = np.array([r, r * u, r * w, r * 0.0, max(p_inf, 1e-8) / (gamma - 1.0) + ke])
    return np.array([rho, u, w, 0.0, p_inf]), cons


def residual_config(farfield, gamma=1.4, flux="hllc", limiter="minmod", second_order=True, low_mach_fix=False,
                    low_mach_cutoff=0.3, stab_mach_floor=0.0, precond_farfield_bc=False, precond=False,
                    mach_ref=0.15, mach_min=1e-3, beta_min=0.2, beta_max=1.0, all_speed_fix=False,
                    all_speed_cutoff=0.25, all_speed_f_min=0.05, all_speed_ramp=1.0,
                    limiter_frozen=False) -> _lib.ResidualConfig:
    """EulerResidualAssemblyConfig scalars (solvers/euler/residual_assembly.hpp:12-31)."""
    f = lambda v: _lib.FLUX[v] if isinstance(v, str) else int(v)
    l = lambda v: _lib.LIMITER[v] if isinstance(v, str) else int(v)
    return _lib.ResidualConfig((C.c_double * 5)(*[float(x) for x in farfield]), gamma, f(flux), l(limiter),
                               int(bool(second_order)), int(bool(low_mach_fix)), low_mach_cutoff, stab_mach_floor,
                               int(bool(precond_farfield_bc)), int(bool(precond)), mach_ref, mach_min, beta_min,
                               beta_max, int(bool(all_speed_fix)), all_speed_cutoff, all_speed_f_min, all_speed_ramp,
                               int(bool(limiter_frozen)))


def _pd(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def hip_available() -> bool:
    err = _lib.errbuf()
    return bool(_lib.lib().aero_hip_available(err, len(err)))


class HipMesh:
    """A mesh resident on the GPU with its state buffers and stream."""

    GRADIENTS = {"green_gauss": 0, "least_squares": 1}

    ACCUMULATIONS = {"auto": 0, "gather": 1, "colour": 2}

    def __init__(self, mesh: Mesh, renumber: bool = True, block_cells: int = 256, gradient: str = "green_gauss",
                 accumulation: str = "auto"):
        """block_cells 128, 256 or 512, or 0 for the layout probe (aero_hip_options); accumulation "gather",
        "colour" or "auto" (gather, or the probe's choice with block_cells = 0)."""
        self.lib = _lib.lib()
        self.mesh = mesh
        self.N, self.F = mesh.num_cells, mesh.num_faces
        self._desc = mesh.desc()
        if gradient not in self.GRADIENTS:
            raise ValueError(f"gradient must be one of {sorted(self.GRADIENTS)}")
        if accumulation not in self.ACCUMULATIONS:
            raise ValueError(f"accumulation must be one of {sorted(self.ACCUMULATIONS)}")
        self.gradient = gradient
        opts = _lib.Options(1 if renumber else 0, block_cells, self.GRADIENTS[gradient],
                            self.ACCUMULATIONS[accumulation])
        h = C.c_void_p()
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_mesh_create(C.byref(self._desc), C.byref(opts), C.byref(h), err, len(err)),
                   err, "aero_hip_mesh_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.aero_hip_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _call(self, name, *args):
        err = _lib.errbuf()
        _lib.check(getattr(self.lib, name)(self.h, *args, err, len(err)), err, name)

    def residual(self, U: np.ndarray, cfg: _lib.ResidualConfig):
        """One residual evaluation: U[N,5] -> (R[N,5], lambda[N], diag dict)."""
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.N, 5)
        R = np.zeros((self.N, 5))
        lam = np.zeros(self.N)
        d = _lib.ResidualDiag()
        self._call("aero_hip_residual", C.byref(cfg), _pd(U), _pd(R), _pd(lam), C.byref(d))
        return R, lam, dict(first_order_fallback=d.first_order_fallback, diffusive_fallback=d.diffusive_fallback,
                            first_failure_face=d.first_failure_face)

    def residual_device(self, U_ptr: int, cfg: _lib.ResidualConfig, R_ptr: int, lam_ptr: int = 0) -> dict:
        """The residual of device arrays (aero_hip_residual_device): U_ptr, R_ptr (and lam_ptr, 0 = none)
        are device addresses of [N, 5] / [N] FP64 arrays on this handle's GPU, inputs ready."""
        d = _lib.ResidualDiag()
        self._call("aero_hip_residual_device", C.byref(cfg), C.c_void_p(U_ptr), C.c_void_p(R_ptr),
                   C.c_void_p(lam_ptr) if lam_ptr else None, C.byref(d))
        return dict(first_order_fallback=d.first_order_fallback, diffusive_fallback=d.diffusive_fallback,
                    first_failure_face=d.first_failure_face)

    def residual_vjp_device(self, cfg: _lib.ResidualConfig, psi_ptr: int, U_ptr: int, out_ptr: int):
        """R_U^T psi of device arrays (aero_hip_residual_vjp_device), [N, 5] FP64 device addresses."""
        self._call("aero_hip_residual_vjp_device", C.byref(cfg), C.c_void_p(U_ptr), C.c_void_p(psi_ptr),
                   C.c_void_p(out_ptr))

    def set_state(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.N, 5)
        self._call("aero_hip_set_state", _pd(U))

    def residual_async(self, cfg: _lib.ResidualConfig):
        self._call("aero_hip_residual_async", C.byref(cfg))

    def get_residual(self):
        R = np.zeros((self.N, 5))
        lam = np.zeros(self.N)
        d = _lib.ResidualDiag()
        self._call("aero_hip_get_residual", _pd(R), _pd(lam), C.byref(d))
        return R, lam, dict(first_order_fallback=d.first_order_fallback, diffusive_fallback=d.diffusive_fallback,
                            first_failure_face=d.first_failure_face)

    def state(self) -> np.ndarray:
        """The handle's current conserved state [N, 5] (reference numbering)."""
        U = np.zeros((self.N, 5))
        self._call("aero_hip_get_state", _pd(U))
        return U

    def gradients(self) -> np.ndarray:
        g = np.zeros((self.N, 5, 2))
        self._call("aero_hip_get_gradients", _pd(g))
        return g

    def timing_begin(self, max_evals: int):
        self._call("aero_hip_timing_begin", max_evals)

    def timing_end(self) -> dict:
        g, f, sp = C.c_double(), C.c_double(), C.c_double()
        n = C.c_int32()
        self._call("aero_hip_timing_end", C.byref(g), C.byref(f), C.byref(sp), C.byref(n))
        return dict(gradient_ms=g.value, flux_ms=f.value, span_ms=sp.value, n=n.value)

    def layout(self) -> dict:
        a = (C.c_int64 * 8)()
        self.lib.aero_hip_layout(self.h, a)
        acc = {1: "gather", 2: "colour"}.get(a[6], "gather")
        return dict(cells=a[0], faces=a[1], blocks=a[2], block_cells=a[3], slots=a[4], max_slots=a[5], grid=a[7],
                    accumulation=acc)

    def layout_probe(self) -> dict:
        """The layout probe's ms per residual evaluation (block_cells = 0), None where not probed."""
        ms = np.zeros(4)
        self.lib.aero_hip_layout_probe(self.h, _pd(ms))
        keys = ("gather128", "gather256", "gather512", "colour")
        return {k: (float(v) if v >= 0 else None) for k, v in zip(keys, ms)}

    def stream(self) -> int:
        return self.lib.aero_hip_stream(self.h) or 0

    def stage(self, cfg: _lib.ResidualConfig, stage: int, last_stage: bool, cfl: float, lts: bool = True,
              write_residual: bool = False):
        self._call("aero_hip_stage", C.byref(cfg), stage, int(last_stage), cfl, int(lts), int(write_residual))

    def run_case(self, case: CaseConfig, hist_capacity: int | None = None, diagnostics: bool = False) -> dict:
        """run_euler_airfoil_case on the device (euler_solver.cpp:463-1208); diagnostics: also the
        per-iteration device diagnostics of euler_diagnostics.csv ([n_iter, DIAG_FIELDS])."""
        cap = hist_capacity or case.iterations
        hist = np.zeros((cap, _lib.HIST_FIELDS))
        diag = np.zeros((cap, _lib.DIAG_FIELDS)) if diagnostics else None
        U = np.zeros((self.N, 5))
        R = np.zeros((self.N, 5))
        lam = np.zeros(self.N)
        forces = np.zeros(8)
        cp = np.zeros(max(self.mesh.wall_count, 1))
        final = _lib.ResidualConfig()
        out = _lib.RunOutput(cap, 0, 0, _pd(hist), _pd(U), _pd(R), _pd(lam), _pd(forces), _pd(cp), 0.0,
                             C.pointer(final), _pd(diag) if diagnostics else None)
        cc = case.to_c()
        self._call("aero_hip_run_case", C.byref(cc), C.byref(out))
        n = out.n_iter
        return dict(history=hist[:n].copy(), n_iter=n, nonfinite=bool(out.nonfinite), conserved=U, residual=R,
                    spectral_radius=lam, forces=forces, cl=forces[5], cd=forces[6], cm=forces[7],
                    wall_cp=cp[: self.mesh.wall_count], seconds=out.device_seconds, final_config=final,
                    diagnostics=diag[:n].copy() if diagnostics else None)

    def face_scales(self, cfg: _lib.ResidualConfig, U=None) -> tuple[np.ndarray, np.ndarray]:
        """The post-run scaling diagnostics at U (None: the handle's state) under cfg (aero_hip_face_scales):
        per-face records [FS_FIELDS][F] in face order and per-cell (beta, local Mach) [2][N]."""
        _, pu = self._state_ptr(U)
        fs = np.zeros((_lib.FS_FIELDS, self.mesh.num_faces))
        cs = np.zeros((2, self.N))
        self._call("aero_hip_face_scales", C.byref(cfg), pu, _pd(fs), _pd(cs))
        return fs, cs


    def forces(self, rho_inf: float, p_inf: float, aoa_deg: float, speed_inf: float, chord: float = 1.0,
               x_ref: float = 0.25, y_ref: float = 0.0, gauge: bool = True, gamma: float = 1.4, U=None) -> dict:
        """Wall pressure forces and Cp at U (None: the handle's state), integrate_pressure_forces /
        extract_wall_cp (aero_hip_forces)."""
        _, pu = self._state_ptr(U)
        ref = _lib.FreestreamRef(rho_inf, p_inf, aoa_deg, speed_inf, chord, x_ref, y_ref)
        f = np.zeros(8)
        cp = np.zeros(max(self.mesh.wall_count, 1))
        self._call("aero_hip_forces", pu, C.byref(ref), int(bool(gauge)), gamma, _pd(f), _pd(cp))
        return dict(fx=f[0], fy=f[1], lift=f[2], drag=f[3], moment=f[4], cl=f[5], cd=f[6], cm=f[7],
                    wall_cp=cp[: self.mesh.wall_count])

    def freeze_limiter(self, cfg: _lib.ResidualConfig, U=None) -> np.ndarray:
        """Capture the limiter ratios at U (None: the handle's state); residual configs with
        limiter_frozen = 1 use them.  Returns the [8, F] ratios."""
        _, pu = self._state_ptr(U)
        out = np.zeros((8, self.F))
        self._call("aero_hip_freeze_limiter", C.byref(cfg), pu, _pd(out))
        return out

    def iterate(self, cfg: _lib.ResidualConfig, n_iters: int, cfl: float, rk_stages: int = 3,
                history: bool = True, noalloc: bool = False):
        """n fixed-configuration pseudo-time iterations on the handle's state (aero_hip_iterate); returns the
        density-residual L2 after each iteration, or None.  noalloc: the allocation-free entry point
        (aero_hip_iterate_noalloc, at most AERO_NOALLOC_MAX_ITERS iterations with a history)."""
        h = np.zeros(max(n_iters, 1)) if history else None
        name = "aero_hip_iterate_noalloc" if noalloc else "aero_hip_iterate"
        self._call(name, C.byref(cfg), n_iters, rk_stages, cfl, _pd(h) if history else None)
        return h[:n_iters] if history else None

    def set_limiter_ratios(self, ratios: np.ndarray):
        """Install frozen limiter ratios [8, F] (e.g. a batch case's)."""
        r = np.ascontiguousarray(ratios, dtype=np.float64).reshape(8, self.F)
        self._call("aero_hip_set_limiter_ratios", _pd(r))

    # ---- tangent / adjoint ------------------------------------------------------------------------
    def _state_ptr(self, U):
        if U is None:
            return None, None
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.N, 5)
        return U, _pd(U)

    def residual_jvp(self, cfg: _lib.ResidualConfig, U=None, dU=None, seed: "Seed | None" = None):
        """(R, dR) with dR = R_U dU + R_p . seed at U (None: the handle's current state)."""
        _, pu = self._state_ptr(U)
        dUa, pdu = self._state_ptr(dU)
        R = np.zeros((self.N, 5))
        dR = np.zeros((self.N, 5))
        self._call("aero_hip_residual_jvp", C.byref(cfg), pu, pdu, seed.ptr() if seed else None, _pd(R), _pd(dR))
        return R, dR

    def residual_vjp(self, cfg: _lib.ResidualConfig, psi, U=None):
        """R_U^T psi at U (None: the handle's current state)."""
        _, pu = self._state_ptr(U)
        psi = np.ascontiguousarray(psi, dtype=np.float64).reshape(self.N, 5)
        out = np.zeros((self.N, 5))
        self._call("aero_hip_residual_vjp", C.byref(cfg), pu, _pd(psi), _pd(out))
        return out

    def objective(self, obj: _lib.Objective, U=None, seed: "Seed | None" = None, gradient: bool = True) -> dict:
        """J = w_cl CL + w_cd CD + w_cm CM of the gauge wall pressure forces, dJ/dU and explicit partials."""
        _, pu = self._state_ptr(U)
        J, da, dq, ds = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        g = np.zeros((self.N, 5)) if gradient else None
        self._call("aero_hip_objective", C.byref(obj), pu, C.byref(J), _pd(g) if gradient else None, C.byref(da),
                   C.byref(dq), seed.ptr() if seed else None, C.byref(ds))
        return dict(J=J.value, dJdU=g, dJ_daoa=da.value, dJ_dq=dq.value, dJ_seed=ds.value)

    def adjoint_solve(self, cfg: _lib.ResidualConfig, rhs, U=None, cfl: float = 1.0, rk_stages: int = 3,
                      max_iters: int = 5000, tol: float = 1e-10, check_every: int = 100, gmres_restart: int = 0,
                      gmres_sweeps: int = 1):
        """psi with R_U^T psi = rhs (transposed pseudo-time iteration, GMRES-accelerated when
        gmres_restart > 0)."""
        _, pu = self._state_ptr(U)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(self.N, 5)
        ctl = _lib.LinearSolve(rk_stages, cfl, max_iters, tol, check_every, gmres_restart, gmres_sweeps, 0, 0.0,
                               0.0, 0.0)
        psi = np.zeros((self.N, 5))
        hist = np.zeros(max_iters)
        self._call("aero_hip_adjoint_solve", C.byref(cfg), pu, _pd(rhs), C.byref(ctl), _pd(psi), _pd(hist))
        return psi, dict(iterations=ctl.iterations, res0=ctl.res0, res=ctl.res, seconds=ctl.seconds,
                         history=hist[: ctl.iterations].copy())

    def tangent_solve(self, cfg: _lib.ResidualConfig, seed: "Seed", U=None, cfl: float = 1.0, rk_stages: int = 3,
                      max_iters: int = 5000, tol: float = 1e-10, check_every: int = 100, gmres_restart: int = 0,
                      gmres_sweeps: int = 1):
        """v with R_U v = -R_p . seed (linearised pseudo-time iteration, GMRES-accelerated when
        gmres_restart > 0)."""
        _, pu = self._state_ptr(U)
        ctl = _lib.LinearSolve(rk_stages, cfl, max_iters, tol, check_every, gmres_restart, gmres_sweeps, 0, 0.0,
                               0.0, 0.0)
        v = np.zeros((self.N, 5))
        hist = np.zeros(max_iters)
        self._call("aero_hip_tangent_solve", C.byref(cfg), pu, seed.ptr(), C.byref(ctl), _pd(v), _pd(hist))
        return v, dict(iterations=ctl.iterations, res0=ctl.res0, res=ctl.res, seconds=ctl.seconds,
                       history=hist[: ctl.iterations].copy())


def adjoint_solve_batch(solves: list, cfl: float = 1.0, rk_stages: int = 3, max_iters: int = 5000,
                        tol: float = 1e-10, iters_per_sync: int = 200, progress=None):
    """Many adjoint solves advanced together (aero_hip_adjoint_solve_batch).  solves: (HipMesh, residual
    config, rhs [N,5][, U [N,5] or None[, cfl]]); solves on one HipMesh share its linearisation; a solve's
    own cfl replaces `cfl`.
    progress(iterations, active) is called at every host check (a true return stops the solve).
    Returns ([psi], [info]), info per solve, and the wall time of the iterations."""
    lib = _lib.lib()
    keep = []
    arr = (_lib.AdjointCaseC * len(solves))()
    psis = []
    for i, sv in enumerate(solves):
        hm, cfg, rhs = sv[0], sv[1], sv[2]
        U = sv[3] if len(sv) > 3 else None
        arr[i].cfl = float(sv[4]) if len(sv) > 4 and sv[4] is not None else 0.0
        rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(hm.N, 5)
        psi = np.zeros((hm.N, 5))
        keep += [rhs, cfg]
        arr[i].mesh = hm.h.value
        arr[i].cfg = C.addressof(cfg)
        if U is not None:
            U = np.ascontiguousarray(U, dtype=np.float64).reshape(hm.N, 5)
            keep.append(U)
            arr[i].conserved = U.ctypes.data
        arr[i].rhs = rhs.ctypes.data
        arr[i].psi = psi.ctypes.data
        psis.append(psi)
    ctl = _lib.LinearSolve(rk_stages, cfl, max_iters, tol, 1, 0, 1, 0, 0.0, 0.0, 0.0)
    cb = _lib.PROGRESS_FN(lambda _u, it, act: int(bool(progress(it, act))) if progress else 0)
    err = _lib.errbuf()
    _lib.check(lib.aero_hip_adjoint_solve_batch(arr, len(solves), C.byref(ctl), iters_per_sync, cb, None, err,
                                                len(err)), err, "aero_hip_adjoint_solve_batch")
    infos = [dict(iterations=a.iterations, converged=bool(a.converged), failed=bool(a.failed), res0=a.res0, res=a.res)
             for a in arr]
    return psis, infos, ctl.seconds


class Seed:
    """A parameter direction (aero_tangent_seed) that keeps its numpy arrays alive."""

    def __init__(self, d_farfield=None, d_stab_floor=0.0, d_mach_ref=0.0, d_mach_min=0.0, d_beta_min=0.0,
                 d_beta_max=0.0, d_face_normal=None, d_face_center=None, d_face_area=None, d_cell_center=None,
                 d_cell_volume=None):
        self._keep = {}

        def arr(name, a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            self._keep[name] = a
            return _pd(a)

        ff = (C.c_double * 5)(*(list(d_farfield) if d_farfield is not None else [0.0] * 5))
        self.c = _lib.TangentSeed(ff, d_stab_floor, d_mach_ref, d_mach_min, d_beta_min, d_beta_max,
                                  arr("n", d_face_normal), arr("fc", d_face_center), arr("A", d_face_area),
                                  arr("cc", d_cell_center), arr("V", d_cell_volume))

    @classmethod
    def from_c(cls, c: _lib.TangentSeed) -> "Seed":
        s = cls(d_farfield=list(c.d_farfield), d_stab_floor=c.d_stab_floor, d_mach_ref=c.d_mach_ref,
                d_mach_min=c.d_mach_min, d_beta_min=c.d_beta_min, d_beta_max=c.d_beta_max)
        return s

    def ptr(self):
        return C.byref(self.c)


def case_linearization(case: CaseConfig, iteration: int | None = None, param: str = "none"):
    """(residual config at `iteration` (default: the last), parameter Seed, Objective reference values,
    d q_inf / d param) for a case (aero_case_linearization)."""
    lib = _lib.lib()
    cc = case.to_c()
    it = case.iterations - 1 if iteration is None else iteration
    rc = _lib.ResidualConfig()
    seed = _lib.TangentSeed()
    obj = _lib.Objective()
    dq = C.c_double()
    err = _lib.errbuf()
    _lib.check(lib.aero_case_linearization(C.byref(cc), it, _lib.PARAMS[param], C.byref(rc), C.byref(seed),
                                           C.byref(obj), C.byref(dq), err, len(err)), err, "aero_case_linearization")
    return rc, Seed.from_c(seed), obj, dq.value


class HipBatch:
    """Batched solves over device geometries (HipMesh handles built with the same block_cells): one
    launch per RK stage over every (case, block), device-side iteration control, hipGraph replay."""

    def __init__(self, meshes: list, cases: list, base: CaseConfig, hist_capacity: int = 0):
        """cases: (mesh_index, mach, aoa_deg[, precond, all_speed_flux_fix[, config]]): the switches are
        0 = as the case's configuration, 1 = on, -1 = off; config (a CaseConfig, or None for `base`) gives
        the case its own schedules and reference values (aero_batch_case.config)."""
        self.lib = _lib.lib()
        self.meshes = list(meshes)
        self.cases = [tuple(c) + (0, 0, None)[max(0, len(c) - 3):] for c in cases]
        self._case_cfgs = [None if c[5] is None else c[5].to_c() for c in self.cases]
        arr = (C.c_void_p * len(self.meshes))(*[m.h.value for m in self.meshes])
        cs = (_lib.BatchCaseC * len(self.cases))(*[
            _lib.BatchCaseC(int(i), float(mach), float(aoa), int(pc), int(af),
                            None if cc is None else C.addressof(cc))
            for (i, mach, aoa, pc, af, _), cc in zip(self.cases, self._case_cfgs)])
        self._cfg = base.to_c()
        self.hist_capacity = hist_capacity
        h = C.c_void_p()
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_create(arr, len(self.meshes), cs, len(self.cases), C.byref(self._cfg),
                                                  hist_capacity, C.byref(h), err, len(err)),
                   err, "aero_hip_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.aero_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, max_iters: int, rho_drop: float = 1e-8, fixed_iters: bool = False, iters_per_sync: int = 50,
            use_graph: bool = True) -> dict:
        res = (_lib.BatchResult * len(self.cases))()
        sec = C.c_double()
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_run(self.h, max_iters, rho_drop, int(fixed_iters), iters_per_sync,
                                               int(use_graph), res, C.byref(sec), err, len(err)),
                   err, "aero_hip_batch_run")
        rows = [dict(iterations=r.iterations, converged=bool(r.converged), failed=bool(r.failed), l2=r.l2,
                     rho_l2=r.rho_l2, rho_ratio=r.rho_ratio, cl=r.cl, cd=r.cd, cm=r.cm) for r in res]
        return dict(cases=rows, seconds=sec.value)

    def state(self, i: int) -> np.ndarray:
        g = self.meshes[self.cases[i][0]]
        U = np.zeros((g.N, 5))
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_get_state(self.h, i, _pd(U), err, len(err)), err, "batch state")
        return U

    def history(self, i: int) -> np.ndarray:
        out = np.zeros((max(self.hist_capacity, 1), 4))
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_history(self.h, i, _pd(out), self.hist_capacity, err, len(err)), err,
                   "batch history")
        return out[: self.hist_capacity]

    def set_state(self, i: int, U: np.ndarray):
        """Warm start: case i's next iteration starts from U [N, 5] (aero_hip_batch_set_state)."""
        g = self.meshes[self.cases[i][0]]
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(g.N, 5)
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_set_state(self.h, i, _pd(U), err, len(err)), err, "batch set_state")

    def set_limiter_ratios(self, i: int, ratios: np.ndarray):
        """Install case i's frozen limiter ratios [8, F]; the batch counts as frozen from its next iteration
        (aero_hip_batch_set_limiter_ratios: install them for every case)."""
        mesh = self.meshes[self.cases[i][0]]
        r = np.ascontiguousarray(ratios, dtype=np.float64).reshape(8, mesh.F)
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_set_limiter_ratios(self.h, i, _pd(r), err, len(err)), err,
                   "batch set_limiter_ratios")

    def set_groups(self, group_bytes: int = 0, group_items: int = 0) -> int:
        """Iterate the active cases in groups of about group_bytes hot bytes with at least group_items work
        items each (0: the defaults, sized to the Infinity Cache and the CU count); returns the number of
        groups (aero_hip_batch_set_groups)."""
        n = C.c_int32()
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_set_groups(self.h, int(group_bytes), int(group_items), C.byref(n), err,
                                                      len(err)), err, "batch set_groups")
        return n.value

    def limiter_ratios(self, i: int) -> np.ndarray:
        """Case i's frozen limiter ratios [8, F] (batches created with a freeze iteration)."""
        mesh = self.meshes[self.cases[i][0]]
        out = np.zeros((8, mesh.F))
        err = _lib.errbuf()
        _lib.check(self.lib.aero_hip_batch_limiter_ratios(self.h, i, _pd(out), err, len(err)), err,
                   "batch limiter ratios")
        return out
