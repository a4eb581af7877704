"""Shared helpers for the tests: meshes, configs and comparison metrics."""
from __future__ import annotations

import ctypes as C

import numpy as np

import cases
import oracle_lib as ol
from aero_hip import generate_mesh
from aero_hip import _lib as hl
from aero_hip.solver import CaseConfig, freestream


def mesh(name: str, precision: int = 8, **kw):
    spec = dict(cases.MESHES[name])
    spec.update(kw)
    return generate_mesh(precision=precision, **spec)


def res_cfg(vals: dict, dtype=np.float64) -> hl.ResidualConfig:
    """A C residual config; for the FP32 oracle every value is rounded to float32 first, exactly what the
    reference receives."""
    cast = (lambda x: float(np.float32(x))) if np.dtype(dtype) == np.float32 else float
    names = [n for n, _ in hl.ResidualConfig._fields_]
    args = []
    for n, t in hl.ResidualConfig._fields_:
        v = vals.get(n, 0)
        if n == "farfield":
            args.append((C.c_double * 5)(*[cast(x) for x in v]))
        elif t is C.c_double:
            args.append(cast(v))
        else:
            args.append(int(v))
    assert len(args) == len(names)
    return hl.ResidualConfig(*args)


def case_cfg(over: dict, env_precond: bool, dtype=np.float64) -> hl.CaseConfigC:
    c = CaseConfig(**over)
    c.precond_experimental_enabled = 1 if env_precond else 0
    cc = c.to_c()
    if np.dtype(dtype) == np.float32:
        for n, t in hl.CASE_FIELDS:
            if t is C.c_double:
                setattr(cc, n, float(np.float32(getattr(cc, n))))
    return cc


def perturbed_state(cell_center, cons_inf, amp, seed):
    """Same construction as tests/golden/make_golden.py, in FP64."""
    rng = np.random.default_rng(seed)
    x, y = cell_center[:, 0].astype(np.float64), cell_center[:, 1].astype(np.float64)
    g = 1.4
    rho0 = float(cons_inf[0])
    u0, v0 = float(cons_inf[1]) / rho0, float(cons_inf[2]) / rho0
    p0 = (g - 1.0) * (float(cons_inf[4]) - 0.5 * rho0 * (u0 * u0 + v0 * v0))
    ph = rng.uniform(0.0, 2.0 * np.pi, (4, 2))
    s = lambda k: np.sin(2 * np.pi * x + ph[k, 0]) * np.cos(2 * np.pi * y + ph[k, 1])
    vref = max(abs(u0), abs(v0), 1.0)
    rho = rho0 * (1.0 + amp * s(0))
    u = u0 + amp * vref * s(1)
    v = v0 + amp * vref * s(2)
    p = p0 * (1.0 + amp * s(3))
    U = np.zeros((len(x), 5))
    U[:, 0] = rho
    U[:, 1] = rho * u
    U[:, 2] = rho * v
    U[:, 4] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def flux_scale(om, cfg, U) -> np.ndarray:
    """Per cell and component, sum over its faces of |F_k A| from the oracle (the scale of the scaled
    criterion |dR| <= tol * sum_f |F A|, which stays well posed when R cancels towards zero)."""
    U = np.ascontiguousarray(U, om.dtype).reshape(om.N, 5)
    out = np.zeros((om.N, 5), om.dtype)
    getattr(ol.oracle(), f"ao_flux_abs_sum_{ol.suffix(om.dtype)}")(C.byref(om.c), C.byref(cfg), ol.ptr(U),
                                                                   ol.ptr(out))
    return out


def scaled_err(R_test, R_ref, scale):
    return np.abs(R_test - R_ref) / np.maximum(scale, 1e-300)
