"""Case files of the reference's native runner, read for the HIP backend.

`read_native_case` reads the key=value file the reference's Python layer writes (`native_case.cfg`,
python/cfd/case.py:313-542) with the parsing rules of parse_case_kv (numerics/residual_stub.cpp:160-189):
blank lines and lines starting with '#' are skipped, the first '=' (else ':') separates key and value,
both trimmed.  `euler_case_from_kv` turns it into a CaseConfig and the O-grid mesh settings with the
semantics of residual_stub.cpp:402-559 (defaults of EulerAirfoilCaseConfig, the precond / low_mach_fix
coupling, aliases and clamps).  `native_kv_from_yaml` maps a case YAML (the reference's sections
geometry / mesh / flow / solver / numerics / post) onto the same keys, so either file drives a run.
"""
from __future__ import annotations

import math
import re
from pathlib import Path

from .solver import CaseConfig

_NUM = re.compile(r"^\s*[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")
_INT = re.compile(r"^\s*[+-]?\d+")


def read_native_case(path: str | Path) -> dict[str, str]:
    kv: dict[str, str] = {}
    p = Path(path)
    if not p.exists():
        return kv
    for line in p.read_text(encoding="utf-8").splitlines():
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        sep = s.find("=")
        if sep < 0:
            sep = s.find(":")
        if sep < 0:
            continue
        key, value = s[:sep].strip(), s[sep + 1:].strip()
        if key:
            kv[key] = value
    return kv


def _get_str(kv, key, fallback: str) -> str:
    v = kv.get(key)
    return fallback if v is None or v == "" else v


def _get_float(kv, key, fallback: float) -> float:
    """Leading decimal prefix of the value (std::stof); the fallback when absent or unparsable."""
    if key not in kv:
        return fallback
    m = _NUM.match(kv[key])
    if m:
        return float(m.group(0))
    low = kv[key].strip().lower()
    if low in ("inf", "+inf", "infinity"):
        return math.inf
    if low in ("-inf", "-infinity"):
        return -math.inf
    return fallback


def _get_int(kv, key, fallback: int) -> int:
    """Leading integer prefix of the value (std::stoi)."""
    if key not in kv:
        return fallback
    m = _INT.match(kv[key])
    return int(m.group(0)) if m else fallback


def _on_off(value: str, fallback: bool) -> bool:
    v = value.strip().lower()
    if v in ("on", "true", "yes", "1"):
        return True
    if v in ("off", "false", "no", "0"):
        return False
    return fallback


def euler_case_from_kv(kv: dict[str, str]) -> tuple[CaseConfig, dict]:
    """(CaseConfig, mesh settings) of an euler_airfoil_2d native case (residual_stub.cpp:402-559)."""
    c = CaseConfig()
    c.iterations = _get_int(kv, "iterations", c.iterations)
    c.min_iterations = _get_int(kv, "min_iterations", c.min_iterations)
    c.cfl_start = _get_float(kv, "cfl_start", c.cfl_start)
    c.cfl_max = _get_float(kv, "cfl_max", c.cfl_max)
    c.cfl_ramp_iters = _get_int(kv, "cfl_ramp_iters", c.cfl_ramp_iters)
    c.residual_reduction_target = _get_float(kv, "residual_reduction_target", c.residual_reduction_target)
    c.force_stability_tol = _get_float(kv, "force_stability_tol", c.force_stability_tol)
    c.force_stability_window = _get_int(kv, "force_stability_window", c.force_stability_window)
    c.force_mean_drift_tol = _get_float(kv, "force_mean_drift_tol", c.force_mean_drift_tol)
    c.startup_first_order_iters = _get_int(kv, "startup_first_order_iters", c.startup_first_order_iters)
    c.rk_stages = _get_int(kv, "rk_stages", c.rk_stages)
    c.local_time_stepping = int(_on_off(_get_str(kv, "local_time_stepping", "on" if c.local_time_stepping else "off"),
                                        bool(c.local_time_stepping)))
    c.gamma = _get_float(kv, "gamma", c.gamma)
    c.gas_constant = _get_float(kv, "gas_constant", c.gas_constant)
    c.mach = _get_float(kv, "mach", c.mach)
    c.aoa_deg = _get_float(kv, "aoa_deg", c.aoa_deg)
    c.p_inf = _get_float(kv, "p_inf", c.p_inf)
    c.t_inf = _get_float(kv, "t_inf", c.t_inf)
    c.rho_inf = _get_float(kv, "rho_inf", c.rho_inf)
    c.x_ref = _get_float(kv, "x_ref", c.x_ref)
    c.y_ref = _get_float(kv, "y_ref", c.y_ref)
    flux = _get_str(kv, "flux", "rusanov").strip().lower()
    c.flux_scheme = {"hllc": 1, "rusanov": 0}.get(flux, c.flux_scheme)
    lim = _get_str(kv, "limiter", "minmod").strip().lower()
    c.limiter = {"venkat": 1, "venkatakrishnan": 1, "minmod": 0}.get(lim, c.limiter)
    # precond defaults to the low_mach_fix token; low_mach_fix defaults to precond (residual_stub.cpp:435-451, :477)
    has_lmf = "low_mach_fix" in kv
    lmf_token = _on_off(_get_str(kv, "low_mach_fix", "on" if c.low_mach_fix else "off"), bool(c.low_mach_fix))
    if "precond" in kv or "precond_on" in kv:
        tok = _get_str(kv, "precond", "on") if "precond" in kv else _get_str(kv, "precond_on", "on")
        c.precond_on = int(_on_off(tok, bool(c.precond_on)))
    else:
        c.precond_on = int(lmf_token)
    c.precond_mach_ref = _get_float(kv, "precond_mach_ref", _get_float(kv, "mach_ref", c.precond_mach_ref))
    c.precond_mach_min = _get_float(kv, "precond_mach_min", c.precond_mach_min)
    c.precond_beta_min = _get_float(kv, "precond_beta_min", c.precond_beta_min)
    c.precond_beta_max = _get_float(kv, "precond_beta_max", c.precond_beta_max)
    c.precond_farfield_bc = int(_on_off(_get_str(kv, "precond_farfield_bc", "on" if c.precond_farfield_bc else "off"),
                                        bool(c.precond_farfield_bc)))
    floor_fallback = _get_float(kv, "stabilization_mach_floor", c.stab_floor_target)
    c.stab_floor_start = _get_float(kv, "stabilization_mach_floor_start", floor_fallback)
    c.stab_floor_target = _get_float(kv, "stabilization_mach_floor_target", floor_fallback)
    c.stab_floor_k_start = _get_float(kv, "stabilization_mach_floor_k_start", c.stab_floor_k_start)
    c.stab_floor_k_target = _get_float(kv, "stabilization_mach_floor_k_target", c.stab_floor_k_target)
    c.stab_ramp_iters = _get_int(kv, "stabilization_ramp_iters", c.stab_ramp_iters)
    c.low_mach_fix = int(lmf_token) if has_lmf else c.precond_on
    c.mach_cutoff = _get_float(kv, "mach_cutoff", c.mach_cutoff)
    c.all_speed_flux_fix = int(_on_off(_get_str(kv, "all_speed_flux_fix", "on" if c.all_speed_flux_fix else "off"),
                                       bool(c.all_speed_flux_fix)))
    c.all_speed_mach_cutoff = _get_float(kv, "all_speed_mach_cutoff", c.all_speed_mach_cutoff)
    if c.all_speed_mach_cutoff <= 0.0:
        c.all_speed_mach_cutoff = 0.25
    c.all_speed_f_min = _get_float(kv, "all_speed_f_min", _get_float(kv, "all_speed_theta_min", c.all_speed_f_min))
    if c.all_speed_f_min <= 0.0:
        c.all_speed_f_min = 0.05
    c.all_speed_f_min = min(max(c.all_speed_f_min, 1.0e-6), 1.0)
    c.all_speed_ramp_start_iter = max(_get_int(kv, "all_speed_ramp_start_iter", c.all_speed_ramp_start_iter), 0)
    c.all_speed_ramp_iters = max(_get_int(kv, "all_speed_ramp_iters", c.all_speed_ramp_iters), 0)
    c.all_speed_staged_controller = int(_on_off(
        _get_str(kv, "all_speed_staged_controller", "on" if c.all_speed_staged_controller else "off"),
        bool(c.all_speed_staged_controller)))
    c.all_speed_stage_a_min_iters = max(_get_int(kv, "all_speed_stage_a_min_iters", c.all_speed_stage_a_min_iters), 0)
    c.all_speed_stage_a_max_iters = max(_get_int(kv, "all_speed_stage_a_max_iters", c.all_speed_stage_a_max_iters), 0)
    c.all_speed_stage_f_min_high = _get_float(kv, "all_speed_stage_f_min_high", c.all_speed_stage_f_min_high)
    c.all_speed_stage_f_min_mid = _get_float(kv, "all_speed_stage_f_min_mid", c.all_speed_stage_f_min_mid)
    c.all_speed_stage_f_min_low = _get_float(kv, "all_speed_stage_f_min_low", c.all_speed_stage_f_min_low)
    c.all_speed_pjump_wall_target = _get_float(kv, "all_speed_pjump_wall_target", c.all_speed_pjump_wall_target)
    c.all_speed_pjump_spike_factor = _get_float(kv, "all_speed_pjump_spike_factor", c.all_speed_pjump_spike_factor)
    c.all_speed_pjump_hold_iters = max(_get_int(kv, "all_speed_pjump_hold_iters", c.all_speed_pjump_hold_iters), 1)
    c.all_speed_freeze_iters = max(_get_int(kv, "all_speed_freeze_iters", c.all_speed_freeze_iters), 0)
    c.all_speed_cfl_drop_factor = _get_float(kv, "all_speed_cfl_drop_factor", c.all_speed_cfl_drop_factor)
    c.all_speed_cfl_restore_factor = _get_float(kv, "all_speed_cfl_restore_factor", c.all_speed_cfl_restore_factor)
    c.all_speed_cfl_min_scale = _get_float(kv, "all_speed_cfl_min_scale", c.all_speed_cfl_min_scale)
    c.aoa0_symmetry_enforce = int(_on_off(_get_str(kv, "aoa0_symmetry_enforce",
                                                   "on" if c.aoa0_symmetry_enforce else "off"),
                                          bool(c.aoa0_symmetry_enforce)))
    c.aoa0_symmetry_enforce_interval = max(_get_int(kv, "aoa0_symmetry_enforce_interval",
                                                    c.aoa0_symmetry_enforce_interval), 0)
    if c.force_mean_drift_tol <= 0.0:
        c.force_mean_drift_tol = c.force_stability_tol
    mesh = dict(airfoil_source=_get_str(kv, "airfoil_source", "naca4"), naca=_get_str(kv, "naca_code", "0012"),
                airfoil_file=_get_str(kv, "airfoil_file", ""), chord=_get_float(kv, "chord", 1.0),
                num_circumferential=_get_int(kv, "num_circumferential", 160), num_radial=_get_int(kv, "num_radial", 48),
                farfield_radius=_get_float(kv, "farfield_radius", 15.0),
                radial_stretch=_get_float(kv, "radial_stretch", 1.4))
    c.chord = mesh["chord"]
    return c, mesh


def _fmt(v) -> str:
    if isinstance(v, bool):
        return "on" if v else "off"
    if isinstance(v, float):
        return f"{v:.9g}"
    return str(v)


# YAML section.key -> native key, for the keys that are renamed between the two files
_YAML_RENAMES = {("mesh", "circumferential"): "num_circumferential", ("mesh", "radial"): "num_radial",
                 ("geometry", "airfoil_file"): "airfoil_file", ("geometry", "chord"): "chord"}


def native_kv_from_yaml(case: dict) -> dict[str, str]:
    """Native keys of a case YAML mapping: geometry.profile -> naca_code (its first four digits),
    mesh.circumferential / radial -> num_circumferential / num_radial, everything else in geometry,
    mesh, flow, solver, numerics and post under its own name."""
    kv: dict[str, str] = {"case_type": str(case.get("case_type", "euler_airfoil_2d"))}
    geometry = case.get("geometry", {}) or {}
    digits = "".join(ch for ch in str(geometry.get("profile", "NACA0012")) if ch.isdigit())
    kv["naca_code"] = digits[:4] if len(digits) >= 4 else "0012"
    kv["airfoil_source"] = "file" if str(geometry.get("airfoil_file", "")).strip() else "naca4"
    for section in ("geometry", "mesh", "flow", "solver", "numerics", "post"):
        for key, value in (case.get(section, {}) or {}).items():
            if section == "geometry" and key in ("type", "profile"):
                continue
            if section == "mesh" and key == "generator":
                continue
            if section == "solver" and key == "equations":
                continue
            kv[_YAML_RENAMES.get((section, key), key)] = _fmt(value)
    kv.setdefault("iterations", "300")  # the YAML layer's default (python/cfd/case.py:47-51); the native one is 400
    return kv


def load_case(path: str | Path) -> dict[str, str]:
    """Native keys of a case file: a key=value native case, or a YAML case (.yaml / .yml)."""
    p = Path(path)
    if p.suffix.lower() in (".yaml", ".yml"):
        import yaml
        data = yaml.safe_load(p.read_text(encoding="utf-8")) or {}
        if not isinstance(data, dict):
            raise ValueError(f"{p}: a case YAML holds a top-level mapping")
        return native_kv_from_yaml(data)
    if not p.exists():
        raise FileNotFoundError(f"case file not found: {p}")
    return read_native_case(p)
