"""The FP32 instantiation of the oracle reproduces the compiled reference's golden vectors bit for bit.

This pins every function of the restatement (pointwise numerics, gradient, residual assembly, stage tail,
forces and the full driver history) to the reference; the FP64 instantiation is the same code with
`Real = double`.
"""
import ctypes as C

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol

F32 = np.float32


def _eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def test_limiter_tables(golden, oracle_lib):
    g = golden("pointwise")
    a, b = g["lim_a"], g["lim_b"]
    for lim in (0, 1):
        out = np.zeros_like(a)
        ol.oracle().ao_limiter_f32(lim, ol.ptr(a), ol.ptr(b), ol.ptr(out), len(a))
        assert _eq(out, g[f"lim_{lim}"]), f"limiter {lim}"


@pytest.mark.parametrize("scheme,asf,w", [(0, 0, 1.0), (1, 0, 1.0), (1, 1, 0.0), (1, 1, 0.5), (1, 1, 1.0)])
def test_face_flux_tables(golden, oracle_lib, scheme, asf, w):
    g = golden("pointwise")
    UL, UR, n, s = g["flux_UL"], g["flux_UR"], g["flux_n"], g["flux_scale"]
    N = len(s)
    F = np.zeros((N, 5), F32)
    lam = np.zeros(N, F32)
    fb = np.zeros(N, np.int32)
    ol.oracle().ao_flux_f32(N, scheme, ol.ptr(UL), ol.ptr(UR), ol.ptr(n), ol.ptr(s), C.c_double(float(F32(1.4))),
                            asf, C.c_double(float(F32(0.25))), C.c_double(float(F32(0.05))), C.c_double(w),
                            ol.ptr(F), ol.ptr(lam), ol.ptr(fb))
    tag = f"{scheme}_{asf}_{w}"
    assert _eq(F, g[f"flux_F_{tag}"])
    assert _eq(lam, g[f"flux_lam_{tag}"])
    assert _eq(fb, g[f"flux_fb_{tag}"])
    if scheme == 1:
        assert fb.sum() > 0, "the table must exercise the HLLC -> Rusanov fallbacks"


@pytest.mark.parametrize("bc,cbc,scheme", [(2, 0, 0), (2, 0, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)])
def test_boundary_flux_tables(golden, oracle_lib, bc, cbc, scheme):
    g = golden("pointwise")
    UL, n, s, ff = g["flux_UL"], g["flux_n"], g["flux_scale"], g["bflux_ff"]
    N = len(s)
    F = np.zeros((N, 5), F32)
    lam = np.zeros(N, F32)
    fb = np.zeros(N, np.int32)
    ol.oracle().ao_boundary_flux_f32(N, bc, scheme, ol.ptr(UL), ol.ptr(ff), ol.ptr(n), ol.ptr(s),
                                     C.c_double(float(F32(1.4))), cbc, 0, C.c_double(float(F32(0.25))),
                                     C.c_double(float(F32(0.05))), C.c_double(1.0), ol.ptr(F), ol.ptr(lam), ol.ptr(fb))
    tag = f"{bc}_{cbc}_{scheme}"
    assert _eq(F, g[f"bflux_F_{tag}"])
    assert _eq(lam, g[f"bflux_lam_{tag}"])
    assert _eq(fb, g[f"bflux_fb_{tag}"])


def _golden_keys(g, mname, si):
    pre = f"res_{mname}_U{si}_"
    return sorted({k[: -len("_R")] for k in g.files if k.startswith(pre) and k.endswith("_R")})


@pytest.mark.parametrize("mname,si", [("40x8", 0), ("40x8", 1), ("64x14", 0)])
def test_residual_and_gradients(golden, oracle_lib, mname, si):
    g = golden("residual")
    m = helpers.mesh(mname, precision=4)
    om = ol.OracleMesh(m, F32)
    U = g[f"res_{mname}_U{si}"]
    assert _eq(ol.gradients(om, U, float(F32(1.4))), g[f"grad_{mname}_U{si}"])
    keys = _golden_keys(g, mname, si)
    assert len(keys) >= 6
    fallbacks = 0
    for key in keys:
        combo = tuple(int(c) for c in key.rsplit("_", 1)[1])
        vals = cases.residual_cfg_values(combo, _ff32())
        R, lam, diag = ol.residual(om, helpers.res_cfg(vals, F32), U)
        assert _eq(R, g[key + "_R"]), key
        assert _eq(lam, g[key + "_lam"]), key
        assert _eq(diag, g[key + "_diag"]), key
        fallbacks += int(diag[0]) + int(diag[1])
    if si == 1:
        assert fallbacks > 0, "the strong perturbation must exercise the first-order / diffusive fallbacks"


def _ff32():
    import make_golden
    return make_golden.freestream32(0.3, 2.0)


def test_stage_tail(golden, oracle_lib):
    g = golden("residual")
    un, us, res, lam = g["stage_un"], g["stage_us"], g["stage_res"], g["stage_lam"]
    n = len(lam)
    for pc in (0, 1):
        vals = cases.residual_cfg_values((1, 1, 1, 0, pc, 0, 0), _ff32())
        cfg = helpers.res_cfg(vals, F32)
        for lts in (0, 1):
            for st in (0, 1, 2):
                nxt = np.zeros((n, 5), F32)
                ol.oracle().ao_stage_tail_f32(n, st, ol.ptr(un), ol.ptr(us), ol.ptr(res), ol.ptr(lam),
                                              C.c_double(float(F32(0.7))), lts, C.byref(cfg), ol.ptr(nxt))
                assert _eq(nxt, g[f"stage_{pc}_{lts}_{st}"]), (pc, lts, st)


def test_forces(golden, oracle_lib):
    g = golden("residual")
    om = ol.OracleMesh(helpers.mesh("40x8", precision=4), F32)
    for sub in (0, 1):
        o = np.zeros(8, F32)
        ol.oracle().ao_forces_f32(C.byref(om.c), ol.ptr(g["forces_p"]), ol.ptr(g["forces_ref7"]), sub, ol.ptr(o))
        assert _eq(o, g[f"forces_{sub}"]), sub


@pytest.mark.parametrize("name", list(cases.RUNS))
def test_driver_history(golden, oracle_lib, name):
    g = golden("runs")
    run = cases.RUNS[name]
    m = helpers.mesh(run["mesh"], precision=4)
    om = ol.OracleMesh(m, F32)
    cfg = helpers.case_cfg(run["cfg"], run["env_precond"], F32)
    out = ol.run_case(om, cfg, run["cfg"]["iterations"])
    ref_hist = g[f"run_{name}_hist"]
    h = out["history"]
    assert h.shape[0] == ref_hist.shape[0]
    cols = [0, 1, 2, 3, 5, 6, 7]  # iter, l1, l2, linf, cl, cd, cm
    assert _eq(h[:, cols], ref_hist), name
    assert _eq(out["conserved"], g[f"run_{name}_U"])
    assert _eq(out["residual"], g[f"run_{name}_R"])
    assert _eq(out["spectral_radius"], g[f"run_{name}_lam"])
    assert _eq(out["forces"], g[f"run_{name}_forces"])
    assert _eq(out["cp"], g[f"run_{name}_cp"])


@pytest.mark.parametrize("name", list(cases.RUNS))
def test_driver_diagnostics_csv(golden, oracle_lib, name):
    """The FP32 oracle driver's per-iteration diagnostics, assembled into rows and printed by the runner
    (aero_hip.outputs.diagnostics_lines), reproduce the reference run's euler_diagnostics.csv: the header
    and every line, character for character."""
    from aero_hip import CaseConfig
    from aero_hip import outputs
    g = golden("runs")
    run = cases.RUNS[name]
    m = helpers.mesh(run["mesh"], precision=4)
    om = ol.OracleMesh(m, F32)
    out = ol.run_case(om, helpers.case_cfg(run["cfg"], run["env_precond"], F32), run["cfg"]["iterations"],
                      diagnostics=True)
    case = CaseConfig(**run["cfg"])
    lines = outputs.diagnostics_lines(m, case, out)
    ref = [str(x) for x in g[f"run_{name}_diagnostics"]]
    assert lines[0] == ref[0], "header"
    assert len(lines) == len(ref), (len(lines), len(ref))
    header = ref[0].split(",")
    bad = {}
    for i, (a, b) in enumerate(zip(lines[1:], ref[1:])):
        for k, (x, y) in enumerate(zip(a.split(","), b.split(","))):
            if x != y:
                bad.setdefault(header[k], (i, x, y))
    assert not bad, bad
