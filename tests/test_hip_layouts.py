"""Layout choices of a mesh handle (aero_hip_options.block_cells / accumulation), through the C ABI: the
face-coloured accumulation against the FP64 oracle at the scaled 1e-12 tolerance (its per-cell sums run in
colour order) with the fallback counters and the first failing face exact, for every scheme, limiter and
option combination and with the frozen limiter; and the layout probe (block_cells = 0), which must pick a
buildable layout, report its timings, reuse them for the same mesh, and give a residual within the same
tolerance of the oracle whatever it picked."""
import sys

import numpy as np
import pytest

import cases
import helpers
import oracle_lib as ol

pytestmark = pytest.mark.gpu
COMP = [0, 1, 2, 4]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    import aero_hip
    assert aero_hip.hip_available()
    return aero_hip


def check(Rh, lh, dh, om, cfg, U, label, frozen=None):
    if frozen is None:
        Ro, lo, do = ol.residual(om, cfg, U)
    else:
        Ro, lo, do = ol.residual_frozen(om, cfg, U, frozen)
    scale = helpers.flux_scale(om, cfg, U)
    err = helpers.scaled_err(Rh[:, COMP], Ro[:, COMP], scale[:, COMP]).max()
    assert err <= 1e-12, (label, err)
    assert (np.abs(lh - lo) / np.abs(lo)).max() <= 1e-12, label
    assert (dh["first_order_fallback"], dh["diffusive_fallback"], dh["first_failure_face"]) == tuple(do), (label, dh, do)


@pytest.mark.parametrize("mname,kw", [("64x14", dict(cells="mixed", quad_layers=4, jitter=0.1, seed=5, rcm=True)),
                                      ("40x8", {})])
def test_colour_accumulation_matches_oracle(hip, mname, kw):
    m = helpers.mesh(mname, **kw)
    om = ol.OracleMesh(m, np.float64)
    ff = hip.freestream(0.5, 2.0)[1]
    with hip.HipMesh(m, accumulation="colour") as hm:
        assert hm.layout()["accumulation"] == "colour"
        for amp, seed in ((0.05, 3), (0.9, 4)):
            U = helpers.perturbed_state(m.cell_center, ff, amp, seed)
            for combo in cases.RESIDUAL_COMBOS:
                cfg = helpers.res_cfg(cases.residual_cfg_values(combo, ff))
                Rh, lh, dh = hm.residual(U, cfg)
                check(Rh, lh, dh, om, cfg, U, f"{mname} colour {combo} amp {amp}")
        # frozen limiter: the ratios the device captured, which the oracle's capture reproduces exactly
        U = helpers.perturbed_state(m.cell_center, ff, 0.05, 3)
        for scheme in (0, 1):
            cfg = helpers.res_cfg(cases.residual_cfg_values((scheme, 1, 1, 1, 0, 0, 0), ff))
            r = hm.freeze_limiter(cfg, U)
            assert np.array_equal(r, ol.limiter_capture(om, cfg, U))
            cfg.limiter_frozen = 1
            Rh, lh, dh = hm.residual(U, cfg)
            check(Rh, lh, dh, om, cfg, U, f"{mname} colour frozen {scheme}", frozen=r)


def test_colour_and_gather_runs_agree(hip):
    """A driver run on a coloured handle (its stages gather; its residual evaluations colour) equals the run
    on a gather handle bit for bit."""
    m = helpers.mesh("64x14")
    run = cases.RUNS["case_file_numerics"]
    case = hip.CaseConfig(**run["cfg"])
    case.precond_experimental_enabled = 0
    with hip.HipMesh(m, accumulation="colour") as a, hip.HipMesh(m, accumulation="gather") as b:
        ra, rb = a.run_case(case), b.run_case(case)
    assert np.array_equal(ra["history"], rb["history"]) and np.array_equal(ra["conserved"], rb["conserved"])


def test_layout_probe(hip):
    m = helpers.mesh("256x64", cells="mixed", quad_layers=12, jitter=0.1, seed=12345, rcm=True)
    om = ol.OracleMesh(m, np.float64)
    ff = hip.freestream(0.3, 2.0)[1]
    U = helpers.perturbed_state(m.cell_center, ff, 0.02, 11)
    cfg = helpers.res_cfg(cases.residual_cfg_values((1, 1, 1, 1, 0, 0, 0), ff))
    with hip.HipMesh(m, block_cells=0) as hm:
        lay, probe = hm.layout(), hm.layout_probe()
        say(f"[layout] probe {probe} -> block_cells {lay['block_cells']} {lay['accumulation']}")
        assert all(v is not None and v > 0 for v in probe.values()), probe
        # The choice: the gather layout with the largest blocks timed within 2 % of the fastest layout,
        # else the fastest (colour).
        fastest = min(probe.values())
        near = [k for k in ("gather512", "gather256", "gather128") if probe[k] <= 1.02 * fastest]
        expect = near[0] if near else min(probe, key=probe.get)
        assert expect == (lay["accumulation"] if lay["accumulation"] == "colour" else f"gather{lay['block_cells']}")
        Rh, lh, dh = hm.residual(U, cfg)
        check(Rh, lh, dh, om, cfg, U, "probed layout")
    with hip.HipMesh(m, block_cells=0) as hm2:  # the same mesh: the cached choice, no second probe
        assert hm2.layout_probe() == probe and hm2.layout()["block_cells"] == lay["block_cells"]
    with hip.HipMesh(m, block_cells=0, accumulation="gather") as hm3:
        p3 = hm3.layout_probe()
        assert p3["colour"] is None and hm3.layout()["accumulation"] == "gather"
