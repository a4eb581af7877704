"""Reduced-size, same-shape copies of the batched workloads, through the C ABI:

* the AoA x Mach sweep: 64 cases (AoA linspace(-4, 8, 16) x M {0.3, 0.5, 0.6, 0.7}) on a 16k-cell jittered
  mixed mesh, HLLC + venkat, CFL ramp 0.2 -> 2.4, limiter frozen at iteration 150, 300 iterations; every
  case bitwise equal to its own one-case batch (graph-replayed) and to the sequential plain solve (one
  case, direct launches, host sync every iteration) for a spread of cases, and two cases against the FP64
  oracle driver (final state, CFL and L-infinity histories bit for bit; L2 norms at 1e-9);
* the uneven batch: three mesh sizes (64 x 14, 128 x 30 and 128 x 78 mixed O-grids), each case its own
  Hicks-Henne geometry (20 bumps, U(+-0.005c)), M from {0.05 with the low-Mach preconditioner and the
  all-speed HLLC fix on, 0.4, 0.7}, 200 iterations with the limiter frozen at 100; every case bitwise equal
  to its solo batch and to the oracle driver (as above), then the batched CL / CD adjoints equal to the cases'
  single adjoint solves.
"""
import sys
from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mixed(nc, nr):
    from aero_hip import generate_mesh
    return generate_mesh(precision=8, num_circumferential=nc, num_radial=nr, farfield_radius=20.0,
                         radial_stretch=200.0, cells="mixed", quad_layers=min(30, nr // 2), jitter=0.1, seed=12345,
                         rcm=True)


def oracle_history(mesh, case, iters):
    """The oracle driver's run of the same case (threaded stage residuals, bit for bit the serial run)."""
    om = ol.OracleMesh(mesh, np.float64)
    return ol.run_case(om, replace(case, iterations=iters).to_c(), iters, threaded=True)


def compare_oracle(hist, U, ref, label):
    """Against the oracle driver: the final state and the CFL and L-infinity histories bit for bit; the L2
    and density-L2 histories (sums in another order) at 1e-9."""
    hr = ref["history"]
    assert hist.shape[0] == hr.shape[0], (label, hist.shape, hr.shape)
    for bcol, ocol in ((0, 2), (1, 8)):
        rel = np.abs(hist[:, bcol] - hr[:, ocol]) / np.maximum(np.abs(hr[:, ocol]), 1e-300)
        assert rel.max() <= 1e-9, (label, bcol, rel.max())
    assert np.array_equal(hist[:, 2], hr[:, 9]), (label, "cfl history")
    assert np.array_equal(hist[:, 3], hr[:, 3]), (label, "L-infinity history")
    assert np.array_equal(U, ref["conserved"]), (label, "final state")


def test_sweep64_reduced(hip_mod):
    hip = hip_mod
    mesh = mixed(128, 78)
    aoas = np.linspace(-4.0, 8.0, 16)
    cases = [(0, float(m), float(a)) for m in (0.3, 0.5, 0.6, 0.7) for a in aoas]
    iters = 300
    base = hip.CaseConfig(flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=2.4, cfl_ramp_iters=120,
                          startup_first_order_iters=25, stab_floor_start=0.1, stab_floor_target=0.02,
                          stab_ramp_iters=400, precond_experimental_enabled=0, limiter_freeze_iter=150,
                          residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, iterations=iters)
    g = hip.HipMesh(mesh)
    say(f"[sweep64] 64 cases x {mesh.num_cells} cells, {iters} iterations")
    with hip.HipBatch([g], cases, base, hist_capacity=iters) as b:
        r = b.run(iters, fixed_iters=True, iters_per_sync=100)
        states = [b.state(i) for i in range(len(cases))]
        hists = [b.history(i) for i in range(len(cases))]
    assert all(c["iterations"] == iters and not c["failed"] for c in r["cases"])
    for i, c in enumerate(cases):
        seq = i % 16 == 5  # a spread of cases as the sequential plain solve, the rest graph-replayed alone
        with hip.HipBatch([g], [c], base) as b1:
            b1.run(iters, fixed_iters=True, iters_per_sync=1 if seq else 100, use_graph=not seq)
            assert np.array_equal(b1.state(0), states[i]), f"case {i} differs from its solo run"
    for i in (0, 63):
        say(f"[sweep64] oracle driver, case {i}")
        ref = oracle_history(mesh, replace(base, mach=cases[i][1], aoa_deg=cases[i][2]), iters)
        compare_oracle(hists[i], states[i], ref, f"sweep case {i}")
    g.close()


def test_uneven_batch_reduced(hip_mod):
    hip = hip_mod
    from aero_hip import deform_mesh
    from aero_hip.adjoint import sensitivities, sensitivities_batch
    rng = np.random.default_rng(5150)
    iters = 200
    base_cfg = hip.CaseConfig(flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=2.4, cfl_ramp_iters=120,
                              startup_first_order_iters=25, stab_floor_start=0.1, stab_floor_target=0.02,
                              stab_ramp_iters=400, precond_on=0, precond_experimental_enabled=0, precond_mach_ref=0.0,
                              precond_mach_min=1e-3, precond_beta_min=1e-4, precond_beta_max=1.0,
                              all_speed_mach_cutoff=0.25, all_speed_f_min=0.05, all_speed_ramp_start_iter=100,
                              all_speed_ramp_iters=200, limiter_freeze_iter=100, residual_reduction_target=0.0,
                              force_stability_tol=0.0, min_iterations=1, iterations=iters)
    low_cfg = replace(base_cfg, cfl_start=0.1, cfl_max=0.6, cfl_ramp_iters=180, stab_floor_k_start=5.0,
                      stab_floor_k_target=1.0, stab_ramp_iters=700, precond_on=1, precond_experimental_enabled=1,
                      all_speed_flux_fix=1)
    specs, handles, cases = [], [], []
    for nc, nr in ((64, 14), (128, 30), (128, 78)):
        base = mixed(nc, nr)
        for mach in (0.05, 0.4, 0.7):
            amp = rng.uniform(-0.005, 0.005, 20)
            aoa = float(rng.uniform(0.0, 6.0))
            m, _ = deform_mesh(base, amp, tangents=False)
            handles.append(hip.HipMesh(m))
            low = mach < 0.1
            specs.append(dict(base=base, amp=amp, mesh=m, mach=mach, aoa=aoa, cfg=low_cfg if low else base_cfg))
            cases.append((len(handles) - 1, mach, aoa, 0, 0, low_cfg if low else None))
    say(f"[uneven] {len(cases)} cases on 3 mesh sizes, {iters} iterations")
    with hip.HipBatch(handles, cases, base_cfg, hist_capacity=iters) as b:
        r = b.run(iters, fixed_iters=True, iters_per_sync=50)
        states = [b.state(i) for i in range(len(cases))]
        hists = [b.history(i) for i in range(len(cases))]
        ratios = [b.limiter_ratios(i) for i in range(len(cases))]
    assert all(c["iterations"] == iters and not c["failed"] for c in r["cases"])
    items = []
    for i, (c, sp) in enumerate(zip(cases, specs)):
        with hip.HipBatch([handles[i]], [(0,) + c[1:]], base_cfg) as b1:
            b1.run(iters, fixed_iters=True, iters_per_sync=50)
            assert np.array_equal(b1.state(0), states[i]), f"uneven case {i} differs from its solo run"
        say(f"[uneven] oracle driver, case {i} (M {sp['mach']})")
        case = replace(sp["cfg"], mach=sp["mach"], aoa_deg=sp["aoa"])
        ref = oracle_history(sp["mesh"], case, iters)
        compare_oracle(hists[i], states[i], ref, f"uneven case {i}")
        if sp["mach"] >= 0.1:
            handles[i].set_limiter_ratios(ratios[i])
            _, tan = deform_mesh(sp["base"], sp["amp"])
            items.append(dict(hm=handles[i], case=replace(case, iterations=iters), U=states[i], shape_tangents=tan))
    # batched CL / CD adjoints of the non-preconditioned cases against each case's single adjoint solves
    say(f"[uneven] batched adjoint, {2 * len(items)} solves")
    sb = sensitivities_batch(items, cfl=1.0, max_iters=20000, tol=1e-9)
    for it, cb in zip(items, sb["cases"]):
        for obj, w in (("cl", (1.0, 0.0, 0.0)), ("cd", (0.0, 1.0, 0.0))):
            if not cb["adjoint"][obj]["converged"]:
                continue
            s1 = sensitivities(it["hm"], it["case"], it["U"], weights=w, shape_tangents=it["shape_tangents"],
                               cfl=1.0, max_iters=20000, tol=1e-9)
            for p, v in s1["d"].items():
                assert abs(cb["d"][obj][p] - v) <= 1e-8 * max(abs(v), 1e-12), (obj, p, cb["d"][obj][p], v)
    for h in handles:
        h.close()


@pytest.fixture(scope="module")
def hip_mod():
    import aero_hip
    assert aero_hip.hip_available()
    return aero_hip
