"""Batched solves (many cases in one launch per RK stage, device-side control, hipGraph replay) against
solo runs of the same cases: at a fixed iteration count every case's state and residual history is
bitwise identical to its solo run, and graph replay equals eager launches."""
import sys

import numpy as np
import pytest

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


BASE = dict(flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=1.1, cfl_ramp_iters=40, startup_first_order_iters=5,
            residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, low_mach_fix=1,
            precond_on=1, precond_experimental_enabled=0)


def solo(hip, mesh, mach, aoa, iters):
    case = hip.CaseConfig(**dict(BASE, iterations=iters, mach=mach, aoa_deg=aoa))
    with hip.HipMesh(mesh) as hm:
        out = hm.run_case(case)
    return out


def test_batch_matches_solo_runs(hip):
    m1 = helpers.mesh("40x8")
    m2 = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    cases = [(0, 0.3, 1.0), (1, 0.5, 2.5), (0, 0.6, -2.0), (1, 0.2, 4.0)]
    iters = 30
    g1, g2 = hip.HipMesh(m1), hip.HipMesh(m2)
    say("[batch] fixed-iteration batch of 4 cases on 2 geometries")
    with hip.HipBatch([g1, g2], cases, hip.CaseConfig(**BASE), hist_capacity=iters) as b:
        r = b.run(iters, fixed_iters=True, iters_per_sync=7, use_graph=True)
        states = [b.state(i) for i in range(len(cases))]
        hists = [b.history(i) for i in range(len(cases))]
    for i, (gi, mach, aoa) in enumerate(cases):
        say(f"[batch] solo case {i}")
        ref = solo(hip, [m1, m2][gi], mach, aoa, iters)
        assert r["cases"][i]["iterations"] == iters
        assert np.array_equal(states[i], ref["conserved"]), f"case {i} state differs from its solo run"
        assert np.array_equal(hists[i][:, 0], ref["history"][:, 2]), f"case {i} L2 history differs"
        assert np.array_equal(hists[i][:, 1], ref["history"][:, 8]), f"case {i} density L2 history differs"
        assert r["cases"][i]["cl"] == pytest.approx(ref["cl"], rel=1e-12, abs=1e-14)
    g1.close()
    g2.close()


def test_graph_replay_equals_eager(hip):
    m = helpers.mesh("64x14")
    cases = [(0, 0.4, 1.0), (0, 0.7, 3.0)]
    g = hip.HipMesh(m)
    out = []
    for use_graph, sync in ((True, 10), (False, 1), (True, 3)):
        say(f"[batch] graph={use_graph} sync={sync}")
        with hip.HipBatch([g], cases, hip.CaseConfig(**BASE)) as b:
            b.run(25, fixed_iters=True, iters_per_sync=sync, use_graph=use_graph)
            out.append([b.state(i) for i in range(len(cases))])
    for o in out[1:]:
        for a, bb in zip(o, out[0]):
            assert np.array_equal(a, bb)
    g.close()


@pytest.mark.parametrize("rk", [3, 1])
def test_converged_cases_are_masked(hip, rk):
    """A case that reaches its drop stops iterating while the other continues; its state is frozen, and the
    batch returns it bitwise equal to the solo run of as many iterations.  Under RK1 (two ring buffers)
    the stopped case's state must be found whatever the parity of the iterations the other case runs after
    it stopped: the batch is rerun to one and to two iterations past the convergence."""
    m = helpers.mesh("64x14")
    g = hip.HipMesh(m)
    first_order = dict(BASE, startup_first_order_iters=100000, rk_stages=rk)
    if rk == 1:
        first_order["cfl_max"] = 0.8
    # first order at M 0.3 and M 0.6: the two cases reach the drop hundreds of iterations apart
    cases = [(0, 0.3, 0.0), (0, 0.6, 3.0)]
    cap = 1500
    say(f"[batch] masked convergence rk={rk}")
    with hip.HipBatch([g], cases, hip.CaseConfig(**first_order), hist_capacity=cap) as b:
        r = b.run(cap, rho_drop=1e-2, iters_per_sync=20)
        its = [c["iterations"] for c in r["cases"]]
        a = int(np.argmin(its))  # the case that stopped first
        o = 1 - a
        ca = r["cases"][a]
        assert ca["converged"] and ca["rho_ratio"] <= 1e-2 and ca["iterations"] < cap
        n = ca["iterations"]
        h = b.history(a)
        assert h[n - 1, 1] / h[0, 1] <= 1e-2 < h[n - 2, 1] / h[0, 1]
        assert its[o] > n + 2, "the other case must still run after the first converged"
        sa, so, no = b.state(a), b.state(o), its[o]
    say(f"[batch] case {a} stopped at {n}, case {o} at {no}")

    def solo_state(i, iters):
        case = hip.CaseConfig(**dict(first_order, iterations=iters, mach=cases[i][1], aoa_deg=cases[i][2]))
        with hip.HipMesh(m) as hm:
            return hm.run_case(case)["conserved"]

    ref = solo_state(a, n)
    assert np.array_equal(sa, ref), "a masked case must keep its state from the iteration it converged"
    # the case still running after the other stopped (launches then cover its blocks only) is unaffected
    assert np.array_equal(so, solo_state(o, no)), "the continuing case must equal its solo run"
    # one and two iterations of the other case past the convergence, with graph replay and without
    for extra, graph in ((1, True), (2, False)):
        with hip.HipBatch([g], cases, hip.CaseConfig(**first_order)) as b:
            r = b.run(n + extra, rho_drop=1e-2, iters_per_sync=7, use_graph=graph)
            assert r["cases"][a]["iterations"] == n and r["cases"][o]["iterations"] == n + extra
            assert np.array_equal(b.state(a), ref), (rk, extra)
    g.close()


def test_batch_frozen_limiter_matches_solo_runs(hip):
    """With a freeze iteration (an option beyond the reference) the batch captures every case's limiter
    ratios at the same iteration as the solo driver, and stays bitwise identical to the solo runs, also
    when the batch runs in two calls around the freeze iteration."""
    m1 = helpers.mesh("40x8")
    m2 = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    cases = [(0, 0.3, 1.0), (1, 0.5, 2.5), (1, 0.6, -2.0)]
    iters, freeze = 40, 17
    base = dict(BASE, limiter_freeze_iter=freeze)
    g1, g2 = hip.HipMesh(m1), hip.HipMesh(m2)
    say("[batch] frozen-limiter batch (two runs around the freeze)")
    with hip.HipBatch([g1, g2], cases, hip.CaseConfig(**base), hist_capacity=iters) as b:
        b.run(11, fixed_iters=True, iters_per_sync=4, use_graph=True)
        b.run(iters - 11, fixed_iters=True, iters_per_sync=5, use_graph=True)
        states = [b.state(i) for i in range(len(cases))]
        hists = [b.history(i) for i in range(len(cases))]
    for i, (gi, mach, aoa) in enumerate(cases):
        case = hip.CaseConfig(**dict(base, iterations=iters, mach=mach, aoa_deg=aoa))
        with hip.HipMesh([m1, m2][gi]) as hm:
            ref = hm.run_case(case)
        assert np.array_equal(states[i], ref["conserved"]), f"case {i} state differs from its solo run"
        assert np.array_equal(hists[i][:, 1], ref["history"][:, 8]), f"case {i} density L2 history differs"
    g1.close()
    g2.close()


def test_case_groups_are_bitwise_one_group(hip):
    """Iterating the active cases in cache-sized groups (each group runs a sync interval's iterations in
    turn) gives every case the same states, histories and stopping iterations as one group: with one case
    per group, with pairs, and with every case in one group; across a freeze, a convergence stop that
    regroups the rest, and graph replay and eager launches."""
    m1 = helpers.mesh("40x8")
    m2 = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    cases = [(0, 0.3, 1.0), (1, 0.5, 2.5), (1, 0.6, -2.0), (0, 0.45, 0.0), (1, 0.3, 3.0)]
    cap, freeze = 300, 30  # three cases reach the 0.1 drop between iterations 210 and 255, two run to the cap
    base = hip.CaseConfig(**dict(BASE, limiter_freeze_iter=freeze, precond_on=0, low_mach_fix=0))
    g1, g2 = hip.HipMesh(m1), hip.HipMesh(m2)
    pair = 2 * max(g.N * 160 + g.layout()["slots"] * 16 for g in (g1, g2))
    out = {}
    for label, budget, graph in (("one", 1 << 40, True), ("single", 1, True), ("pairs", pair, False),
                                 ("default", 0, True)):
        say(f"[batch] case groups: {label}")
        with hip.HipBatch([g1, g2], cases, base, hist_capacity=cap) as b:
            ng = b.set_groups(budget, 1)
            r = b.run(cap, rho_drop=0.1, iters_per_sync=13, use_graph=graph)
            out[label] = (ng, [c["iterations"] for c in r["cases"]], [b.state(i) for i in range(len(cases))],
                          [b.history(i) for i in range(len(cases))])
    assert out["one"][0] == 1 and out["single"][0] == len(cases) and 1 < out["pairs"][0] < len(cases)
    its = out["one"][1]
    assert min(its) < max(its), "the test needs a case that stops before another"
    for label in ("single", "pairs", "default"):
        ng, it2, st2, h2 = out[label]
        assert it2 == its, label
        for i in range(len(cases)):
            assert np.array_equal(st2[i], out["one"][2][i]), (label, i)
            assert np.array_equal(h2[i], out["one"][3][i]), (label, i)
    g1.close()
    g2.close()


def test_batch_per_case_switches_match_solo_runs(hip):
    """Per-case low-Mach preconditioner and all-speed switches: a batch mixing preconditioned all-speed,
    preconditioned and plain cases equals the solo runs of the same cases configured that way."""
    m1 = helpers.mesh("40x8")
    m2 = helpers.mesh("64x14", cells="mixed", quad_layers=4, jitter=0.1, seed=3, rcm=True)
    base = dict(BASE, all_speed_ramp_start_iter=5, all_speed_ramp_iters=10, precond_beta_min=1e-4,
                precond_mach_ref=0.0)
    # a case with its own configuration: CFL, floor and all-speed schedules, preconditioner, p_inf
    own = dict(base, cfl_start=0.1, cfl_max=0.6, cfl_ramp_iters=18, stab_floor_k_start=5.0, stab_floor_k_target=1.0,
               stab_ramp_iters=20, precond_on=1, precond_experimental_enabled=1, all_speed_flux_fix=1,
               all_speed_ramp_start_iter=3, all_speed_ramp_iters=7, p_inf=90000.0, chord=1.3)
    # (mesh, mach, aoa, precond, all-speed fix[, config]): the base has the preconditioner gated off
    cases = [(0, 0.05, 1.0, 1, 1), (1, 0.5, 2.5, 0, 0), (1, 0.1, 2.0, 1, 0), (0, 0.3, -1.0, -1, 1),
             (1, 0.2, 1.5, 0, 0, hip.CaseConfig(**own))]
    iters = 30
    g1, g2 = hip.HipMesh(m1), hip.HipMesh(m2)
    say("[batch] per-case preconditioner / all-speed switches and configurations")
    with hip.HipBatch([g1, g2], cases, hip.CaseConfig(**base), hist_capacity=iters) as b:
        r = b.run(iters, fixed_iters=True, iters_per_sync=9, use_graph=True)
        states = [b.state(i) for i in range(len(cases))]
        hists = [b.history(i) for i in range(len(cases))]
    for i, c in enumerate(cases):
        gi, mach, aoa, pc, af = c[:5]
        if len(c) > 5:
            kw = dict(own, iterations=iters, mach=mach, aoa_deg=aoa)
        else:
            kw = dict(base, iterations=iters, mach=mach, aoa_deg=aoa)
            kw["precond_experimental_enabled"] = 1 if pc == 1 else 0
            kw["precond_on"] = 1 if pc == 1 else 0
            kw["all_speed_flux_fix"] = 1 if af == 1 else 0
        say(f"[batch] solo case {i}")
        with hip.HipMesh([m1, m2][gi]) as hm:
            ref = hm.run_case(hip.CaseConfig(**kw))
        assert np.array_equal(states[i], ref["conserved"]), f"case {i} state differs from its solo run"
        assert np.array_equal(hists[i][:, 1], ref["history"][:, 8]), f"case {i} density L2 history differs"
        assert r["cases"][i]["cl"] == pytest.approx(ref["cl"], rel=1e-12, abs=1e-14)
    g1.close()
    g2.close()


def test_batch_rejects_bad_switch(hip):
    g = hip.HipMesh(helpers.mesh("40x8"))
    with pytest.raises(RuntimeError, match="switch"):
        hip.HipBatch([g], [(0, 0.3, 1.0, 2, 0)], hip.CaseConfig(**BASE))
    with pytest.raises(RuntimeError, match="differs from the base"):
        hip.HipBatch([g], [(0, 0.3, 1.0, 0, 0, hip.CaseConfig(**dict(BASE, limiter=0)))], hip.CaseConfig(**BASE))
    g.close()

