#!/usr/bin/env python3
"""FP64 residual-evaluation throughput of the HIP backend on one MI355X.

Workload (the single-large-mesh configuration): a ~1.02M-cell NACA 0012 O-grid, 30 wall-normal quad
layers then triangles, interior nodes jittered by seeded U(+-10%) of local spacing (seed 12345), cells
renumbered by reverse Cuthill-McKee; HLLC + Venkatakrishnan, M = 0.3, AoA = 2 deg, second order.
Synthetic state: freestream with a seeded smooth perturbation.  One step = one residual evaluation
(Green-Gauss gradient kernel + flux/accumulation kernel) on the HBM-resident state.

Prints one JSON line (rank 0).  `roofline` uses the algorithmic bytes of the dominant (flux) kernel,
48 B per face + 152 B per cell, over its average duration from HIP events recorded on the kernel's
stream inside the timed region.  `cpu_baseline` times the FP64 CPU oracle (a serial restatement of the
reference CPU path) on a bounded sample of the same workload; `cpu_reference` times the compiled
reference's own CPU residual (FP32) on the same mesh and state, where oracle/_ref is built.  The same line
carries the other two headline quantities, measured after the timed residual region: `sweep` (the 64-case
batched AoA x Mach sweep to an 8-decade drop, solves/hr and the speed-up over sequential solves timed on
--default-seq-cases of the cases) and `adjoint` (adjoint / primal time of the dCL, dCD gradients on the
Hicks-Henne-deformed 66k mesh).  Other modes: --sweep (every sequential case timed), --adjoint (with
finite-difference and tangent checks), --uneven, --vjp.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "aero-cfd-solver_amd"))

METRIC = "FP64 Mcell-res/s (%HBM peak); sweep solves/hr; adjoint/primal time ratio"
HBM_PEAK_GBS = 8000.0
VALU_SIMDS = 1024          # MI355X: 256 CUs x 4 SIMDs
VALU_CLOCK_HZ = 2.4e9      # peak engine clock (FP64 vector peak 78.6 TFLOP/s = 1024 x 16 lanes x 2 x 2.4 GHz)

WORKLOADS = {
    # name: (mesh kwargs, flux, limiter, mach, aoa)
    "mixed1m_hllc_venkat": (dict(num_circumferential=1024, num_radial=512, farfield_radius=20.0, radial_stretch=163.9,
                                 cells="mixed", quad_layers=30, jitter=0.1, seed=12345, rcm=True),
                            "hllc", "venkat", 0.3, 2.0),
    "mixed1m_rusanov_minmod": (dict(num_circumferential=1024, num_radial=512, farfield_radius=20.0,
                                    radial_stretch=163.9, cells="mixed", quad_layers=30, jitter=0.1, seed=12345,
                                    rcm=True), "rusanov", "minmod", 0.3, 2.0),
    "quad256x64_hllc_venkat": (dict(num_circumferential=256, num_radial=64, farfield_radius=20.0,
                                    radial_stretch=445.79, cells="quad"), "hllc", "venkat", 0.5, 1.25),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def perturbed_state(cc, cons_inf, amp=0.02, seed=20260101):
    import numpy as np
    rng = np.random.default_rng(seed)
    x, y = cc[:, 0], cc[:, 1]
    g = 1.4
    rho0 = cons_inf[0]
    u0, v0 = cons_inf[1] / rho0, cons_inf[2] / rho0
    p0 = (g - 1.0) * (cons_inf[4] - 0.5 * rho0 * (u0 * u0 + v0 * v0))
    ph = rng.uniform(0.0, 2.0 * np.pi, (4, 2))
    s = lambda k: np.sin(2 * np.pi * x + ph[k, 0]) * np.cos(2 * np.pi * y + ph[k, 1])
    vref = max(abs(u0), abs(v0), 1.0)
    rho, u, v, p = rho0 * (1 + amp * s(0)), u0 + amp * vref * s(1), v0 + amp * vref * s(2), p0 * (1 + amp * s(3))
    U = np.zeros((len(x), 5))
    U[:, 0], U[:, 1], U[:, 2] = rho, rho * u, rho * v
    U[:, 4] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def cpu_baseline(mesh, cfg_vals, U, seconds_target: float) -> dict | None:
    """FP64 oracle (serial restatement of the reference CPU path) on the same mesh and state."""
    try:
        sys.path.insert(0, str(ROOT / "tests"))
        sys.path.insert(0, str(ROOT / "tests" / "golden"))
        import numpy as np
        import oracle_lib as ol
        import helpers
    except Exception as e:  # the oracle is test infrastructure; report its absence instead of failing
        log(f"cpu baseline unavailable: {e}")
        return None
    om = ol.OracleMesh(mesh, np.float64)
    cfg = helpers.res_cfg(cfg_vals)
    t1 = ol.time_residual(om, cfg, U, 1)
    n = max(1, int(seconds_target / max(t1, 1e-6)))
    t = ol.time_residual(om, cfg, U, n)
    out = {"value": round(mesh.num_cells * n / t / 1e6, 4), "unit": "Mcell-res/s", "cores": 1, "kind": "port",
           "sample": f"{n} back-to-back FP64 oracle residual evaluations (primitive + Green-Gauss + assembly, "
                     f"serial like the reference CPU path) on the same {mesh.num_cells}-cell mesh, "
                     f"{t:.1f} s on the host CPU"}
    # the same evaluation on all OpenMP threads (per-face values, per-cell sums in face order: the serial
    # result bit for bit)
    threads = ol.max_threads()
    t1 = ol.time_residual_threaded(om, cfg, U, 1)
    n = max(1, int(seconds_target / max(t1, 1e-6)))
    t = ol.time_residual_threaded(om, cfg, U, n)
    out["threaded"] = {"value": round(mesh.num_cells * n / t / 1e6, 4), "unit": "Mcell-res/s", "cores": threads,
                       "kind": "port", "sample": f"{n} evaluations of the same FP64 oracle on {threads} OpenMP "
                                                 f"threads, {t:.1f} s"}
    return out


def cpu_reference(mesh, cfg_vals, U, seconds_target: float) -> dict | None:
    """The compiled reference's own CPU residual (oracle/_ref: its primitive conversion, Green-Gauss gradient
    and residual assembly, FP32, serial) on the same mesh and state, built from the mesh arrays."""
    try:
        sys.path.insert(0, str(ROOT / "tests"))
        sys.path.insert(0, str(ROOT / "tests" / "golden"))
        import numpy as np
        import oracle_lib as ol
        ref = ol.ref()
    except Exception as e:
        log(f"reference cpu baseline unavailable: {e}")
        return None
    if ref is None:
        log("reference cpu baseline unavailable: oracle/_ref/libaero_ref.so is not built")
        return None
    U32 = np.ascontiguousarray(U, np.float32)
    with ol.RefMesh.from_mesh(ref, mesh) as rm:
        cfg = ol.ref_residual_cfg(cfg_vals)
        t0 = time.perf_counter()
        rm.residual(cfg, U32)
        t1 = time.perf_counter() - t0
        n = max(1, int(seconds_target / max(t1, 1e-6)))
        t0 = time.perf_counter()
        for _ in range(n):
            rm.residual(cfg, U32)
        t = time.perf_counter() - t0
    return {"value": round(mesh.num_cells * n / t / 1e6, 4), "unit": "Mcell-res/s", "cores": 1, "kind": "reference",
            "dtype": "f32", "sample": f"{n} back-to-back residual evaluations of the compiled reference (FP32, its "
                                      f"driver's stage body: primitives, Green-Gauss gradient, assembly) on the same "
                                      f"{mesh.num_cells}-cell mesh, {t:.1f} s on the host CPU"}


def copy_ceiling(mib: int = 1024, reps: int = 20) -> float:
    """Achieved bandwidth of a plain device-to-device copy of `mib` MiB (read + write bytes over the time of
    `reps` copies, HIP events): the practical HBM ceiling of this GPU, beside the datasheet peak."""
    import torch
    n = mib * 1024 * 1024 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * n * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def host_cpu() -> dict:
    """The host CPU the CPU baselines ran on: model name and logical CPUs."""
    model = ""
    try:
        for line in open("/proc/cpuinfo", encoding="utf-8"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def time_residual(hm, cfg, steps: int, warmup: int, stream):
    """(total ms from HIP events on the handle's stream, wall seconds, per-kernel timing) of `steps`
    back-to-back residual evaluations after `warmup` untimed ones.  The timed region holds only the
    evaluations; the per-kernel durations (a HIP event pair around every launch, on the same stream) come
    from a second pass over the same `steps` evaluations, since recording events between the launches
    adds about 10 us to each evaluation."""
    import torch
    for _ in range(warmup):
        hm.residual_async(cfg)
    torch.cuda.synchronize()
    log(f"[bench] timing {steps} steps")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if torch.distributed.is_available() and torch.distributed.is_initialized():
        torch.distributed.barrier()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        hm.residual_async(cfg)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    log(f"[bench] per-kernel timing pass, {steps} steps")
    hm.timing_begin(steps)
    for _ in range(steps):
        hm.residual_async(cfg)
    torch.cuda.synchronize()
    tm = hm.timing_end()
    return e0.elapsed_time(e1), wall, tm


def variant_configs(flux: str, lim: str, mach: float, aoa: float):
    """(name, mach, farfield -> residual config) of the numerics reported beside the headline workload."""
    from aero_hip import residual_config
    out = []
    if (flux, lim) != ("rusanov", "minmod"):
        out.append(("rusanov_minmod", mach, lambda ff: residual_config(ff, flux="rusanov", limiter="minmod",
                                                                       second_order=True, stab_mach_floor=0.02)))
    if (flux, lim) != ("hllc", "venkat"):
        out.append(("hllc_venkat", mach, lambda ff: residual_config(ff, flux="hllc", limiter="venkat",
                                                                    second_order=True, stab_mach_floor=0.02)))
    # low-Mach preconditioned all-speed run (the reference's low-Mach test settings,
    # tests/cpp/test_euler_low_mach_asymptotic.cpp:135-191, at M 0.01)
    out.append(("lowmach_m0.01_precond_allspeed", 0.01,
                lambda ff: residual_config(ff, flux=flux, limiter=lim, second_order=True, stab_mach_floor=0.0,
                                           precond=True, mach_ref=0.01, mach_min=1e-3, beta_min=1e-4, beta_max=1.0,
                                           all_speed_fix=True, all_speed_cutoff=0.25, all_speed_f_min=0.05,
                                           all_speed_ramp=1.0)))
    return out


FLUX_KERNEL_SOURCES = ("aero_kernels.hip", "face_eval.hpp", "euler_device.hpp", "aero_internal.hpp")


def kernel_source_digest() -> str:
    """sha256 of the flux kernel's sources, which a committed traffic measurement is keyed to."""
    import hashlib
    h = hashlib.sha256()
    for name in FLUX_KERNEL_SOURCES:
        h.update((ROOT / "aero-cfd-solver_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()


def counters_for(workload: str, layout: dict) -> dict:
    """The k_flux counter record of profiles/traffic.json for the gather layout's block size: HBM bytes per
    launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and VALU wave
    instructions per launch (SQ_INSTS_VALU), only when measured on the current kernel sources; else {}."""
    p = ROOT / "profiles" / "traffic.json"
    try:
        rec = json.loads(p.read_text()).get(workload)
    except Exception:
        return {}
    if not isinstance(rec, dict) or rec.get("kernel_sources_sha256") != kernel_source_digest():
        return {}
    if layout.get("accumulation") != "gather":
        return {}
    return rec.get("by_block_cells", {}).get(str(layout.get("block_cells")), {})


SWEEP_MESH = dict(num_circumferential=256, num_radial=144, farfield_radius=20.0, radial_stretch=200.0, cells="mixed",
                  quad_layers=30, jitter=0.1, seed=12345, rcm=True)


def measure_sweep(args, seq_cases: int) -> dict:
    """64-case AoA x Mach sweep on one ~66k-cell jittered mixed mesh: batched (one launch per RK stage over
    all cases, device-side control, hipGraph replay, host sync every 500 iterations) against the same cases
    solved one after another (batch of 1, no graph, host sync every iteration), `seq_cases` of them timed
    (all 64: measured; fewer: spread over the sweep and extrapolated by iteration count).  A case stops at an
    8-decade density-residual drop or at the iteration cap.  Returns the line's fields."""
    import numpy as np
    import torch
    import aero_hip
    from aero_hip import CaseConfig, HipBatch, HipMesh, generate_mesh

    if not aero_hip.hip_available():
        raise SystemExit("bench.py: no gfx950 device visible")
    mkw = SWEEP_MESH
    log("[sweep] generating mesh")
    mesh = generate_mesh(precision=8, **mkw)
    g = HipMesh(mesh, block_cells=args.batch_block_cells)
    aoas = np.linspace(-4.0, 8.0, 16)
    machs = [0.3, 0.5, 0.6, 0.7]
    cases = [(0, float(m), float(a)) for m in machs for a in aoas]
    base = CaseConfig(flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=args.sweep_cfl, cfl_ramp_iters=120,
                      startup_first_order_iters=25, stab_floor_start=0.1, stab_floor_target=0.02,
                      stab_ramp_iters=400, precond_experimental_enabled=0, limiter_freeze_iter=args.freeze_iter)
    cap = args.sweep_iters
    log(f"[sweep] batched: {len(cases)} cases x {mesh.num_cells} cells, cap {cap} iterations, "
        f"cfl {args.sweep_cfl}, limiter frozen at {args.freeze_iter}")
    with HipBatch([g], cases, base) as b:
        n_groups = b.set_groups()  # the default Infinity-Cache-sized case groups
        torch.cuda.synchronize()
        rb = b.run(cap, 1e-8, iters_per_sync=500, use_graph=not args.sweep_no_graph)
    t_batch = rb["seconds"]
    log(f"[sweep] batched: {t_batch:.1f} s")
    nseq = min(seq_cases, len(cases))
    log(f"[sweep] sequential: {nseq} cases one after another")
    t_seq = 0.0
    seq_rows = []
    for i in range(nseq):
        j = (i * len(cases)) // nseq  # spread the timed cases over the sweep
        with HipBatch([g], [cases[j]], base) as b1:
            r1 = b1.run(cap, 1e-8, iters_per_sync=1, use_graph=False)
        t_seq += r1["seconds"]
        seq_rows.append((j, r1["cases"][0]))
        log(f"[sweep] sequential case {j}: {r1['seconds']:.2f} s, {r1['cases'][0]['iterations']} iterations")
    # each sequential case ran the same iterations as in the batch (identical per-case arithmetic)
    same = all(r["iterations"] == rb["cases"][j]["iterations"] and r["rho_l2"] == rb["cases"][j]["rho_l2"]
               for j, r in seq_rows)
    it_timed = sum(r["iterations"] for _, r in seq_rows)
    it_all = sum(r["iterations"] for r in rb["cases"])
    measured = nseq == len(cases)
    t_seq_all = t_seq if measured else t_seq / max(it_timed, 1) * it_all
    g.close()
    rows = rb["cases"]
    n_conv = sum(int(r["converged"]) for r in rows)
    return {"value": round(n_conv / t_batch * 3600.0, 1), "unit": "solves/hr", "cap": cap, "cells": mesh.num_cells,
            "ms_per_step": round(t_batch / max(max(r["iterations"] for r in rows), 1) * 1e3, 4),
            "config": {"workload": "sweep64_mixed66k_hllc_venkat", "cells": mesh.num_cells, "cases": len(cases),
                       "mesh": mkw, "aoa_deg": [float(a) for a in aoas], "mach": machs, "iteration_cap": cap,
                       "rho_drop_target": 1e-8, "cfl_max": args.sweep_cfl, "limiter_freeze_iter": args.freeze_iter,
                       "frozen_ratio_resolution": 2.0 ** -15},
            "sweep": {"solves_per_hr": round(n_conv / t_batch * 3600.0, 1),
                      "batched_seconds": round(t_batch, 3), "sequential_seconds": round(t_seq_all, 3),
                      "cases_per_hr": round(len(cases) / t_batch * 3600.0, 1),
                      "sequential_cases_timed": nseq,
                      "case_groups": n_groups,
                      "sequential_baseline": ("measured: every case solved alone, one after another" if measured else
                                              "seconds per iteration of the timed cases x all cases' iterations"),
                      "speedup_vs_sequential": round(t_seq_all / t_batch, 2) if t_seq_all > 0 else None,
                      "iterations_max": max(r["iterations"] for r in rows),
                      "iterations_median": float(np.median([r["iterations"] for r in rows])),
                      "sequential_solves_per_hr": round(n_conv / t_seq_all * 3600.0, 1) if t_seq_all > 0 else None,
                      "sequential_matches_batched": same,
                      "converged_8_decades": n_conv,
                      "failed_nonfinite": sum(r["failed"] for r in rows),
                      "median_rho_drop": float(np.median([r["rho_ratio"] for r in rows])),
                      "cl": [round(r["cl"], 6) for r in rows], "cd": [round(r["cd"], 6) for r in rows]}}


def run_sweep(args) -> int:
    """The sweep as its own line (measure_sweep with --sweep-seq-cases sequential cases, default all 64)."""
    m = measure_sweep(args, args.sweep_seq_cases)
    out = {"metric": METRIC, "value": m["value"], "unit": m["unit"], "n_gpus": 1, "steps": m["cap"], "warmup": 0,
           "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic", "config": m["config"], "sweep": m["sweep"]}
    print(json.dumps(out), flush=True)
    return 0


def measure_adjoint_pair(args) -> dict:
    """The config-4 workload's gradients without the finite-difference checks: the deformed 66k mesh's
    primal through the torch Function's forward (a one-case batch to an 8-decade drop) and one backward that
    solves the CL and CD adjoints as a batched pair (a second backward reuses them).  ratio = adjoint time
    for both gradients / 2 / primal time."""
    import numpy as np
    import torch
    import aero_hip
    from aero_hip import CaseConfig, generate_mesh
    from aero_hip.torch_ops import last_run, steady_coefficients

    if not aero_hip.hip_available():
        raise SystemExit("bench.py: no gfx950 device visible")
    log("[adjoint] generating the mesh")
    base = generate_mesh(precision=8, **SWEEP_MESH)
    amp = np.random.default_rng(4242).uniform(-0.002, 0.002, 20)
    cap = args.adjoint_iters
    case = CaseConfig(mach=0.6, aoa_deg=2.0, flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=args.adjoint_cfl,
                      cfl_ramp_iters=120, startup_first_order_iters=25, stab_floor_start=0.1, stab_floor_target=0.02,
                      stab_ramp_iters=400, precond_experimental_enabled=0, limiter_freeze_iter=args.freeze_iter,
                      residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, iterations=cap)
    params = torch.tensor(np.concatenate([[case.aoa_deg, case.mach], amp]), requires_grad=True)
    solve_kw = dict(primal="batch", rho_drop=args.adjoint_primal_drop, tol=args.adjoint_tol, max_iters=4 * cap,
                    jacobian=True, cfl=args.adjoint_solve_cfl)
    log("[adjoint] torch Function forward (primal)")
    coef = steady_coefficients(params, base, case, solve_kw=solve_kw)
    run = dict(last_run)
    log(f"[adjoint] primal {run['primal_seconds']:.2f} s, {run['primal_iterations']} iterations; backward (CL, CD)")
    (g_cl,) = torch.autograd.grad(coef[0], params, retain_graph=True)
    t_adj = last_run["adjoint_seconds"]
    adj_iters = last_run["adjoint_iterations"]
    (g_cd,) = torch.autograd.grad(coef[1], params)
    names = ["aoa", "mach"] + [f"shape{i}" for i in range(len(amp))]
    ratio = t_adj / 2.0 / run["primal_seconds"]
    return {"workload": "adjoint_mixed66k_hh20_hllc_venkat", "ratio_per_gradient": round(ratio, 3),
            "primal_rho_drop_target": args.adjoint_primal_drop, "unit": "adjoint/primal time per gradient", "primal_seconds": round(run["primal_seconds"], 3),
            "primal_iterations": run["primal_iterations"], "primal_converged": bool(run["primal_converged"]),
            "primal_rho_drop": run["primal_rho_ratio"], "adjoint_seconds_both": round(t_adj, 3),
            "adjoint_iterations_max": adj_iters, "adjoint_tol": args.adjoint_tol,
            "linear_solve_cfl": args.adjoint_solve_cfl, "cfl_max": args.adjoint_cfl,
            "cl": float(coef[0].detach()), "cd": float(coef[1].detach()),
            "dcl": dict(zip(names, [round(x, 10) for x in g_cl.tolist()])),
            "dcd": dict(zip(names, [round(x, 10) for x in g_cd.tolist()]))}


UNEVEN_MESHES = {  # nominal size: (circumferential, radial) of a jittered mixed O-grid, 30 quad wall layers
    "16k": (128, 78), "65k": (256, 144), "262k": (512, 271)}


def run_uneven(args) -> int:
    """Batched primal + adjoint over uneven meshes: --uneven-cases cases on each of the 16k, 65k and 262k
    jittered mixed meshes, every case its own seeded Hicks-Henne geometry (20 bumps, amplitudes
    U(+-0.005c), IDW deformation), M drawn from {0.05 (low-Mach preconditioner + all-speed HLLC fix),
    0.4, 0.7}, AoA U(0, 6) deg; HLLC + venkat, limiter frozen at --freeze-iter.  All cases run as one
    batch to an 8-decade density drop; then one batched adjoint solve for CL and CD of every case gives
    dJ/d(20 amplitudes, AoA, Mach).  value = solves/hr of the primal batch; the adjoint/primal time ratio is
    reported beside it, with a tangent-solve check of one case."""
    from dataclasses import replace
    import numpy as np
    import aero_hip
    from aero_hip import CaseConfig, HipBatch, HipMesh, case_linearization, deform_mesh, generate_mesh
    from aero_hip.adjoint import sensitivities_batch

    if not aero_hip.hip_available():
        raise SystemExit("bench.py: no gfx950 device visible")
    sizes = [x for x in args.uneven_meshes.split(",") if x]
    rng = np.random.default_rng(args.uneven_seed)
    base_cfg = CaseConfig(flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=args.sweep_cfl, cfl_ramp_iters=120,
                          startup_first_order_iters=25, stab_floor_start=0.1, stab_floor_target=0.02,
                          stab_ramp_iters=400, precond_on=0, precond_experimental_enabled=0, precond_mach_ref=0.0,
                          precond_mach_min=1e-3, precond_beta_min=1e-4, precond_beta_max=1.0,
                          all_speed_mach_cutoff=0.25, all_speed_f_min=0.05, all_speed_ramp_start_iter=100,
                          all_speed_ramp_iters=200, limiter_freeze_iter=args.freeze_iter,
                          residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1)
    # the low-Mach cases' own schedules: the reference's low-Mach test settings
    # (tests/cpp/test_euler_low_mach_asymptotic.cpp:135-191): CFL 0.1 -> 0.6 over 180 iterations, floor
    # k 5 -> 1 over 700, preconditioner at M_ref = M, all-speed HLLC fix ramped over iterations 100..300
    precond = args.uneven_lowmach == "precond"
    low_cfg = replace(base_cfg, cfl_start=0.1, cfl_max=0.6, cfl_ramp_iters=180, stab_floor_k_start=5.0,
                      stab_floor_k_target=1.0, stab_ramp_iters=700, precond_on=int(precond),
                      precond_experimental_enabled=int(precond), all_speed_flux_fix=1)
    if not precond:  # without the preconditioner the low-Mach cases run the plain schedules
        low_cfg = replace(base_cfg, all_speed_flux_fix=1)
    specs, handles, cases, mesh_cells = [], [], [], {}
    t0 = time.time()
    for size in sizes:
        nc, nr = UNEVEN_MESHES[size]
        mkw = dict(num_circumferential=nc, num_radial=nr, farfield_radius=20.0, radial_stretch=200.0, cells="mixed",
                   quad_layers=30, jitter=0.1, seed=12345, rcm=True)
        base = generate_mesh(precision=8, **mkw)
        mesh_cells[size] = base.num_cells
        log(f"[uneven] {size}: {base.num_cells} cells, deforming {args.uneven_cases} geometries")
        for _ in range(args.uneven_cases):
            amp = rng.uniform(-0.005, 0.005, 20)
            mach = float(rng.choice([0.05, 0.4, 0.7]))
            aoa = float(rng.uniform(0.0, 6.0))
            m, _ = deform_mesh(base, amp, tangents=False)
            handles.append(HipMesh(m, block_cells=args.batch_block_cells))
            low = mach < 0.1
            specs.append(dict(size=size, base=base, amp=amp, mach=mach, aoa=aoa, low_mach=low))
            cases.append((len(handles) - 1, mach, aoa, 0, 0, low_cfg if low else None))
    total_cells = sum(h.N for h in handles)
    log(f"[uneven] {len(cases)} cases, {total_cells} cells in all ({time.time() - t0:.1f} s of setup)")
    cap, chunk = args.uneven_iters, 5000
    t_primal = 0.0
    with HipBatch(handles, cases, base_cfg) as b:
        done = 0
        while done < cap:
            n = min(chunk, cap - done)
            r = b.run(n, 1e-8, iters_per_sync=500, use_graph=True)
            t_primal += r["seconds"]
            done += n
            active = sum(1 for c in r["cases"] if not c["converged"] and not c["failed"] and c["iterations"] >= done)
            log(f"[uneven] primal: {done} iterations, {active} cases still running, {t_primal:.1f} s")
            if active == 0:
                break
        rows = r["cases"]
        states = [b.state(i) for i in range(len(cases))]
        ratios = [b.limiter_ratios(i) for i in range(len(cases))]
    for i, (sp, row) in enumerate(zip(specs, rows)):
        log(f"[uneven] case {i}: {sp['size']} M {sp['mach']} AoA {sp['aoa']:.2f} iterations {row['iterations']} "
            f"rho drop {row['rho_ratio']:.2e} converged {row['converged']} failed {row['failed']}")
    log("[uneven] adjoint set-up (converged cases): shape tangents and frozen ratios")
    items, adj_index = [], []
    for i, sp in enumerate(specs):
        if not rows[i]["converged"]:
            continue
        adj_index.append(i)
        handles[i].set_limiter_ratios(ratios[i])
        _, tan = deform_mesh(sp["base"], sp["amp"])
        case = replace(low_cfg if sp["low_mach"] else base_cfg, mach=sp["mach"], aoa_deg=sp["aoa"],
                       iterations=rows[i]["iterations"])
        items.append(dict(hm=handles[i], case=case, U=states[i], shape_tangents=tan))
    if not items:
        raise SystemExit("bench.py --uneven: no case reached the density drop; raise --uneven-iters")
    log(f"[uneven] batched adjoint: {2 * len(items)} solves (CL, CD)")
    def progress(it, act):
        if it % 10000 == 0:
            log(f"[uneven] adjoint: {it} iterations, {act} solves active")
    sb = sensitivities_batch(items, tol=args.uneven_adjoint_tol, max_iters=4 * cap, iters_per_sync=500,
                             progress=progress)
    for i, c in zip(adj_index, sb["cases"]):
        log(f"[uneven] adjoint case {i}: " + ", ".join(f"{k}: {v['iterations']} it, conv {v['converged']}, "
                                                      f"failed {v['failed']}" for k, v in c["adjoint"].items()))
    # dual consistency on the first adjoint case: tangent solve for AoA against the adjoint
    it0 = items[0]
    rc, seed, obj, _ = case_linearization(it0["case"], param="aoa")
    obj.w_cl = 1.0
    ob = it0["hm"].objective(obj, it0["U"])
    log("[uneven] tangent check (case 0, AoA)")
    try:
        v, ti = it0["hm"].tangent_solve(rc, seed, it0["U"], cfl=it0["case"].cfl_max, max_iters=4 * cap,
                                        tol=args.uneven_adjoint_tol, check_every=500)
        d_tan = ob["dJ_daoa"] + float(np.sum(ob["dJdU"] * v))
    except aero_hip._lib.HipError as e:
        log(f"[uneven] tangent check failed: {e}")
        ti, d_tan = {"iterations": None}, float("nan")
    d_adj = sb["cases"][0]["d"]["cl"]["aoa"]
    for h in handles:
        h.close()
    t_adj = sb["solve_seconds"]
    adj_rows = [c["adjoint"] for c in sb["cases"]]
    adj_of = dict(zip(adj_index, adj_rows))
    per_size = {}
    for i, (sp, row) in enumerate(zip(specs, rows)):
        d = per_size.setdefault(sp["size"], dict(cells=mesh_cells[sp["size"]], cases=0, converged=0, iterations=[],
                                                 adjoint_iterations=[], low_mach_cases=0, low_mach_converged=0))
        d["cases"] += 1
        d["converged"] += int(row["converged"])
        d["low_mach_cases"] += int(sp["low_mach"])
        d["low_mach_converged"] += int(sp["low_mach"] and row["converged"])
        d["iterations"].append(row["iterations"])
        if i in adj_of:
            d["adjoint_iterations"] += [adj_of[i]["cl"]["iterations"], adj_of[i]["cd"]["iterations"]]
    for d in per_size.values():
        d["iterations_median"] = float(np.median(d.pop("iterations")))
        ai = d.pop("adjoint_iterations")
        d["adjoint_iterations_median"] = float(np.median(ai)) if ai else None
    value = sum(int(r["converged"]) for r in rows) / t_primal * 3600.0  # cases that reached the 8-decade drop
    out = {"metric": METRIC, "value": round(value, 1), "unit": "solves/hr", "n_gpus": 1, "steps": done, "warmup": 0,
           "ms_per_step": round(t_primal / max(max(r["iterations"] for r in rows), 1) * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "uneven48_hh20_mixed16k65k262k_hllc_venkat", "cases": len(cases),
                      "meshes": {k: UNEVEN_MESHES[k] for k in sizes}, "cells_total": total_cells,
                      "seed": args.uneven_seed, "bump_amplitude": 0.005, "mach_choices": [0.05, 0.4, 0.7],
                      "aoa_range_deg": [0.0, 6.0],
                      "low_mach": ("M 0.05: precond on (M_ref = M, beta_min 1e-4), all-speed HLLC fix on (ramp "
                                   "100..300), CFL 0.1 -> 0.6, floor k 5 -> 1 over 700 iterations") if precond else
                                  "M 0.05: preconditioner off, all-speed HLLC fix on, the other cases' schedules",
                      "iteration_cap": cap, "rho_drop_target": 1e-8, "cfl_max": args.sweep_cfl,
                      "limiter_freeze_iter": args.freeze_iter, "adjoint_tol": args.uneven_adjoint_tol},
           "uneven": {"primal_seconds": round(t_primal, 3), "adjoint_seconds": round(t_adj, 3),
                      "cases_per_hr": round(len(cases) / t_primal * 3600.0, 1),
                      "adjoint_solves": 2 * len(items), "adjoint_cases": adj_index,
                      "adjoint_over_primal": round(t_adj / t_primal, 3),
                      "adjoint_over_primal_per_objective": round(t_adj / t_primal / 2.0, 3),
                      "gradient_assembly_seconds": round(sb["assembly_seconds"], 3),
                      "converged_8_decades": sum(int(r["converged"]) for r in rows),
                      "failed_nonfinite": sum(int(r["failed"]) for r in rows),
                      "adjoint_converged": sum(int(a[k]["converged"]) for a in adj_rows for k in a),
                      "adjoint_failed_nonfinite": sum(int(a[k]["failed"]) for a in adj_rows for k in a),
                      "cases": [dict(mesh=sp["size"], mach=sp["mach"], aoa_deg=round(sp["aoa"], 4),
                                     iterations=row["iterations"], rho_drop=row["rho_ratio"],
                                     converged=bool(row["converged"]), cl=row["cl"], cd=row["cd"],
                                     adjoint_iterations={k: v["iterations"] for k, v in adj_of[i].items()}
                                     if i in adj_of else None,
                                     adjoint_converged={k: v["converged"] for k, v in adj_of[i].items()}
                                     if i in adj_of else None)
                                for i, (sp, row) in enumerate(zip(specs, rows))],
                      "per_mesh": per_size,
                      "tangent_check_case0_aoa": {"adjoint": d_adj, "tangent": d_tan,
                                                  "rel_diff": abs(d_adj - d_tan) / max(abs(d_tan), 1e-300),
                                                  "tangent_iterations": ti["iterations"]}}}
    print(json.dumps(out), flush=True)
    return 0


def _frozen_continuation(hm, rc, U0, iters, cfl):
    """Pseudo-time iterations from U0 with the handle's frozen limiter ratios (aero_hip_iterate); returns the
    density-residual L2 after the first and the last iteration."""
    hm.set_state(U0)
    h = hm.iterate(rc, iters, cfl)
    return float(h[0]), float(h[-1])


def run_adjoint(args) -> int:
    """Discrete adjoint on the ~66k-cell jittered mixed mesh deformed by 20 Hicks-Henne bumps (10 per
    surface, peaks at x/c = 0.05..0.95, amplitudes U(+-0.002c) from seed 4242, IDW volume deformation);
    M 0.6, AoA 2 deg, HLLC + venkat with the limiter frozen at --freeze-iter.  The gradients come from the
    torch.autograd.Function over the converged (CL, CD) (aero_hip.torch_ops.steady_coefficients): its
    forward runs the primal as a one-case batch (device control, hipGraph) to an 8-decade density drop;
    the backward of CL solves the CL and CD adjoints together (one batched pair) and the backward of CD
    reuses them, giving d(CL, CD)/d(20 amplitudes, AoA, Mach).  value = adjoint time for both gradients /
    2 / primal time.  Checks (at a second primal of the same case, the same state): central differences of
    converged solutions for all 22 parameters and both objectives (the 44 perturbed solutions continued
    from the converged state in one batch with the same frozen ratios, aero_hip.adjoint.
    fd_sensitivities_batch) and tangent solves for AoA, Mach and two amplitudes."""
    from dataclasses import replace
    import numpy as np
    import aero_hip
    from aero_hip import CaseConfig, HipBatch, HipMesh, case_linearization, deform_mesh, generate_mesh
    from aero_hip.adjoint import geometry_seed

    if not aero_hip.hip_available():
        raise SystemExit("bench.py: no gfx950 device visible")
    mkw = dict(num_circumferential=256, num_radial=144, farfield_radius=20.0, radial_stretch=200.0, cells="mixed",
               quad_layers=30, jitter=0.1, seed=12345, rcm=True)
    log("[adjoint] generating and deforming the mesh")
    base = generate_mesh(precision=8, **mkw)
    amp = np.random.default_rng(4242).uniform(-0.002, 0.002, 20)
    mesh, tan = deform_mesh(base, amp)
    cap = args.adjoint_iters
    case = CaseConfig(mach=0.6, aoa_deg=2.0, flux_scheme=1, limiter=1, cfl_start=0.2, cfl_max=args.adjoint_cfl,
                      cfl_ramp_iters=120, startup_first_order_iters=25, stab_floor_start=0.1, stab_floor_target=0.02,
                      stab_ramp_iters=400, precond_experimental_enabled=0, limiter_freeze_iter=args.freeze_iter,
                      residual_reduction_target=0.0, force_stability_tol=0.0, min_iterations=1, iterations=cap)
    hm = HipMesh(mesh, block_cells=args.batch_block_cells)
    log(f"[adjoint] primal: {mesh.num_cells} cells to a {args.adjoint_primal_drop:.0e} density drop (cap {cap})")
    with HipBatch([hm], [(0, case.mach, case.aoa_deg)], case) as b:
        r = b.run(cap, args.adjoint_primal_drop, iters_per_sync=500, use_graph=True)
        U = b.state(0)
        ratios = b.limiter_ratios(0)
        hm.set_limiter_ratios(ratios)
    row = r["cases"][0]
    iters = row["iterations"]
    done = replace(case, iterations=iters)
    # The measured gradients: the torch.autograd.Function over the converged coefficients (its forward is the
    # same one-case batch primal; each backward is one adjoint solve), once for CL and once for CD.
    import torch
    from aero_hip.torch_ops import last_run, steady_coefficients
    params = torch.tensor(np.concatenate([[case.aoa_deg, case.mach], amp]), requires_grad=True)
    solve_kw = dict(primal="batch", rho_drop=args.adjoint_primal_drop, tol=args.adjoint_tol, max_iters=4 * cap,
                    jacobian=True, cfl=args.adjoint_solve_cfl)
    log("[adjoint] torch Function forward (primal)")
    coef = steady_coefficients(params, base, case, solve_kw=solve_kw)
    t_primal, t_iters = last_run["primal_seconds"], last_run["primal_iterations"]
    out, t_adj = {}, {}
    names = ["aoa", "mach"] + [f"shape{i}" for i in range(len(amp))]
    for k, name in enumerate(("cl", "cd")):
        log(f"[adjoint] torch Function backward ({name})")
        (grad,) = torch.autograd.grad(coef[k], params, retain_graph=(k == 0))
        t_adj[name] = last_run["adjoint_seconds"]
        out[name] = dict(J=float(coef[k]), d=dict(zip(names, grad.tolist())),
                         adjoint=dict(iterations=last_run["adjoint_iterations"], solves=last_run["adjoint_solves"]))
    del coef
    t_both = sum(t_adj.values())  # the first backward solves both adjoints (paired), the second reuses them
    # checks: central differences of converged solutions for all 22 parameters and both objectives (the 44
    # perturbed solutions continued from the converged state in one batch, with the same frozen ratios),
    # and tangent solves for four parameters (one solve serves CL and CD)
    from aero_hip.adjoint import fd_sensitivities_batch
    log("[adjoint] 44 finite-difference solutions in one batch")
    fd = fd_sensitivities_batch(base, amp, done, U, ratios, cfl=args.fd_cfl, max_iters=args.fd_iters)
    checks = {"fd_seconds": round(fd["seconds"], 3), "fd_iterations": fd["iterations"], "per_parameter": {}}
    worst = {"cl": 0.0, "cd": 0.0}
    for pname, v in fd["d"].items():
        prow = {"fd_converged": v["converged"], "h": v["h"]}
        for obj in ("cl", "cd"):
            a = out[obj]["d"][pname]
            rel = abs(a - v[obj]) / abs(v[obj])
            worst[obj] = max(worst[obj], rel)
            prow[obj] = dict(adjoint=a, fd=v[obj], rel_adjoint_fd=rel)
        checks["per_parameter"][pname] = prow
    checks["max_rel_adjoint_fd"] = worst
    rc, _, obj, _ = case_linearization(done)
    tworst = {"cl": 0.0, "cd": 0.0}
    for pname in ("aoa", "mach", "shape3", "shape14"):
        log(f"[adjoint] tangent solve {pname}")
        if pname in ("aoa", "mach"):
            _, seed, _, dq = case_linearization(done, param=pname)
            seed_geo = None
        else:
            seed = seed_geo = geometry_seed(tan, int(pname[5:]))
        v, ti = hm.tangent_solve(rc, seed, U, cfl=args.adjoint_solve_cfl, max_iters=4 * cap, tol=args.adjoint_tol,
                                 check_every=500)
        prow = checks["per_parameter"][pname]
        prow["tangent_iterations"] = ti["iterations"]
        for name, w in (("cl", 1.0), ("cd", 0.0)):
            ob = hm.objective(replace_obj(obj, w, 1.0 - w), U, seed=seed_geo)
            j_p = ob["dJ_seed"] if seed_geo is not None else (ob["dJ_daoa"] if pname == "aoa" else ob["dJ_dq"] * dq)
            d_tan = j_p + float(np.sum(ob["dJdU"] * v))
            a = out[name]["d"][pname]
            rel = abs(a - d_tan) / abs(a)
            tworst[name] = max(tworst[name], rel)
            prow[name].update(tangent=d_tan, rel_adjoint_tangent=rel)
    checks["max_rel_adjoint_tangent"] = tworst
    hm.close()
    ratio = t_both / 2.0 / t_primal  # per gradient: dCL and dCD come from one paired solve
    res = {"metric": METRIC, "value": round(ratio, 3), "unit": "adjoint/primal time per gradient", "n_gpus": 1,
           "steps": t_iters,
           "warmup": 0, "ms_per_step": round(t_primal / max(t_iters, 1) * 1e3, 4), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "adjoint_mixed66k_hh20_hllc_venkat", "cells": mesh.num_cells, "mesh": mkw,
                      "bumps": 20, "amplitude_range": 0.002, "seed": 4242, "mach": 0.6, "aoa_deg": 2.0,
                      "cfl_max": args.adjoint_cfl, "linear_solve_cfl": args.adjoint_solve_cfl,
                      "limiter_freeze_iter": args.freeze_iter, "rho_drop_target": args.adjoint_primal_drop,
                      "adjoint_tol": args.adjoint_tol},
           "adjoint": {"gradients": "torch.autograd.grad through aero_hip.torch_ops.steady_coefficients",
                       "primal_seconds": round(t_primal, 3), "primal_iterations": t_iters,
                       "check_primal_iterations": iters,
                       "primal_rho_drop": row["rho_ratio"], "primal_converged": bool(row["converged"]),
                       "adjoint_seconds_both": round(t_both, 3),
                       "adjoint_over_primal_both": round(t_both / t_primal, 3),
                       "adjoint_seconds": {k: round(v, 3) for k, v in t_adj.items()},
                       "adjoint_iterations_max": out["cl"]["adjoint"]["iterations"],
                       "cl": out["cl"]["J"], "cd": out["cd"]["J"],
                       "dcl": out["cl"]["d"], "dcd": out["cd"]["d"], "checks": checks}}
    print(json.dumps(res), flush=True)
    return 0


def replace_obj(obj, w_cl, w_cd=0.0, w_cm=0.0):
    """A copy of an aero_objective with other weights."""
    o = type(obj).from_buffer_copy(obj)
    o.w_cl, o.w_cd, o.w_cm = w_cl, w_cd, w_cm
    return o


def run_vjp(args) -> int:
    """The adjoint product R_U^T psi (aero_hip_residual_vjp: face Jacobians, then the gather and gradient
    transpose kernels) on the jittered mixed mesh of --vjp's size, HLLC + venkat at a perturbed M 0.6
    state; --vjp-calls calls (--vjp-frozen: with the limiter frozen at that state, so that the face
    Jacobians take their factored form).  Meant to run under rocprofv3 for the per-kernel times; the line
    reports the wall time per call."""
    import numpy as np
    import aero_hip
    from aero_hip import HipMesh, freestream, generate_mesh, residual_config

    if not aero_hip.hip_available():
        raise SystemExit("bench.py: no gfx950 device visible")
    nc, nr = UNEVEN_MESHES[args.vjp]
    m = generate_mesh(precision=8, num_circumferential=nc, num_radial=nr, farfield_radius=20.0,
                      radial_stretch=200.0, cells="mixed", quad_layers=30, jitter=0.1, seed=12345, rcm=True)
    ff = freestream(0.6, 2.0)[1]
    U = perturbed_state(m.cell_center, ff)
    cfg = residual_config(ff, flux="hllc", limiter="venkat", second_order=True, stab_mach_floor=0.02)
    psi = np.random.default_rng(1).standard_normal((m.num_cells, 5))
    walls = []
    with HipMesh(m, block_cells=args.batch_block_cells) as hm:
        hm.set_state(U)
        if args.vjp_frozen:
            hm.freeze_limiter(cfg)
            cfg = residual_config(ff, flux="hllc", limiter="venkat", second_order=True, stab_mach_floor=0.02,
                                  limiter_frozen=True)
        for i in range(args.vjp_calls):
            t0 = time.time()
            out = hm.residual_vjp(cfg, psi)
            walls.append(time.time() - t0)
            log(f"[vjp] call {i}: {walls[-1] * 1e3:.2f} ms")
    assert np.all(np.isfinite(out))
    print(json.dumps({"metric": "adjoint_product_ms", "value": round(1e3 * float(np.median(walls)), 4), "unit": "ms",
                      "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": f"vjp_mixed{args.vjp}_hllc_venkat" + ("_frozen" if args.vjp_frozen else ""),
                                 "cells": m.num_cells,
                                 "faces": m.num_faces, "calls": args.vjp_calls}}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="mixed1m_hllc_venkat", choices=sorted(WORKLOADS))
    ap.add_argument("--block-cells", type=int, default=0,
                    help="cells per block of the headline residual workload (0: the layout probe chooses block size "
                         "and accumulation for the mesh)")
    ap.add_argument("--batch-block-cells", type=int, default=256,
                    help="cells per block of the batched workloads (sweep, adjoint, uneven)")
    ap.add_argument("--no-renumber", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU-baseline sample length (0 = skip)")
    ap.add_argument("--no-variants", dest="variants", action="store_false",
                    help="skip the other schemes / low-Mach lines timed beside the headline workload")
    ap.add_argument("--sweep", action="store_true", help="run the 64-case batched sweep instead")
    ap.add_argument("--sweep-iters", type=int, default=60000, help="iteration cap per case in --sweep")
    ap.add_argument("--sweep-seq-cases", type=int, default=64,
                    help="cases timed sequentially in --sweep (64: every case, no extrapolation)")
    ap.add_argument("--no-sweep", dest="with_sweep", action="store_false",
                    help="default mode: skip the batched-sweep object")
    ap.add_argument("--no-adjoint", dest="with_adjoint", action="store_false",
                    help="default mode: skip the adjoint object")
    ap.add_argument("--default-seq-cases", type=int, default=4,
                    help="default mode: sequential cases timed for the sweep object's baseline")
    ap.add_argument("--sweep-no-graph", action="store_true",
                    help="--sweep: launch the batch's kernels directly instead of replaying hipGraphs (for "
                         "rocprofv3 kernel statistics; slower)")
    ap.add_argument("--sweep-cfl", type=float, default=2.4,
                    help="CFL after the ramp in --sweep / --uneven (2.4: the most cases converged of 2.0, 2.4, "
                         "2.8 and 3.2 on the sweep)")
    ap.add_argument("--freeze-iter", type=int, default=1000,
                    help="--sweep / --adjoint: iteration at which the limiter is frozen (-1: never)")
    ap.add_argument("--adjoint", action="store_true", help="run the discrete-adjoint workload instead")
    ap.add_argument("--uneven", action="store_true",
                    help="run the batched primal + adjoint workload over uneven meshes instead")
    ap.add_argument("--uneven-cases", type=int, default=16, help="--uneven: cases per mesh size")
    ap.add_argument("--uneven-meshes", default="16k,65k,262k", help="--uneven: mesh sizes (of 16k, 65k, 262k)")
    ap.add_argument("--uneven-iters", type=int, default=120000, help="--uneven: iteration cap of the primal batch")
    ap.add_argument("--uneven-seed", type=int, default=5150, help="--uneven: seed of shapes, Mach and AoA")
    ap.add_argument("--uneven-lowmach", choices=["precond", "plain"], default="precond",
                    help="--uneven: M 0.05 cases with the low-Mach preconditioner (its long runs diverge) or without")
    ap.add_argument("--uneven-adjoint-tol", type=float, default=1e-8,
                    help="--uneven: linear residual drop of the batched adjoint solves (the primal's 8-decade drop: "
                         "an adjoint solved below its linearisation point's own convergence adds no accuracy)")
    ap.add_argument("--adjoint-tol", type=float, default=1e-11,
                    help="--adjoint: linear residual drop of the adjoint and tangent solves")
    ap.add_argument("--adjoint-iters", type=int, default=150000, help="--adjoint: iteration cap of the primal")
    ap.add_argument("--adjoint-primal-drop", type=float, default=1e-10,
                    help="--adjoint: density-residual drop of the primal the gradients are taken at (1e-6 "
                         "gradients need a primal converged beyond 8 decades)")
    ap.add_argument("--adjoint-solve-cfl", type=float, default=2.4,
                    help="--adjoint: CFL of the linear (adjoint / tangent) pseudo-time solves; with the limiter "
                         "frozen the linearised iteration on this case converges at 2.5 and diverges from 3.0")
    ap.add_argument("--adjoint-cfl", type=float, default=2.0,
                    help="--adjoint: CFL of the primal and the adjoint (at 2.4 the AoA / Mach finite-difference "
                         "continuations no longer converge to 1e-7)")
    ap.add_argument("--fd-iters", type=int, default=200000,
                    help="--adjoint: iteration cap of the batched finite-difference continuation")
    ap.add_argument("--fd-cfl", type=float, default=2.0,
                    help="--adjoint: fixed CFL of the FD continuation runs (they have no CFL adaptation)")
    ap.add_argument("--vjp", choices=sorted(UNEVEN_MESHES), default=None,
                    help="instead of the benchmark, time --vjp-calls adjoint products on that mixed mesh (for "
                         "rocprofv3 kernel statistics)")
    ap.add_argument("--vjp-calls", type=int, default=12)
    ap.add_argument("--vjp-frozen", action="store_true", help="--vjp: with the limiter frozen at the state")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run 5 plain device copies of MIB MiB (the reference kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 1024 * 1024 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8}))
        return 0
    if args.vjp:
        return run_vjp(args)
    if args.sweep:
        return run_sweep(args)
    if args.adjoint:
        return run_adjoint(args)
    if args.uneven:
        return run_uneven(args)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); the path has no "
              "multi-GPU execution in this build", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import aero_hip
    from aero_hip import HipMesh, freestream, generate_mesh, residual_config

    if not aero_hip.hip_available():
        raise SystemExit("bench.py: no gfx950 device visible")
    mkw, flux, lim, mach, aoa = WORKLOADS[args.workload]
    log(f"[bench] generating mesh {args.workload}")
    t0 = time.time()
    mesh = generate_mesh(precision=8, **mkw)
    N, F = mesh.num_cells, mesh.num_faces
    log(f"[bench] mesh N={N} F={F} ({time.time() - t0:.1f} s)")
    ff = freestream(mach, aoa)[1]
    U = perturbed_state(mesh.cell_center, ff)
    cfg_vals = dict(farfield=list(ff), gamma=1.4, flux_scheme=aero_hip._lib.FLUX[flux],
                    limiter=aero_hip._lib.LIMITER[lim], second_order=1, low_mach_fix=0, low_mach_cutoff=0.3,
                    stab_mach_floor=0.02, precond_farfield_bc=0, precond_enabled=0, precond_mach_ref=0.15,
                    precond_mach_min=1e-3, precond_beta_min=0.2, precond_beta_max=1.0, all_speed_fix=0,
                    all_speed_cutoff=0.25, all_speed_f_min=0.05, all_speed_ramp=1.0)
    cfg = residual_config(ff, flux=flux, limiter=lim, second_order=True, stab_mach_floor=0.02)

    t_probe = time.time()
    hm = HipMesh(mesh, renumber=not args.no_renumber, block_cells=args.block_cells)
    layout = hm.layout()
    layout["probe_ms"] = {k: (round(v, 5) if v is not None else None) for k, v in hm.layout_probe().items()}
    layout["setup_seconds"] = round(time.time() - t_probe, 2)
    log(f"[bench] layout {layout['block_cells']} cells per block, {layout['accumulation']} (probe {layout['probe_ms']})")
    hm.set_state(U)
    stream = torch.cuda.ExternalStream(hm.stream())
    log(f"[bench] warmup {args.warmup}")
    total_ms, wall, tm = time_residual(hm, cfg, args.steps, args.warmup, stream)
    ms_per_step = total_ms / args.steps
    value = N * args.steps / (total_ms / 1e3) / 1e6
    R, lam, diag = hm.get_residual()
    assert np.all(np.isfinite(R)), "non-finite residual in the timed workload"

    flux_bytes = 48 * F + 152 * N
    grad_bytes = 24 * F + 104 * N
    eval_bytes = 72 * F + 256 * N
    n = max(tm["n"], 1)
    flux_ms = tm["flux_ms"] / n
    grad_ms = tm["gradient_ms"] / n
    achieved = flux_bytes / (flux_ms * 1e-3) / 1e9
    ceiling = copy_ceiling()
    counters = counters_for(args.workload, layout)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": counters.get("k_flux_bytes"),
                "kernel": ("k_flux (reconstruction + Riemann flux + face->cell gather)"
                           if layout["accumulation"] == "gather" else
                           "k_flux_colour (every colour's launch of one evaluation, and the zero fill of R)"),
                "algorithmic_bytes_per_launch": flux_bytes, "avg_launch_ms": round(flux_ms, 5),
                "launch_timing": "HIP event pair around every launch on the handle's stream, over a second "
                                 "pass of the same steps (the timed region records no per-launch events)",
                "copy_ceiling_gbs": round(ceiling, 1), "frac_of_copy_ceiling": round(achieved / ceiling, 4)}
    if counters.get("k_flux_valu_wave_insts"):
        # FP64 VALU issue: a wave64 FP64 instruction holds a 16-lane SIMD for 4 cycles; 1024 SIMDs at 2.4 GHz
        issue = counters["k_flux_valu_wave_insts"] * 4 / (VALU_SIMDS * VALU_CLOCK_HZ * flux_ms * 1e-3)
        roofline["valu_issue_frac"] = round(issue, 4)
        roofline["valu_wave_insts_per_launch"] = counters["k_flux_valu_wave_insts"]
    extra = {"gradient_kernel_ms": round(grad_ms, 5),
             "gradient_kernel_gbs": round(grad_bytes / (grad_ms * 1e-3) / 1e9, 1) if grad_ms > 0 else None,
             "eval_hbm_frac": round(eval_bytes * args.steps / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
             "eval_algorithmic_bytes": eval_bytes, "wall_s": round(wall, 4), "layout": layout,
             "fallbacks": diag}
    # The workload's other numerics on the same mesh: the (Rusanov, minmod) scheme pair and the low-Mach
    # preconditioned all-speed run (M 0.01, precond on, all-speed HLLC fix f_min 0.05), as reported lines
    # beside the headline one.
    variants = {}
    if args.variants:
        for vname, vmach, vcfg in variant_configs(flux, lim, mach, aoa):
            log(f"[bench] variant {vname}")
            vff = freestream(vmach, aoa)[1]
            hm.set_state(perturbed_state(mesh.cell_center, vff))
            v_ms, _, vtm = time_residual(hm, vcfg(vff), args.steps, args.warmup, stream)
            vR, _, vdiag = hm.get_residual()
            vn = max(vtm["n"], 1)
            v_flux = vtm["flux_ms"] / vn
            variants[vname] = {"value": round(N * args.steps / (v_ms / 1e3) / 1e6, 2), "unit": "Mcell-res/s",
                               "ms_per_step": round(v_ms / args.steps, 5), "mach": vmach,
                               "flux_kernel_ms": round(v_flux, 5),
                               "flux_kernel_hbm_frac": round(flux_bytes / (v_flux * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                               "finite": bool(np.all(np.isfinite(vR))), "fallbacks": vdiag}
    hm.close()
    # The other two headline quantities, bounded so that the default run stays within a few minutes: the
    # 64-case batched sweep (its sequential baseline on --default-seq-cases spread cases) and the adjoint /
    # primal time of the config-4 gradients (no finite-difference checks).
    sweep = adjoint = None
    if args.with_sweep:
        m = measure_sweep(args, args.default_seq_cases)
        sweep = dict(m["sweep"], value=m["value"], unit=m["unit"], config=m["config"])
        sweep.pop("cl"), sweep.pop("cd")
    if args.with_adjoint:
        adjoint = measure_adjoint_pair(args)
    cpu = cpu_ref = None
    if args.cpu_seconds > 0:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(mesh, cfg_vals, U, args.cpu_seconds)
        log("[bench] reference cpu path")
        cpu_ref = cpu_reference(mesh, cfg_vals, U, args.cpu_seconds)
    out = {"metric": METRIC, "value": round(value, 2), "unit": "Mcell-res/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": args.workload, "cells": N, "faces": F, "flux": flux, "limiter": lim,
                      "mach": mach, "aoa_deg": aoa, "mesh": mkw, "block_cells": layout["block_cells"],
                      "renumber": not args.no_renumber},
           "roofline": roofline, "cpu_baseline": cpu, "cpu_reference": cpu_ref, "cpu_host": host_cpu(),
           "variants": variants, "sweep": sweep, "adjoint": adjoint,
           "detail": extra}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
