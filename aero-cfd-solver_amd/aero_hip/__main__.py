"""Command line of the HIP backend: `python -m aero_hip run CASE [--out DIR] [--mesh FILE]` and
`python -m aero_hip sweep SWEEP.yaml [--out DIR] [--batched] [--mesh FILE]`, the `cfd run` / `cfd sweep`
commands of python/cfd/cli.py:23-80 with the backend fixed to hip (case files: the reference's YAML or
native key=value format), and `python -m aero_hip mesh CASE FILE`, which writes the case's O-grid to a
binary mesh file that --mesh reads back."""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m aero_hip", description="AeroCFD Euler runs on the HIP backend")
    sub = ap.add_subparsers(dest="command")
    r = sub.add_parser("run", help="Run one case file (YAML or native key=value).")
    r.add_argument("case", type=Path)
    r.add_argument("--out", type=Path, default=Path("results/run_hip"))
    r.add_argument("--block-cells", type=int, default=256)
    r.add_argument("--mesh", type=Path, default=None, help="binary mesh file to run on (python -m aero_hip mesh)")
    mk = sub.add_parser("mesh", help="Write a case file's O-grid to a binary mesh file.")
    mk.add_argument("case", type=Path)
    mk.add_argument("file", type=Path)
    s = sub.add_parser("sweep", help="Run a sweep YAML (base_case + sweep.aoa_deg / sweep.mach).")
    s.add_argument("sweep", type=Path)
    s.add_argument("--out", type=Path, default=Path("results/sweep_hip"))
    s.add_argument("--jobs", type=int, default=1, help="recorded like the reference's; runs go to one GPU")
    s.add_argument("--batched", action="store_true", help="all cases in one device batch")
    s.add_argument("--block-cells", type=int, default=256)
    s.add_argument("--mesh", type=Path, default=None, help="binary mesh file to run on (python -m aero_hip mesh)")
    args = ap.parse_args(argv)
    if args.command is None:
        ap.print_help()
        return 2
    if args.command == "mesh":  # host only: no device needed
        from .case_file import load_case
        from .mesh import save_mesh
        from .outputs import case_and_mesh
        try:
            _, m = case_and_mesh(load_case(args.case))
            save_mesh(m, args.file)
        except Exception as exc:
            print(f"Mesh generation failed: {exc}")
            return 1
        print(f"Wrote {args.file}: {m.num_cells} cells, {m.num_faces} faces")
        return 0
    from . import hip_available
    from .outputs import run_case_file, run_sweep_file
    if not hip_available():
        print("HIP backend requested but no gfx950 device is visible.")
        return 2
    try:
        if args.command == "run":
            summary = run_case_file(args.case, args.out, block_cells=args.block_cells, mesh_path=args.mesh)
            print(f"Run completed. Outputs: {args.out}")
            print(f"Final coefficients: Cl={summary['cl']:.6f}, Cd={summary['cd']:.6f}, Cm={summary['cm']:.6f}")
            print(json.dumps(summary))
        else:
            res = run_sweep_file(args.sweep, args.out, batched=args.batched, block_cells=args.block_cells,
                                 jobs=args.jobs, progress=lambda i, n: print(f"run {i}/{n} done", flush=True),
                                 mesh_path=args.mesh)
            print(json.dumps(res))
    except Exception as exc:  # the CLI reports failures like the reference's (cli.py:62-64)
        print(f"Run failed: {exc}")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
