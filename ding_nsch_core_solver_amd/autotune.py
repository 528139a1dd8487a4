"""Autotuner for the sweep wave roles (pp_problem.ring_depth) on the local GPU.

For each grid shape the FP64 solve runs a fixed number of PCG iterations (tolerance 0, so every
candidate does the same work) with every choice of wave roles, and the fastest wins: the staged
sweeps with each valid ring depth (how far the 8 - S loader waves run ahead of the S sweep waves)
and the direct sweeps (ring_depth -1: no loaders, every sweep wave loads its own operands). None of
them changes the arithmetic, so every candidate returns the same bits (checked here too). The table is written
as JSON (default: tuning_gfx950.json beside this module), which Engine installs at creation.

    python -m ding_nsch_core_solver_amd.autotune [--iterations 40] [--repeats 3] [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import engine as E
from . import workloads as W

DEFAULT_SHAPES = ((64, 64), (96, 192), (128, 128), (192, 192), (256, 256), (192, 384), (64, 256), (320, 160))
MAX_STAGED_WAVES = 6


DIRECT = -1  # pp_problem.ring_depth for direct sweeps: every sweep wave loads its own operands


def candidates(nx: int) -> list[int]:
    """Wave-role choices for a grid nx wide: the staged sweeps with each accepted ring depth (8 - S
    loader waves), and the direct sweeps (no loaders). Wider grids than 6 waves only run direct."""
    S = (nx + 63) // 64
    if S > MAX_STAGED_WAVES:
        return []
    return [r for r in (8, 16, 32, 64) if S * r <= 64] + [DIRECT]


def tune_shape(eng: E.Engine, nx: int, ny: int, iterations: int, repeats: int, log=print) -> dict:
    case = W.droplet_case(nx, ny, seed=0, density_ratio=100.0, tolerance=0.0, max_iterations=iterations)
    timings, reference = {}, None
    for ring in candidates(nx):
        best = float("inf")
        for _ in range(repeats):
            p, rep = eng.solve(case, E.PP_ICPCG, ring_depth=ring)
            if reference is None:
                reference = p
            elif not np.array_equal(p, reference):
                raise RuntimeError(f"ring depth {ring} changed the result at {nx}x{ny}")
            best = min(best, rep["kernel_ms"])
        timings[ring] = best / iterations
        log(f"[tune] {nx}x{ny} ring {ring:2d}: {timings[ring] * 1e3:8.2f} us/iteration")
    if not timings:
        return {}
    ring = min(timings, key=timings.get)
    return {"nx": nx, "ny": ny, "ring_depth": ring,
            "us_per_iteration": {str(k): round(v * 1e3, 3) for k, v in sorted(timings.items())}}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--iterations", type=int, default=40)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", default=E.TUNING_FILE)
    ap.add_argument("--shapes", default="", help="comma list of NXxNY (default: the built-in set)")
    args = ap.parse_args(argv)
    shapes = DEFAULT_SHAPES
    if args.shapes:
        shapes = tuple(tuple(int(v) for v in s.split("x")) for s in args.shapes.split(","))
    eng = E.Engine(0, tuning=False)
    entries = []
    t0 = time.time()
    for nx, ny in shapes:
        e = tune_shape(eng, nx, ny, args.iterations, args.repeats, log=lambda m: print(m, flush=True))
        if e:
            entries.append(e)
    table = {"device": "gfx950", "metric": "device us per PCG iteration (FP64 icpcg, droplet dr=100)",
             "iterations": args.iterations, "repeats": args.repeats, "seconds": round(time.time() - t0, 1),
             "entries": entries}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"[tune] wrote {args.out} ({len(entries)} shapes)", flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
