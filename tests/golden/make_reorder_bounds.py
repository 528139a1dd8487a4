"""Per-case bounds on how far a reordered summation may move the FP64 pressure.

The engine keeps every element-wise operation of the reference and differs only in the order its
reductions (dot products, means) are summed. This script measures, with the oracle on the CPU, how
far that alone moves the solution of each parity case: the max-abs relative difference between the
reference's sequential order and two other orders, pairwise and the engine's canonical tree. The bound
written for each (case, scheme) is three times the larger of the two spreads (floor 1e-15), so a
change that makes the engine's order drift further from the reference than reordering explains
fails the GPU parity test. Run with the oracle built (make -C oracle oracle):
    python tests/golden/make_reorder_bounds.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import cpu  # noqa: E402
from tests.cases import SCHEMES, large_cases, small_cases  # noqa: E402
from ding_nsch_core_solver_amd import workloads as W  # noqa: E402

FACTOR, FLOOR = 3.0, 1.0e-15
OUT = os.path.join(HERE, "reorder_bounds.json")


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def history_case(case):
    """The 192x192 split case with the pressure history the GPU test builds (icpcg solves of seeds 1
    and 2, equal bit for bit to the canonical-order oracle)."""
    case.p_n, _, _ = cpu.pressure_solve("icpcg", W.droplet_case(192, 192, seed=1), cpu.CANONICAL)
    case.p_nm1, _, _ = cpu.pressure_solve("icpcg", W.droplet_case(192, 192, seed=2), cpu.CANONICAL)
    return case


def main():
    table = {}
    groups = [(small_cases(), SCHEMES), (large_cases(), ("icpcg", "ildlt_pcg", "split_icpcg"))]
    for cases, schemes in groups:
        for name, case in cases:
            if name == "192x192_split_history":
                case = history_case(case)
            for scheme in schemes:
                try:
                    p_seq, _, _ = cpu.pressure_solve(scheme, case, cpu.SEQUENTIAL)
                except RuntimeError:
                    continue  # the factorization fails in every order: the test checks the error text
                p_pw, _, _ = cpu.pressure_solve(scheme, case, cpu.PAIRWISE)
                p_can, _, _ = cpu.pressure_solve(scheme, case, cpu.CANONICAL)
                pw, can = rel(p_pw, p_seq), rel(p_can, p_seq)
                table[f"{name}/{scheme}"] = {"pairwise": pw, "canonical": can,
                                             "bound": FACTOR * max(pw, can, FLOOR)}
                print(name, scheme, table[f"{name}/{scheme}"], flush=True)
    with open(OUT, "w") as f:
        json.dump({"factor": FACTOR, "floor": FLOOR, "cases": table}, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
