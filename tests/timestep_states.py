"""Pressure-solve states captured from the reference's own full runs (tests/golden/timestep_*.npz,
written by tests/golden/make_timestep_golden.py): the 64x64 static droplet's 10 steps, the 192x384
rising bubble at steps 1 and 18, and the published 128x128 validation run one step past its
step-5000 restart snapshot."""
import glob
import hashlib
import os

import numpy as np

from ding_nsch_core_solver_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def state_files():
    return sorted(glob.glob(os.path.join(GOLDEN, "timestep_*.npz")))


def state_id(path):
    return os.path.basename(path)[len("timestep_"):-len(".npz")]


def load_state(path):
    """(PressureCase, reference results {scheme: (iterations, residual, sha256 of p)})."""
    z = np.load(path, allow_pickle=False)
    case = W.PressureCase(nx=int(z["nx"]), ny=int(z["ny"]), lx=float(z["lx"]), ly=float(z["ly"]), dt=float(z["dt"]),
                          periodic_x=bool(z["periodic_x"]), periodic_y=bool(z["periodic_y"]),
                          max_iterations=int(z["max_iterations"]), tolerance=float(z["tolerance"]),
                          density_ratio=float(z["density_ratio"]), step_index=int(z["step_index"]),
                          rho=z["rho"].copy(), u=z["u"].copy(), v=z["v"].copy(), p_n=z["p_n"].copy(),
                          p_nm1=z["p_nm1"].copy())
    results = {s: (int(z[f"{s}_iterations"]), float(z[f"{s}_residual"]), str(z[f"{s}_sha"]))
               for s in ("icpcg", "split_icpcg")}
    return case, results


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
