"""Pressure-solve inputs captured from real timesteps of the reference's own full runs.

Run where the reference sources exist (``make -C oracle ref`` first):
    python tests/golden/make_timestep_golden.py
The reference Solver (compiled from its own sources) runs its CH -> momentum predictor -> pressure
loop (solver.cpp:224-283) through oracle.ref.RefRun; at the chosen steps the inputs of the pressure
step are saved: rho_mid_, u_star_, v_star_, pressure_ and pressure_previous_step_ (interiors), the
step index, and the results of the reference's solve on exactly that state: iterations, relative
residual and the SHA-256 of the interior of pressure_correction_, for the run's own icpcg scheme and
for split_icpcg on the same state (with its real pressure history).

Runs:
* static_droplet_64: 64x64 periodic static droplet (the CPU-runnable workload), all 10 steps;
* rising_bubble_192x384: the 192x384 rising bubble at steps 1 and 18. The reference's own run
  of this workload diverges at step 20: the chemical potential at the interface grows every step
  (|mu| 1.7e3, 3.3e3, 5.4e3 at steps 18-20), c jumps to [-0.82, 1.67], the density goes negative
  and its ICC factorization of the pressure matrix fails. The same happens within 20-25 steps at
  dt 2.5e-5 and 1e-5, with 80 or 200 CH iterations, at Pe 1000 and at Cn 0.02, so step 18 is the
  last state kept;
* static_droplet_128_restart: the reference's published 128x128 validation run, restarted from its
  step-5000 snapshot (176Result/.../restart_latest.bin, read by the reference's own
  load_restart_snapshot, io.cpp:345-437), one step further. Its history.csv records 133 ICC-PCG
  iterations on every one of its 5000 steps at tolerance 1e-10.
Files hold data only (no code) and load with numpy.load(allow_pickle=False).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from ding_nsch_core_solver_amd import workloads as W  # noqa: E402
from oracle import ref  # noqa: E402

RESTART = "/root/reference/176Result/static_droplet_surface_tension_validation_128/restart_latest.bin"


def interior(run, name):
    f = run.field(name)
    return np.ascontiguousarray(f[3:-3, 3:-3]).reshape(-1).copy()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def capture(run, spec):
    state = {
        "nx": spec["nx"], "ny": spec["ny"], "lx": spec["lx"], "ly": spec["ly"], "dt": spec["dt"],
        "periodic_x": spec["periodic_x"], "periodic_y": spec["periodic_y"],
        "density_ratio": spec["density_ratio"], "tolerance": spec["pressure_tolerance"],
        "max_iterations": spec["poisson_iterations"], "step_index": run.step,
        "rho": interior(run, "rho_mid"), "u": interior(run, "u_star"), "v": interior(run, "v_star"),
        "p_n": interior(run, "pressure"), "p_nm1": interior(run, "pressure_previous_step"),
    }
    case = W.PressureCase(nx=spec["nx"], ny=spec["ny"], lx=spec["lx"], ly=spec["ly"], dt=spec["dt"],
                          periodic_x=bool(spec["periodic_x"]), periodic_y=bool(spec["periodic_y"]),
                          max_iterations=spec["poisson_iterations"], tolerance=spec["pressure_tolerance"],
                          density_ratio=spec["density_ratio"], step_index=run.step, rho=state["rho"],
                          u=state["u"], v=state["v"], p_n=state["p_n"], p_nm1=state["p_nm1"])
    p, it, res = ref.pressure_solve("split_icpcg", case)
    state.update(split_icpcg_iterations=it, split_icpcg_residual=res, split_icpcg_sha=sha(p))
    return state


def run_and_capture(name, spec, steps, keep, restart=None):
    run = ref.RefRun(spec, "icpcg", restart)
    for _ in range(steps):
        step = run.step
        run.phase_momentum()
        state = capture(run, spec) if step in keep else None
        it, res, _ = run.pressure()
        if state is not None:
            state.update(icpcg_iterations=it, icpcg_residual=res, icpcg_sha=sha(interior(run, "pressure_correction")))
            path = os.path.join(HERE, f"timestep_{name}_step{step}.npz")
            np.savez_compressed(path, **state)
            print(f"{path}: icpcg {it} iterations, residual {res:.6e}; split {state['split_icpcg_iterations']}",
                  flush=True)
        run.finish()
    run.close()


def main():
    run_and_capture("static_droplet_64", W.STATIC_DROPLET_64, W.STATIC_DROPLET_64_STEPS, set(range(10)))
    run_and_capture("rising_bubble_192x384", W.RISING_BUBBLE_192X384, 18, {0, 17})
    run_and_capture("static_droplet_128_restart", W.STATIC_DROPLET_128_RESTART, 1, {5000}, restart=RESTART)


if __name__ == "__main__":
    main()
