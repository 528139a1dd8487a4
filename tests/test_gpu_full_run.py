"""Full runs with the HIP pressure step in the reference's own timestep loop (CH -> momentum
predictor -> pressure correction, solver.cpp:224-283), against the same loop with the reference's
pressure step:

* the 64x64 periodic static droplet, all 10 steps, FP64 hip_icpcg and mixed precision;
* the 192x384 rising bubble (density ratio 1000, viscosity ratio 100, gravity, pressure to 1e-10)
  over its first 5 steps, FP64 and mixed.

FP64: the same iteration count on every step and final pressure, velocity and centroid within
reordering noise of the reference's. Mixed: within the solver tolerance's reach. Deviations are
printed. Needs the reference checker library (oracle/_ref); skipped where it was not built.
"""
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W
from tests.full_runs import deviation, hybrid_run, reference_run

pytestmark = pytest.mark.gpu

RUNS = [("static_droplet_64", W.STATIC_DROPLET_64, W.STATIC_DROPLET_64_STEPS),
        ("rising_bubble_192x384", W.RISING_BUBBLE_192X384, 5)]


@pytest.mark.parametrize("name,spec,steps", RUNS, ids=[r[0] for r in RUNS])
def test_full_run_with_hip_pressure_step(engine, ref_lib, name, spec, steps):
    ref_out, ref_log = reference_run(ref_lib, spec, steps)
    fp64_out, fp64_log = hybrid_run(ref_lib, engine, spec, steps)
    mix_out, mix_log = hybrid_run(ref_lib, engine, spec, steps, precision=E.PP_MIXED)
    d64, dmix = deviation(fp64_out, ref_out), deviation(mix_out, ref_out)
    t_ref = sum(e["pressure_s"] for e in ref_log) / steps * 1e3
    t64 = sum(e["pressure_s"] for e in fp64_log) / steps * 1e3
    tmix = sum(e["pressure_s"] for e in mix_log) / steps * 1e3
    print(f"[full run] {name}, {steps} steps: iterations reference {[e['iterations'] for e in ref_log]}, "
          f"hip FP64 {[e['iterations'] for e in fp64_log]}, mixed {[e['iterations'] for e in mix_log]}")
    print(f"[full run] {name}: FP64 deviation {d64}; mixed deviation {dmix}")
    print(f"[full run] {name}: pressure step per timestep: reference {t_ref:.2f} ms, hip FP64 {t64:.2f} ms "
          f"(host buffers, transfers included), mixed {tmix:.2f} ms")
    assert [e["iterations"] for e in fp64_log] == [e["iterations"] for e in ref_log]
    tol = spec["pressure_tolerance"]
    assert max(d64["pressure"], d64["u"], d64["v"]) <= 1e-7
    assert d64["centroid"] <= 1e-12
    assert max(dmix["pressure"], dmix["u"], dmix["v"]) <= 1e3 * tol
    assert dmix["centroid"] <= 1e-9


@pytest.mark.parametrize("name,spec,steps", RUNS, ids=[r[0] for r in RUNS])
def test_full_run_with_hip_momentum_and_pressure(engine, ref_lib, name, spec, steps):
    """Both the momentum predictor and the pressure step on the device inside the reference's loop."""
    ref_out, ref_log = reference_run(ref_lib, spec, steps)
    out, log = hybrid_run(ref_lib, engine, spec, steps, momentum=True)
    d = deviation(out, ref_out)
    print(f"[full run] {name} with hip momentum + pressure: momentum iterations reference "
          f"{[e['momentum_iterations'] for e in ref_log]}, hip {[e['momentum_iterations'] for e in log]}; "
          f"deviation {d}; momentum step reference "
          f"{sum(e['momentum_s'] for e in ref_log) / steps * 1e3:.2f} ms, hip device "
          f"{sum(e['momentum_kernel_ms'] for e in log) / steps:.2f} ms")
    assert [e["momentum_iterations"] for e in log] == [e["momentum_iterations"] for e in ref_log]
    assert [e["iterations"] for e in log] == [e["iterations"] for e in ref_log]
    assert max(d["pressure"], d["u"], d["v"]) <= 1e-7 and d["centroid"] <= 1e-12


@pytest.mark.parametrize("name,spec,steps", RUNS, ids=[r[0] for r in RUNS])
def test_full_run_with_hip_ch_momentum_and_pressure(engine, ref_lib, name, spec, steps):
    """The CH linear solve, the momentum predictor and the pressure step on the device inside the
    reference's loop: the same iteration counts of all three solvers on every step, and the final
    state within reordering noise of the reference-only run."""
    ref_out, ref_log = reference_run(ref_lib, spec, steps)
    out, log = hybrid_run(ref_lib, engine, spec, steps, momentum=True, ch=True)
    d = deviation(out, ref_out)
    print(f"[full run] {name} with hip CH + momentum + pressure: CH iterations reference "
          f"{[e['ch_iterations'] for e in ref_log]}, hip {[e['ch_iterations'] for e in log]}; deviation {d}; "
          f"CH device {sum(e['ch_kernel_ms'] for e in log) / steps:.2f} ms/step")
    assert [e["ch_iterations"] for e in log] == [e["ch_iterations"] for e in ref_log]
    assert [e["momentum_iterations"] for e in log] == [e["momentum_iterations"] for e in ref_log]
    assert [e["iterations"] for e in log] == [e["iterations"] for e in ref_log]
    assert max(d["pressure"], d["u"], d["v"]) <= 1e-7 and d["centroid"] <= 1e-12
