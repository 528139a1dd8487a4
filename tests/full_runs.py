"""Full runs of the reference's timestep loop with the HIP pressure step plugged in.

oracle.ref.RefRun steps the reference Solver (compiled from its own sources) in three parts; between
its CH + momentum stage and the rest of the timestep, `hybrid_run` calls the engine's
pp_pressure_step on the Solver's own Field2D buffers (rho_mid_, u_star_, v_star_ in; u_, v_,
pressure_, pressure_previous_step_ and pressure_correction_ out), which is what a reference build
with a hip_* pressure scheme does at solver.cpp:271-272. `reference_run` is the same loop with the
reference's own pressure step.
"""
import time

import numpy as np

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W


def pressure_case(spec, step):
    return W.PressureCase(nx=spec["nx"], ny=spec["ny"], lx=spec["lx"], ly=spec["ly"], dt=spec["dt"],
                          periodic_x=bool(spec["periodic_x"]), periodic_y=bool(spec["periodic_y"]),
                          max_iterations=spec["poisson_iterations"], tolerance=spec["pressure_tolerance"],
                          density_ratio=spec["density_ratio"], step_index=step)


def centroid(run):
    """Centroid of the c = 1 phase (the droplet / bubble) over the interior cells."""
    spec = run.spec
    c = run.field("c")[3:-3, 3:-3]
    dx, dy = spec["lx"] / spec["nx"], spec["ly"] / spec["ny"]
    x = (np.arange(spec["nx"]) + 0.5) * dx
    y = (np.arange(spec["ny"]) + 0.5) * dy
    m = c.sum()
    return float((c.sum(axis=1) @ x) / m), float((c.sum(axis=0) @ y) / m)


def snapshot(run):
    return {k: run.field(k)[3:-3, 3:-3].copy() for k in ("pressure", "u", "v", "c")} | {"centroid": centroid(run)}


def reference_run(ref, spec, steps, scheme="icpcg", on_step=None):
    """The reference's own loop; on_step(k, log_entry) is called after each step."""
    run = ref.RefRun(spec, scheme)
    log = []
    for _ in range(steps):
        run.phase()
        msec = run.momentum()
        it, res, sec = run.pressure()
        info = run.finish()
        log.append({"iterations": it, "residual": res, "pressure_s": sec, "momentum_s": msec,
                    "momentum_iterations": info["momentum_iterations"], "ch_iterations": info["ch_iterations"]})
        if on_step is not None:
            on_step(len(log) - 1, log[-1])
    out = snapshot(run)
    run.close()
    return out, log


def hip_momentum(run, eng):
    """The step's momentum predictor on the device (pp_momentum_predictor) on the Solver's own
    fields: u_star_, v_star_ and the stored convection fields are written back."""
    mc = run.momentum_case()
    u_star, v_star, u_prev, v_prev, rep = eng.momentum_predictor(mc)
    for name, new in (("u_star", u_star), ("v_star", v_star), ("momentum_u_rhs_prev", u_prev),
                      ("momentum_v_rhs_prev", v_prev)):
        f = run.field(name)
        f[...] = new.reshape(f.shape)
    run.set_momentum_report(rep["absolute_residual"], rep["iterations"])
    return rep


def ch_case_of(spec):
    """The run's CH linear-solve settings as a workloads.ChCase (alpha0 set per step)."""
    return W.ChCase(nx=spec["nx"], ny=spec["ny"], lx=spec["lx"], ly=spec["ly"], periodic_x=bool(spec["periodic_x"]),
                    periodic_y=bool(spec["periodic_y"]), cn=spec["cn"], pe=spec["pe"],
                    max_iterations=spec["ch_inner_iterations"], tolerance=spec["ch_tolerance"])


def hybrid_run(ref, eng, spec, steps, scheme=E.PP_ICPCG, precision=E.PP_FP64, ref_scheme="icpcg", momentum=False,
               ch=False, on_step=None):
    """The reference's loop with the HIP pressure step (with momentum=True also the HIP momentum
    predictor, with ch=True also the HIP CH linear solve); on_step(k, log_entry) is called after
    each step."""
    import dataclasses
    run = ref.RefRun(spec, ref_scheme)
    vbc = W.VelocityBC()
    log = []
    ch_log = []
    cc0 = ch_case_of(spec)

    def ch_solve(alpha0, rhs, c):
        cc = dataclasses.replace(cc0, alpha0=alpha0)
        c_new, rep = eng.ch_solve(cc, rhs=rhs, c0=c)
        c[...] = c_new
        ch_log.append(rep)
        return rep["iterations"], rep["iterate_residual"], rep["equation_residual"]

    for _ in range(steps):
        step = run.step
        if ch:
            run.phase_external_ch(ch_solve)
        else:
            run.phase()
        if momentum:
            mom_rep = hip_momentum(run, eng)
        else:
            run.momentum()
        case = pressure_case(spec, step)
        f = {k: run.field(k) for k in ("rho_mid", "u_star", "v_star", "u", "v", "pressure", "pressure_previous_step",
                                        "pressure_correction")}
        t0 = time.perf_counter()
        p_corr, u, v, p, prev, _, rep = eng.pressure_step(case, vbc, f["rho_mid"], f["u_star"], f["v_star"], f["u"],
                                                          f["v"], f["pressure"], f["pressure_previous_step"],
                                                          scheme=scheme, precision=precision, ghost=3)
        sec = time.perf_counter() - t0
        for name, new in (("pressure_correction", p_corr), ("u", u), ("v", v), ("pressure", p),
                          ("pressure_previous_step", prev)):
            f[name][...] = new.reshape(f[name].shape)
        run.finish()
        entry = {"iterations": rep["iterations"], "residual": rep["relative_residual"], "pressure_s": sec,
                 "kernel_ms": rep["kernel_ms"]}
        if momentum:
            entry["momentum_iterations"] = mom_rep["iterations"]
            entry["momentum_kernel_ms"] = mom_rep["kernel_ms"]
        if ch:
            entry["ch_iterations"] = ch_log[-1]["iterations"]
            entry["ch_kernel_ms"] = ch_log[-1]["kernel_ms"]
        log.append(entry)
        if on_step is not None:
            on_step(len(log) - 1, entry)
    out = snapshot(run)
    run.close()
    return out, log


def deviation(a, b):
    """Max-abs relative differences of pressure, u, v and the centroid shift (absolute)."""
    rel = lambda x, y: float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))  # noqa: E731
    return {"pressure": rel(a["pressure"], b["pressure"]), "u": rel(a["u"], b["u"]), "v": rel(a["v"], b["v"]),
            "centroid": float(np.hypot(a["centroid"][0] - b["centroid"][0], a["centroid"][1] - b["centroid"][1]))}
