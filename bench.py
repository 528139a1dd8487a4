#!/usr/bin/env python3
"""Pressure-solve throughput on one MI355X.

Workload (one step = one pressure correction): 192x192 cells on [0,1]^2, Neumann pressure on all
four walls, a droplet of radius 0.2 at the centre with Cn = 0.01 and density ratio 1000
(rho = (1+phi)/2*1000 + (1-phi)/2), RHS = -div(u*)/dt of a seeded N(0,1) face velocity field with
dt = 1e-4, ICC(0)-preconditioned CG to relative residual 1e-8 (the reference's icpcg backend).
Inputs are resident in HBM when the timed region starts; every step runs assembly, factorization,
PCG and the output mean subtraction.

Prints one JSON line with solves/s, the roofline of the solve kernel (algorithmic bytes from
HIP-event kernel time) and the reference CPU path timed on this host's cores. Beside it: the same
workload in mixed precision, the velocity correction that follows the solve (and the whole device
pressure step), and the 256-system uneven batch, each with its own roofline and CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from ding_nsch_core_solver_amd import engine as E  # noqa: E402
from ding_nsch_core_solver_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(REPO, "profiles", "fp64_solve_192x192_dr1000_traffic.txt")
MIXED_TRAFFIC_FILE = os.path.join(REPO, "profiles", "mixed_solve_192x192_dr1000_traffic.txt")
BATCH_TRAFFIC_FILE = os.path.join(REPO, "profiles", "batched_256_uneven_mixed_traffic.txt")
BATCH64_TRAFFIC_FILE = os.path.join(REPO, "profiles", "batched_256_uneven_fp64_traffic.txt")
CORR_TRAFFIC_FILE = os.path.join(REPO, "profiles", "velocity_correction_192x192_traffic.txt")
BICG_TRAFFIC_FILE = os.path.join(REPO, "profiles", "bicgstab_momentum_192x192_traffic.txt")
BICG1024_TRAFFIC_FILE = os.path.join(REPO, "profiles", "bicgstab_momentum_1024x1024_traffic.txt")
MOMENTUM_TRAFFIC_FILE = os.path.join(REPO, "profiles", "momentum_predictor_192x384_traffic.txt")
CH_TRAFFIC_FILE = os.path.join(REPO, "profiles", "ch_solve_192x384_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup_bytes(n):
    """Algorithmic bytes of one solve's setup: assembly, regularization, factorization."""
    return 120 * n


def iteration_bytes(n, fp32_inner=False):
    """Algorithmic bytes of one PCG iteration: 25 n-length streams (SpMV, x/r update, two sweeps, p update)."""
    return (100 if fp32_inner else 200) * n


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (profiles/)."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, seconds):
    """Reference icpcg (compiled from the reference's sources) on one pinned core, bounded sample."""
    kind = "reference"
    try:
        from oracle import ref as impl
        if not impl.available():
            raise FileNotFoundError
        solve = lambda: impl.pressure_solve("icpcg", case)  # noqa: E731
    except (ImportError, FileNotFoundError):
        from oracle import cpu as impl
        if not impl.available():
            log("[bench] WARNING: no CPU baseline: neither oracle/_ref/libdingref.so (the compiled reference) "
                "nor oracle/liboracle.so is built")
            return None, None
        log("[bench] WARNING: oracle/_ref/libdingref.so (the reference compiled from its own sources) is not "
            "present; the CPU baseline falls back to the repository's single-threaded restatement "
            "(oracle/liboracle.so, kind \"port\")")
        kind = "port"
        solve = lambda: impl.pressure_solve("icpcg", case, impl.SEQUENTIAL)  # noqa: E731
    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        result = solve()  # warm-up, also the parity reference
        times = []
        t_end = time.perf_counter() + seconds
        while time.perf_counter() < t_end and len(times) < 200:
            t0 = time.perf_counter()
            solve()
            times.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, old)
    med = float(np.median(times))
    return {
        "value": 1.0 / med, "unit": "solves/s", "cores": 1, "kind": kind,
        "sample": f"{len(times)} icpcg solves of the same 192x192 dr=1000 seed-0 input, median "
                  f"{med * 1e3:.1f} ms; host {cpu_model()}, nproc {os.cpu_count()}",
    }, result


def mixed_bytes(n, iterations, cycles):
    """Algorithmic bytes of one mixed-precision solve: FP64 setup, FP32 operator/factor copies, FP32
    inner iterations, FP64 outer residual and correction once per cycle (and the final check)."""
    return setup_bytes(n) + 72 * n + iterations * iteration_bytes(n, fp32_inner=True) + (cycles + 1) * 56 * n


def batched_leg(eng, torch, dev, repeats):
    """The 256-system uneven batch (grids from 64x64 to 256x256, density ratios 1..1000, icpcg to
    1e-8 with FP32 inner PCG inside FP64 refinement), inputs resident in HBM; one
    pp_solve_batched_device call per repeat. The same batch in plain FP64 is timed beside it."""
    cases = W.batched_cases(256, seed=0)
    tens = []
    for c in cases:
        tens.append({k: torch.from_numpy(np.ascontiguousarray(getattr(c, k)).reshape(-1)).to(dev)
                     for k in ("rho", "u", "v")})
        tens[-1]["out"] = torch.zeros(c.n, dtype=torch.float64, device=dev)
    ptr = lambda k: [t[k].data_ptr() for t in tens]  # noqa: E731
    stream = torch.cuda.current_stream(dev).cuda_stream

    def timed(precision):
        def run():
            return eng.solve_batched_device(cases, ptr("rho"), ptr("u"), ptr("v"), ptr("out"), scheme=E.PP_ICPCG,
                                            precision=precision, stream=stream)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kernel_ms = []
        for _ in range(repeats):
            reps = run()
            kernel_ms.append(reps[0]["kernel_ms"])
        torch.cuda.synchronize()
        return reps, (time.perf_counter() - t0) / repeats, float(np.mean(kernel_ms))

    cells = sum(c.n for c in cases)
    reps64, elapsed64, kms64 = timed(E.PP_FP64)
    p64 = [t["out"].cpu().numpy() for t in tens]
    alg64 = sum(setup_bytes(c.n) + r["iterations"] * iteration_bytes(c.n) for c, r in zip(cases, reps64))
    reps, elapsed, kms = timed(E.PP_MIXED)
    diff = max(float(np.abs(t["out"].cpu().numpy() - q).max() / np.abs(q).max()) for t, q in zip(tens, p64))
    alg = sum(mixed_bytes(c.n, r["iterations"], r["outer_cycles"]) for c, r in zip(cases, reps))
    achieved = alg / (kms * 1e-3) / 1e9
    # CU occupancy: each system runs on one CU for the sum of its device phase timers, so busy CU
    # time over (batch kernel time x CUs) is the share of the chip the batch keeps working
    cus = torch.cuda.get_device_properties(dev).multi_processor_count

    def occupancy(rs, batch_ms):
        busy = [sum(r["phase_ms"].values()) for r in rs]
        return {"cus": cus, "mean_system_ms": float(np.mean(busy)), "max_system_ms": float(np.max(busy)),
                "busy_fraction": float(sum(busy) / (batch_ms * cus))}
    out = {"systems": len(cases), "cells": cells, "precision": "fp32 inner / fp64 refinement",
           "solves_per_s": len(cases) / elapsed, "ms_per_batch": elapsed * 1e3, "kernel_ms": kms,
           "inner_iterations_total": int(sum(r["iterations"] for r in reps)),
           "outer_cycles_total": int(sum(r["outer_cycles"] for r in reps)),
           "max_relative_residual": float(max(r["relative_residual"] for r in reps)),
           "max_abs_diff_vs_fp64_batch_over_max_abs": diff,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(BATCH_TRAFFIC_FILE),
                        "algorithmic_bytes_per_launch": alg},
           "cu_occupancy": occupancy(reps, kms),
           "fp64": {"solves_per_s": len(cases) / elapsed64, "ms_per_batch": elapsed64 * 1e3, "kernel_ms": kms64,
                    "cu_occupancy": occupancy(reps64, kms64),
                    "iterations_total": int(sum(r["iterations"] for r in reps64)),
                    "roofline": {"bound": "hbm", "achieved": alg64 / (kms64 * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": alg64 / (kms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                 "traffic": read_traffic(BATCH64_TRAFFIC_FILE),
                                 "algorithmic_bytes_per_launch": alg64}}}
    return out, cases


def correction_bytes(n, split=False):
    """Algorithmic bytes of one velocity correction on n cells: rho, p_corr, u*, v* and the old
    pressure read once; u, v, the new pressure and the previous pressure written once (the split
    scheme also reads the previous pressure)."""
    return (80 if split else 72) * n


def correction_leg(eng, torch, dev, case, steps, step_repeats, cpu_seconds, cpu=True):
    """The step after the solve on the same workload: velocity correction and pressure-history shift
    (Solver::apply_pressure_velocity_correction) with 3 ghost layers and no-slip velocity walls,
    non-split icpcg. pressure_correction_ comes from pp_solve_device_field on the same Field2D
    buffers (solve + apply_pressure_bc on the device). All fields resident in HBM; one
    pp_velocity_correction_device call per step. The whole device pressure step (field solve +
    correction) is timed beside it."""
    nx, ny, G = case.nx, case.ny, 3
    rng = np.random.default_rng(7)
    pad = lambda a, shape: np.pad(np.asarray(a, dtype=np.float64).reshape(shape), G, mode="symmetric")  # noqa: E731
    host = {"rho_mid": pad(case.rho, (nx, ny)), "u_star": pad(case.u, (nx + 1, ny)),
            "v_star": pad(case.v, (nx, ny + 1)), "p_corr": np.zeros(W.ghosted_shape(nx, ny, G)),
            "u": np.zeros(W.ghosted_shape(nx + 1, ny, G)), "v": np.zeros(W.ghosted_shape(nx, ny + 1, G)),
            "pressure": rng.standard_normal(W.ghosted_shape(nx, ny, G)),
            "pressure_prev": rng.standard_normal(W.ghosted_shape(nx, ny, G))}
    host = {k: np.ascontiguousarray(a).reshape(-1) for k, a in host.items()}
    dt = {k: torch.from_numpy(a.copy()).to(dev) for k, a in host.items()}
    order = ("rho_mid", "u_star", "v_star", "p_corr", "u", "v", "pressure", "pressure_prev")
    vbc = W.VelocityBC()
    stream = torch.cuda.current_stream(dev).cuda_stream

    def field_solve():
        return eng.solve_device_field(case, dt["rho_mid"].data_ptr(), dt["u_star"].data_ptr(),
                                      dt["v_star"].data_ptr(), dt["pressure"].data_ptr(),
                                      dt["pressure_prev"].data_ptr(), dt["p_corr"].data_ptr(), scheme=E.PP_ICPCG,
                                      ghost=G, stream=stream)

    def run():
        return eng.velocity_correction_device(case, vbc, *(dt[k].data_ptr() for k in order), scheme=E.PP_ICPCG,
                                              ghost=G, stream=stream)

    field_solve()
    host["p_corr"] = dt["p_corr"].cpu().numpy()

    def reset():
        for k in ("u", "v", "pressure", "pressure_prev"):
            dt[k].copy_(torch.from_numpy(host[k]).to(dev))

    def first_call_and_timing():
        """Outputs of one call on the initial fields, then the mean time of `steps` calls."""
        reset()
        sp = run()
        res = {k: dt[k].cpu().numpy() for k in ("u", "v", "pressure", "pressure_prev")}
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        kernel_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
            kernel_ms.append(eng.last_kernel_ms())
        torch.cuda.synchronize()
        return sp, res, (time.perf_counter() - t0) / steps, float(np.mean(kernel_ms))

    # the reference's sequential pressure mean (pp_set_exact_mean), then the default tree order
    eng.set_exact_mean(True)
    speeds_x, first_x, elapsed_x, kms_x = first_call_and_timing()
    eng.set_exact_mean(False)
    speeds, first, elapsed, kms = first_call_and_timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(step_repeats):
        field_solve()
        run()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) / step_repeats * 1e3
    alg = correction_bytes(case.n)
    out = {"value": 1.0 / elapsed, "unit": "corrections/s", "ms_per_step": elapsed * 1e3, "kernel_ms": kms,
           "device_pressure_step_ms": step_ms, "mean_order": "tree (default)",
           "kernels": "correction_faces_kernel + correction_history_kernel + correction_finish_kernel",
           "roofline": {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": read_traffic(CORR_TRAFFIC_FILE),
                        "algorithmic_bytes_per_launch": alg,
                        "note": "three small launches over 0.1 M faces and cells: bound by launch and ghost-pass "
                                "latency, not bytes"},
           "exact_mean": {"value": 1.0 / elapsed_x, "ms_per_step": elapsed_x * 1e3, "kernel_ms": kms_x,
                          "mean_chain_terms": case.n,
                          "note": "pp_set_exact_mean(1): the reference's sequential mean, one dependent FP64 add "
                                  "per cell"}}
    if cpu:
        try:
            from oracle import ref
            if not ref.available():
                raise FileNotFoundError
        except (ImportError, FileNotFoundError):
            return out
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
        try:
            expect = ref.velocity_correction("icpcg", case, vbc, G, **host)
            reps = max(10, min(2000, int(cpu_seconds / 1e-3)))
            timing = []
            ref.velocity_correction("icpcg", case, vbc, G, **host, repeats=reps, timing=timing)
        finally:
            os.sched_setaffinity(0, old)
        names = ("u", "v", "pressure", "pressure_prev")
        out["parity"] = {
            "exact_mean_bitwise_equal_to_reference": bool(
                all(np.array_equal(first_x[k], e) for k, e in zip(names, expect[:4]))
                and tuple(speeds_x) == tuple(expect[4])),
            "tree_mean_max_rel_deviation": float(max(np.abs(first[k] - e).max() / max(np.abs(e).max(), 1e-300)
                                                     for k, e in zip(names, expect[:4]))),
            "tolerance": 1e-12}
        out["cpu_baseline"] = {"value": 1.0 / timing[0], "unit": "corrections/s", "cores": 1, "kind": "reference",
                               "sample": f"{reps} calls of apply_pressure_velocity_correction on the same inputs, "
                                         f"median {timing[0] * 1e6:.0f} us (solver set-up excluded)"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def bicgstab_bytes(n, nnz):
    """Algorithmic HBM bytes of one BiCGSTAB iteration (FP64 vectors, int32 CSR indices), three
    passes: A reads r, p_old, v_old, M^-1, r_hat and the CSR, writes p, p_hat, v; B reads r, v,
    M^-1, p_hat and the CSR, writes s, update, s_hat, t; C reads update, s_hat, x, s, t, r_hat and
    writes update, x, r. Each array counts once per pass (neighbour gathers hit cache)."""
    csr = 4 * (n + 1) + 12 * nnz
    return (8 + 8 + 9) * 8 * n + 2 * csr


def momentum_leg(eng, torch, dev, steps, cpu_seconds, cpu=True, side=192):
    """The momentum predictor's linear solve (build_diagonal_preconditioner +
    solve_preconditioned_bicgstab with the reference's momentum settings: 50 iterations, absolute
    tolerance 1e-8, x0 = 0) on a synthetic side x side momentum-shaped CSR system (73 344 unknowns
    at 192), resident in HBM; one pp_bicgstab_device call per step, x zeroed before each."""
    s = W.momentum_system(side, side, seed=0)
    t = {k: torch.from_numpy(getattr(s, k)).to(dev) for k in ("row_ptr", "col_idx", "values", "rhs")}
    x = torch.zeros(s.n, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def run():
        x.zero_()
        return eng.bicgstab_device(s.n, t["row_ptr"].data_ptr(), t["col_idx"].data_ptr(), t["values"].data_ptr(),
                                   t["rhs"].data_ptr(), x.data_ptr(), stream=stream)

    for _ in range(3):
        rep = run()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        rep = run()
        kernel_ms.append(rep["kernel_ms"])
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t0) / steps
    kms = float(np.mean(kernel_ms))
    nnz = int(s.row_ptr[-1])
    alg = bicgstab_bytes(s.n, nnz) * rep["iterations"]
    out = {"value": 1.0 / elapsed, "unit": "solves/s", "ms_per_step": elapsed * 1e3, "device_ms": kms,
           "iterations": rep["iterations"], "absolute_residual": rep["absolute_residual"], "unknowns": s.n,
           "nonzeros": nnz, "data": "synthetic momentum-shaped system (droplet density and viscosity)",
           "kernels": "bicg_setup_kernel + per iteration bicg_pass_a/b/c_kernel",
           "roofline": {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": read_traffic(BICG_TRAFFIC_FILE),
                        "algorithmic_bytes_per_solve": alg,
                        "note": "whole solve (all passes and the host's state checks); traffic per solve"}}
    if side == 1024:
        out["roofline"]["traffic"] = read_traffic(BICG1024_TRAFFIC_FILE)
    elif side != 192:
        out["roofline"]["traffic"] = None
    if cpu:
        try:
            from oracle import ref
            if not ref.available():
                raise FileNotFoundError
        except (ImportError, FileNotFoundError):
            return out
        x_gpu = x.cpu().numpy()
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
        try:
            x_r, it_r, ab_r, _, _ = ref.bicgstab(s, max_iterations=50, tolerance=1e-8)
            timing = []
            reps = max(1, min(500, int(cpu_seconds / (0.03 * (side / 192.0) ** 2))))
            ref.bicgstab(s, max_iterations=50, tolerance=1e-8, repeats=reps, timing=timing)
        finally:
            os.sched_setaffinity(0, old)
        out["parity"] = {"iterations_gpu": rep["iterations"], "iterations_cpu": it_r,
                         "max_abs_diff_over_max_abs": float(np.abs(x_gpu - x_r).max() / np.abs(x_r).max())}
        out["cpu_baseline"] = {"value": 1.0 / timing[0], "unit": "solves/s", "cores": 1, "kind": "reference",
                               "sample": f"{reps} solves of the same system, mean {timing[0] * 1e3:.2f} ms"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def momentum_predictor_bytes(n, nnz, iterations, cells_u, cells_v):
    """Algorithmic HBM bytes of one device momentum predictor (pp_momentum_predictor_device) with n
    unknowns, nnz matrix entries and ghosted face fields of cells_u / cells_v doubles:
    assembly reads u, v, rho, eta, eta_prev, the surface force and the previous convection field of
    each face (9 doubles per unknown) and writes the row slots (9 columns + 9 values), the count,
    rhs and x0; the count / scan / compact passes read the counts twice and the slots once and write
    row_ptr and the CSR; the convection pass reads u, v and writes u_prev, v_prev; BiCGSTAB moves
    bicgstab_bytes per iteration plus its setup (b, x, CSR, M^-1, r, r_hat); the scatter reads x, u,
    v and writes u_star, v_star."""
    assembly = n * (9 * 8 + 9 * 12 + 4 + 8 + 8)
    compact = n * (2 * 4 + 9 * 12) + 4 * (n + 1) + 12 * nnz
    convection = 2 * 8 * (cells_u + cells_v)
    setup = 12 * nnz + 4 * (n + 1) + 6 * 8 * n
    scatter = 8 * n + 2 * 8 * (cells_u + cells_v)
    return assembly + compact + convection + setup + iterations * bicgstab_bytes(n, nnz) + scatter


def momentum_predictor_leg(eng, torch, dev, steps, cpu_seconds, cpu=True):
    """The whole momentum predictor on the device (Solver::compute_momentum_fluxes +
    solve_momentum_predictor, momentum.cpp:55-502): assembly of the implicit u-v system from the
    Field2D buffers, BiCGSTAB from x0 = u / v (50 iterations, absolute 1e-8), scatter and ghost fill
    of u_star_ / v_star_. The grid and fluid are the rising-bubble run's: 192x384 on [0,1]x[0,2],
    density ratio 1000, viscosity ratio 100, Re 10, gravity -0.98, dt 5e-5, no-slip walls; seeded
    velocity, surface-force and convection fields. One pp_momentum_predictor_device call per step
    (the stored convection fields restored before each)."""
    mc = W.momentum_case(192, 384, ly=2.0, density_ratio=1000.0, viscosity_ratio=100.0, re=10.0, dt=5.0e-5,
                         body_force_y=-0.98, radius=0.25, velocity_scale=0.05)
    names = ("u", "v", "rho_mid", "eta", "eta_prev", "surface_fx_u", "surface_fy_v", "u_prev", "v_prev")
    d = {k: torch.from_numpy(np.ascontiguousarray(getattr(mc, k), dtype=np.float64)).to(dev) for k in names}
    keep = {k: d[k].clone() for k in ("u_prev", "v_prev")}
    d["u_star"], d["v_star"] = torch.zeros_like(d["u"]), torch.zeros_like(d["v"])
    stream = torch.cuda.current_stream(dev).cuda_stream

    def run():
        d["u_prev"].copy_(keep["u_prev"])
        d["v_prev"].copy_(keep["v_prev"])
        return eng.momentum_predictor_device(mc, *(d[k].data_ptr() for k in names), d["u_star"].data_ptr(),
                                             d["v_star"].data_ptr(), stream=stream)

    for _ in range(3):
        rep = run()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        rep = run()
        kernel_ms.append(rep["kernel_ms"])
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t0) / steps
    kms = float(np.mean(kernel_ms))
    n = mc.unknowns
    nnz = 9 * n  # upper bound: 5 own-component and 4 cross entries per row
    alg = momentum_predictor_bytes(n, nnz, rep["iterations"], d["u"].numel(), d["v"].numel())
    out = {"value": 1.0 / elapsed, "unit": "predictor steps/s", "ms_per_step": elapsed * 1e3, "device_ms": kms,
           "iterations": rep["iterations"], "absolute_residual": rep["absolute_residual"], "unknowns": n,
           "grid": "192x384 rising-bubble fluid (density ratio 1000, viscosity ratio 100, Re 10, dt 5e-5)",
           "data": "synthetic fields of the rising-bubble run's shape",
           "kernels": "momentum_assemble/count/block_scan/compact/convection/ghost + BiCGSTAB passes + "
                      "momentum_scatter/ghost",
           "roofline": {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": read_traffic(MOMENTUM_TRAFFIC_FILE),
                        "algorithmic_bytes_per_step": alg,
                        "note": "whole predictor (assembly, solve, scatter); nnz taken as its bound 9 n"}}
    if cpu:
        try:
            from oracle import ref
            if not ref.available():
                raise FileNotFoundError
        except (ImportError, FileNotFoundError):
            log("[bench] WARNING: momentum predictor CPU baseline skipped: oracle/_ref/libdingref.so is not present")
            return out
        u_gpu, v_gpu = d["u_star"].cpu().numpy(), d["v_star"].cpu().numpy()
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
        try:
            u_r, v_r, _, _, it_r, _ = ref.momentum_predictor(mc)
            timing = []
            reps = max(3, int(cpu_seconds / 0.25))
            ref.momentum_predictor(mc, repeats=reps, timing=timing)
        finally:
            os.sched_setaffinity(0, old)
        scale = max(np.abs(u_r).max(), np.abs(v_r).max())
        out["parity"] = {"iterations_gpu": rep["iterations"], "iterations_cpu": it_r,
                         "max_abs_diff_over_max_abs": float(max(np.abs(u_gpu - u_r).max(), np.abs(v_gpu - v_r).max())
                                                            / scale)}
        out["cpu_baseline"] = {"value": 1.0 / timing[0], "unit": "predictor steps/s", "cores": 1, "kind": "reference",
                               "sample": f"{reps} calls of compute_momentum_fluxes + solve_momentum_predictor on the "
                                         f"same fields, median {timing[0] * 1e3:.1f} ms"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def ch_diag_bytes(n, nnz, iterations):
    """Algorithmic HBM bytes of one CH solve with the diagonal preconditioner as full-chip passes:
    setup (CSR 12 nnz + row_ptr, x, b, M^-1 read; r, z, p written); per iteration pass A (CSR, z and
    p_old read, p and Ap written) and pass B (p, x, r, Ap, M^-1 read; x, r, z written); the closing
    residual pass (CSR, x, b)."""
    csr = 12 * nnz + 4 * (n + 1)
    setup = csr + 3 * 8 * n + 3 * 8 * n
    per_it = (csr + 2 * 8 * n + 2 * 8 * n) + (5 * 8 * n + 3 * 8 * n)
    return setup + iterations * per_it + csr + 2 * 8 * n


def ch_leg(eng, torch, dev, steps, cpu_seconds, cpu=True):
    """The Cahn-Hilliard semi-implicit linear solve (Solver::solve_phase_linear_system_eq25,
    phase_field.cpp:437-503) on the device: PCG on the 13-point operator alpha0 I - beta1 Lap +
    beta2 Lap^2 with the ILDLT(0) preconditioner, or the diagonal one where ILDLT(0) fails (as it does
    on this grid, in the reference too), from x0 = c, 30 iterations to 1e-8 (the reference's
    ch_inner_iterations / ch_tolerance defaults). Grid 192x384 on [0,1]x[0,2] (the rising-bubble
    run), Cn 0.01, Pe 100, dt 5e-5 (alpha0 = 3 / (2 dt)); seeded right-hand side around the bubble's
    phase field. The operator and factor are built once and cached, as the reference's solver caches
    them across steps; one pp_ch_solve_device call per step, c restored before each."""
    cc = W.ch_case(192, 384, ly=2.0, radius=0.25, dt=5.0e-5, seed=0)
    d_rhs = torch.from_numpy(np.ascontiguousarray(cc.rhs)).to(dev)
    c0 = torch.from_numpy(np.ascontiguousarray(cc.c0)).to(dev)
    d_c = c0.clone()
    stream = torch.cuda.current_stream(dev).cuda_stream

    def run():
        d_c.copy_(c0)
        return eng.ch_solve_device(cc, d_rhs.data_ptr(), d_c.data_ptr(), stream=stream)

    for _ in range(2):
        rep = run()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        rep = run()
        kernel_ms.append(rep["kernel_ms"])
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t0) / steps
    kms = float(np.mean(kernel_ms))
    n = cc.nx * cc.ny
    nnz = 13 * n  # interior rows: the 13-point stencil (walls trim a few)
    alg = ch_diag_bytes(n, nnz, rep["iterations"])
    out = {"value": 1.0 / elapsed, "unit": "CH solves/s", "ms_per_step": elapsed * 1e3, "device_ms": kms,
           "iterations": rep["iterations"], "preconditioner": rep["solver_name"],
           "equation_residual": rep["equation_residual"], "levels": rep["levels"], "unknowns": n,
           "data": "synthetic phase field and right-hand side on the rising-bubble grid",
           "kernels": "ch_diag_setup + per iteration ch_diag_pass_a / ch_diag_pass_b + ch_diag_eq / finish",
           "roofline": {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": read_traffic(CH_TRAFFIC_FILE),
                        "algorithmic_bytes_per_solve": alg,
                        "note": "ILDLT(0) of this operator fails at this grid, as in the reference, so the solve "
                                "is the diagonal-preconditioned CG (full-chip passes); whole solve over HIP-event "
                                "time"}}
    if cpu:
        try:
            from oracle import ref
            if not ref.available():
                raise FileNotFoundError
        except (ImportError, FileNotFoundError):
            log("[bench] WARNING: CH CPU baseline skipped: oracle/_ref/libdingref.so is not present")
            return out
        c_gpu = d_c.cpu().numpy()
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(old)})
        try:
            rch = ref.RefCh(cc)
            c_r, it_r, _, _, name_r = rch.solve(cc.alpha0, cc.rhs, cc.c0)
            timing = []
            reps = max(3, int(cpu_seconds / 0.1))
            rch.solve(cc.alpha0, cc.rhs, cc.c0, repeats=reps, timing=timing)
            rch.close()
        finally:
            os.sched_setaffinity(0, old)
        out["parity"] = {"iterations_gpu": rep["iterations"], "iterations_cpu": it_r, "solver_cpu": name_r,
                         "max_abs_diff_over_max_abs": float(np.abs(c_gpu - c_r).max() / np.abs(c_r).max())}
        out["cpu_baseline"] = {"value": 1.0 / timing[0], "unit": "CH solves/s", "cores": 1, "kind": "reference",
                               "sample": f"{reps} solves with the reference's cached operator, median "
                                         f"{timing[0] * 1e3:.1f} ms"}
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def batched_cpu_baseline(cases, threads):
    """Reference icpcg on every system of the batch, one system per thread (the reference is
    single-threaded; independent solves run concurrently), wall time of the whole batch."""
    from concurrent.futures import ThreadPoolExecutor
    try:
        from oracle import ref as impl
        if not impl.available():
            return None
    except ImportError:
        return None
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda c: impl.pressure_solve("icpcg", c), cases))
    wall = time.perf_counter() - t0
    return {"value": len(cases) / wall, "unit": "solves/s", "cores": threads, "kind": "reference",
            "sample": f"all {len(cases)} systems of the batch, {threads} threads, wall {wall:.1f} s"}


TOLERANCE_TARGET = 1.0e-12  # relative pressure agreement with the reference asked of the FP64 drop-in


def parity_leg(eng, seeds=(0, 1, 2)):
    """FP64 drop-in parity on the 192x192 dr=1000 workload: every seed x scheme (icpcg, ildlt_pcg,
    split_icpcg) solved on the GPU and by the compiled reference (oracle/_ref) on the host; reports the
    iteration counts and the max-abs relative pressure difference against the 1e-12 target."""
    try:
        from oracle import ref
        if not ref.available():
            raise FileNotFoundError
    except (ImportError, FileNotFoundError):
        log("[bench] WARNING: parity leg skipped: oracle/_ref/libdingref.so is not present")
        return None
    rows, worst, same_iters = [], 0.0, True
    for seed in seeds:
        case = W.droplet_case(192, 192, seed=seed, density_ratio=1000.0, tolerance=1e-8, max_iterations=20000)
        for name, code in (("icpcg", E.PP_ICPCG), ("ildlt_pcg", E.PP_ILDLT_PCG), ("split_icpcg", E.PP_SPLIT_ICPCG)):
            p_gpu, rep = eng.solve(case, code)
            p_ref, it_ref, _ = ref.pressure_solve(name, case)
            dev = float(np.abs(p_gpu - p_ref).max() / np.abs(p_ref).max())
            rows.append({"seed": seed, "scheme": name, "iterations_gpu": rep["iterations"], "iterations_cpu": it_ref,
                         "max_abs_diff_over_max_abs": dev})
            worst = max(worst, dev)
            same_iters = same_iters and rep["iterations"] == it_ref
    return {"tolerance_target": TOLERANCE_TARGET, "max_deviation": worst,
            "met": bool(worst <= TOLERANCE_TARGET and same_iters), "iterations_equal": bool(same_iters),
            "cases": rows,
            "note": "the engine sums reductions in a fixed tree order and is bit-identical to the oracle replaying "
                    "that order; the deviation from the reference's sequential order is the reordering spread of "
                    "the reference's own arithmetic"}


def copy_calibration(nbytes=256 << 20, repeats=3):
    """A plain device-to-device copy of known size, run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    to calibrate the HBM byte counters: nbytes read and nbytes written per copy."""
    import torch
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(repeats):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copied {nbytes} bytes x{repeats}")


def _read_csv(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find_csv(directory, suffix):
    for root, _, files in os.walk(directory):
        for name in sorted(files):
            if name.endswith(suffix):
                return os.path.join(root, name)
    raise FileNotFoundError(f"no *{suffix} under {directory}")


def _short_name(name):
    return name.replace("void ", "").replace("(anonymous namespace)::", "").split("(")[0]


def summarize_trace(directory, match=None):
    """Per-kernel dispatch statistics of a `rocprofv3 --kernel-trace --output-format csv` run, grouped
    by kernel name and grid size: calls, total / average / min / max duration in ms, and the kernel's
    resources (LDS, scratch, VGPR, SGPR, workgroup)."""
    rows = _read_csv(_find_csv(directory, "kernel_trace.csv"))
    groups = {}
    for r in rows:
        name = r["Kernel_Name"]
        if match and not any(m in name for m in match):
            continue
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        g = groups.setdefault((name, grid), {"d": [], "r": r})
        g["d"].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6)
    lines = [f"{'kernel':<62} {'grid':>8} {'calls':>6} {'total_ms':>10} {'avg_ms':>10} {'min_ms':>10} {'max_ms':>10}"]
    for (name, grid), g in sorted(groups.items(), key=lambda kv: -sum(kv[1]["d"])):
        d = g["d"]
        short = _short_name(name)
        lines.append(f"{short:<62} {grid:>8} {len(d):>6} {sum(d):>10.3f} {sum(d) / len(d):>10.4f} {min(d):>10.4f} "
                     f"{max(d):>10.4f}")
    lines.append("")
    for (name, grid), g in sorted(groups.items(), key=lambda kv: -sum(kv[1]["d"])):
        r = g["r"]
        short = _short_name(name)
        lines.append(f"{short} grid {grid}: workgroup {r['Workgroup_Size_X']}, LDS {r['LDS_Block_Size']} B, "
                     f"scratch {r['Scratch_Size']} B/lane, VGPR {r['VGPR_Count']}, SGPR {r['SGPR_Count']}")
    return "\n".join(lines)


def summarize_traffic(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, match=None, grids=None, calls=None):
    """HBM bytes per dispatch from two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs (KiB per
    dispatch), each scaled by what the counter reads for the --calibrate-copy run's 256 MiB device
    copy (its largest dispatches read and write 256 MiB), grouped by kernel name and grid size."""
    def per_kernel(directory):
        out = {}
        for r in _read_csv(_find_csv(directory, "counter_collection.csv")):
            name = r["Kernel_Name"]
            if match and not any(m in name for m in match):
                continue
            if grids and int(r["Grid_Size"]) not in grids:
                continue
            out.setdefault((name, int(r["Grid_Size"])), []).append(float(r["Counter_Value"]))
        return out

    def copy_value(directory):
        vals = [float(r["Counter_Value"]) for r in _read_csv(_find_csv(directory, "counter_collection.csv"))]
        return max(vals)

    mib256_kib = 256 * 1024.0
    f_scale = mib256_kib / copy_value(copy_fetch_dir)
    w_scale = mib256_kib / copy_value(copy_write_dir)
    fetch, write = per_kernel(fetch_dir), per_kernel(write_dir)
    lines = [f"copy calibration (256 MiB device copy): FETCH_SIZE {copy_value(copy_fetch_dir):.1f} KiB, "
             f"WRITE_SIZE {copy_value(copy_write_dir):.1f} KiB -> scale factors fetch x{f_scale:.4f}, "
             f"write x{w_scale:.4f} (bytes moved = counter x factor)",
             f"{'kernel':<62} {'grid':>8} {'calls':>6} {'FETCH_KiB':>12} {'WRITE_KiB':>12} {'bytes/dispatch':>16}"]
    for key in sorted(fetch, key=lambda k: -sum(fetch[k])):
        if key not in write:
            continue
        fk, wk = np.mean(fetch[key]), np.mean(write[key])
        b = (fk * f_scale + wk * w_scale) * 1024.0
        short = _short_name(key[0])
        lines.append(f"{short:<62} {key[1]:>8} {len(fetch[key]):>6} {fk:>12.1f} {wk:>12.1f} {b:>16.0f}")
    if calls:
        total = sum(sum(fetch[k]) * f_scale + sum(write.get(k, [0.0])) * w_scale for k in fetch) * 1024.0
        lines.append(f"all listed dispatches: {total:.0f} bytes over {calls} calls -> {total / calls:.0f} bytes per call")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-system batch leg")
    ap.add_argument("--batched-repeats", type=int, default=3)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device copy (HBM counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print the per-kernel summary of a rocprofv3 --kernel-trace --output-format csv run")
    ap.add_argument("--summarize-traffic", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="print HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE csv runs")
    ap.add_argument("--kernels", nargs="*", help="with --summarize-*: only kernels whose name contains one of these")
    ap.add_argument("--grids", nargs="*", type=int, help="with --summarize-traffic: only dispatches of these grid sizes")
    ap.add_argument("--calls", type=int, help="with --summarize-traffic: also print the listed dispatches' bytes per call")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace, args.kernels))
        return
    if args.summarize_traffic:
        print(summarize_traffic(*args.summarize_traffic, args.kernels, args.grids, args.calls))
        return

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this engine runs one pressure system per "
              "solve on one GPU (independent systems would shard by system); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        copy_calibration()
        return

    import torch
    dev = torch.device("cuda:0")
    case = W.droplet_case(192, 192, seed=args.seed, density_ratio=1000.0, tolerance=1e-8, max_iterations=20000)
    n = case.n
    d_rho = torch.from_numpy(case.rho.reshape(-1)).to(dev)
    d_u = torch.from_numpy(case.u.reshape(-1)).to(dev)
    d_v = torch.from_numpy(case.v.reshape(-1)).to(dev)
    d_out = torch.zeros(n, dtype=torch.float64, device=dev)
    eng = E.Engine(0)
    stream = torch.cuda.current_stream(dev)

    def step(precision=E.PP_FP64):
        return eng.solve_device(case, d_rho.data_ptr(), d_u.data_ptr(), d_v.data_ptr(), None, None,
                                d_out.data_ptr(), scheme=E.PP_ICPCG, precision=precision,
                                stream=stream.cuda_stream)

    log(f"[bench] warmup {args.warmup} steps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] timing {args.steps} FP64 steps")
    reports = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reports.append(step())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    p_gpu = d_out.cpu().numpy()
    iters = reports[-1]["iterations"]
    kernel_ms = float(np.mean([r["kernel_ms"] for r in reports]))
    alg_bytes = setup_bytes(n) + iters * iteration_bytes(n)
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    log(f"[bench] {args.steps / elapsed:.2f} solves/s, {elapsed / args.steps * 1e3:.3f} ms/solve, "
        f"kernel {kernel_ms:.3f} ms, iterations {iters}")

    line = {
        "metric": "pressure solves/s to 1e-8 rel resid at 192^2 dr=1000 (FP64 vs mixed); ms/step",
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single variable-coefficient pressure solve, 192x192, Neumann walls, droplet "
                               "R=0.2 Cn=0.01 density ratio 1000, seeded N(0,1) u*/v*, dt=1e-4, hip_icpcg "
                               "FP64 to rel. residual 1e-8",
                   "nx": 192, "ny": 192, "scheme": "hip_icpcg", "tolerance": 1e-8, "seed": args.seed,
                   "iterations": iters, "relative_residual": reports[-1]["relative_residual"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "kernel": "fp64_solve_kernel", "kernel_ms": kernel_ms,
                     "algorithmic_bytes_per_launch": alg_bytes},
    }

    # mixed precision on the same workload (FP32 inner PCG inside FP64 refinement, same tolerance)
    log(f"[bench] timing {args.steps} mixed-precision steps")
    for _ in range(args.warmup):
        step(E.PP_MIXED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mreps = [step(E.PP_MIXED) for _ in range(args.steps)]
    torch.cuda.synchronize()
    m_elapsed = time.perf_counter() - t0
    p_mixed = d_out.cpu().numpy()
    m_kernel = float(np.mean([r["kernel_ms"] for r in mreps]))
    m_it, m_cyc = mreps[-1]["iterations"], mreps[-1]["outer_cycles"]
    m_bytes = mixed_bytes(n, m_it, m_cyc)
    line["mixed"] = {"value": args.steps / m_elapsed, "unit": "solves/s", "ms_per_step": m_elapsed / args.steps * 1e3,
                     "kernel_ms": m_kernel, "inner_iterations": m_it, "outer_cycles": m_cyc,
                     "relative_residual": mreps[-1]["relative_residual"],
                     "speedup_vs_fp64": (args.steps / m_elapsed) / line["value"],
                     "max_abs_diff_vs_fp64_over_max_abs": float(np.abs(p_mixed - p_gpu).max() / np.abs(p_gpu).max()),
                     "roofline": {"bound": "hbm", "achieved": m_bytes / (m_kernel * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": m_bytes / (m_kernel * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "traffic": read_traffic(MIXED_TRAFFIC_FILE), "algorithmic_bytes_per_launch": m_bytes}}
    log(f"[bench] mixed: {line['mixed']['value']:.2f} solves/s, kernel {m_kernel:.2f} ms, {m_it} inner iterations in "
        f"{m_cyc} cycles")

    log("[bench] velocity correction after the solve")
    line["velocity_correction"] = correction_leg(eng, torch, dev, case, args.steps * 5, 3, 2.0,
                                                 cpu=not args.no_cpu_baseline)
    log(f"[bench] correction: {line['velocity_correction']['value']:.0f} /s, kernels "
        f"{line['velocity_correction']['kernel_ms'] * 1e3:.0f} us")

    log("[bench] momentum predictor BiCGSTAB")
    line["momentum_bicgstab"] = momentum_leg(eng, torch, dev, args.steps * 5, 3.0, cpu=not args.no_cpu_baseline)
    log(f"[bench] bicgstab: {line['momentum_bicgstab']['device_ms']:.3f} ms, "
        f"{line['momentum_bicgstab']['iterations']} iterations")
    log("[bench] momentum predictor BiCGSTAB, 1024x1024 (2.1 M unknowns)")
    line["momentum_bicgstab"]["large_1024"] = momentum_leg(eng, torch, dev, 5, 5.0, cpu=not args.no_cpu_baseline,
                                                           side=1024)
    log(f"[bench] bicgstab 1024: {line['momentum_bicgstab']['large_1024']['device_ms']:.3f} ms, "
        f"{line['momentum_bicgstab']['large_1024']['roofline']['frac'] * 100:.1f}% of HBM")

    log("[bench] momentum predictor on the device (assembly + BiCGSTAB + scatter), 192x384")
    line["momentum_predictor"] = momentum_predictor_leg(eng, torch, dev, args.steps * 2, 3.0,
                                                        cpu=not args.no_cpu_baseline)
    log(f"[bench] momentum predictor: {line['momentum_predictor']['device_ms']:.3f} ms, "
        f"{line['momentum_predictor']['iterations']} iterations")
    log("[bench] CH linear solve on the device, 192x384")
    line["ch_solve"] = ch_leg(eng, torch, dev, args.steps, 3.0, cpu=not args.no_cpu_baseline)
    log(f"[bench] CH solve: {line['ch_solve']['device_ms']:.3f} ms, {line['ch_solve']['iterations']} iterations")

    batch_cases = None
    if not args.no_batched:
        log("[bench] 256-system batch")
        line["batched"], batch_cases = batched_leg(eng, torch, dev, args.batched_repeats)
        log(f"[bench] batch: {line['batched']['solves_per_s']:.1f} solves/s, kernel {line['batched']['kernel_ms']:.1f} ms, "
            f"{line['batched']['roofline']['achieved']:.0f} GB/s algorithmic, CUs busy "
            f"{line['batched']['cu_occupancy']['busy_fraction'] * 100:.0f}% of the batch")

    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline (~{args.cpu_seconds:.0f} s on one core)")
        base, ref_result = cpu_baseline(case, args.cpu_seconds)
        line["cpu_baseline"] = base
        if ref_result is not None:
            p_ref, it_ref, res_ref = ref_result
            line["parity"] = {"iterations_gpu": iters, "iterations_cpu": it_ref,
                              "max_abs_diff_over_max_abs": float(np.abs(p_gpu - p_ref).max() / np.abs(p_ref).max())}
            log("[bench] parity leg: seeds 0-2 x icpcg / ildlt_pcg / split_icpcg against the reference")
            leg = parity_leg(eng)
            if leg is not None:
                line["parity"].update(leg)
            line["mixed"]["max_abs_diff_vs_cpu_over_max_abs"] = float(np.abs(p_mixed - p_ref).max() / np.abs(p_ref).max())
            line["speedup_vs_cpu"] = line["value"] / base["value"]
        if batch_cases is not None:
            threads = min(16, len(os.sched_getaffinity(0)))
            log(f"[bench] CPU baseline of the batch ({threads} threads)")
            bb = batched_cpu_baseline(batch_cases, threads)
            if bb is not None:
                line["batched"]["cpu_baseline"] = bb
                line["batched"]["speedup_vs_cpu"] = line["batched"]["solves_per_s"] / bb["value"]
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
