"""Write golden vectors for the pressure path from the reference compiled from its own sources.

Run in a container where the reference sources exist (``make -C oracle ref`` first):
    python tests/golden/make_golden.py
Each fixture is an .npz with the case inputs, the reference's CSR operator and regularized RHS
(pressure.cpp:123-212 + ch_sparse_krylov.hpp:105-133), its ICC(0)/ILDLT(0) factor
(ch_sparse_krylov.hpp:252-419), one SpMV and one preconditioner application on a seeded probe
vector (235-250, 444-513), the scalar CGMonitor log of the reference's PCG on that system
(ch_sparse_krylov.hpp:47-58, 515-660) and the full Solver::solve_pressure_correction result
(pressure.cpp:1031). The velcorr_*.npz fixtures hold the inputs and outputs of
Solver::apply_pressure_velocity_correction (operators.cpp:480-534) on whole ghosted buffers, and
of apply_pressure_bc (143-181) on a random ghosted field. The bicgstab_*.npz fixtures hold a
momentum-shaped CSR system with the result of the reference's build_diagonal_preconditioner +
solve_preconditioned_bicgstab (ch_sparse_krylov.hpp:421-431, 839-963) on it. The momentum_*.npz
fixtures hold the inputs and outputs of Solver::compute_momentum_fluxes + solve_momentum_predictor
(momentum.cpp:55-85, 157-502) on whole ghosted buffers.
The files hold data only (no code) and load with numpy.load(allow_pickle=False).
"""
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from ding_nsch_core_solver_amd import workloads as W  # noqa: E402
from oracle import ref  # noqa: E402

U, D, N = W.BC_UNSET, W.BC_DIRICHLET, W.BC_NEUMANN


def cases():
    out = []
    out.append(("icpcg_12x10_neumann_dr1000", "icpcg", W.droplet_case(12, 10, seed=21, density_ratio=1000.0,
                                                                       radius=0.3)))
    out.append(("ildlt_pcg_10x12_dirichlet", "ildlt_pcg",
                W.droplet_case(10, 12, seed=22, density_ratio=100.0, bc_type=(D, N, N, D),
                               bc_value=(1.0, 0.5, -0.25, -2.0))))
    c = W.droplet_case(14, 9, seed=23, density_ratio=250.0, step_index=2)
    rng = np.random.default_rng(123)
    c.p_n = rng.standard_normal(c.n) * 10.0
    c.p_nm1 = rng.standard_normal(c.n) * 10.0
    out.append(("split_icpcg_14x9_history", "split_icpcg", c))
    out.append(("icpcg_16x8_periodic_x", "icpcg", W.droplet_case(16, 8, seed=24, density_ratio=30.0,
                                                                  periodic_x=True)))
    out.append(("icpcg_8x16_periodic_y", "icpcg", W.droplet_case(8, 16, seed=25, density_ratio=30.0,
                                                                  periodic_y=True)))
    out.append(("ildlt_pcg_12x12_periodic_xy", "ildlt_pcg", W.droplet_case(12, 12, seed=26, density_ratio=1.0,
                                                                             periodic_x=True, periodic_y=True)))
    c = W.droplet_case(9, 13, seed=27, density_ratio=600.0, step_index=0, periodic_x=True)
    c.p_n = np.random.default_rng(7).standard_normal(c.n)
    out.append(("split_ildlt_pcg_9x13_periodic_x", "split_ildlt_pcg", c))
    return out


def write(name, scheme, case):
    use_split = scheme.startswith("split")
    kind = 2 if "ildlt" in scheme else 1
    rs = ref.RefSystem(case, use_split)
    ok, lower, diag, fl = rs.factor(kind)
    probe = np.random.default_rng(2024).standard_normal(rs.n)
    p, it, res = ref.pressure_solve(scheme, case)
    monitor = rs.pcg_monitored(case.max_iterations, case.tolerance)[4] if ok else np.zeros(0)
    empty = np.zeros(0)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        scheme=np.array(scheme), nx=case.nx, ny=case.ny, lx=case.lx, ly=case.ly, dt=case.dt,
        periodic_x=int(case.periodic_x), periodic_y=int(case.periodic_y),
        bc_type=np.array(case.bc_type, dtype=np.int32), bc_value=np.array(case.bc_value, dtype=np.float64),
        max_iterations=case.max_iterations, tolerance=case.tolerance, density_ratio=case.density_ratio,
        step_index=case.step_index, rho=case.rho, u=case.u, v=case.v,
        p_n=case.p_n if case.p_n is not None else empty, p_nm1=case.p_nm1 if case.p_nm1 is not None else empty,
        row_ptr=rs.row_ptr, col_idx=rs.col_idx, values=rs.values, rhs=rs.rhs,
        factor_ok=int(ok), floor=fl, lower=lower, diag=diag,
        probe=probe, probe_Ax=rs.apply_matrix(probe), probe_Mx=rs.apply_precond(probe),
        iterations=it, residual=res, pressure=p, monitor=monitor)
    print(f"{name}: n={rs.n} nnz={rs.nnz} iterations={it} residual={res:.3e} monitor records={len(monitor)}")


def velcorr_cases():
    """(name, case, velocity BCs, scheme, ghost) of the correction fixtures."""
    out = []
    out.append(("velcorr_11x7_neumann_lid", W.droplet_case(11, 7, seed=31, density_ratio=80.0, lx=1.1, ly=0.7),
                W.VelocityBC(top_wall_velocity_x=1.0), "icpcg", 3))
    out.append(("velcorr_9x10_split_periodic_x",
                W.droplet_case(9, 10, seed=32, density_ratio=300.0, periodic_x=True, step_index=4,
                               bc_type=(U, U, D, N), bc_value=(0.0, 0.0, 1.25, -0.5)),
                W.VelocityBC(u_type=(U, U, N, D), u_value=(0.0, 0.0, 0.2, 0.6), v_type=(U, U, D, N),
                             v_value=(0.0, 0.0, -0.1, 0.3)), "split_icpcg", 3))
    out.append(("velcorr_8x8_dirichlet_g2_periodic_y",
                W.droplet_case(8, 8, seed=33, density_ratio=20.0, periodic_y=True, bc_type=(D, N, U, U),
                               bc_value=(2.0, 0.75, 0.0, 0.0)),
                W.VelocityBC(u_type=(D, D, U, U), u_value=(0.4, -0.3, 0.0, 0.0), v_type=(D, N, U, U),
                             v_value=(0.1, 0.05, 0.0, 0.0)), "split_icpcg", 2))
    return out


def write_velcorr(name, case, vbc, scheme, g):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    nx, ny = case.nx, case.ny
    cell, ush, vsh = W.ghosted_shape(nx, ny, g), W.ghosted_shape(nx + 1, ny, g), W.ghosted_shape(nx, ny + 1, g)
    f = dict(rho_mid=(1.0 + 60.0 * rng.random(cell)).reshape(-1), u_star=rng.standard_normal(ush).reshape(-1),
             v_star=rng.standard_normal(vsh).reshape(-1), p_corr=(rng.standard_normal(cell) * 1e2).reshape(-1),
             u=rng.standard_normal(ush).reshape(-1), v=rng.standard_normal(vsh).reshape(-1),
             pressure=(rng.standard_normal(cell) * 1e2).reshape(-1),
             pressure_prev=(rng.standard_normal(cell) * 1e2).reshape(-1))
    u, v, pr, prev, speeds = ref.velocity_correction(scheme, case, vbc, g, **f)
    ghost_in = rng.standard_normal(cell).reshape(-1)
    ghost_out = ref.pressure_ghosts(case, ghost_in.copy(), g)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        scheme=np.array(scheme), ghost=g, nx=nx, ny=ny, lx=case.lx, ly=case.ly, dt=case.dt,
        periodic_x=int(case.periodic_x), periodic_y=int(case.periodic_y),
        bc_type=np.array(case.bc_type, dtype=np.int32), bc_value=np.array(case.bc_value, dtype=np.float64),
        density_ratio=case.density_ratio, step_index=case.step_index,
        u_type=np.array(vbc.u_type, dtype=np.int32), u_value=np.array(vbc.u_value, dtype=np.float64),
        v_type=np.array(vbc.v_type, dtype=np.int32), v_value=np.array(vbc.v_value, dtype=np.float64),
        bottom_wall_velocity_x=vbc.bottom_wall_velocity_x, top_wall_velocity_x=vbc.top_wall_velocity_x,
        **{"in_" + k: a for k, a in f.items()},
        out_u=u, out_v=v, out_pressure=pr, out_pressure_prev=prev, out_speeds=np.array(speeds),
        ghost_in=ghost_in, ghost_out=ghost_out)
    print(f"{name}: speeds={speeds}")


def bicgstab_cases():
    """(name, momentum_system kwargs, max_iterations, tolerance, use_relative, random x0 seed or None)"""
    return [("bicgstab_momentum_12x10_abs", dict(nx=12, ny=10, seed=41), 50, 1e-8, False, None),
            ("bicgstab_momentum_20x16_rel_x0", dict(nx=20, ny=16, seed=42, viscosity_ratio=300.0), 50, 1e-11, True, 7),
            ("bicgstab_momentum_16x16_limit3", dict(nx=16, ny=16, seed=43, dt=1e-2), 3, 1e-15, False, None)]


def write_bicgstab(name, kw, max_it, tol, rel, x0_seed):
    s = W.momentum_system(**kw)
    x0 = np.zeros(s.n) if x0_seed is None else np.random.default_rng(x0_seed).standard_normal(s.n)
    x, it, ab, rr, itr = ref.bicgstab(s, x0, max_iterations=max_it, tolerance=tol, use_relative=rel)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), row_ptr=s.row_ptr, col_idx=s.col_idx, values=s.values,
                        rhs=s.rhs, x0=x0, max_iterations=max_it, tolerance=tol, use_relative=int(rel), out_x=x,
                        out_iterations=it, out_absolute=ab, out_relative=rr, out_iterate=itr)
    print(f"{name}: {it} iterations, |r| {ab:.3e}")


def momentum_golden_cases():
    """Momentum predictor steps (workloads.momentum_case) covering walls with wall speed and gravity,
    Neumann and Dirichlet values, and a periodic stiff step needing many BiCGSTAB iterations."""
    VB = W.VelocityBC
    return [("momentum_12x10_lid_gravity", W.momentum_case(12, 10, seed=51, vbc=VB(top_wall_velocity_x=1.0),
                                                           body_force_y=-0.98, re=10.0)),
            ("momentum_10x8_neumann_values_step0", W.momentum_case(
                10, 8, seed=52, step_index=0, vbc=VB(u_type=(N, N, D, N), u_value=(0.3, -0.2, 0.5, 0.1),
                                                     v_type=(D, N, N, D), v_value=(0.2, 0.4, -0.3, 0.6)))),
            ("momentum_9x7_periodic_xy_stiff", W.momentum_case(9, 7, seed=53, periodic_x=True, periodic_y=True,
                                                               dt=1e-2, re=0.5, density_ratio=20.0))]


MOMENTUM_SCALARS = ("nx", "ny", "lx", "ly", "dt", "re", "body_force_x", "body_force_y", "periodic_x", "periodic_y",
                    "step_index", "max_iterations", "tolerance", "ghost")
MOMENTUM_ARRAYS = ("u", "v", "rho_mid", "eta", "eta_prev", "surface_fx_u", "surface_fy_v", "u_prev", "v_prev")


def write_momentum(name, mc):
    u_star, v_star, u_prev, v_prev, it, ab = ref.momentum_predictor(mc)
    vb = mc.vbc
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"), **{k: getattr(mc, k) for k in MOMENTUM_SCALARS},
        **{"in_" + k: getattr(mc, k) for k in MOMENTUM_ARRAYS},
        u_type=np.array(vb.u_type, dtype=np.int32), u_value=np.array(vb.u_value, dtype=np.float64),
        v_type=np.array(vb.v_type, dtype=np.int32), v_value=np.array(vb.v_value, dtype=np.float64),
        bottom_wall_velocity_x=vb.bottom_wall_velocity_x, top_wall_velocity_x=vb.top_wall_velocity_x,
        out_u_star=u_star, out_v_star=v_star, out_u_prev=u_prev, out_v_prev=v_prev, out_iterations=it,
        out_absolute=ab)
    print(f"{name}: {it} iterations, |r| {ab:.3e}")


if __name__ == "__main__":
    if not ref.available():
        sys.exit("build the reference checker first: make -C oracle ref")
    for name, scheme, case in cases():
        write(name, scheme, case)
    for name, case, vbc, scheme, g in velcorr_cases():
        write_velcorr(name, case, vbc, scheme, g)
    for args in bicgstab_cases():
        write_bicgstab(*args)
    for name, mc in momentum_golden_cases():
        write_momentum(name, mc)
