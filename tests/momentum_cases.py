"""Momentum predictor cases: every velocity-BC kind the reference supports (no-slip defaults, wall
speeds, Dirichlet and Neumann values on each edge, periodic axes), step 0 and later steps, and
time-step / Reynolds-number settings where the BiCGSTAB needs from 1 to dozens of iterations."""
from ding_nsch_core_solver_amd import workloads as W

U, D, N = W.BC_UNSET, W.BC_DIRICHLET, W.BC_NEUMANN


def small_momentum_cases():
    VB = W.VelocityBC
    out = [
        ("12x10_walls", W.momentum_case(12, 10, seed=1)),
        ("12x10_periodic_xy", W.momentum_case(12, 10, seed=2, periodic_x=True, periodic_y=True, density_ratio=1.0,
                                              viscosity_ratio=1.0)),
        ("9x11_periodic_x", W.momentum_case(9, 11, seed=3, periodic_x=True)),
        ("11x9_periodic_y_step0", W.momentum_case(11, 9, seed=4, periodic_y=True, step_index=0)),
        ("10x8_neumann_dirichlet_values", W.momentum_case(
            10, 8, seed=5, vbc=VB(u_type=(N, N, D, N), u_value=(0.3, -0.2, 0.5, 0.1), v_type=(D, N, N, D),
                                  v_value=(0.2, 0.4, -0.3, 0.6)))),
        ("16x16_lid_gravity", W.momentum_case(16, 16, seed=6, vbc=VB(top_wall_velocity_x=1.0), body_force_y=-0.98,
                                              re=10.0)),
        ("14x12_stiff_dt", W.momentum_case(14, 12, seed=7, dt=5e-2, re=0.05, density_ratio=10.0, viscosity_ratio=50.0)),
        ("13x9_stiff_periodic_x_neumann_top", W.momentum_case(
            13, 9, seed=8, dt=2e-2, re=0.1, periodic_x=True, vbc=VB(u_type=(U, U, D, N), u_value=(0, 0, 0.0, 0.25),
                                                                    v_type=(U, U, D, N), v_value=(0, 0, 0.0, -0.5)))),
        ("40x36_stiff_limit", W.momentum_case(40, 36, seed=9, dt=1e-1, re=0.02, density_ratio=100.0,
                                              viscosity_ratio=100.0, max_iterations=7)),
        ("3x3_smallest", W.momentum_case(3, 3, seed=10)),
        ("6x5_ghost2", W.momentum_case(6, 5, seed=11, ghost=2, dt=1e-2, re=1.0)),
        ("7x6_ghost1_periodic_xy", W.momentum_case(7, 6, seed=12, ghost=1, periodic_x=True, periodic_y=True,
                                                   dt=1e-2, re=0.5)),
    ]
    return out


def load_momentum_golden(path):
    """(MomentumCase, expected (u_star, v_star, u_prev, v_prev, iterations, absolute residual)) from a
    tests/golden/momentum_*.npz fixture written by make_golden.py from the compiled reference."""
    import numpy as np
    z = np.load(path, allow_pickle=False)
    vbc = W.VelocityBC(u_type=tuple(int(t) for t in z["u_type"]), u_value=tuple(float(t) for t in z["u_value"]),
                       v_type=tuple(int(t) for t in z["v_type"]), v_value=tuple(float(t) for t in z["v_value"]),
                       bottom_wall_velocity_x=float(z["bottom_wall_velocity_x"]),
                       top_wall_velocity_x=float(z["top_wall_velocity_x"]))
    ints = ("nx", "ny", "step_index", "max_iterations", "ghost")
    kw = {k: (int(z[k]) if k in ints else (bool(z[k]) if k.startswith("periodic") else float(z[k])))
          for k in ("nx", "ny", "lx", "ly", "dt", "re", "body_force_x", "body_force_y", "periodic_x", "periodic_y",
                    "step_index", "max_iterations", "tolerance", "ghost")}
    arrays = {k: z["in_" + k].copy() for k in ("u", "v", "rho_mid", "eta", "eta_prev", "surface_fx_u", "surface_fy_v",
                                               "u_prev", "v_prev")}
    mc = W.MomentumCase(vbc=vbc, **kw, **arrays)
    expect = (z["out_u_star"], z["out_v_star"], z["out_u_prev"], z["out_v_prev"], int(z["out_iterations"]),
              float(z["out_absolute"]))
    return mc, expect
