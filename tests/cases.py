"""Shared pressure-solve cases for the parity tests.

Small cases cover every boundary combination the reference supports for pressure (unset, Neumann
with a value, Dirichlet with a value, periodic in x and/or y), all schemes and the split
extrapolation at step 0 and later steps. Row counts above 64 and 128 exercise the hand-off between
sweep waves of the HIP engine.
"""
import numpy as np

from ding_nsch_core_solver_amd import workloads as W

U, D, N = W.BC_UNSET, W.BC_DIRICHLET, W.BC_NEUMANN


def _with_history(case, seed):
    rng = np.random.default_rng(seed)
    case.p_n = rng.standard_normal(case.n) * 50.0
    case.p_nm1 = rng.standard_normal(case.n) * 50.0
    return case


def small_cases():
    """(name, case) pairs sized for the CPU suite (a few ms each in the oracle)."""
    out = []
    out.append(("6x5_neumann_dr10", W.droplet_case(6, 5, seed=11, density_ratio=10.0, radius=0.3)))
    out.append(("8x8_neumann_dr1000", W.droplet_case(8, 8, seed=1, density_ratio=1000.0, radius=0.3)))
    out.append(("12x9_dirichlet_left", W.droplet_case(12, 9, seed=2, density_ratio=50.0,
                                                      bc_type=(D, U, U, U), bc_value=(3.0, 0.0, 0.0, 0.0))))
    out.append(("10x14_mixed_values", W.droplet_case(10, 14, seed=3, density_ratio=7.0,
                                                     bc_type=(N, D, N, D), bc_value=(0.5, -1.0, 2.0, 0.25))))
    out.append(("9x11_neumann_values", W.droplet_case(9, 11, seed=4, density_ratio=3.0,
                                                      bc_type=(N, N, N, N), bc_value=(1.0, -2.0, 0.5, 0.75))))
    out.append(("16x12_periodic_x", W.droplet_case(16, 12, seed=5, density_ratio=20.0, periodic_x=True)))
    out.append(("12x16_periodic_y_dirichlet", W.droplet_case(12, 16, seed=6, density_ratio=20.0, periodic_y=True,
                                                             bc_type=(U, D, U, U), bc_value=(0.0, 1.5, 0.0, 0.0))))
    out.append(("16x16_periodic_xy_dr1", W.droplet_case(16, 16, seed=7, density_ratio=1.0, periodic_x=True,
                                                        periodic_y=True)))
    out.append(("4x4_periodic_xy", W.droplet_case(4, 4, seed=8, density_ratio=5.0, periodic_x=True, periodic_y=True)))
    out.append(("1x6_column", W.droplet_case(1, 6, seed=9, density_ratio=2.0)))
    out.append(("70x20_two_sweep_waves", W.droplet_case(70, 20, seed=10, density_ratio=100.0)))
    out.append(("130x12_periodic_x_three_waves", W.droplet_case(130, 12, seed=12, density_ratio=30.0, periodic_x=True)))
    out.append(("20x24_rect_dx_ne_dy", W.droplet_case(20, 24, seed=13, density_ratio=40.0, lx=1.0, ly=2.0)))
    out.append(("split_step1_14x10", _with_history(W.droplet_case(14, 10, seed=14, density_ratio=500.0,
                                                                  step_index=1), 99)))
    out.append(("split_step0_10x14_py", _with_history(W.droplet_case(10, 14, seed=15, density_ratio=500.0,
                                                                     periodic_y=True, step_index=0), 98)))
    out.append(("split_step3_70x9_px", _with_history(W.droplet_case(70, 9, seed=16, density_ratio=80.0,
                                                                    periodic_x=True, step_index=3), 97)))
    return out


SCHEMES = ("icpcg", "ildlt_pcg", "split_icpcg", "split_ildlt_pcg")


def large_cases():
    """Full-size parity cases (GPU tests)."""
    out = []
    for seed in (0, 1, 2):
        out.append((f"192x192_dr1000_seed{seed}", W.droplet_case(192, 192, seed=seed)))
    out.append(("64x64_periodic_dr1", W.droplet_case(64, 64, seed=0, density_ratio=1.0, periodic_x=True,
                                                     periodic_y=True)))
    out.append(("256x256_dr100", W.droplet_case(256, 256, seed=0, density_ratio=100.0)))
    out.append(("96x192_dr300", W.droplet_case(96, 192, seed=5, density_ratio=300.0, radius=0.15)))
    out.append(("200x72_periodic_x_dirichlet_top", W.droplet_case(200, 72, seed=6, density_ratio=50.0,
                                                                  periodic_x=True, bc_type=(U, U, U, D),
                                                                  bc_value=(0.0, 0.0, 0.0, 2.0))))
    out.append(("420x40_seven_sweep_waves", W.droplet_case(420, 40, seed=3, density_ratio=100.0)))
    split = W.droplet_case(192, 192, seed=1)
    split.step_index = 1
    split.p_n = np.zeros(split.n)
    split.p_nm1 = np.zeros(split.n)
    out.append(("192x192_split_history", split))
    return out
