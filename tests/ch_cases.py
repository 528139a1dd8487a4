"""Cahn-Hilliard linear-solve cases: walls and every periodic combination, first-step and BDF2
alpha0, and stabilization / time-step settings where ILDLT(0) builds (a few to the full 30 PCG
iterations) and where it fails (the diagonal preconditioner)."""
from ding_nsch_core_solver_amd import workloads as W


def small_ch_cases():
    return [
        ("16x12_walls", W.ch_case(16, 12, seed=1)),
        ("16x12_periodic_x", W.ch_case(16, 12, seed=2, periodic_x=True)),
        ("12x16_periodic_y", W.ch_case(12, 16, seed=3, periodic_y=True)),
        ("10x10_periodic_xy", W.ch_case(10, 10, seed=4, periodic_x=True, periodic_y=True)),
        ("20x14_first_step", W.ch_case(20, 14, seed=5, first_step=True, dt=1e-3)),
        ("24x20_many_iterations", W.ch_case(24, 20, seed=6, dt=1.0, pe=1.0)),
        ("24x20_iteration_limit", W.ch_case(24, 20, seed=7, dt=1e2, pe=0.1)),
        ("16x16_ildlt_fails_diagonal", W.ch_case(16, 16, seed=8, cn=0.05, pe=1.0, dt=1e-2)),
        ("18x14_ildlt_fails_periodic_x", W.ch_case(18, 14, seed=9, cn=0.3, pe=100.0, dt=1e3, periodic_x=True)),
        ("33x70_rectangular", W.ch_case(33, 70, seed=10, lx=1.0, ly=2.0, dt=1e-3, pe=10.0)),
        # periodic y chains the level schedule, so these take the batched factorization and sweeps
        ("37x29_periodic_y_odd_batches", W.ch_case(37, 29, seed=12, periodic_y=True)),
        ("24x20_periodic_xy_many_iterations", W.ch_case(24, 20, seed=16, dt=1.0, pe=1.0, periodic_x=True,
                                                        periodic_y=True)),
        ("16x16_ildlt_fails_periodic_y", W.ch_case(16, 16, seed=8, cn=0.05, pe=1.0, dt=1e-2, periodic_y=True)),
        ("96x72_periodic_y_largest_batched", W.ch_case(96, 72, seed=15, periodic_y=True)),
    ]
