"""The Cahn-Hilliard linear solve on the device (pp_ch_solve[_device]) against the oracle, which
equals the reference's solve_phase_linear_system_eq25 bit for bit (test_ch_oracle.py), run with its
reductions in the engine's order: the solution, the iteration count, the iterate / absolute /
relative residuals and the preconditioner kind bit for bit. Against the reference's sequential order:
the same iteration count and the solution within rounding. Also the operator cache across a run's
alpha0 values, the 192x384 rising-bubble grid, the device entry and the breakdown texts."""
import dataclasses

import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W
from tests.ch_cases import small_ch_cases

pytestmark = pytest.mark.gpu

CASES = small_ch_cases()


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _check(engine, oracle_lib, cc, precond=-1, operator_alpha0=None):
    """operator_alpha0: the alpha0 the cached operator was built with, when the context reuses it."""
    c, rep = engine.ch_solve(cc)
    occ = cc if operator_alpha0 is None else dataclasses.replace(cc, alpha0=operator_alpha0)
    want = oracle_lib.ch_solve(occ, cc.rhs, cc.c0, oracle_lib.CH_ENGINE, precond)
    assert np.array_equal(c, want[0]), f"max diff {np.abs(c - want[0]).max():.3e}"
    assert rep["iterations"] == want[1]
    assert rep["absolute_residual"] == want[2] and rep["relative_residual"] == want[3]
    assert rep["iterate_residual"] == want[4]
    assert rep["preconditioner"] == want[5]
    return c, rep


@pytest.mark.parametrize("name,cc", CASES, ids=[c[0] for c in CASES])
def test_ch_solve_bitwise_vs_oracle(engine, oracle_lib, name, cc):
    c, rep = _check(engine, oracle_lib, cc)
    seq = oracle_lib.ch_solve(cc, cc.rhs, cc.c0, oracle_lib.SEQUENTIAL)
    assert rep["iterations"] == seq[1]
    converged = rep["relative_residual"] < cc.tolerance or rep["absolute_residual"] < cc.tolerance
    assert _rel(c, seq[0]) < (1e-10 if converged else 1e-8)  # an unconverged stop keeps reordering noise
    print(f"[ch] {name}: {rep['iterations']} iterations, {rep['solver_name']}, {rep['levels']} levels, "
          f"{rep['kernel_ms']:.3f} ms, deviation from the reference order {_rel(c, seq[0]):.1e}")


@pytest.mark.parametrize("periodic_y", [False, True], ids=["walls", "periodic_y_batched_factor"])
def test_operator_cache_follows_alpha0(engine, oracle_lib, periodic_y):
    """A run's sequence: 1/dt at step 0, then 3/(2 dt) (a rebuild), then the cached operator again.
    With periodic y the factor is built by batches of rows (a chained level schedule)."""
    base = W.ch_case(40, 30, seed=11, first_step=True, dt=1e-3, pe=10.0, periodic_y=periodic_y)
    cached = None
    for k, alpha0 in enumerate((1e3, 1.5e3, 1.5e3, 1.5e3 * (1 + 1e-15))):
        # ensure_ch_linear_system keeps the operator while alpha0 stays within 1e-14 max(1, |alpha0|)
        if cached is None or abs(alpha0 - cached) > 1e-14 * max(1.0, abs(alpha0)):
            cached = alpha0
        cc = dataclasses.replace(base, alpha0=alpha0, rhs=base.rhs * (1.0 + 0.1 * k))
        _check(engine, oracle_lib, cc, operator_alpha0=cached)


@pytest.mark.parametrize("periodic_y", [False, True], ids=["walls", "periodic_y_batched_factor"])
def test_ildlt_failure_then_cache_keeps_diagonal(engine, oracle_lib, periodic_y):
    cc = W.ch_case(16, 16, seed=8, cn=0.05, pe=1.0, dt=1e-2, periodic_y=periodic_y)
    _, rep = _check(engine, oracle_lib, cc)
    assert rep["preconditioner"] == E.PP_CH_DIAGONAL
    _, rep = _check(engine, oracle_lib, cc, precond=0)  # the cached operator keeps the diagonal one
    assert rep["preconditioner"] == E.PP_CH_DIAGONAL


def test_rising_bubble_grid(engine, oracle_lib):
    cc = W.ch_case(192, 384, seed=12, radius=0.25, ly=2.0, dt=5e-5, cn=0.01, pe=100.0)
    c, rep = _check(engine, oracle_lib, cc)
    print(f"[ch] 192x384: {rep['iterations']} iterations, {rep['levels']} levels, {rep['kernel_ms']:.2f} ms")


def test_device_entry(engine, oracle_lib):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    cc = W.ch_case(64, 64, seed=13, periodic_x=True, periodic_y=True, dt=1e-4)
    d_rhs = torch.from_numpy(cc.rhs.copy()).to(dev)
    d_c = torch.from_numpy(cc.c0.copy()).to(dev)
    rep = engine.ch_solve_device(cc, d_rhs.data_ptr(), d_c.data_ptr(),
                                 stream=torch.cuda.current_stream(dev).cuda_stream)
    want = oracle_lib.ch_solve(cc, cc.rhs, cc.c0, oracle_lib.CH_ENGINE)
    assert rep["iterations"] == want[1] and np.array_equal(d_c.cpu().numpy(), want[0])


def test_breakdown_falls_back_then_reports_the_reference_text(engine, oracle_lib):
    """A NaN right-hand side breaks the ILDLT PCG; the diagonal retry breaks too and its error is the
    reference's text (the second solve is outside the reference's catch)."""
    cc = W.ch_case(12, 12, seed=14)
    cc.rhs = cc.rhs.copy()
    cc.rhs[5] = np.nan
    with pytest.raises(RuntimeError) as e_or:
        oracle_lib.ch_solve(cc, cc.rhs, cc.c0, oracle_lib.CH_ENGINE)
    with pytest.raises(E.EngineError) as e_gpu:
        engine.ch_solve(cc)
    assert str(e_gpu.value) == str(e_or.value)
