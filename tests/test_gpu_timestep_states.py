"""The HIP engine on pressure-solve states captured from real timesteps of the reference's own runs
(the 64x64 static droplet's 10 steps, the 192x384 rising bubble at steps 1 and 18, and the published
128x128 validation run one step past its step-5000 snapshot).

FP64 (hip_icpcg, hip_ildlt_pcg, hip_split_icpcg with the state's real pressure history): the
reference's iteration count, bit-for-bit equality with the oracle in the engine's summation order,
and a pressure within ten times the spread that an independent reordering (pairwise sums) gives the
reference's own sequential result. The oracle in sequential order equals the reference's solve on
these states bit for bit (test_timestep_oracle.py). Mixed precision on the rising-bubble states,
refined to the run's 1e-10: the true FP64 residual meets the tolerance and the error against a
1e-13 solution stays within ten times the FP64 path's; both errors are printed per state.
"""
import dataclasses

import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from tests.timestep_states import load_state, state_files, state_id

pytestmark = pytest.mark.gpu

FILES = state_files()
CODES = {"icpcg": E.PP_ICPCG, "ildlt_pcg": E.PP_ILDLT_PCG, "split_icpcg": E.PP_SPLIT_ICPCG}


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("path", FILES, ids=[state_id(p) for p in FILES])
@pytest.mark.parametrize("scheme", ["icpcg", "ildlt_pcg", "split_icpcg"])
def test_fp64_on_timestep_state(engine, oracle_lib, path, scheme):
    case, results = load_state(path)
    p_gpu, rep = engine.solve(case, CODES[scheme])
    p_can, it_can, res_can = oracle_lib.pressure_solve(scheme, case, oracle_lib.CANONICAL)
    p_seq, it_seq, _ = oracle_lib.pressure_solve(scheme, case, oracle_lib.SEQUENTIAL)
    p_pw, _, _ = oracle_lib.pressure_solve(scheme, case, oracle_lib.PAIRWISE)
    assert rep["iterations"] == it_can and rep["relative_residual"] == res_can
    assert np.array_equal(p_gpu, p_can)
    if scheme in results:
        assert it_seq == results[scheme][0]
    assert rep["iterations"] == it_seq
    dev, spread = _rel(p_gpu, p_seq), _rel(p_pw, p_seq)
    print(f"[timestep] {state_id(path)} {scheme}: {rep['iterations']} iterations (reference {it_seq}), "
          f"deviation from the reference {dev:.2e}, pairwise-order spread {spread:.2e}")
    assert dev <= 10.0 * max(spread, 1e-14)


def test_restart_state_takes_the_published_133_iterations(engine):
    path = next(p for p in FILES if "128_restart" in p)
    case, results = load_state(path)
    _, rep = engine.solve(case, E.PP_ICPCG)
    assert rep["iterations"] == 133 == results["icpcg"][0]
    assert 8.8e-11 <= rep["relative_residual"] < 1e-10


def _true_residual(oracle_lib, case, scheme, p):
    s = oracle_lib.OracleSystem(case, use_split=scheme.startswith("split"))
    r = s.rhs - s.apply_matrix(p)
    if not s.has_dirichlet:
        r = r - r.mean()
    return np.linalg.norm(r) / np.linalg.norm(s.rhs)


BUBBLE = [p for p in FILES if "rising_bubble" in p]


@pytest.mark.parametrize("path", BUBBLE, ids=[state_id(p) for p in BUBBLE])
def test_mixed_on_rising_bubble_state(engine, oracle_lib, path):
    case, _ = load_state(path)
    assert case.tolerance == 1e-10
    p_mix, rep = engine.solve(case, E.PP_ICPCG, precision=E.PP_MIXED)
    p64, rep64 = engine.solve(case, E.PP_ICPCG)
    assert rep["relative_residual"] < case.tolerance
    assert _true_residual(oracle_lib, case, "icpcg", p_mix) < 2.0 * case.tolerance
    tight = dataclasses.replace(case, tolerance=1e-13)
    p_ex, _, _ = oracle_lib.pressure_solve("icpcg", tight, oracle_lib.SEQUENTIAL)
    e_mix, e64 = _rel(p_mix, p_ex), _rel(p64, p_ex)
    print(f"[timestep mixed] {state_id(path)}: {rep['iterations']} inner iterations in {rep['outer_cycles']} "
          f"cycles vs FP64 {rep64['iterations']}; error vs 1e-13 solution: mixed {e_mix:.2e}, FP64 {e64:.2e}; "
          f"mixed vs FP64 {_rel(p_mix, p64):.2e}; kernel {rep['kernel_ms']:.1f} ms vs {rep64['kernel_ms']:.1f} ms")
    assert e_mix <= 10.0 * max(e64, 1e-13)
