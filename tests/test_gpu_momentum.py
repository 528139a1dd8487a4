"""The momentum predictor on the device (pp_momentum_predictor[_device]) against the oracle
(or_momentum_predictor, equal to the reference's compute_momentum_fluxes + solve_momentum_predictor
bit for bit, test_momentum_oracle.py) with the BiCGSTAB dot products in the engine's chunked order:
u_star, v_star and the stored convection fields (ghost layers included) bit for bit, the same
iteration count and absolute residual. Against the reference's sequential order: the same iteration
count and the predictor within 1e-9 relative (the BiCGSTAB stops at an absolute residual of 1e-8; the
dot-product order moves the iterates by up to 1e-10 on the stiff cases). Also on states taken from the reference's own full run, on
a 192x192 droplet through the device entry twice on one context, and the error texts."""
import glob
import os

import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W
from tests.momentum_cases import load_momentum_golden, small_momentum_cases

pytestmark = pytest.mark.gpu

CASES = small_momentum_cases()


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _check(engine, oracle_lib, mc):
    got = engine.momentum_predictor(mc)
    want = oracle_lib.momentum_predictor(mc, oracle_lib.CHUNKED)
    for x, y, what in zip(got[:4], want[:4], ("u_star", "v_star", "u_rhs_prev", "v_rhs_prev")):
        assert np.array_equal(x, y), f"{what}: max diff {np.abs(x - y).max():.3e}"
    rep = got[4]
    assert rep["iterations"] == want[4] and rep["absolute_residual"] == want[5]
    return got


@pytest.mark.parametrize("name,mc", CASES, ids=[c[0] for c in CASES])
def test_momentum_predictor_bitwise_vs_oracle(engine, oracle_lib, name, mc):
    got = _check(engine, oracle_lib, mc)
    seq = oracle_lib.momentum_predictor(mc, oracle_lib.SEQUENTIAL)
    assert got[4]["iterations"] == seq[4]
    assert np.array_equal(got[2], seq[2]) and np.array_equal(got[3], seq[3])  # no reduction in the convection
    if seq[5] < mc.tolerance:  # converged in both orders: the predictors agree within the solver tolerance's reach
        assert _rel(got[0], seq[0]) < 1e-9 and _rel(got[1], seq[1]) < 1e-9


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "momentum_*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_momentum_golden(engine, oracle_lib, path):
    mc, expect = load_momentum_golden(path)
    got = _check(engine, oracle_lib, mc)
    assert got[4]["iterations"] == expect[4]
    assert np.array_equal(got[2], expect[2]) and np.array_equal(got[3], expect[3])
    assert _rel(got[0], expect[0]) < 1e-9 and _rel(got[1], expect[1]) < 1e-9


def test_device_entry_192x192_twice(engine, oracle_lib):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    for seed in (0, 1):
        mc = W.momentum_case(192, 192, seed=seed, dt=1e-3, re=10.0)
        t = {k: torch.from_numpy(np.ascontiguousarray(getattr(mc, k))).to(dev) for k in E.Engine.MOMENTUM_INPUTS}
        prev = {k: torch.from_numpy(np.array(getattr(mc, k), copy=True)).to(dev) for k in ("u_prev", "v_prev")}
        star = {k: torch.zeros_like(prev[p]) for k, p in (("u_star", "u_prev"), ("v_star", "v_prev"))}
        rep = engine.momentum_predictor_device(mc, *[t[k].data_ptr() for k in E.Engine.MOMENTUM_INPUTS],
                                               prev["u_prev"].data_ptr(), prev["v_prev"].data_ptr(),
                                               star["u_star"].data_ptr(), star["v_star"].data_ptr(), stream=stream)
        want = oracle_lib.momentum_predictor(mc, oracle_lib.CHUNKED)
        assert rep["iterations"] == want[4] and rep["absolute_residual"] == want[5] and rep["kernel_ms"] > 0
        for got, exp in zip((star["u_star"], star["v_star"], prev["u_prev"], prev["v_prev"]), want[:4]):
            assert np.array_equal(got.cpu().numpy(), exp)


def test_states_of_the_reference_full_run(engine, ref_lib):
    """The 64x64 static droplet and the 192x384 rising bubble: at each step the reference's own
    phase stage, then the device predictor on the Solver's fields against the reference's predictor."""
    for spec, steps in ((W.STATIC_DROPLET_64, 4), (W.RISING_BUBBLE_192X384, 2)):
        run = ref_lib.RefRun(spec)
        for _ in range(steps):
            run.phase()
            mc = run.momentum_case()
            u_star, v_star, u_prev, v_prev, rep = engine.momentum_predictor(mc)
            run.momentum()
            assert np.array_equal(u_prev, run.field("momentum_u_rhs_prev").reshape(-1))
            assert np.array_equal(v_prev, run.field("momentum_v_rhs_prev").reshape(-1))
            assert _rel(u_star, run.field("u_star").reshape(-1)) < 1e-9
            assert _rel(v_star, run.field("v_star").reshape(-1)) < 1e-9
            run.pressure()
            info = run.finish()
            assert rep["iterations"] == info["momentum_iterations"]
        run.close()


def test_error_texts(engine):
    mc = W.momentum_case(8, 8, seed=0, periodic_x=True, vbc=W.VelocityBC(v_type=(0, 2, 0, 0)))
    with pytest.raises(E.EngineError, match="v_bc_right cannot be set when the direction is periodic"):
        engine.momentum_predictor(mc)
    mc = W.momentum_case(2, 8, seed=0)
    with pytest.raises(E.EngineError, match="nx, ny >= 3"):
        engine.momentum_predictor(mc)
