"""Several pressure steps in a row on device-resident Field2D buffers, as Solver::step runs them
(solver.cpp:271-272: solve_pressure_correction, then apply_pressure_velocity_correction), with the
state carried from step to step: corrected u, v, the pressure and its history (the split schemes
read both). The HIP chain (pp_solve_device_field -> pp_velocity_correction_device) equals the
oracle's chain in the engine's reduction order bit for bit at every step; against the compiled
reference's own chain the iteration counts are equal and the fields agree to rounding."""
import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W

pytestmark = pytest.mark.gpu

G = 3


def _interior(f, rows, cols):
    return f.reshape(rows + 2 * G, cols + 2 * G)[G:-G, G:-G].reshape(-1).copy()


def _oracle_step(lib, scheme, case, vbc, state, u_star, v_star, reduction=None):
    """One step through the oracle (or the reference checker when reduction is None)."""
    nx, ny = case.nx, case.ny
    step = W.PressureCase(**{k: getattr(case, k) for k in ("nx", "ny", "lx", "ly", "dt", "periodic_x", "periodic_y",
                                                          "bc_type", "bc_value", "max_iterations", "tolerance",
                                                          "density_ratio", "step_index")})
    step.rho = _interior(state["rho_mid"], nx, ny)
    step.u = _interior(u_star, nx + 1, ny)
    step.v = _interior(v_star, nx, ny + 1)
    step.p_n = _interior(state["pressure"], nx, ny)
    step.p_nm1 = _interior(state["pressure_prev"], nx, ny)
    if reduction is None:
        p, it, res = lib.pressure_solve(scheme, step)
    else:
        p, it, res = lib.pressure_solve(scheme, step, reduction)
    pc = np.zeros(W.ghosted_shape(nx, ny, G))
    pc[G:-G, G:-G] = p.reshape(nx, ny)
    pc = lib.pressure_ghosts(step, pc.reshape(-1), G)
    if reduction is None:
        out = lib.velocity_correction(scheme, step, vbc, G, state["rho_mid"], u_star, v_star, pc, state["u"],
                                      state["v"], state["pressure"], state["pressure_prev"])
    else:
        out = lib.velocity_correction(step, vbc, scheme.startswith("split"), G, state["rho_mid"], u_star, v_star, pc,
                                      state["u"], state["v"], state["pressure"], state["pressure_prev"])
    new = dict(state)
    new["u"], new["v"], new["pressure"], new["pressure_prev"] = out[:4]
    return new, it, out[4]


@pytest.mark.parametrize("scheme,kw", [("split_icpcg", dict(periodic_x=True)),
                                       ("icpcg", dict(bc_type=(0, 0, 0, 1), bc_value=(0, 0, 0, 0.5)))])
def test_four_pressure_steps(engine, oracle_lib, ref_lib, scheme, kw):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    nx, ny = 48, 40
    base = W.droplet_case(nx, ny, seed=9, density_ratio=500.0, **kw)
    vbc = W.VelocityBC(top_wall_velocity_x=0.5)
    rng = np.random.default_rng(42)
    rho = np.pad(base.rho.reshape(nx, ny), G, mode="edge").reshape(-1)
    state = dict(rho_mid=rho, u=np.zeros(W.ghosted_shape(nx + 1, ny, G)).reshape(-1),
                 v=np.zeros(W.ghosted_shape(nx, ny + 1, G)).reshape(-1),
                 pressure=np.zeros(W.ghosted_shape(nx, ny, G)).reshape(-1),
                 pressure_prev=np.zeros(W.ghosted_shape(nx, ny, G)).reshape(-1))
    oracle_state, ref_state = dict(state), dict(state)
    d = {k: torch.from_numpy(a.copy()).to(dev) for k, a in state.items()}
    d_pc = torch.zeros(W.ghosted_shape(nx, ny, G)[0] * W.ghosted_shape(nx, ny, G)[1], dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    code = {"icpcg": E.PP_ICPCG, "split_icpcg": E.PP_SPLIT_ICPCG}[scheme]
    for k in range(4):
        case = W.droplet_case(nx, ny, seed=9, density_ratio=500.0, step_index=k, **kw)
        # a synthetic predictor: the current velocity plus a seeded forcing increment
        du = rng.standard_normal(W.ghosted_shape(nx + 1, ny, G)).reshape(-1)
        dv = rng.standard_normal(W.ghosted_shape(nx, ny + 1, G)).reshape(-1)
        u_star = d["u"].cpu().numpy() + du
        v_star = d["v"].cpu().numpy() + dv
        d_us, d_vs = torch.from_numpy(u_star).to(dev), torch.from_numpy(v_star).to(dev)
        rep = engine.solve_device_field(case, d["rho_mid"].data_ptr(), d_us.data_ptr(), d_vs.data_ptr(),
                                        d["pressure"].data_ptr(), d["pressure_prev"].data_ptr(), d_pc.data_ptr(),
                                        scheme=code, ghost=G, stream=stream)
        speeds = engine.velocity_correction_device(case, vbc, d["rho_mid"].data_ptr(), d_us.data_ptr(),
                                                   d_vs.data_ptr(), d_pc.data_ptr(), d["u"].data_ptr(),
                                                   d["v"].data_ptr(), d["pressure"].data_ptr(),
                                                   d["pressure_prev"].data_ptr(), scheme=code, ghost=G, stream=stream)
        got = {key: d[key].cpu().numpy() for key in ("u", "v", "pressure", "pressure_prev")}
        # the oracle's chain from the oracle's own state: bitwise at every step
        ou = oracle_state["u"] + du
        ov = oracle_state["v"] + dv
        oracle_state, it_o, sp_o = _oracle_step(oracle_lib, scheme, case, vbc, oracle_state, ou, ov,
                                                oracle_lib.CANONICAL)
        assert rep["iterations"] == it_o, k
        for key in got:
            assert np.array_equal(got[key], oracle_state[key]), (k, key)
        assert tuple(speeds) == tuple(sp_o)
        # the reference's chain from its own state: same iterations, fields within rounding
        ru = ref_state["u"] + du
        rv = ref_state["v"] + dv
        ref_state, it_r, _ = _oracle_step(ref_lib, scheme, case, vbc, ref_state, ru, rv)
        assert rep["iterations"] == it_r, k
        for key in got:
            scale = max(np.abs(ref_state[key]).max(), 1e-30)
            assert np.abs(got[key] - ref_state[key]).max() <= 1e-7 * scale, (k, key)
        print(f"[timestep] {scheme} step {k}: {rep['iterations']} iterations, bitwise vs oracle", flush=True)


@pytest.mark.parametrize("scheme", ["icpcg", "split_ildlt_pcg"])
def test_pressure_step_entry_equals_separate_calls(engine, oracle_lib, scheme):
    """pp_pressure_step (host buffers, one upload and download) against the oracle's chain on the
    same inputs, bit for bit: pressure_correction with ghosts, u, v, pressure, pressure_prev, speeds."""
    nx, ny = 40, 36
    case = W.droplet_case(nx, ny, seed=12, density_ratio=200.0, step_index=1, periodic_x=True)
    rng = np.random.default_rng(8)
    state = dict(rho_mid=np.pad(case.rho.reshape(nx, ny), G, mode="edge").reshape(-1),
                 u=rng.standard_normal(W.ghosted_shape(nx + 1, ny, G)).reshape(-1),
                 v=rng.standard_normal(W.ghosted_shape(nx, ny + 1, G)).reshape(-1),
                 pressure=rng.standard_normal(W.ghosted_shape(nx, ny, G)).reshape(-1) * 10.0,
                 pressure_prev=rng.standard_normal(W.ghosted_shape(nx, ny, G)).reshape(-1) * 10.0)
    u_star = state["u"] + rng.standard_normal(state["u"].shape)
    v_star = state["v"] + rng.standard_normal(state["v"].shape)
    vbc = W.VelocityBC(bottom_wall_velocity_x=-0.25)
    code = {"icpcg": E.PP_ICPCG, "split_ildlt_pcg": E.PP_SPLIT_ILDLT_PCG}[scheme]
    p_corr, u, v, p, prev, speeds, rep = engine.pressure_step(case, vbc, state["rho_mid"], u_star, v_star, state["u"],
                                                              state["v"], state["pressure"], state["pressure_prev"],
                                                              scheme=code, ghost=G)
    expect, it_o, sp_o = _oracle_step(oracle_lib, scheme, case, vbc, state, u_star, v_star, oracle_lib.CANONICAL)
    assert rep["iterations"] == it_o and speeds == tuple(sp_o)
    for key, a in (("u", u), ("v", v), ("pressure", p), ("pressure_prev", prev)):
        assert np.array_equal(a, expect[key]), key
    step = W.PressureCase(**{k: getattr(case, k) for k in ("nx", "ny", "lx", "ly", "dt", "periodic_x", "periodic_y",
                                                          "bc_type", "bc_value", "density_ratio", "step_index")})
    pc = np.zeros(W.ghosted_shape(nx, ny, G))
    step.rho = _interior(state["rho_mid"], nx, ny)
    step.u, step.v = _interior(u_star, nx + 1, ny), _interior(v_star, nx, ny + 1)
    step.p_n, step.p_nm1 = _interior(state["pressure"], nx, ny), _interior(state["pressure_prev"], nx, ny)
    pc[G:-G, G:-G] = oracle_lib.pressure_solve(scheme, step, oracle_lib.CANONICAL)[0].reshape(nx, ny)
    assert np.array_equal(p_corr, oracle_lib.pressure_ghosts(step, pc.reshape(-1), G))


def test_pressure_step_failure_leaves_fields(engine):
    nx, ny = 8, 8
    case = W.droplet_case(nx, ny, seed=0, density_ratio=10.0)
    rho = case.rho.reshape(nx, ny).copy()
    rho[:3, :3] = -5.0
    rng = np.random.default_rng(1)
    u = rng.standard_normal(W.ghosted_shape(nx + 1, ny, G)).reshape(-1)
    v = rng.standard_normal(W.ghosted_shape(nx, ny + 1, G)).reshape(-1)
    p = rng.standard_normal(W.ghosted_shape(nx, ny, G)).reshape(-1)
    with pytest.raises(E.EngineError, match="ICC preconditioner construction failed"):
        engine.pressure_step(case, W.VelocityBC(), np.pad(rho, G, mode="edge").reshape(-1), u, v, u, v, p, p)
