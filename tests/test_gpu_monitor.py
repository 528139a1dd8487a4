"""HIP PCG monitor log (pp_solve_monitored / pp_solve_device_monitored) against the oracle's log in
the engine's reduction order, bit for bit; the oracle's log is pinned to the reference's CGMonitor
in tests/test_monitor_oracle.py. Against the reference directly: same number of records and the
early iterations within rounding (later ones drift with the reduction order, as the iterates do)."""
import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W
from tests.cases import SCHEMES, small_cases

pytestmark = pytest.mark.gpu

CODES = {"icpcg": E.PP_ICPCG, "ildlt_pcg": E.PP_ILDLT_PCG, "split_icpcg": E.PP_SPLIT_ICPCG,
         "split_ildlt_pcg": E.PP_SPLIT_ILDLT_PCG}
CASES = small_cases()


def _ids(cases):
    return [name for name, _ in cases]


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("name,case", CASES, ids=_ids(CASES))
def test_records_match_oracle_bitwise(engine, oracle_lib, name, case, scheme):
    try:
        p_o, it_o, res_o, rec_o = oracle_lib.pressure_solve_monitored(scheme, case, oracle_lib.CANONICAL)
    except RuntimeError as e:
        with pytest.raises(E.EngineError):
            engine.solve_monitored(case, CODES[scheme])
        pytest.skip(f"oracle raises: {e}")
    p, rep, rec = engine.solve_monitored(case, CODES[scheme])
    assert rep["iterations"] == it_o and rep["monitor_calls"] == len(rec_o)
    assert rec.tobytes() == rec_o.tobytes()
    assert np.array_equal(p, p_o)


@pytest.mark.parametrize("scheme", ["icpcg", "ildlt_pcg", "split_icpcg"])
def test_192x192_dr1000_log(engine, oracle_lib, ref_lib, scheme):
    case = W.droplet_case(192, 192, seed=0)
    if scheme.startswith("split"):
        case.step_index = 2
        case.p_n = np.random.default_rng(1).standard_normal(case.n) * 20.0
        case.p_nm1 = np.random.default_rng(2).standard_normal(case.n) * 20.0
    p, rep, rec = engine.solve_monitored(case, CODES[scheme])
    p_o, it_o, res_o, rec_o = oracle_lib.pressure_solve_monitored(scheme, case, oracle_lib.CANONICAL)
    print(f"[monitor] {scheme} 192x192: {rep['iterations']} iterations, {len(rec)} records", flush=True)
    assert rec.tobytes() == rec_o.tobytes()
    # monitoring changes nothing in the solve
    p_plain, rep_plain = engine.solve(case, CODES[scheme])
    assert np.array_equal(p, p_plain) and rep_plain["iterations"] == rep["iterations"]
    # against the reference's own PCG on its own system: same count, early records within rounding
    r = ref_lib.RefSystem(case, use_split=scheme.startswith("split"))
    assert r.factor(2 if "ildlt" in scheme else 1)[0]
    x, it_r, rel_r, ab_r, rec_r, err = r.pcg_monitored(case.max_iterations, case.tolerance)
    assert err is None and it_r == rep["iterations"] and len(rec_r) == len(rec)
    for f in ("alpha", "beta", "absolute_residual", "relative_residual", "preconditioned_residual"):
        a, b = rec[f][:20], rec_r[f][:20]
        assert np.allclose(a, b, rtol=1e-10, atol=0), f


def test_capacity_and_device_entry(engine, oracle_lib):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    case = W.droplet_case(64, 48, seed=4, density_ratio=200.0, periodic_x=True)
    p_full, rep_full, rec_full = engine.solve_monitored(case, E.PP_ICPCG)
    p_cut, rep_cut, rec_cut = engine.solve_monitored(case, E.PP_ICPCG, capacity=5)
    assert len(rec_cut) == 5 and rep_cut["monitor_calls"] == len(rec_full)
    assert rec_cut.tobytes() == rec_full[:5].tobytes() and np.array_equal(p_cut, p_full)
    t = {k: torch.from_numpy(np.ascontiguousarray(getattr(case, k)).reshape(-1)).to(dev) for k in ("rho", "u", "v")}
    out = torch.zeros(case.n, dtype=torch.float64, device=dev)
    cap = len(rec_full) + 3
    d_rec = torch.zeros(cap * E.MONITOR_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    rep = engine.solve_device_monitored(case, t["rho"].data_ptr(), t["u"].data_ptr(), t["v"].data_ptr(), None, None,
                                        out.data_ptr(), d_rec.data_ptr(), cap, scheme=E.PP_ICPCG,
                                        stream=torch.cuda.current_stream(dev).cuda_stream)
    got = np.frombuffer(d_rec.cpu().numpy().tobytes(), dtype=E.MONITOR_DTYPE)[:rep["monitor_calls"]]
    assert rep["monitor_calls"] == len(rec_full)
    assert got.tobytes() == rec_full.tobytes()
    assert np.array_equal(out.cpu().numpy(), p_full)


def test_mixed_precision_rejected(engine):
    case = W.droplet_case(16, 16, seed=0)
    with pytest.raises(E.EngineError, match="FP64"):
        engine.solve_monitored(case, E.PP_ICPCG, precision=E.PP_MIXED)
