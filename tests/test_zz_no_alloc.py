"""The allocation-free solve entry (pp_solve_device_noalloc) against pp_solve_device, through the C
ABI: bit-identical pressure, iteration count and residuals on two calls in a row, the second on a
smaller grid (the shared scratch array keeps the first call's contents), in FP64 and mixed
precision, plus the size limit."""
import ctypes

import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W

pytestmark = pytest.mark.gpu


def _device_inputs(torch, case):
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(getattr(case, k)).reshape(-1)).to(dev) for k in ("rho", "u", "v")}
    for k in ("p_n", "p_nm1"):
        a = getattr(case, k)
        t[k] = None if a is None else torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).to(dev)
    return t


def _noalloc(lib, case, t, out, scheme, precision, stream):
    pb = E.make_problem(case, scheme, precision)
    rep = E.PPReport()
    ptr = lambda x: None if x is None else x.data_ptr()  # noqa: E731
    rc = lib.pp_solve_device_noalloc(ctypes.byref(pb), ptr(t["rho"]), ptr(t["u"]), ptr(t["v"]), ptr(t["p_n"]),
                                     ptr(t["p_nm1"]), out.data_ptr(), ctypes.byref(rep), stream)
    return rc, rep


@pytest.mark.parametrize("precision", [E.PP_FP64, E.PP_MIXED])
@pytest.mark.parametrize("scheme", [E.PP_ICPCG, E.PP_ILDLT_PCG, E.PP_SPLIT_ICPCG])
def test_noalloc_matches_device_entry_two_calls(engine, scheme, precision):
    torch = pytest.importorskip("torch")
    lib = E.lib()
    stream = torch.cuda.current_stream(torch.device("cuda:0")).cuda_stream
    big = W.droplet_case(192, 192, seed=4, density_ratio=1000.0)
    small = W.droplet_case(70, 52, seed=5, density_ratio=30.0, periodic_x=True)
    if scheme == E.PP_SPLIT_ICPCG:
        rng = np.random.default_rng(1)
        for c in (big, small):
            c.step_index = 2
            c.p_n = rng.standard_normal(c.n) * 10.0
            c.p_nm1 = rng.standard_normal(c.n) * 10.0
    for case in (big, small):
        t = _device_inputs(torch, case)
        out_ref = torch.full((case.n,), -1.0, dtype=torch.float64, device="cuda:0")
        out = torch.full((case.n,), -2.0, dtype=torch.float64, device="cuda:0")
        ptr = lambda x: None if x is None else x.data_ptr()  # noqa: E731
        rep_ref = engine.solve_device(case, ptr(t["rho"]), ptr(t["u"]), ptr(t["v"]), ptr(t["p_n"]), ptr(t["p_nm1"]),
                                      out_ref.data_ptr(), scheme=scheme, precision=precision, stream=stream)
        rc, rep = _noalloc(lib, case, t, out, scheme, precision, stream)
        assert rc == 0, rep.message
        assert rep.iterations == rep_ref["iterations"] and rep.outer_cycles == rep_ref["outer_cycles"]
        assert rep.relative_residual == rep_ref["relative_residual"]
        assert rep.absolute_residual == rep_ref["absolute_residual"]
        assert torch.equal(out, out_ref)


def test_noalloc_rejects_grids_above_its_limit(engine):
    torch = pytest.importorskip("torch")
    lib = E.lib()
    case = W.droplet_case(400, 400, seed=0, density_ratio=10.0)
    t = _device_inputs(torch, case)
    out = torch.zeros(case.n, dtype=torch.float64, device="cuda:0")
    rc, rep = _noalloc(lib, case, t, out, E.PP_ICPCG, E.PP_FP64, None)
    assert rc == E.PP_ERR_ARG and b"allocation-free solve supports" in rep.message
    # the failing call touched nothing: a valid call right after it is still exact
    case = W.droplet_case(40, 36, seed=1, density_ratio=10.0)
    t = _device_inputs(torch, case)
    out = torch.zeros(case.n, dtype=torch.float64, device="cuda:0")
    rc, rep = _noalloc(lib, case, t, out, E.PP_ICPCG, E.PP_FP64, None)
    p_ref, rep_ref = engine.solve(case, E.PP_ICPCG)
    assert rc == 0 and rep.iterations == rep_ref["iterations"]
    assert np.array_equal(out.cpu().numpy(), p_ref)
