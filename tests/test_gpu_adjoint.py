"""HIP adjoint (pp_adjoint / pp_solve_coeffs through the C ABI) against the oracle's FP64 adjoint (bit
for bit, up to the 256x256 density-ratio-100 adjoint workload itself), against central finite
differences at that workload, and through torch.autograd."""
import numpy as np
import pytest

from ding_nsch_core_solver_amd import engine as E
from ding_nsch_core_solver_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    "neumann_24x20": dict(nx=24, ny=20),
    "dirichlet_mixed_30x17": dict(nx=30, ny=17, bc_type=(1, 2, 0, 1), bc_value=(1.5, 0.4, 0.0, -0.7)),
    "neumann_values_16x33": dict(nx=16, ny=33, bc_type=(2, 2, 2, 2), bc_value=(0.3, -0.2, 0.5, 0.1)),
    "periodic_x_40x12": dict(nx=40, ny=12, periodic_x=True),
    "periodic_xy_20x20": dict(nx=20, ny=20, periodic_x=True, periodic_y=True),
    "droplet_96x80": dict(nx=96, ny=80),
    "droplet_130x70_periodic_y": dict(nx=130, ny=70, periodic_y=True),
    # the adjoint workload itself (256x256, density ratio 100): ~400 PCG iterations each way
    "droplet_256x256_adjoint_workload": dict(nx=256, ny=256),
}


def _case(nx, ny, seed=0, dr=100.0, tol=1e-10, **kw):
    return W.droplet_case(nx, ny, seed=seed, density_ratio=dr, tolerance=tol, max_iterations=20000, **kw)


@pytest.mark.parametrize("scheme", [E.PP_ICPCG, E.PP_ILDLT_PCG])
@pytest.mark.parametrize("name", sorted(CASES))
def test_coefficient_solve_and_adjoint_bitwise(engine, oracle_lib, name, scheme):
    case = _case(**CASES[name])
    ku, kv = W.face_inverse_density(case)
    src = W.divergence_source(case)
    sname = "icpcg" if scheme == E.PP_ICPCG else "ildlt_pcg"
    p, rep = engine.solve_coeffs(case, ku, kv, src, scheme=scheme)
    p_o, it_o, res_o = oracle_lib.solve_coeffs(sname, case, ku, kv, src, oracle_lib.CANONICAL)
    assert rep["iterations"] == it_o and rep["relative_residual"] == res_o
    assert np.array_equal(p, p_o)
    g = p - np.random.default_rng(5).standard_normal(case.n) * np.abs(p).max()
    lam, dku, dkv, arep = engine.adjoint(case, ku, kv, p, g, scheme=scheme)
    lam_o, dku_o, dkv_o, ait_o, _ = oracle_lib.adjoint(sname, case, ku, kv, p, g, oracle_lib.CANONICAL)
    assert arep["iterations"] == ait_o
    assert np.array_equal(lam, lam_o)
    assert np.array_equal(dku, dku_o)
    assert np.array_equal(dkv, dkv_o)


def test_split_scheme_rejected(engine):
    case = _case(8, 8)
    ku, kv = W.face_inverse_density(case)
    with pytest.raises(E.EngineError, match="PP_ICPCG or PP_ILDLT_PCG"):
        engine.solve_coeffs(case, ku, kv, W.divergence_source(case), scheme=E.PP_SPLIT_ICPCG)


def test_finite_differences_at_64_faces_256x256(engine):
    """L = 1/2 |p - p_target|^2 on 256x256, density ratio 100, R = 0.2: RHS from seed 0, target the
    pressure of seed 1; central differences with relative step 1e-6 at 64 faces drawn with seed 2;
    every solve to relative residual 1e-12."""
    case = _case(256, 256, seed=0, dr=100.0, tol=1e-12)
    target_case = _case(256, 256, seed=1, dr=100.0, tol=1e-12)
    ku, kv = W.face_inverse_density(case)
    src = W.divergence_source(case)
    p_target, _ = engine.solve_coeffs(target_case, ku, kv, W.divergence_source(target_case))
    p, _ = engine.solve_coeffs(case, ku, kv, src)
    _, dku, dkv, _ = engine.adjoint(case, ku, kv, p, p - p_target)
    grad = np.concatenate([dku, dkv])
    k = np.concatenate([ku.reshape(-1), kv.reshape(-1)])
    nu = ku.size
    rng = np.random.default_rng(2)
    faces = rng.choice(k.size, 64, replace=False)

    def loss(kk):
        pp, _ = engine.solve_coeffs(case, kk[:nu], kk[nu:], src)
        return 0.5 * float(np.sum((pp - p_target) ** 2))

    scale = np.abs(grad).max()
    worst = 0.0
    for f in faces:
        h = 1e-6 * abs(k[f])
        kp, km = k.copy(), k.copy()
        kp[f] += h
        km[f] -= h
        fd = (loss(kp) - loss(km)) / (2 * h)
        err = abs(fd - grad[f])
        worst = max(worst, err / scale)
        assert err <= 1e-6 * scale + 1e-4 * abs(grad[f]), (int(f), fd, grad[f])
    print(f"[fd] 64 faces, max |fd - adjoint| / max|grad| = {worst:.3e}")


def test_autograd_function_matches_engine_adjoint(engine):
    torch = pytest.importorskip("torch")
    from ding_nsch_core_solver_amd import autograd as A
    case = _case(48, 40, periodic_x=True)
    ku, kv = W.face_inverse_density(case)
    src = W.divergence_source(case)
    dev = torch.device("cuda:0")
    tku = torch.tensor(ku.reshape(-1), device=dev, requires_grad=True)
    tkv = torch.tensor(kv.reshape(-1), device=dev, requires_grad=True)
    tsrc = torch.tensor(src.reshape(-1), device=dev, requires_grad=True)
    target = torch.zeros(case.n, dtype=torch.float64, device=dev)
    p = A.pressure_solve(tku, tkv, tsrc, case)
    loss = 0.5 * ((p - target) ** 2).sum()
    loss.backward()
    p_host, _ = engine.solve_coeffs(case, ku, kv, src)
    assert np.array_equal(p.detach().cpu().numpy(), p_host)
    lam, dku, dkv, _ = engine.adjoint(case, ku, kv, p_host, p_host)
    assert np.array_equal(tku.grad.cpu().numpy(), dku)
    assert np.array_equal(tkv.grad.cpu().numpy(), dkv)
    assert np.array_equal(tsrc.grad.cpu().numpy(), lam)


def test_autograd_gradcheck_small_grid():
    torch = pytest.importorskip("torch")
    from ding_nsch_core_solver_amd import autograd as A
    case = _case(6, 5, dr=10.0, tol=1e-14, bc_type=(1, 0, 0, 2), bc_value=(0.5, 0.0, 0.0, 0.2))
    ku, kv = W.face_inverse_density(case)
    src = W.divergence_source(case) * 1e-4
    dev = torch.device("cuda:0")
    inputs = tuple(torch.tensor(a.reshape(-1), device=dev, requires_grad=True) for a in (ku, kv, src))
    assert torch.autograd.gradcheck(lambda a, b, c: A.pressure_solve(a, b, c, case), inputs, eps=1e-6, atol=1e-5,
                                    rtol=1e-4)


def test_autograd_chain_to_density():
    """Gradients reach rho through the differentiable face average."""
    torch = pytest.importorskip("torch")
    from ding_nsch_core_solver_amd import autograd as A
    case = _case(20, 16, dr=10.0, tol=1e-13)
    dev = torch.device("cuda:0")
    rho = torch.tensor(case.rho.reshape(-1), device=dev, requires_grad=True)
    src = torch.tensor(W.divergence_source(case).reshape(-1) * 1e-4, device=dev)

    def loss_of(r):
        ku, kv = A.face_inverse_density(r, case)
        p = A.pressure_solve(ku, kv, src, case)
        return 0.5 * (p ** 2).sum()

    loss_of(rho).backward()
    g = rho.grad.cpu().numpy()
    rng = np.random.default_rng(3)
    with torch.no_grad():
        for c in rng.choice(case.n, 6, replace=False):
            h = 1e-6 * float(rho[c])
            rp, rm = rho.detach().clone(), rho.detach().clone()
            rp[c] += h
            rm[c] -= h
            fd = (float(loss_of(rp)) - float(loss_of(rm))) / (2 * h)
            assert abs(fd - g[c]) <= 1e-5 * np.abs(g).max() + 1e-4 * abs(g[c]), (int(c), fd, g[c])


def test_mixed_precision_adjoint_close_to_fp64_and_meets_tolerance(engine, oracle_lib):
    """FP32-inner adjoint: the adjoint solve meets its tolerance in FP64 and its face gradients match
    the FP64 adjoint's on the 256x256 density-ratio-100 workload (both at relative residual 1e-10)."""
    case = _case(256, 256, seed=0, dr=100.0, tol=1e-10)
    target_case = _case(256, 256, seed=1, dr=100.0, tol=1e-10)
    ku, kv = W.face_inverse_density(case)
    src = W.divergence_source(case)
    p_target, _ = engine.solve_coeffs(target_case, ku, kv, W.divergence_source(target_case))
    p64, _ = engine.solve_coeffs(case, ku, kv, src)
    pm, prep = engine.solve_coeffs(case, ku, kv, src, precision=E.PP_MIXED)
    assert prep["outer_cycles"] >= 1 and prep["relative_residual"] < 1e-10
    assert np.abs(pm - p64).max() <= 1e-6 * np.abs(p64).max()
    lam64, dku64, dkv64, r64 = engine.adjoint(case, ku, kv, p64, p64 - p_target)
    lam, dku, dkv, rep = engine.adjoint(case, ku, kv, pm, pm - p_target, precision=E.PP_MIXED)
    assert rep["outer_cycles"] >= 1 and rep["relative_residual"] < 1e-10
    g64 = np.concatenate([dku64, dkv64])
    g = np.concatenate([dku, dkv])
    err_g = np.abs(g - g64).max() / np.abs(g64).max()
    err_l = np.abs(lam - lam64).max() / np.abs(lam64).max()
    print(f"[adjoint mixed] iterations {rep['iterations']} ({rep['outer_cycles']} cycles) vs FP64 {r64['iterations']}; "
          f"max grad diff / max grad {err_g:.2e}, max lambda diff / max lambda {err_l:.2e}")
    assert err_g < 1e-5 and err_l < 1e-5


def test_mixed_precision_finite_differences_at_64_faces_256x256(engine):
    """The FP32-inner adjoint gradient against central differences of the FP64 loss (64 faces, seed 2)."""
    case = _case(256, 256, seed=0, dr=100.0, tol=1e-12)
    target_case = _case(256, 256, seed=1, dr=100.0, tol=1e-12)
    ku, kv = W.face_inverse_density(case)
    src = W.divergence_source(case)
    p_target, _ = engine.solve_coeffs(target_case, ku, kv, W.divergence_source(target_case))
    p, _ = engine.solve_coeffs(case, ku, kv, src, precision=E.PP_MIXED)
    _, dku, dkv, _ = engine.adjoint(case, ku, kv, p, p - p_target, precision=E.PP_MIXED)
    grad = np.concatenate([dku, dkv])
    k = np.concatenate([ku.reshape(-1), kv.reshape(-1)])
    nu = ku.size
    faces = np.random.default_rng(2).choice(k.size, 64, replace=False)

    def loss(kk):
        pp, _ = engine.solve_coeffs(case, kk[:nu], kk[nu:], src)
        return 0.5 * float(np.sum((pp - p_target) ** 2))

    scale = np.abs(grad).max()
    worst = 0.0
    for f in faces:
        h = 1e-6 * abs(k[f])
        kp, km = k.copy(), k.copy()
        kp[f] += h
        km[f] -= h
        fd = (loss(kp) - loss(km)) / (2 * h)
        err = abs(fd - grad[f])
        worst = max(worst, err / scale)
        assert err <= 1e-6 * scale + 1e-4 * abs(grad[f]), (int(f), fd, grad[f])
    print(f"[fd mixed] 64 faces, max |fd - adjoint| / max|grad| = {worst:.3e}")


def test_autograd_mixed_precision():
    torch = pytest.importorskip("torch")
    from ding_nsch_core_solver_amd import autograd as A
    case = _case(64, 48, dr=50.0, tol=1e-10)
    ku, kv = W.face_inverse_density(case)
    src = W.divergence_source(case)
    dev = torch.device("cuda:0")
    grads = {}
    for prec in ("fp64", "mixed"):
        t = [torch.tensor(a.reshape(-1), device=dev, requires_grad=True) for a in (ku, kv, src)]
        p = A.pressure_solve(*t, case, precision=prec)
        (0.5 * (p ** 2).sum()).backward()
        grads[prec] = [x.grad.cpu().numpy() for x in t]
    for g64, gm in zip(grads["fp64"], grads["mixed"]):
        assert np.abs(gm - g64).max() <= 1e-5 * np.abs(g64).max()
