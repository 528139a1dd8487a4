"""Differentiable pressure solve for PyTorch: forward = pp_solve_coeffs_device, backward =
pp_adjoint_device (one adjoint PCG solve plus the face-gradient kernel).

    p = pressure_solve(k_u, k_v, src, case)      # float64 CUDA tensors, p of shape (nx*ny,)
    loss = 0.5 * ((p - p_target) ** 2).sum(); loss.backward()

precision="mixed" runs both the forward and the adjoint solve with FP32 inner PCG inside FP64
iterative refinement (same tolerance, FP64 results).

k_u ((nx+1)*ny) and k_v (nx*(ny+1)) are face inverse densities 1/rho_f, src (nx*ny) the source
term (the reference's -div(u*)/dt); geometry, boundary conditions and solver settings come from
`case` (a PressureCase; its rho/u/v are not used). Gradients flow to all three tensors;
face_inverse_density and divergence_source build them differentiably from rho, u* and v*.
"""
from __future__ import annotations

import torch

from . import engine as E

_engines: dict[int, E.Engine] = {}


def _engine(device: torch.device) -> E.Engine:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _engines:
        _engines[idx] = E.Engine(idx)
    return _engines[idx]


def _checked(t: torch.Tensor, numel: int, name: str) -> torch.Tensor:
    if t.dtype != torch.float64 or not t.is_cuda:
        raise TypeError(f"{name} must be a float64 CUDA tensor")
    if t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")
    return t.contiguous()


class PressureSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k_u, k_v, src, case, scheme, precision):
        nx, ny = case.nx, case.ny
        k_u = _checked(k_u, (nx + 1) * ny, "k_u")
        k_v = _checked(k_v, nx * (ny + 1), "k_v")
        src = _checked(src, nx * ny, "src")
        eng = _engine(k_u.device)
        p = torch.empty(nx * ny, dtype=torch.float64, device=k_u.device)
        stream = torch.cuda.current_stream(k_u.device).cuda_stream
        ctx.report = eng.solve_coeffs_device(case, k_u.data_ptr(), k_v.data_ptr(), src.data_ptr(), p.data_ptr(),
                                             scheme=scheme, stream=stream, precision=precision)
        ctx.case, ctx.scheme, ctx.precision = case, scheme, precision
        ctx.save_for_backward(k_u, k_v, p)
        return p

    @staticmethod
    def backward(ctx, grad_p):
        k_u, k_v, p = ctx.saved_tensors
        case = ctx.case
        g = _checked(grad_p, case.n, "grad_p")
        lam = torch.empty_like(p)
        dku = torch.empty_like(k_u)
        dkv = torch.empty_like(k_v)
        stream = torch.cuda.current_stream(p.device).cuda_stream
        _engine(p.device).adjoint_device(case, k_u.data_ptr(), k_v.data_ptr(), p.data_ptr(), g.data_ptr(),
                                         lam.data_ptr(), dku.data_ptr(), dkv.data_ptr(), scheme=ctx.scheme,
                                         stream=stream, precision=ctx.precision)
        return dku, dkv, lam, None, None, None


_PRECISIONS = {"fp64": E.PP_FP64, "mixed": E.PP_MIXED}


def pressure_solve(k_u, k_v, src, case, scheme: str | int = "hip_icpcg", precision: str | int = "fp64"):
    code = E.SCHEME_CODES[scheme] if isinstance(scheme, str) else int(scheme)
    prec = _PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    return PressureSolve.apply(k_u, k_v, src, case, code, prec)


def face_inverse_density(rho: torch.Tensor, case) -> tuple[torch.Tensor, torch.Tensor]:
    """Differentiable k_f = 1/rho_f with the reference's face averages (operators.cpp:330-350)."""
    nx, ny = case.nx, case.ny
    r = rho.reshape(nx, ny)
    if case.periodic_x:
        wrap = 0.5 * (r[nx - 1] + r[0])
        ru = torch.cat([wrap[None], 0.5 * (r[:-1] + r[1:]), wrap[None]], 0)
    else:
        ru = torch.cat([r[:1], 0.5 * (r[:-1] + r[1:]), r[-1:]], 0)
    if case.periodic_y:
        wrap = 0.5 * (r[:, ny - 1] + r[:, 0])
        rv = torch.cat([wrap[:, None], 0.5 * (r[:, :-1] + r[:, 1:]), wrap[:, None]], 1)
    else:
        rv = torch.cat([r[:, :1], 0.5 * (r[:, :-1] + r[:, 1:]), r[:, -1:]], 1)
    return (1.0 / ru).reshape(-1), (1.0 / rv).reshape(-1)


def divergence_source(u: torch.Tensor, v: torch.Tensor, case) -> torch.Tensor:
    """Differentiable -div(u*)/dt (operators.cpp:388-390, pressure.cpp:150)."""
    nx, ny = case.nx, case.ny
    u = u.reshape(nx + 1, ny)
    v = v.reshape(nx, ny + 1)
    div = (u[1:] - u[:-1]) / case.dx + (v[:, 1:] - v[:, :-1]) / case.dy
    return (-div / case.dt).reshape(-1)
