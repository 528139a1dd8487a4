"""Synthetic pressure-solve inputs (droplet density fields and seeded face velocities).

Every array is float64, i-major with j contiguous (cell (i, j) at ``i*ny + j``), the interior
layout of the reference's ``Field2D`` (src/core/base/field.hpp:23-33) without its ghost ring:

* ``rho``  nx x ny cell densities (the solver's ``rho_mid_``)
* ``u``    (nx+1) x ny x-face velocities (``u_star_``), face i sits between cells i-1 and i
* ``v``    nx x (ny+1) y-face velocities (``v_star_``)

The random stream is numpy's PCG64 seeded per case; u is drawn first (i-major), then v. Wall-normal
faces of non-periodic directions are zeroed, mirroring the Dirichlet-0 no-slip velocity ghosts the
reference imposes (operators.cpp:194-199, 250-255).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BC_UNSET, BC_DIRICHLET, BC_NEUMANN = 0, 1, 2
LEFT, RIGHT, BOTTOM, TOP = 0, 1, 2, 3


@dataclasses.dataclass
class PressureCase:
    """One pressure-correction problem: grid, boundary options, solver controls and fields."""

    nx: int
    ny: int
    lx: float = 1.0
    ly: float = 1.0
    dt: float = 1.0e-4
    periodic_x: bool = False
    periodic_y: bool = False
    bc_type: tuple = (BC_UNSET, BC_UNSET, BC_UNSET, BC_UNSET)
    bc_value: tuple = (0.0, 0.0, 0.0, 0.0)
    max_iterations: int = 20000
    tolerance: float = 1.0e-8
    density_ratio: float = 1000.0
    step_index: int = 0
    rho: np.ndarray | None = None
    u: np.ndarray | None = None
    v: np.ndarray | None = None
    p_n: np.ndarray | None = None
    p_nm1: np.ndarray | None = None

    @property
    def n(self) -> int:
        return self.nx * self.ny

    @property
    def dx(self) -> float:
        return self.lx / max(float(self.nx), 1.0)

    @property
    def dy(self) -> float:
        return self.ly / max(float(self.ny), 1.0)

    def has_dirichlet(self) -> bool:
        """Mirror of pressure_has_dirichlet_boundary (operators.cpp:114-128)."""
        if not self.periodic_x and (self.bc_type[LEFT] == BC_DIRICHLET or self.bc_type[RIGHT] == BC_DIRICHLET):
            return True
        if not self.periodic_y and (self.bc_type[BOTTOM] == BC_DIRICHLET or self.bc_type[TOP] == BC_DIRICHLET):
            return True
        return False


def droplet_density(nx, ny, lx=1.0, ly=1.0, cx=0.5, cy=0.5, radius=0.2, cn=0.01,
                    rho_in=1000.0, rho_out=1.0, width_factor=math.sqrt(2.0)):
    """rho = (1+phi)/2*rho_in + (1-phi)/2*rho_out with phi = tanh((R - r)/(width_factor*Cn))."""
    dx, dy = lx / nx, ly / ny
    x = (np.arange(nx, dtype=np.float64) + 0.5) * dx
    y = (np.arange(ny, dtype=np.float64) + 0.5) * dy
    r = np.sqrt((x[:, None] - cx) ** 2 + (y[None, :] - cy) ** 2)
    phi = np.tanh((radius - r) / (width_factor * cn))
    return np.ascontiguousarray((1.0 + phi) / 2.0 * rho_in + (1.0 - phi) / 2.0 * rho_out)


def random_faces(nx, ny, seed, periodic_x=False, periodic_y=False):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((nx + 1) * ny).reshape(nx + 1, ny)
    v = rng.standard_normal(nx * (ny + 1)).reshape(nx, ny + 1)
    if not periodic_x:
        u[0, :] = 0.0
        u[nx, :] = 0.0
    if not periodic_y:
        v[:, 0] = 0.0
        v[:, ny] = 0.0
    return np.ascontiguousarray(u), np.ascontiguousarray(v)


def droplet_case(nx=192, ny=192, seed=0, density_ratio=1000.0, radius=0.2, cx=0.5, cy=0.5, cn=0.01,
                 lx=1.0, ly=1.0, dt=1.0e-4, tolerance=1.0e-8, max_iterations=20000,
                 periodic_x=False, periodic_y=False, **kw) -> PressureCase:
    """A droplet of density ``density_ratio`` in a unit-density bath with a seeded random u*, v*.

    With the defaults this is the 192x192, density-ratio-1000 single-solve workload: Neumann
    pressure on all four walls, droplet radius 0.2 centred, Cn = 0.01, dt = 1e-4, tol 1e-8.
    """
    rho = droplet_density(nx, ny, lx, ly, cx, cy, radius, cn, rho_in=density_ratio, rho_out=1.0)
    u, v = random_faces(nx, ny, seed, periodic_x, periodic_y)
    return PressureCase(nx=nx, ny=ny, lx=lx, ly=ly, dt=dt, periodic_x=periodic_x, periodic_y=periodic_y,
                        tolerance=tolerance, max_iterations=max_iterations, density_ratio=density_ratio,
                        rho=rho, u=u, v=v, **kw)


BATCH_GRIDS = ((64, 64), (96, 192), (128, 128), (192, 192), (256, 256))


def batched_cases(count=256, seed=0, tolerance=1.0e-8):
    """Uneven batch: grid from BATCH_GRIDS, centre U[0.3,0.7]^2, radius U[0.1,0.25],
    density ratio 10**U[0,3], per-system RHS seed."""
    rng = np.random.default_rng(seed)
    cases = []
    for k in range(count):
        nx, ny = BATCH_GRIDS[int(rng.integers(0, len(BATCH_GRIDS)))]
        cx, cy = rng.uniform(0.3, 0.7, size=2)
        radius = rng.uniform(0.1, 0.25)
        dr = 10.0 ** rng.uniform(0.0, 3.0)
        cases.append(droplet_case(nx, ny, seed=1000 + k, density_ratio=dr, radius=radius, cx=cx, cy=cy,
                                  tolerance=tolerance))
    return cases


def face_inverse_density(case):
    """k_f = 1/rho_f on x-faces ((nx+1) x ny) and y-faces (nx x (ny+1)) with the reference's face
    averages: mean of the two cells, the adjacent cell on a wall, the wrap mean on a periodic edge
    (operators.cpp:330-350)."""
    nx, ny = case.nx, case.ny
    r = np.asarray(case.rho, dtype=np.float64).reshape(nx, ny)
    ru = np.empty((nx + 1, ny))
    ru[1:nx] = 0.5 * (r[:-1] + r[1:])
    if case.periodic_x:
        ru[0] = ru[nx] = 0.5 * (r[nx - 1] + r[0])
    else:
        ru[0], ru[nx] = r[0], r[nx - 1]
    rv = np.empty((nx, ny + 1))
    rv[:, 1:ny] = 0.5 * (r[:, :-1] + r[:, 1:])
    if case.periodic_y:
        rv[:, 0] = rv[:, ny] = 0.5 * (r[:, ny - 1] + r[:, 0])
    else:
        rv[:, 0], rv[:, ny] = r[:, 0], r[:, ny - 1]
    return np.ascontiguousarray(1.0 / ru), np.ascontiguousarray(1.0 / rv)


def divergence_source(case):
    """-div(u*)/dt per cell (pressure.cpp:150 with operators.cpp:388-390), nx x ny."""
    nx, ny = case.nx, case.ny
    u = np.asarray(case.u, dtype=np.float64).reshape(nx + 1, ny)
    v = np.asarray(case.v, dtype=np.float64).reshape(nx, ny + 1)
    div = (u[1:] - u[:-1]) / case.dx + (v[:, 1:] - v[:, :-1]) / case.dy
    return np.ascontiguousarray(-div / case.dt)


@dataclasses.dataclass
class VelocityBC:
    """Velocity boundary conditions of the correction step (the reference Config's u_bc_* / v_bc_*
    per edge in left, right, bottom, top order, and the wall speeds of unset u bottom/top edges)."""
    u_type: tuple = (BC_UNSET, BC_UNSET, BC_UNSET, BC_UNSET)
    u_value: tuple = (0.0, 0.0, 0.0, 0.0)
    v_type: tuple = (BC_UNSET, BC_UNSET, BC_UNSET, BC_UNSET)
    v_value: tuple = (0.0, 0.0, 0.0, 0.0)
    bottom_wall_velocity_x: float = 0.0
    top_wall_velocity_x: float = 0.0


def ghosted_shape(rows: int, cols: int, ghost: int = 3) -> tuple:
    """Shape of a Field2D buffer (rows x cols interior, `ghost` layers on every side)."""
    return (rows + 2 * ghost, cols + 2 * ghost)


@dataclasses.dataclass
class CsrSystem:
    """A sparse system in the form finalize_row_maps produces (ch_sparse_krylov.hpp:105-133):
    int32 row_ptr (n + 1) and col_idx with ascending columns per row, one diagonal entry per row,
    no entry below 1e-18 in magnitude; float64 values and right-hand side."""
    row_ptr: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray
    rhs: np.ndarray

    @property
    def n(self) -> int:
        return len(self.row_ptr) - 1


def csr_from_coo(n, rows, cols, vals, rhs) -> CsrSystem:
    """Sums duplicate (row, col) entries, drops |value| < 1e-18 as finalize_row_maps does."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key, vals = key[order], vals[order]
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(vals, start) if len(vals) else np.zeros(0)
    keep = np.abs(summed) >= 1.0e-18
    uniq, summed = uniq[keep], summed[keep]
    r, c = uniq // n, uniq % n
    row_ptr = np.zeros(n + 1, dtype=np.int32)
    np.add.at(row_ptr, r + 1, 1)
    return CsrSystem(np.cumsum(row_ptr).astype(np.int32), c.astype(np.int32), summed,
                     np.ascontiguousarray(rhs, dtype=np.float64))


def momentum_system(nx=192, ny=192, seed=0, density_ratio=1000.0, viscosity_ratio=100.0, dt=1.0e-3, re=100.0,
                    lx=1.0, ly=1.0) -> CsrSystem:
    """A synthetic system of the momentum predictor's shape (momentum.cpp:157-472): the u unknowns
    of the interior x-faces, then the v unknowns of the interior y-faces (walls on every side,
    numbered face by face in row-major order), each row rho_face u + alpha (full viscous tensor) with
    a droplet's density and viscosity: 5 entries on its own component and 4 cross entries on the
    other, alpha = dt / (2 Re). Right-hand side: rho_face times a seeded velocity plus forcing.
    Not the reference's assembly, the same structure and size (2 nx ny - nx - ny unknowns)."""
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    alpha = 0.5 * dt / re
    rho = droplet_density(nx, ny, lx, ly, rho_in=density_ratio, rho_out=1.0)
    phase = (rho - rho.min()) / max(rho.max() - rho.min(), 1e-30)
    eta_c = 1.0 + (viscosity_ratio - 1.0) * phase  # cell centres
    eta_pad = np.pad(eta_c, 1, mode="edge")
    eta_n = 0.25 * (eta_pad[:-1, :-1] + eta_pad[1:, :-1] + eta_pad[:-1, 1:] + eta_pad[1:, 1:])  # corners (nx+1, ny+1)
    u_id = -np.ones((nx + 1, ny), dtype=np.int64)
    u_id[1:nx, :] = np.arange((nx - 1) * ny).reshape(nx - 1, ny)
    nu = (nx - 1) * ny
    v_id = -np.ones((nx, ny + 1), dtype=np.int64)
    v_id[:, 1:ny] = nu + np.arange(nx * (ny - 1)).reshape(nx, ny - 1)
    n = nu + nx * (ny - 1)
    rows, cols, vals = [], [], []
    rhs = np.zeros(n)

    def add(r, c, v):
        m = (r >= 0) & (c >= 0)
        rows.append(r[m]); cols.append(c[m]); vals.append(v[m])

    cxy = 1.0 / (dx * dy)
    # u rows at faces (i, j), i = 1..nx-1
    I, J = np.meshgrid(np.arange(1, nx), np.arange(ny), indexing="ij")
    r = u_id[I, J]
    rho_f = 0.5 * (rho[I - 1, J] + rho[I, J])
    ce, cw = 2.0 * eta_c[I, J] / dx**2, 2.0 * eta_c[I - 1, J] / dx**2
    cn, cs = eta_n[I, J + 1] / dy**2, eta_n[I, J] / dy**2
    wall = np.where(J == ny - 1, cn, 0.0) + np.where(J == 0, cs, 0.0)  # no-slip walls: reflected ghost
    add(r, u_id[I - 1, J], -alpha * cw)
    add(r, np.where(J > 0, u_id[I, np.maximum(J - 1, 0)], -1), -alpha * cs)
    add(r, r, rho_f + alpha * (ce + cw + cn + cs + wall))
    add(r, np.where(J < ny - 1, u_id[I, np.minimum(J + 1, ny - 1)], -1), -alpha * cn)
    add(r, u_id[np.minimum(I + 1, nx), J], -alpha * ce)
    add(r, v_id[I - 1, J], alpha * eta_n[I, J] * cxy)
    add(r, v_id[I, J], -alpha * eta_n[I, J] * cxy)
    add(r, v_id[I - 1, J + 1], -alpha * eta_n[I, J + 1] * cxy)
    add(r, v_id[I, J + 1], alpha * eta_n[I, J + 1] * cxy)
    rhs[r.ravel()] = (rho_f * rng.standard_normal(rho_f.shape) + dt * rng.standard_normal(rho_f.shape)).ravel()
    # v rows at faces (i, j), j = 1..ny-1
    I, J = np.meshgrid(np.arange(nx), np.arange(1, ny), indexing="ij")
    r = v_id[I, J]
    rho_f = 0.5 * (rho[I, J - 1] + rho[I, J])
    cn, cs = 2.0 * eta_c[I, J] / dy**2, 2.0 * eta_c[I, J - 1] / dy**2
    ce, cw = eta_n[I + 1, J] / dx**2, eta_n[I, J] / dx**2
    wall = np.where(I == nx - 1, ce, 0.0) + np.where(I == 0, cw, 0.0)
    add(r, np.where(I > 0, v_id[np.maximum(I - 1, 0), J], -1), -alpha * cw)
    add(r, u_id[I, J - 1], alpha * eta_n[I, J] * cxy)
    add(r, u_id[I, J], -alpha * eta_n[I, J] * cxy)
    add(r, v_id[I, J - 1], -alpha * cs)
    add(r, r, rho_f + alpha * (ce + cw + cn + cs + wall))
    add(r, v_id[I, J + 1], -alpha * cn)
    add(r, u_id[I + 1, J - 1], -alpha * eta_n[I + 1, J] * cxy)
    add(r, u_id[I + 1, J], alpha * eta_n[I + 1, J] * cxy)
    add(r, np.where(I < nx - 1, v_id[np.minimum(I + 1, nx - 1), J], -1), -alpha * ce)
    rhs[r.ravel()] = (rho_f * rng.standard_normal(rho_f.shape) + dt * rng.standard_normal(rho_f.shape)).ravel()
    rows = np.concatenate([a.ravel() for a in rows])
    cols = np.concatenate([a.ravel() for a in cols])
    vals = np.concatenate([a.ravel() for a in vals])
    return csr_from_coo(n, rows, cols, vals, rhs)


# Full-run settings (the reference's Config fields as oracle.ref.RefRun takes them). "width" is the
# interface width of the initial phase c = 0.5 - 0.5 tanh((r - radius) / width): sqrt(2) Cn gives the
# profile phi = tanh((R - r) / (sqrt(2) Cn)), c = (1 + phi) / 2, that the workloads state (the
# reference's own initializer uses 2 sqrt(2) Cn). Unstated values are the reference's defaults
# (config.hpp: pe 100, ch_inner_iterations 30, ch_tolerance 1e-8, momentum 50 iterations to 1e-8).
_SQRT2 = 2.0 ** 0.5


def _run_spec(**kw):
    spec = dict(nx=64, ny=64, lx=1.0, ly=1.0, dt=1.0e-4, periodic_x=1, periodic_y=1, re=1.0, ca=1.0, pe=100.0,
                cn=0.01, density_ratio=1.0, viscosity_ratio=1.0, surface_tension_multiplier=1.0, body_force_x=0.0,
                body_force_y=0.0, center_x=0.5, center_y=0.5, radius=0.25, width=0.0, ch_inner_iterations=30,
                ch_tolerance=1.0e-8, momentum_iterations=50, momentum_tolerance=1.0e-8, poisson_iterations=400,
                pressure_tolerance=1.0e-8)
    spec.update(kw)
    return spec


# 64x64 periodic static droplet, R = 0.25, Cn = 0.01, density and viscosity ratio 1, surface tension
# 1 (re = ca = 1), dt = 1e-4, icpcg to 1e-8; 10 steps.
STATIC_DROPLET_64 = _run_spec(width=_SQRT2 * 0.01)
STATIC_DROPLET_64_STEPS = 10

# 192x384 planar rising bubble on [0,1]x[0,2]: R = 0.25 at (0.5, 0.5), density ratio 1000, viscosity
# ratio 100, gravity 0.98 downwards, sigma = 1/(re ca) = 1.96 with re = 10, Cn = 0.01, no-slip walls,
# dt = 5e-5, pressure to 1e-10; 500 steps. The initial interface has the reference initializer's
# width 2 sqrt(2) Cn: with sqrt(2) Cn (2.7 cells at this resolution) the first CH step overshoots c to
# 1.004, the density 1 - 1000 (1 - c) turns negative near the bubble and the reference's own ICC
# factorization of the pressure matrix fails ("ICC preconditioner construction failed").
RISING_BUBBLE_192X384 = _run_spec(nx=192, ny=384, ly=2.0, dt=5.0e-5, periodic_x=0, periodic_y=0, re=10.0,
                                  ca=1.0 / (10.0 * 1.96), density_ratio=1000.0, viscosity_ratio=100.0,
                                  body_force_y=-0.98, width=0.0, poisson_iterations=20000,
                                  pressure_tolerance=1.0e-10)
RISING_BUBBLE_STEPS = 500

# The reference's 128x128 static-droplet validation run whose history is published with it
# (testcases/92_static_droplet_surface_tension_validation_128.cfg with the icpcg pressure scheme):
# restarted from its step-5000 snapshot.
STATIC_DROPLET_128_RESTART = _run_spec(nx=128, ny=128, dt=2.0e-7, re=50.0, ca=1.0, pe=150.0, cn=0.02, radius=0.16,
                                       ch_inner_iterations=80, momentum_iterations=200, poisson_iterations=1200,
                                       pressure_tolerance=1.0e-10)


@dataclasses.dataclass
class MomentumCase:
    """Inputs of one momentum predictor step (Solver::compute_momentum_fluxes +
    solve_momentum_predictor): ghosted Field2D buffers with `ghost` layers, flattened i-major:
    u ((nx+1) x ny faces) and v (nx x (ny+1) faces) at step n, rho_mid / eta / eta_prev (nx x ny
    cells: the midpoint density, the viscosity and the previous step's viscosity), the surface
    tension forces on the u and v faces and the previous step's convection fields u_prev / v_prev
    (momentum_u_rhs_prev_, momentum_v_rhs_prev_). dt, re, body forces, periodicity, velocity BCs and
    the BiCGSTAB controls (momentum_iterations, momentum_tolerance) as the reference Config."""
    nx: int
    ny: int
    lx: float = 1.0
    ly: float = 1.0
    dt: float = 1.0e-4
    re: float = 100.0
    body_force_x: float = 0.0
    body_force_y: float = 0.0
    periodic_x: bool = False
    periodic_y: bool = False
    vbc: VelocityBC = dataclasses.field(default_factory=VelocityBC)
    step_index: int = 1
    max_iterations: int = 50
    tolerance: float = 1.0e-8
    ghost: int = 3
    u: np.ndarray | None = None
    v: np.ndarray | None = None
    rho_mid: np.ndarray | None = None
    eta: np.ndarray | None = None
    eta_prev: np.ndarray | None = None
    surface_fx_u: np.ndarray | None = None
    surface_fy_v: np.ndarray | None = None
    u_prev: np.ndarray | None = None
    v_prev: np.ndarray | None = None

    @property
    def unknowns(self) -> int:
        nu = (self.nx if self.periodic_x else self.nx - 1) * self.ny
        nv = self.nx * (self.ny if self.periodic_y else self.ny - 1)
        return nu + nv


def _scalar_ghosts(a: np.ndarray, g: int, periodic_x: bool, periodic_y: bool) -> np.ndarray:
    """apply_scalar_bc (operators.cpp:130-141): periodic wrap or copy of the edge cell, x then y."""
    a = np.pad(a, ((g, g), (0, 0)), mode="wrap" if periodic_x else "edge")
    return np.pad(a, ((0, 0), (g, g)), mode="wrap" if periodic_y else "edge")


def momentum_case(nx=192, ny=192, seed=0, density_ratio=1000.0, viscosity_ratio=100.0, radius=0.2, cn=0.01,
                  lx=1.0, ly=1.0, dt=1.0e-4, re=100.0, body_force_y=0.0, periodic_x=False, periodic_y=False,
                  vbc=None, step_index=1, velocity_scale=0.1, ghost=3, **kw) -> MomentumCase:
    """A droplet (density `density_ratio`, viscosity `viscosity_ratio`, as c + (1 - c) ratio with c = 0
    inside) in a unit bath; u, v, the surface forces and the previous convection fields are seeded
    N(0, velocity_scale) (ghost layers included: the predictor reads them); rho, eta ghost-filled as
    apply_scalar_bc leaves them, eta_prev a slightly perturbed eta."""
    rng = np.random.default_rng(seed)
    g = ghost
    x = (np.arange(nx) + 0.5) * lx / nx
    y = (np.arange(ny) + 0.5) * ly / ny
    r = np.hypot(x[:, None] - 0.5 * lx, y[None, :] - 0.5 * ly)
    c = 0.5 - 0.5 * np.tanh((radius - r) / (2.0 * math.sqrt(2.0) * cn))
    rho = c + (1.0 - c) * density_ratio
    eta = c + (1.0 - c) * viscosity_ratio
    eta_prev = eta * (1.0 + 1e-3 * rng.standard_normal(eta.shape))
    shape_u = (nx + 1 + 2 * g, ny + 2 * g)
    shape_v = (nx + 2 * g, ny + 1 + 2 * g)
    noise = lambda shape: velocity_scale * rng.standard_normal(shape)  # noqa: E731
    return MomentumCase(nx=nx, ny=ny, lx=lx, ly=ly, dt=dt, re=re, body_force_y=body_force_y, periodic_x=periodic_x,
                        periodic_y=periodic_y, vbc=vbc if vbc is not None else VelocityBC(), step_index=step_index,
                        ghost=g, u=noise(shape_u).reshape(-1), v=noise(shape_v).reshape(-1),
                        rho_mid=_scalar_ghosts(rho, g, periodic_x, periodic_y).reshape(-1),
                        eta=_scalar_ghosts(eta, g, periodic_x, periodic_y).reshape(-1),
                        eta_prev=_scalar_ghosts(eta_prev, g, periodic_x, periodic_y).reshape(-1),
                        surface_fx_u=noise(shape_u).reshape(-1), surface_fy_v=noise(shape_v).reshape(-1),
                        u_prev=noise(shape_u).reshape(-1), v_prev=noise(shape_v).reshape(-1), **kw)


def stabilization_coefficients(cn: float, pe: float):
    """beta1 = a1 / Pe and beta2 = a2 / Pe with a1 = 3 sqrt(2) / (16 Cn), a2 = 3 sqrt(2) Cn / 4
    (config.hpp:9-15, phase_field.cpp:129-130)."""
    s2 = 1.4142135623730950488
    a1 = (3.0 * s2 / (16.0 * cn)) if cn > 0.0 else 0.0
    a2 = (3.0 * s2 * cn / 4.0) if cn > 0.0 else 0.0
    return a1 / pe, a2 / pe


@dataclasses.dataclass
class ChCase:
    """One Cahn-Hilliard linear solve (Solver::solve_phase_linear_system_eq25): the operator
    alpha0 I - beta1 Lap + beta2 Lap^2 on an nx x ny grid (beta1, beta2 from Cn and Pe), the
    right-hand side and the initial guess c0 (natural order i * ny + j), ch_inner_iterations and
    ch_tolerance. alpha0 is 1/dt on the first step and 3/(2 dt) after (phase_field.cpp:678)."""
    nx: int
    ny: int
    lx: float = 1.0
    ly: float = 1.0
    periodic_x: bool = False
    periodic_y: bool = False
    cn: float = 0.01
    pe: float = 100.0
    alpha0: float = 1.0e4
    max_iterations: int = 30
    tolerance: float = 1.0e-8
    rhs: np.ndarray | None = None
    c0: np.ndarray | None = None

    @property
    def beta1(self) -> float:
        return stabilization_coefficients(self.cn, self.pe)[0]

    @property
    def beta2(self) -> float:
        return stabilization_coefficients(self.cn, self.pe)[1]


def ch_case(nx=192, ny=192, seed=0, radius=0.2, cn=0.01, pe=100.0, dt=1.0e-4, first_step=False, lx=1.0, ly=1.0,
            periodic_x=False, periodic_y=False, **kw) -> ChCase:
    """A droplet phase field c0 (c = 1 inside, tanh profile of width 2 sqrt(2) Cn) and a right-hand
    side alpha0 c0 + a seeded perturbation, as the BDF2 right-hand side of a step carries."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) * lx / nx
    y = (np.arange(ny) + 0.5) * ly / ny
    r = np.hypot(x[:, None] - 0.5 * lx, y[None, :] - 0.5 * ly)
    c0 = 0.5 - 0.5 * np.tanh((r - radius) / (2.0 * math.sqrt(2.0) * cn))
    alpha0 = 1.0 / dt if first_step else 3.0 / (2.0 * dt)
    rhs = alpha0 * c0 + 50.0 * rng.standard_normal(c0.shape)
    return ChCase(nx=nx, ny=ny, lx=lx, ly=ly, periodic_x=periodic_x, periodic_y=periodic_y, cn=cn, pe=pe,
                  alpha0=alpha0, rhs=rhs.reshape(-1), c0=c0.reshape(-1), **kw)
