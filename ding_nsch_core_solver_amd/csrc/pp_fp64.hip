// FP64 pressure correction on one workgroup per system: assembly, regularization, ICC(0)/ILDLT(0)
// factorization, projected PCG and the final mean subtraction, all inside one launch.
//
// Arithmetic follows the reference operation for operation (compiled with -ffp-contract=off, IEEE
// division and square root):
//   assembly   src/core/coupled/pressure.cpp:123-196 (+ operators.cpp:330-350, 388-390, 411-444)
//   regularize pressure.cpp:198-212           factor  ch_sparse_krylov.hpp:337-419 / 252-335
//   SpMV       ch_sparse_krylov.hpp:235-250   sweeps  ch_sparse_krylov.hpp:444-513
//   PCG        ch_sparse_krylov.hpp:515-660 with the projections of pressure.cpp:225-239
//   output     pressure.cpp:241-248 and subtract_mean (operators.cpp:447-460)
// Only the summation order of reductions differs (canonical tree, see pp_device.hpp).
//
// Projections are applied lazily and bit-identically: x, r and p are stored unprojected and
// "v - mean" is recomputed wherever the reference would read the projected vector, which yields
// the same double as the reference's stored v[idx] -= mean.
#include "pp_device.hpp"
#include "pp_kernels.hpp"

namespace pp {

namespace {

__device__ __forceinline__ double rho_xface(const SysArgs &a, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const double *rho = a.rho;
  if (i <= 0) return a.periodic_x ? 0.5 * (rho[(nx - 1) * ny + j] + rho[j]) : rho[j];
  if (i >= nx) return a.periodic_x ? 0.5 * (rho[(nx - 1) * ny + j] + rho[j]) : rho[(nx - 1) * ny + j];
  return 0.5 * (rho[(i - 1) * ny + j] + rho[i * ny + j]);
}

__device__ __forceinline__ double rho_yface(const SysArgs &a, int i, int j) {
  const int ny = a.ny;
  const double *rho = a.rho;
  if (j <= 0) return a.periodic_y ? 0.5 * (rho[i * ny + ny - 1] + rho[i * ny]) : rho[i * ny];
  if (j >= ny) return a.periodic_y ? 0.5 * (rho[i * ny + ny - 1] + rho[i * ny]) : rho[i * ny + ny - 1];
  return 0.5 * (rho[i * ny + j - 1] + rho[i * ny + j]);
}

__device__ __forceinline__ double divergence(const SysArgs &a, int i, int j) {
  const int ny = a.ny;
  return (a.u[(i + 1) * ny + j] - a.u[i * ny + j]) / a.dx + (a.v[i * (ny + 1) + j + 1] - a.v[i * (ny + 1) + j]) / a.dy;
}

// p_ext = p^n (first step) or 2 p^n - p^{n-1}, with periodic wrap (operators.cpp:413-422).
__device__ __forceinline__ double p_ext(const SysArgs &a, int i, int j) {
  if (i < 0) i += a.nx;
  if (i >= a.nx) i -= a.nx;
  if (j < 0) j += a.ny;
  if (j >= a.ny) j -= a.ny;
  const int k = i * a.ny + j;
  const double pn = a.p_n ? a.p_n[k] : 0.0;
  if (a.step_index == 0) return pn;
  const double pm = a.p_nm1 ? a.p_nm1[k] : 0.0;
  return 2.0 * pn - pm;
}

// operators.cpp:424-444
__device__ double split_explicit(const SysArgs &a, int i, int j) {
  const bool east_open = a.periodic_x || i < a.nx - 1;
  const bool west_open = a.periodic_x || i > 0;
  const bool north_open = a.periodic_y || j < a.ny - 1;
  const bool south_open = a.periodic_y || j > 0;
  const double pc = p_ext(a, i, j);
  const double rr = a.rho_ref;
  const double fe = east_open ? (1.0 - rr / rho_xface(a, i + 1, j)) * ((p_ext(a, i + 1, j) - pc) / a.dx) : 0.0;
  const double fw = west_open ? (1.0 - rr / rho_xface(a, i, j)) * ((pc - p_ext(a, i - 1, j)) / a.dx) : 0.0;
  const double fn = north_open ? (1.0 - rr / rho_yface(a, i, j + 1)) * ((p_ext(a, i, j + 1) - pc) / a.dy) : 0.0;
  const double fs = south_open ? (1.0 - rr / rho_yface(a, i, j)) * ((pc - p_ext(a, i, j - 1)) / a.dy) : 0.0;
  return (fe - fw) / a.dx + (fn - fs) / a.dy;
}

__device__ __forceinline__ double drop_tiny(double value) { return fabs(value) < 1.0e-18 ? 0.0 : value; }

// One row of pressure.cpp:133-195 in the wavefront layout. Returns the diagonal.
__device__ void assemble_cell(const SysArgs &a, int q, int i, int j) {
  const bool has_e = a.periodic_x || i < a.nx - 1;
  const bool has_w = a.periodic_x || i > 0;
  const bool has_n = a.periodic_y || j < a.ny - 1;
  const bool has_s = a.periodic_y || j > 0;
  const double dx = a.dx, dy = a.dy;
  double ce, cw, cn, cs;
  if (a.split) {
    ce = cw = 1.0 / (dx * dx);
    cn = cs = 1.0 / (dy * dy);
  } else {
    ce = 1.0 / (rho_xface(a, i + 1, j) * dx * dx);
    cw = 1.0 / (rho_xface(a, i, j) * dx * dx);
    cn = 1.0 / (rho_yface(a, i, j + 1) * dy * dy);
    cs = 1.0 / (rho_yface(a, i, j) * dy * dy);
  }
  double diag = 0.0;
  double b = a.split ? -(split_explicit(a, i, j) + a.rho_ref * divergence(a, i, j) / a.dt) : -divergence(a, i, j) / a.dt;
  auto boundary = [&](int side, double coeff, double spacing) {
    int type = a.bc_type[side];
    const double value = type == 0 ? 0.0 : a.bc_value[side];
    if (type == 0) type = 2;
    if (type == 1) {
      diag += 2.0 * coeff;
      b += 2.0 * coeff * value;
    } else {
      b += coeff * value * spacing;
    }
  };
  if (has_e) diag += ce; else boundary(1, ce, dx);
  if (has_w) diag += cw; else boundary(0, cw, dx);
  if (has_n) diag += cn; else boundary(3, cn, dy);
  if (has_s) diag += cs; else boundary(2, cs, dy);
  a.aP[q] = 0.0 + diag;
  a.aW[q] = has_w ? drop_tiny(0.0 + -cw) : 0.0;
  a.aS[q] = has_s ? drop_tiny(0.0 + -cs) : 0.0;
  a.b[q] = b;
}

// ---- factorization: forward wavefront over levels ------------------------------------------
template <bool kIldlt>
__device__ void factor_sweep(const SysArgs &a, SweepShared &sh, double floor) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int S = (nx + 63) >> 6;
  if (wave >= S) return;
  const int i = wave * 64 + lane;
  const int i_last = min(wave * 64 + 63, nx - 1);
  const int d_beg = wave * 64, d_end = i_last + ny - 1;
  const double floor_sq = floor * floor;
  double prev = 0.0, first = 0.0;
  bool failed = false;
  for (int d = d_beg; d <= d_end; ++d) {
    const int j = d - i;
    const bool act = i < nx && j >= 0 && j < ny;
    double west = __shfl_up(prev, 1, 64);
    west = forward_receive(west, wave, lane, d - wave * 64, ny, sh);
    double cur = 0.0;
    if (act) {
      const int q = sh.off[d] + i - level_lo(d, ny);
      double dd = a.aP[q];
      if (a.periodic_x && i == nx - 1) {  // entry (nx-1, j) -> (0, j), first in column order
        const double dc = sh.wrap[j];
        const double l = a.aW[cell_pos(sh, 0, j, ny)] / floor_max(dc, floor);
        a.LX[j] = l;
        dd = kIldlt ? dd - (l * l) * dc : dd - l * l;
      }
      if (i > 0) {
        const double l = a.aW[q] / floor_max(west, floor);
        a.Lw[q] = l;
        dd = kIldlt ? dd - (l * l) * west : dd - l * l;
      } else {
        a.Lw[q] = 0.0;
      }
      if (a.periodic_y && j == ny - 1) {  // entry (i, ny-1) -> (i, 0)
        const double l = a.aS[cell_pos(sh, i, 0, ny)] / floor_max(first, floor);
        a.LY[i] = l;
        dd = kIldlt ? dd - (l * l) * first : dd - l * l;
      }
      if (j > 0) {
        const double l = a.aS[q] / floor_max(prev, floor);
        a.Ls[q] = l;
        dd = kIldlt ? dd - (l * l) * prev : dd - l * l;
      } else {
        a.Ls[q] = 0.0;
      }
      if (kIldlt) {
        if (!isfinite(dd) || dd <= floor) failed = true;
        cur = dd;
      } else {
        if (!isfinite(dd) || dd <= floor_sq) failed = true;
        cur = sqrt(dd);
      }
      a.dg[q] = cur;
      if (j == 0) first = cur;
      if (a.periodic_x && i == 0) sh.wrap[j] = cur;
    }
    prev = cur;
    forward_send(cur, wave, lane, S, d - (wave * 64 + 63), ny, sh);
  }
  if (failed) sh.status = kFactorFail;
}

// ---- z = M^{-1} (r - m_r): forward then backward wavefront ----------------------------------
template <bool kIldlt>
__device__ void forward_solve(const SysArgs &a, SweepShared &sh, double m_r, double floor) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int S = (nx + 63) >> 6;
  if (wave >= S) return;
  const int i = wave * 64 + lane;
  const int i_last = min(wave * 64 + 63, nx - 1);
  const int d_beg = wave * 64, d_end = i_last + ny - 1;
  double prev = 0.0, first = 0.0;
  for (int d = d_beg; d <= d_end; ++d) {
    const int j = d - i;
    const bool act = i < nx && j >= 0 && j < ny;
    double west = __shfl_up(prev, 1, 64);
    west = forward_receive(west, wave, lane, d - wave * 64, ny, sh);
    double cur = 0.0;
    if (act) {
      const int q = sh.off[d] + i - level_lo(d, ny);
      double s = a.r[q] - m_r;
      if (a.periodic_x && i == nx - 1) s = s - a.LX[j] * sh.wrap[j];
      if (i > 0) s = s - a.Lw[q] * west;
      if (a.periodic_y && j == ny - 1) s = s - a.LY[i] * first;
      if (j > 0) s = s - a.Ls[q] * prev;
      if (kIldlt) {
        cur = s;
        a.y[q] = s / floor_max(a.dg[q], floor);
      } else {
        cur = s / floor_max(a.dg[q], floor);
        a.y[q] = cur;
      }
      if (j == 0) first = cur;
      if (a.periodic_x && i == 0) sh.wrap[j] = cur;
    }
    prev = cur;
    forward_send(cur, wave, lane, S, d - (wave * 64 + 63), ny, sh);
  }
}

template <bool kIldlt>
__device__ void backward_solve(const SysArgs &a, SweepShared &sh, double floor) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int S = (nx + 63) >> 6;
  if (wave >= S) return;
  const int i = wave * 64 + lane;
  const int i_last = min(wave * 64 + 63, nx - 1);
  const int d_beg = wave * 64, d_end = i_last + ny - 1;
  double prev = 0.0, first = 0.0;
  for (int d = d_end; d >= d_beg; --d) {
    const int j = d - i;
    const bool act = i < nx && j >= 0 && j < ny;
    double east = __shfl_down(prev, 1, 64);
    east = backward_receive(east, wave, lane, S, d - (wave * 64 + 63), ny, sh);
    double cur = 0.0;
    if (act) {
      const int q = sh.off[d] + i - level_lo(d, ny);
      double s = a.y[q];
      if (j + 1 < ny) {
        const int qn = sh.off[d + 1] + i - level_lo(d + 1, ny);
        s = s - a.Ls[qn] * prev;
      }
      if (a.periodic_y && j == 0) s = s - a.LY[i] * first;
      if (i + 1 < nx) {
        const int qe = sh.off[d + 1] + i + 1 - level_lo(d + 1, ny);
        s = s - a.Lw[qe] * east;
      }
      if (a.periodic_x && i == 0) s = s - a.LX[j] * sh.wrap[j];
      cur = kIldlt ? s : s / floor_max(a.dg[q], floor);
      a.z[q] = cur;
      if (j == ny - 1) first = cur;
      if (a.periodic_x && i == nx - 1) sh.wrap[j] = cur;
    }
    prev = cur;
    backward_send(cur, wave, lane, d - wave * 64, ny, sh);
  }
}

// (A p)_q in CSR column order, p read through the lazy projection p - m_p.
__device__ __forceinline__ double spmv_row(const SysArgs &a, const SweepShared &sh, const double *p, double m_p,
                                           int q, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  double s = 0.0;
  if (a.periodic_x && i == nx - 1) {
    const int qw = cell_pos(sh, 0, j, ny);
    s += a.aW[qw] * (p[qw] - m_p);
  }
  if (i >= 1) s += a.aW[q] * (p[cell_pos(sh, i - 1, j, ny)] - m_p);
  if (a.periodic_y && j == ny - 1) {
    const int qs = cell_pos(sh, i, 0, ny);
    s += a.aS[qs] * (p[qs] - m_p);
  }
  if (j >= 1) s += a.aS[q] * (p[cell_pos(sh, i, j - 1, ny)] - m_p);
  s += a.aP[q] * (p[q] - m_p);
  if (j <= ny - 2) {
    const int qn = cell_pos(sh, i, j + 1, ny);
    s += a.aS[qn] * (p[qn] - m_p);
  }
  if (a.periodic_y && j == 0) s += a.aS[q] * (p[cell_pos(sh, i, ny - 1, ny)] - m_p);
  if (i <= nx - 2) {
    const int qe = cell_pos(sh, i + 1, j, ny);
    s += a.aW[qe] * (p[qe] - m_p);
  }
  if (a.periodic_x && i == 0) s += a.aW[q] * (p[cell_pos(sh, nx - 1, j, ny)] - m_p);
  return s;
}

template <bool kIldlt>
__device__ void precondition(const SysArgs &a, SweepShared &sh, double m_r, double floor) {
  reset_handoff(sh);
  __syncthreads();
  forward_solve<kIldlt>(a, sh, m_r, floor);
  __syncthreads();
  reset_handoff(sh);
  __syncthreads();
  backward_solve<kIldlt>(a, sh, floor);
  __syncthreads();
}

template <bool kIldlt>
__device__ void solve_system(const SysArgs &a, SweepShared &sh) {
  const int tid = threadIdx.x;
  const int n = a.n, nx = a.nx, ny = a.ny;
  int parity = 0;
  for (int d = tid; d <= a.nlev; d += kThreads) sh.off[d] = a.lev_off[d];
  if (tid == 0) {
    sh.spin_error = 0;
    sh.status = kOk;
  }
  __syncthreads();

  // assembly (+ missing-diagonal check), rhs regularization
  double bsum = 0.0, dmax = 0.0;
  int nodiag = 0;
  for (int q = tid; q < n; q += kThreads) {
    const uint32_t c = a.cell[q];
    assemble_cell(a, q, static_cast<int>(c >> 16), static_cast<int>(c & 0xffffu));
    const double ap = a.aP[q];
    if (fabs(ap) < 1.0e-18) nodiag = 1;
    const double aa = fabs(ap);
    dmax = dmax < aa ? aa : dmax;
    bsum = bsum + a.b[q];
  }
  if (nodiag) sh.status = kNoDiag;
  const double total = block_sum1(bsum, sh, parity);
  const double max_diag = block_max1(dmax, sh, parity);
  if (sh.status != kOk) return;
  const double nd = static_cast<double>(n) < 1.0 ? 1.0 : static_cast<double>(n);
  const double b_mean = a.projected ? total / nd : 0.0;

  // factorization
  const double base = 1.0 < max_diag ? max_diag : 1.0;
  const double floor = kIldlt ? base * 1.0e-12 : sqrt(base * 1.0e-12);
  reset_handoff(sh);
  __syncthreads();
  factor_sweep<kIldlt>(a, sh, floor);
  __syncthreads();
  if (sh.spin_error) sh.status = kSweepFail;
  if (sh.status != kOk) return;

  // r = b - A*0 (= b), x = 0; projections and norms
  double acc[2] = {0.0, 0.0};
  for (int q = tid; q < n; q += kThreads) {
    const double bq = a.projected ? a.b[q] - b_mean : a.b[q];
    a.b[q] = bq;
    a.r[q] = bq - 0.0;
    a.x[q] = 0.0;
    acc[0] = acc[0] + bq * bq;
    acc[1] = acc[1] + a.r[q];
  }
  block_sum<2>(acc, sh, parity);
  const double bnorm = sqrt(acc[0] < 0.0 ? 0.0 : acc[0]);
  double m_r = a.projected ? acc[1] / nd : 0.0;
  double rr = 0.0;
  for (int q = tid; q < n; q += kThreads) {
    const double t = a.r[q] - m_r;
    rr = rr + t * t;
  }
  rr = block_sum1(rr, sh, parity);
  double abs_res = sqrt(rr < 0.0 ? 0.0 : rr);
  double rel_res = abs_res / (bnorm < 1.0e-30 ? 1.0e-30 : bnorm);
  double iter_res = 0.0;
  int iterations = 0;
  double m_x = 0.0, m_p = 0.0;
  const double tol = a.tolerance;
  double *p_raw = a.pa, *p_prj = a.pb;

  if (!(rel_res < tol || abs_res < tol)) {
    precondition<kIldlt>(a, sh, m_r, floor);
    double zs = 0.0;
    for (int q = tid; q < n; q += kThreads) zs = zs + a.z[q];
    zs = block_sum1(zs, sh, parity);
    double m_z = a.projected ? zs / nd : 0.0;
    double rz = 0.0;
    for (int q = tid; q < n; q += kThreads) {
      const double zq = a.z[q] - m_z;
      rz = rz + (a.r[q] - m_r) * zq;
      p_raw[q] = zq;
    }
    rz = block_sum1(rz, sh, parity);
    if (!isfinite(rz) || rz <= 0.0) {
      if (tid == 0) sh.status = kRz0Fail;
      return;
    }
    const int limit = a.max_iterations < 1 ? 1 : a.max_iterations;
    for (int it = 0; it < limit; ++it) {
      // Ap and p^T A p
      double pap = 0.0;
      for (int q = tid; q < n; q += kThreads) {
        const uint32_t c = a.cell[q];
        const double apq = spmv_row(a, sh, p_raw, m_p, q, static_cast<int>(c >> 16), static_cast<int>(c & 0xffffu));
        const double pq = p_raw[q] - m_p;
        a.y[q] = apq;  // Ap shares storage with the sweep intermediate
        p_prj[q] = pq;
        pap = pap + pq * apq;
      }
      pap = block_sum1(pap, sh, parity);
      if (!isfinite(pap) || fabs(pap) < 1.0e-30) {
        if (tid == 0) sh.status = kPapFail;
        return;
      }
      const double alpha = rz / pap;
      // x += alpha p, r -= alpha Ap (on projected x, r), with the update / iterate norms
      double up[4] = {0.0, 0.0, 0.0, 0.0};
      for (int q = tid; q < n; q += kThreads) {
        const double delta = alpha * p_prj[q];
        const double xq = (a.x[q] - m_x) + delta;
        const double rq = (a.r[q] - m_r) - alpha * a.y[q];
        a.x[q] = xq;
        a.r[q] = rq;
        up[0] = up[0] + xq;
        up[1] = up[1] + rq;
        up[2] = up[2] + delta * delta;
        up[3] = up[3] + xq * xq;
      }
      block_sum<4>(up, sh, parity);
      m_x = a.projected ? up[0] / nd : 0.0;
      m_r = a.projected ? up[1] / nd : 0.0;
      rr = 0.0;
      for (int q = tid; q < n; q += kThreads) {
        const double t = a.r[q] - m_r;
        rr = rr + t * t;
      }
      rr = block_sum1(rr, sh, parity);
      iterations = it + 1;
      iter_res = sqrt(up[2] / (up[3] < 1.0e-30 ? 1.0e-30 : up[3]));
      abs_res = sqrt(rr < 0.0 ? 0.0 : rr);
      rel_res = abs_res / (bnorm < 1.0e-30 ? 1.0e-30 : bnorm);
      if (rel_res < tol || abs_res < tol) break;

      precondition<kIldlt>(a, sh, m_r, floor);
      zs = 0.0;
      for (int q = tid; q < n; q += kThreads) zs = zs + a.z[q];
      zs = block_sum1(zs, sh, parity);
      m_z = a.projected ? zs / nd : 0.0;
      double rzn = 0.0;
      for (int q = tid; q < n; q += kThreads) rzn = rzn + (a.r[q] - m_r) * (a.z[q] - m_z);
      rzn = block_sum1(rzn, sh, parity);
      if (!isfinite(rzn) || rzn <= 0.0) {
        if (tid == 0) sh.status = kRzFail;
        return;
      }
      const double beta = rzn / rz;
      double ps = 0.0;
      for (int q = tid; q < n; q += kThreads) {
        const double pn = (a.z[q] - m_z) + beta * p_prj[q];
        p_raw[q] = pn;
        ps = ps + pn;
      }
      ps = block_sum1(ps, sh, parity);
      m_p = a.projected ? ps / nd : 0.0;
      rz = rzn;
    }
  }
  if (sh.spin_error) {
    if (tid == 0) sh.status = kSweepFail;
    return;
  }

  // scatter the projected x, subtract the interior mean when the system is singular
  double xs = 0.0;
  for (int q = tid; q < n; q += kThreads) xs = xs + (a.x[q] - m_x);
  xs = block_sum1(xs, sh, parity);
  const double mean = a.projected ? xs / static_cast<double>(nx * ny) : 0.0;
  for (int q = tid; q < n; q += kThreads) {
    const uint32_t c = a.cell[q];
    const int i = static_cast<int>(c >> 16), j = static_cast<int>(c & 0xffffu);
    const double xq = a.x[q] - m_x;
    a.p_out[i * ny + j] = a.projected ? xq - mean : xq;
  }
  if (tid == 0) {
    a.rep->iterations = iterations;
    a.rep->outer_cycles = 1;
    a.rep->rel = rel_res;
    a.rep->abs = abs_res;
    a.rep->iter_res = iter_res;
  }
}

}  // namespace

template <bool kIldlt>
__global__ void __launch_bounds__(kThreads) fp64_solve_kernel(const SysArgs *systems) {
  __shared__ SweepShared sh;
  const SysArgs &a = systems[blockIdx.x];
  solve_system<kIldlt>(a, sh);
  __syncthreads();
  if (threadIdx.x == 0) a.rep->status = sh.status;
}

hipError_t launch_fp64_solve(const SysArgs *d_systems, int count, bool ildlt, hipStream_t stream) {
  if (ildlt)
    hipLaunchKernelGGL(fp64_solve_kernel<true>, dim3(count), dim3(kThreads), 0, stream, d_systems);
  else
    hipLaunchKernelGGL(fp64_solve_kernel<false>, dim3(count), dim3(kThreads), 0, stream, d_systems);
  return hipGetLastError();
}

}  // namespace pp
