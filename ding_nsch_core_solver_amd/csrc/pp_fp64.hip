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

// Coefficient mode: inverse density of x-face i / y-face j; along a periodic direction face nx
// (ny) is face 0.
__device__ __forceinline__ double k_xface(const SysArgs &a, int i, int j) {
  if (a.periodic_x && i == a.nx) i = 0;
  return a.ku[i * a.ny + j];
}

__device__ __forceinline__ double k_yface(const SysArgs &a, int i, int j) {
  if (a.periodic_y && j == a.ny) j = 0;
  return a.kv[i * (a.ny + 1) + j];
}

// One row of pressure.cpp:133-195 in the wavefront layout. Returns the diagonal.
__device__ void assemble_cell(const SysArgs &a, int q, int i, int j) {
  const bool has_e = a.periodic_x || i < a.nx - 1;
  const bool has_w = a.periodic_x || i > 0;
  const bool has_n = a.periodic_y || j < a.ny - 1;
  const bool has_s = a.periodic_y || j > 0;
  const double dx = a.dx, dy = a.dy;
  double ce, cw, cn, cs;
  if (a.coeff_mode) {
    ce = k_xface(a, i + 1, j) / (dx * dx);
    cw = k_xface(a, i, j) / (dx * dx);
    cn = k_yface(a, i, j + 1) / (dy * dy);
    cs = k_yface(a, i, j) / (dy * dy);
  } else if (a.split) {
    ce = cw = 1.0 / (dx * dx);
    cn = cs = 1.0 / (dy * dy);
  } else {
    ce = 1.0 / (rho_xface(a, i + 1, j) * dx * dx);
    cw = 1.0 / (rho_xface(a, i, j) * dx * dx);
    cn = 1.0 / (rho_yface(a, i, j + 1) * dy * dy);
    cs = 1.0 / (rho_yface(a, i, j) * dy * dy);
  }
  double diag = 0.0;
  double b = a.coeff_mode ? a.src[i * a.ny + j]
             : a.split    ? -(split_explicit(a, i, j) + a.rho_ref * divergence(a, i, j) / a.dt)
                          : -divergence(a, i, j) / a.dt;
  const bool rhs_bc = !a.coeff_mode || a.rhs_bc;
  auto boundary = [&](int type, double value, double coeff, double spacing) {
    if (type == 0) {  // unset edges are homogeneous Neumann (operators.cpp:36-60)
      type = 2;
      value = 0.0;
    }
    if (type == 1) {
      diag += 2.0 * coeff;
      if (rhs_bc) b += 2.0 * coeff * value;
    } else if (rhs_bc) {
      b += coeff * value * spacing;
    }
  };
  if (has_e) diag += ce; else boundary(a.bc_type[1], a.bc_value[1], ce, dx);
  if (has_w) diag += cw; else boundary(a.bc_type[0], a.bc_value[0], cw, dx);
  if (has_n) diag += cn; else boundary(a.bc_type[3], a.bc_value[3], cn, dy);
  if (has_s) diag += cs; else boundary(a.bc_type[2], a.bc_value[2], cs, dy);
  a.aP[q] = 0.0 + diag;
  a.aW[q] = has_w ? drop_tiny(0.0 + -cw) : 0.0;
  a.aS[q] = has_s ? drop_tiny(0.0 + -cs) : 0.0;
  a.b[q] = b;
}

// ---- wavefront sweeps --------------------------------------------------------------------------
//
// Wave w (w < S) owns rows 64w .. 64w+63 and walks its levels d = 64w .. 64w+63+ny-1 (ascending in
// the forward direction). Per level the position of every lane is base(d) + i with a wave-uniform
// base, operands of level d+P are loaded while level d is computed (register ring of P slots), the
// west/east neighbour comes from the adjacent lane by DPP and, for the boundary lane, from the
// adjacent wave in batches of P hand-off values. Lanes outside the grid compute on padding values;
// only their stores are masked, and their results only ever feed other inactive lanes.
constexpr int kPrefetch = 8;

struct FactorOps {
  double ap, aw, as;
};

template <bool kIldlt>
__device__ __forceinline__ void factor_sweep(const SysArgs &a, SweepShared &sh, double floor) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int S = (nx + 63) >> 6;
  if (wave >= S) return;
  const int i = wave * 64 + lane;
  const int d_beg = wave * 64, d_end = min(wave * 64 + 63, nx - 1) + ny - 1, d_last = nx + ny - 2;
  const bool px = a.periodic_x != 0, py = a.periodic_y != 0;
  const bool row_ok = i < nx, has_w = i > 0, wrap_row = px && i == nx - 1, wrap_src = px && i == 0;
  const global_ptr<const double> aP = as_global<const double>(a.aP);
  const global_ptr<const double> aW = as_global<const double>(a.aW);
  const global_ptr<const double> aS = as_global<const double>(a.aS);
  const global_ptr<double> Lw = as_global(a.Lw);
  const global_ptr<double> Ls = as_global(a.Ls);
  const global_ptr<double> dg = as_global(a.dg);
  const global_ptr<double> LX = as_global(a.LX);
  const global_ptr<double> LY = as_global(a.LY);
  const double floor_sq = floor * floor;
  const double ay = (py && row_ok) ? aS[level_base(i, nx, ny) + i] : 0.0;  // entry (i,ny-1)->(i,0)
  double prev = 0.0, first = 0.0;
  bool failed = false;
  FactorOps buf[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {
    const int q = level_base(min(d_beg + k, d_last), nx, ny) + i;
    buf[k] = {aP[q], aW[q], aS[q]};
  }
  for (int d0 = d_beg; d0 <= d_end; d0 += kPrefetch) {
    const double hv = wave > 0 ? receive_batch<double, kPrefetch>(wave, lane, d0 - wave * 64, ny, sh) : 0.0;
    if (wave + 1 < S) reserve_batch(wave + 1, min(d0 + kPrefetch - 1 - (wave * 64 + 63), ny - 1), sh);
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const int d = d0 + k;
      const FactorOps o = buf[k];
      {
        const int qp = level_base(min(d + kPrefetch, d_last), nx, ny) + i;
        buf[k] = {aP[qp], aW[qp], aS[qp]};
      }
      if (d <= d_end) {
        const int j = d - i;
        const bool act = row_ok && j >= 0 && j < ny;
        double west = wave_shr1(prev);
        if (wave > 0) {
          const double h = lane_value(hv, k);
          west = lane == 0 ? h : west;
        }
        double dd = o.ap;
        if (wrap_row && act) {  // entry (nx-1, j) -> (0, j), first in column order
          const double dc = *lds(&sh.wrap[j]);
          const double l = aW[level_base(j, nx, ny)] / floor_max(dc, floor);
          LX[j] = l;
          dd = kIldlt ? dd - (l * l) * dc : dd - l * l;
        }
        const double lw = o.aw / floor_max(west, floor);
        dd = has_w ? (kIldlt ? dd - (lw * lw) * west : dd - lw * lw) : dd;
        if (py && j == ny - 1) {  // entry (i, ny-1) -> (i, 0)
          const double l = ay / floor_max(first, floor);
          if (act) LY[i] = l;
          dd = kIldlt ? dd - (l * l) * first : dd - l * l;
        }
        const double ls = o.as / floor_max(prev, floor);
        dd = j > 0 ? (kIldlt ? dd - (ls * ls) * prev : dd - ls * ls) : dd;
        double cur;
        if (kIldlt) {
          failed = failed || (act && (!isfinite(dd) || dd <= floor));
          cur = dd;
        } else {
          failed = failed || (act && (!isfinite(dd) || dd <= floor_sq));
          cur = sqrt(dd);
        }
        if (act) {
          const int q = level_base(d, nx, ny) + i;
          Lw[q] = has_w ? lw : 0.0;
          Ls[q] = j > 0 ? ls : 0.0;
          dg[q] = cur;
        }
        first = j == 0 ? cur : first;
        if (wrap_src && act) *lds(&sh.wrap[j]) = cur;
        prev = cur;
        if (wave + 1 < S) publish(wave + 1, lane, 63, d - (wave * 64 + 63), ny, cur, sh);
      }
    }
  }
  if (failed) *lds(&sh.status) = kFactorFail;
}

// ---- triangular sweeps ---------------------------------------------------------------------------
//
// The sweeps run in the element type T of the solve. double: the FP64 path, dividing by the floored
// pivot with IEEE division exactly as the reference. float: the inner solves of mixed precision,
// multiplying by the reciprocal of the floored pivot (computed in FP64, rounded once).

template <typename T>
struct SweepArrays {
  const T *in;        // forward: r (read as r - m_r); backward: y
  const T *Lw, *Ls;   // lower entries west / south
  const T *dg;        // double: pivot; float: reciprocal of the floored pivot
  T *out;             // forward: y; backward: z
  const double *LY;   // periodic-y wrap entries (FP64 factor)
};

// s - c * x: separately rounded in FP64 (the reference's arithmetic), fused in FP32.
template <typename T>
__device__ __forceinline__ T sub_mul(T s, T c, T x) {
  if constexpr (sizeof(T) == 8)
    return s - c * x;
  else
    return __builtin_fmaf(-c, x, s);
}

template <typename T>
__device__ __forceinline__ T pivot_apply(T s, T d, T floor) {
  if constexpr (sizeof(T) == 8)
    return s / floor_max(d, floor);
  else
    return s * d;
}

template <typename T>
struct ForwardOps {
  T r, lw, ls, dg;
};

// Sweep loop shape shared by the solves: whole chunks of kPrefetch levels (levels past the wave's
// range are inactive for every lane), every vector-memory operation unconditional (inactive lanes
// store into the padding), so the compiler can keep the prefetch ring in flight across levels.
__device__ __forceinline__ int chunk_end(int d_beg, int d_end) {
  return d_beg + ((d_end - d_beg + kPrefetch) / kPrefetch) * kPrefetch - 1;
}

// y = L^{-1} (r - m_r) (ICC: divided by the pivot; ILDLT: unit lower, stored divided by D).
// Lower entries in column order: periodic-x wrap (0, j), west (i-1, j), periodic-y wrap (i, 0),
// south (i, j-1).
template <typename T, bool kIldlt>
__device__ __forceinline__ void forward_solve(const SysArgs &a, const SweepArrays<T> &io, SweepShared &sh, T m_r,
                                              T floor) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int S = (nx + 63) >> 6;
  if (wave >= S) return;
  const int i = wave * 64 + lane;
  const int d_beg = wave * 64, d_end = min(wave * 64 + 63, nx - 1) + ny - 1, d_last = nx + ny - 2;
  const int d_stop = chunk_end(d_beg, d_end);
  const bool px = a.periodic_x != 0, py = a.periodic_y != 0;
  const bool row_ok = i < nx, has_w = i > 0, wrap_row = px && i == nx - 1, wrap_src = px && i == 0;
  const global_ptr<const T> r = as_global<const T>(io.in);
  const global_ptr<const T> Lw = as_global<const T>(io.Lw);
  const global_ptr<const T> Ls = as_global<const T>(io.Ls);
  const global_ptr<const T> dg = as_global<const T>(io.dg);
  const global_ptr<T> y = as_global(io.out);
  const int sink = a.n + kPad - 64 + lane;  // store target of inactive lanes
  const T ly = (py && row_ok) ? static_cast<T>(as_global<const double>(io.LY)[i]) : T(0);
  T prev = T(0), first = T(0);
  ForwardOps<T> buf[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {
    const int q = level_base(min(d_beg + k, d_last), nx, ny) + i;
    buf[k] = {r[q], Lw[q], Ls[q], dg[q]};
  }
  for (int d0 = d_beg; d0 <= d_stop; d0 += kPrefetch) {
    const T hv = wave > 0 ? receive_batch<T, kPrefetch>(wave, lane, d0 - wave * 64, ny, sh) : T(0);
    if (wave + 1 < S) reserve_batch(wave + 1, min(d0 + kPrefetch - 1 - (wave * 64 + 63), ny - 1), sh);
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const int d = d0 + k;
      const ForwardOps<T> o = buf[k];
      {
        const int qp = level_base(min(d + kPrefetch, d_last), nx, ny) + i;
        buf[k] = {r[qp], Lw[qp], Ls[qp], dg[qp]};
      }
      const int j = d - i;
      const bool act = row_ok && j >= 0 && j < ny;
      T west = wave_shr1(prev);
      if (wave > 0) {
        const T h = lane_value(hv, k);
        west = lane == 0 ? h : west;
      }
      T s = o.r - m_r;
      if (px) {
        const int jc = min(max(j, 0), ny - 1);
        const T wx = static_cast<T>(*lds(&sh.wrap[jc]));
        s = (wrap_row && act) ? s - static_cast<T>(sh.lx[jc]) * wx : s;
      }
      s = has_w ? sub_mul(s, o.lw, west) : s;  // the staged sweeps' arithmetic (fused in FP32)
      if (py) s = j == ny - 1 ? s - ly * first : s;
      s = j > 0 ? sub_mul(s, o.ls, prev) : s;
      T cur, out;
      if (kIldlt) {
        cur = s;
        out = pivot_apply(s, o.dg, floor);
      } else {
        cur = pivot_apply(s, o.dg, floor);
        out = cur;
      }
      y[act ? level_base(d, nx, ny) + i : sink] = out;
      first = j == 0 ? cur : first;
      if (px && wrap_src && act) *lds(&sh.wrap[j]) = static_cast<double>(cur);
      prev = cur;
      if (wave + 1 < S) publish(wave + 1, lane, 63, d - (wave * 64 + 63), ny, cur, sh);
    }
  }
}

template <typename T>
struct BackwardOps {
  T y, dg, ls_n, lw_e;
};

// z = L^{-T} y, levels descending. Children of (i, j) in ascending row order: (i, j+1), the
// periodic-y wrap (i, ny-1) when j = 0, (i+1, j), the periodic-x wrap (nx-1, j) when i = 0.
template <typename T, bool kIldlt>
__device__ __forceinline__ void backward_solve(const SysArgs &a, const SweepArrays<T> &io, SweepShared &sh, T floor) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int S = (nx + 63) >> 6;
  if (wave >= S) return;
  const int i = wave * 64 + lane;
  const int d_beg = wave * 64, d_end = min(wave * 64 + 63, nx - 1) + ny - 1, d_last = nx + ny - 2;
  const bool px = a.periodic_x != 0, py = a.periodic_y != 0;
  const bool row_ok = i < nx, has_e = i + 1 < nx, wrap_row = px && i == nx - 1, wrap_src = px && i == 0;
  const global_ptr<const T> yv = as_global<const T>(io.in);
  const global_ptr<const T> Lw = as_global<const T>(io.Lw);
  const global_ptr<const T> Ls = as_global<const T>(io.Ls);
  const global_ptr<const T> dg = as_global<const T>(io.dg);
  const global_ptr<T> z = as_global(io.out);
  const int sink = a.n + kPad - 64 + lane;
  const T ly = (py && row_ok) ? static_cast<T>(as_global<const double>(io.LY)[i]) : T(0);
  auto fetch = [&](int d) {
    const int dc = min(max(d, 0), d_last);
    const int q = level_base(dc, nx, ny) + i;
    const int qn = level_base(min(dc + 1, d_last), nx, ny) + i;  // (i, j+1); (i+1, j) follows it
    return BackwardOps<T>{yv[q], dg[q], Ls[qn], Lw[qn + 1]};
  };
  // levels d_end down to d_low (whole chunks; levels below d_beg are inactive for every lane)
  const int d_low = d_end - ((d_end - d_beg + kPrefetch) / kPrefetch) * kPrefetch + 1;
  T prev = T(0), first = T(0);
  BackwardOps<T> buf[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) buf[k] = fetch(d_end - k);
  for (int d0 = d_end; d0 >= d_low; d0 -= kPrefetch) {
    const int j63_first = d0 - (wave * 64 + 63);  // column of lane 63 at level d0
    const T hv = wave + 1 < S ? receive_batch<T, kPrefetch>(wave, lane, ny - 1 - j63_first, ny, sh) : T(0);
    if (wave > 0) reserve_batch(wave - 1, min(ny - 1 - (d0 - kPrefetch + 1 - wave * 64), ny - 1), sh);
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const int d = d0 - k;
      const BackwardOps<T> o = buf[k];
      buf[k] = fetch(d - kPrefetch);
      const int j = d - i;
      const bool act = row_ok && j >= 0 && j < ny;
      T east = wave_shl1(prev);
      if (wave + 1 < S) {
        const T h = lane_value(hv, k);
        east = lane == 63 ? h : east;
      }
      T s = o.y;
      s = j + 1 < ny ? sub_mul(s, o.ls_n, prev) : s;  // the staged sweeps' arithmetic (fused in FP32)
      if (py) s = j == 0 ? s - ly * first : s;
      s = has_e ? sub_mul(s, o.lw_e, east) : s;
      if (px) {
        const int jc = min(max(j, 0), ny - 1);
        const T wx = static_cast<T>(*lds(&sh.wrap[jc]));
        s = (wrap_src && act) ? s - static_cast<T>(sh.lx[jc]) * wx : s;
      }
      const T cur = kIldlt ? s : pivot_apply(s, o.dg, floor);
      z[act ? level_base(d, nx, ny) + i : sink] = cur;
      first = j == ny - 1 ? cur : first;
      if (px && wrap_row && act) *lds(&sh.wrap[j]) = static_cast<double>(cur);
      prev = cur;
      if (wave > 0) publish(wave - 1, lane, 0, ny - 1 - (d - wave * 64), ny, cur, sh);
    }
  }
}

// ---- LDS-staged sweeps ---------------------------------------------------------------------------
//
// With S <= kMaxStagedWaves sweep waves, the remaining waves of the block are loaders: they copy the
// four operands of every (sweep wave, level) into an LDS ring, kStageUnit levels at a time, so the
// sweep waves never wait on global memory. Ring slot of step t is t mod R; ready[w][slot] = t + 1
// once filled (flag written after the data, LDS operations of a wave complete in order), used[w] =
// steps sweep wave w has finished, which bounds how far loaders may run ahead.

// Ring depth R (levels per sweep wave, a power of two, S * R slots of 4 x 64 elements within the
// 128 KB stage) and loader unit U: loaders may run R - U levels ahead of the sweep. FP64 rings are
// tunable (SysArgs::ring); FP32 slots are half the size, so their default is twice as deep.
template <typename T>
__device__ __forceinline__ int stage_ring(const SysArgs &a, int S) {
  if constexpr (sizeof(T) == 8)
    return a.ring > 0 ? a.ring : (S <= 4 ? 16 : 8);
  else
    return S <= 2 ? 64 : (S <= 4 ? 32 : 16);
}
__device__ __forceinline__ int stage_unit(int R) { return R / 2 < kStageUnit ? R / 2 : kStageUnit; }

// Steps every sweep wave walks: its 64 rows span 63 + ny levels, rounded up to whole units.
__device__ __forceinline__ int stage_steps(int ny) { return ((63 + ny + kStageUnit - 1) / kStageUnit) * kStageUnit; }

template <bool kBackward>
__device__ __forceinline__ int step_level(int w, int t, int T) {
  return kBackward ? w * 64 + T - 1 - t : w * 64 + t;
}

template <typename T>
__device__ __forceinline__ volatile __attribute__((address_space(3))) T *stage_slot(SweepShared &sh, int w, int R,
                                                                                   int slot, int lane) {
  return lds(reinterpret_cast<volatile T *>(&sh.stage[0]) + ((w * R + slot) * 4) * 64 + lane);
}

// Loaders take units (sweep wave w, first step t0) in order of when they are needed: a forward
// sweep wave w starts 64 w steps after wave 0, a backward one 64 (S-1-w) steps after wave S-1, so the
// key is that offset plus t0. Everything a unit waits for (ring space of its wave, which in turn needs
// hand-offs from the upstream wave) has a smaller key, so loaders processing their share of units in
// key order cannot wait on each other in a cycle. Unit k of that order goes to loader k mod L.
// Staged operand values: the forward input is r - m_r and, in FP64, the pivot is already floored,
// so the sweep waves do neither (the same operations, moved off the sweep's chain).
template <typename T>
__device__ __forceinline__ T staged_pivot(T d, T floor) {
  if constexpr (sizeof(T) == 8)
    return floor_max(d, floor);
  else
    return d;
}

template <typename T>
__device__ __forceinline__ T staged_apply(T s, T d) {
  if constexpr (sizeof(T) == 8)
    return s / d;
  else
    return s * d;
}

// Walks the units in key order and returns the next one owned by loader `me` of L.
struct UnitCursor {
  int tau, w, k_unit;
};

template <bool kBackward>
__device__ __forceinline__ bool next_unit(UnitCursor &c, int S, int Tn, int U, int L, int me, int &w_out, int &t0_out) {
  const int tau_end = 64 * (S - 1) + Tn;
  while (c.tau < tau_end) {
    const int w = c.w;
    const int t0 = c.tau - 64 * (kBackward ? (S - 1 - w) : w);
    if (++c.w == S) {
      c.w = 0;
      c.tau += U;
    }
    if (t0 < 0 || t0 >= Tn) continue;
    if ((c.k_unit++) % L != me) continue;
    w_out = w;
    t0_out = t0;
    return true;
  }
  return false;
}

// Loaders are software-pipelined over two register buffers: the global loads of a loader's next
// unit are in flight while it waits for ring space and writes the current one to LDS. Only the LDS
// writes are gated by ring space, in the same order as before, so the deadlock argument is unchanged.
// The unit size is a template constant and the next unit's loads are issued unconditionally (the
// last step reloads the current unit, which is never written), so the code between a unit's loads
// and its LDS writes has no branch that holds a load: the compiler then waits only for that unit's
// loads (vmcnt = the other unit's count) instead of draining the queue before every level.
// Level bases are read without clamping: the level-base table's margin (0 outside the grid) covers
// every level a sweep wave steps through, and the positions it yields stay below n + kPad.
template <typename T, bool kBackward, int U>
__device__ __forceinline__ void stage_loader_units(const SysArgs &a, const SweepArrays<T> &io, SweepShared &sh, int S,
                                                   T m_r, T floor) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ny = a.ny;
  const int R = stage_ring<T>(a, S), Tn = stage_steps(ny);
  const int L = kWaves - S, me = wave - S;
  const global_ptr<const T> f0 = as_global<const T>(io.in);
  const global_ptr<const T> dg = as_global<const T>(io.dg);
  const global_ptr<const T> Lw = as_global<const T>(io.Lw);
  const global_ptr<const T> Ls = as_global<const T>(io.Ls);
  const __attribute__((address_space(4))) int *lb = (const __attribute__((address_space(4))) int *)a.lev_base;
  auto load = [&](int w, int t0, T (&v)[U][4]) {
    const int i = w * 64 + lane;
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int d = step_level<kBackward>(w, t0 + k, Tn);
      const unsigned q = static_cast<unsigned>(lb[d] + i);
      if (kBackward) {
        const unsigned qn = static_cast<unsigned>(lb[d + 1] + i);  // (i, j+1); (i+1, j) follows it
        v[k][0] = f0[q];
        v[k][1] = dg[q];
        v[k][2] = Ls[qn];
        v[k][3] = Lw[qn + 1];
      } else {
        v[k][0] = f0[q];
        v[k][1] = Lw[q];
        v[k][2] = Ls[q];
        v[k][3] = dg[q];
      }
    }
  };
  auto store = [&](int w, int t0, const T (&v)[U][4]) {
    wait_at_least(&sh.used[w], t0 + U - R, sh);
#pragma unroll
    for (int k = 0; k < U; ++k) {
      volatile __attribute__((address_space(3))) T *slot = stage_slot<T>(sh, w, R, (t0 + k) & (R - 1), lane);
      if (kBackward) {
        slot[0] = v[k][0];
        slot[64] = staged_pivot(v[k][1], floor);
        slot[128] = v[k][2];
        slot[192] = v[k][3];
      } else {
        slot[0] = v[k][0] - m_r;
        slot[64] = v[k][1];
        slot[128] = v[k][2];
        slot[192] = staged_pivot(v[k][3], floor);
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < U; ++k) *lds(&sh.ready[w * kMaxRing + ((t0 + k) & (R - 1))]) = t0 + k + 1;
    }
  };
  // With at least as many loaders as sweep waves, each loader serves one wave only (its units in
  // step order, interleaved with the wave's other loaders), so a loader waiting for ring space never
  // holds back another wave; spare loaders go to the waves that start first (0 forward, S-1
  // backward): the head wave sets the pace of the whole pipeline, since every later wave runs a fixed
  // lag behind it (measured: 192x192 backward sweep 14.4 -> 13.7 ms per solve against giving them to
  // the waves that finish the sweep). Otherwise loaders take all units in the order of need.
  const bool dedicated = S <= L;
  const int rank = me % S;
  const int w_own = kBackward ? S - 1 - rank : rank;
  const int n_own = (L - 1 - rank) / S + 1, k_own = me / S;
  int m_own = 0;
  UnitCursor cur{0, 0, 0};
  auto next = [&](int &w_out, int &t_out) -> bool {
    if (dedicated) {
      const int t0 = U * (k_own + n_own * m_own);
      if (t0 >= Tn) return false;
      ++m_own;
      w_out = w_own;
      t_out = t0;
      return true;
    }
    return next_unit<kBackward>(cur, S, Tn, U, L, me, w_out, t_out);
  };
  int wA = 0, tA = 0, wB = 0, tB = 0;
  T bufA[U][4], bufB[U][4];
  if (!next(wA, tA)) return;
  load(wA, tA, bufA);
  for (;;) {
    const bool hasB = next(wB, tB);
    if (!hasB) {
      wB = wA;
      tB = tA;
    }
    load(wB, tB, bufB);
    store(wA, tA, bufA);
    if (!hasB) break;
    const bool hasA = next(wA, tA);
    if (!hasA) {
      wA = wB;
      tA = tB;
    }
    load(wA, tA, bufA);
    store(wB, tB, bufB);
    if (!hasA) break;
  }
}

template <typename T, bool kBackward>
__device__ __forceinline__ void stage_loader(const SysArgs &a, const SweepArrays<T> &io, SweepShared &sh, int S, T m_r,
                                             T floor) {
  if (stage_unit(stage_ring<T>(a, S)) == kStageUnit)
    stage_loader_units<T, kBackward, kStageUnit>(a, io, sh, S, m_r, floor);
  else
    stage_loader_units<T, kBackward, kStageUnit / 2>(a, io, sh, S, m_r, floor);
}

// Sweep side of the staging: per chunk of kPrefetch levels, wait once for the chunk's units, read
// all of its operands from LDS, then run the levels with no further LDS waits except the hand-off
// batch. used[w] advances once per chunk.
__device__ __forceinline__ void await_chunk(SweepShared &sh, int w, int R, int t0) {
  const int U = stage_unit(R);
  for (int u = t0; u < t0 + kPrefetch; u += U) {
    const int last = u + U - 1;
    wait_at_least(&sh.ready[w * kMaxRing + (last & (R - 1))], last + 1, sh);
  }
}

// The chunk's units are staged and (when the wave publishes) the receiving wave has consumed enough
// of its hand-off ring, read in one LDS round trip. Counters only grow, so a true answer stays true
// and both waits may be skipped.
__device__ __forceinline__ bool staged_and_room(SweepShared &sh, int w, int R, int t0, int send_w, int send_need) {
  const int last1 = t0 + kPrefetch - 1, last0 = t0 + stage_unit(R) - 1;  // one or two units per chunk
  const int r1 = *lds(&sh.ready[w * kMaxRing + (last1 & (R - 1))]);
  const int r0 = *lds(&sh.ready[w * kMaxRing + (last0 & (R - 1))]);
  const int ak = *lds(&sh.ack[send_w < 0 ? 0 : send_w]);
  return (r1 > last1) & (r0 > last0) & ((send_w < 0) | (ak >= send_need));
}

template <typename T>
struct ChunkOps {
  T f[kPrefetch][4];
};

// Plain loads after a compiler-only fence: they stay behind the flag reads and waits written before
// them (LDS executes a wave's operations in order), and the compiler may pair them into
// ds_read2st64_b64 (fields 64 doubles apart).
template <typename T>
__device__ __forceinline__ void read_chunk(SweepShared &sh, int w, int R, int t0, int lane, ChunkOps<T> &c) {
  __asm__ volatile("" ::: "memory");
  const __attribute__((address_space(3))) T *st =
      (const __attribute__((address_space(3))) T *)(reinterpret_cast<T *>(&sh.stage[0]));
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {
    const __attribute__((address_space(3))) T *p = st + ((w * R + ((t0 + k) & (R - 1))) * 4) * 64 + lane;
#pragma unroll
    for (int f = 0; f < 4; ++f) c.f[k][f] = p[f * 64];
  }
  __asm__ volatile("" ::: "memory");
}

template <typename T, bool kIldlt, bool kPeriodic>
__device__ __forceinline__ void staged_forward(const SysArgs &a, const SweepArrays<T> &io, SweepShared &sh, T m_r,
                                               T floor, int S) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int R = stage_ring<T>(a, S), Tn = stage_steps(ny);
  const int i = wave * 64 + lane;
  const bool px = kPeriodic && a.periodic_x != 0, py = kPeriodic && a.periodic_y != 0;
  const bool row_ok = i < nx, wrap_row = px && i == nx - 1, wrap_src = px && i == 0;
  const global_ptr<T> y = as_global(io.out);
  const __attribute__((address_space(4))) int *lev_base = (const __attribute__((address_space(4))) int *)a.lev_base;
  const int sink = a.n + kPad - 64 + lane;
  const T ly = (py && row_ok) ? static_cast<T>(as_global<const double>(io.LY)[i]) : T(0);
  T prev = T(0), first = T(0);
  for (int t0 = 0; t0 < Tn; t0 += kPrefetch) {
    const int d0 = wave * 64 + t0;
    const int send_last = min(d0 + kPrefetch - 1 - (wave * 64 + 63), ny - 1);
    const bool send = wave + 1 < S && send_last >= 0;
    const bool ready = staged_and_room(sh, wave, R, t0, send ? wave + 1 : -1, send_last + 1 - kRing);
    if (!ready) await_chunk(sh, wave, R, t0);
    ChunkOps<T> c;
    read_chunk<T>(sh, wave, R, t0, lane, c);
    T hk[kPrefetch];  // west inputs of lane 0 (row 64w-1 of wave w-1) at the chunk's levels
    if (wave > 0) {
      receive_chunk<T, kPrefetch>(wave, t0, ny, sh, hk);
    } else {
#pragma unroll
      for (int k = 0; k < kPrefetch; ++k) hk[k] = T(0);
    }
    if (!ready && send) reserve_batch(wave + 1, send_last, sh);
    int base[kPrefetch];  // level bases of the chunk, scalar loads from the constant table
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) base[k] = lev_base[d0 + k];
    T pub = T(0);  // lane k: this wave's row 64w+63 value at level d0 + k (wave w+1's west input)
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const int d = d0 + k;
      const int j = d - i;
      const bool act = row_ok && static_cast<unsigned>(j) < static_cast<unsigned>(ny);
      // west neighbour: lane l-1 of this wave, lane 0 from wave w-1. Row 0 (lane 0 of wave 0) has no
      // west entry: its Lw is +0 and its west is +0, so s - Lw * west = s exactly, no select needed.
      const T west = shr1_fill(prev, hk[k]);
      T s = c.f[k][0];  // r - m_r
      if (kPeriodic && px) {
        const int jc = min(max(j, 0), ny - 1);
        const T wx = static_cast<T>(*lds(&sh.wrap[jc]));
        s = (wrap_row && act) ? s - static_cast<T>(sh.lx[jc]) * wx : s;
      }
      s = sub_mul(s, c.f[k][1], west);
      if (kPeriodic && py) s = j == ny - 1 ? s - ly * first : s;
      s = j > 0 ? sub_mul(s, c.f[k][2], prev) : s;
      T cur, out;
      if (kIldlt) {
        cur = s;
        out = staged_apply(s, c.f[k][3]);
      } else {
        cur = staged_apply(s, c.f[k][3]);
        out = cur;
      }
      y[static_cast<unsigned>(act ? base[k] + i : sink)] = out;
      if (kPeriodic) {
        first = j == 0 ? cur : first;
        if (px && wrap_src && act) *lds(&sh.wrap[j]) = static_cast<double>(cur);
      }
      prev = cur;
      pub = write_lane(pub, lane_value(cur, 63), k);
    }
    if (wave + 1 < S) publish_chunk<kPrefetch>(wave + 1, lane, d0 - (wave * 64 + 63), ny, pub, sh);
    if (lane == 0) *lds(&sh.used[wave]) = t0 + kPrefetch;
  }
}

template <typename T, bool kIldlt, bool kPeriodic>
__device__ __forceinline__ void staged_backward(const SysArgs &a, const SweepArrays<T> &io, SweepShared &sh, T floor,
                                                int S) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny;
  const int R = stage_ring<T>(a, S), Tn = stage_steps(ny);
  const int i = wave * 64 + lane;
  const bool px = kPeriodic && a.periodic_x != 0, py = kPeriodic && a.periodic_y != 0;
  const bool row_ok = i < nx, has_e = i + 1 < nx, wrap_row = px && i == nx - 1, wrap_src = px && i == 0;
  const global_ptr<T> z = as_global(io.out);
  const __attribute__((address_space(4))) int *lev_base = (const __attribute__((address_space(4))) int *)a.lev_base;
  const int sink = a.n + kPad - 64 + lane;
  const T ly = (py && row_ok) ? static_cast<T>(as_global<const double>(io.LY)[i]) : T(0);
  T prev = T(0), first = T(0);
  for (int t0 = 0; t0 < Tn; t0 += kPrefetch) {
    const int d0 = wave * 64 + Tn - 1 - t0;        // levels descend
    const int j63_first = d0 - (wave * 64 + 63);  // column of lane 63 at level d0
    const int send_last = min(ny - 1 - (d0 - kPrefetch + 1 - wave * 64), ny - 1);
    const bool send = wave > 0 && send_last >= 0;
    const bool ready = staged_and_room(sh, wave, R, t0, send ? wave - 1 : -1, send_last + 1 - kRing);
    if (!ready) await_chunk(sh, wave, R, t0);
    ChunkOps<T> c;
    read_chunk<T>(sh, wave, R, t0, lane, c);
    T hk[kPrefetch];  // east inputs of lane 63 (row 64w+64 of wave w+1) at the chunk's levels
    if (wave + 1 < S) {
      receive_chunk<T, kPrefetch>(wave, ny - 1 - j63_first, ny, sh, hk);
    } else {
#pragma unroll
      for (int k = 0; k < kPrefetch; ++k) hk[k] = T(0);
    }
    if (!ready && send) reserve_batch(wave - 1, send_last, sh);
    int base[kPrefetch];  // level bases of the chunk, scalar loads from the constant table
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) base[k] = lev_base[d0 - k];
    T pub = T(0);  // lane k: this wave's row 64w value at level d0 - k (wave w-1's east input)
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const int d = d0 - k;
      const int j = d - i;
      const bool act = row_ok && static_cast<unsigned>(j) < static_cast<unsigned>(ny);
      // east neighbour: lane l+1 of this wave, lane 63 from wave w+1 (0 for the last wave)
      const T east = shl1_fill(prev, hk[k]);
      T s = c.f[k][0];
      s = j + 1 < ny ? sub_mul(s, c.f[k][2], prev) : s;
      if (kPeriodic && py) s = j == 0 ? s - ly * first : s;
      s = has_e ? sub_mul(s, c.f[k][3], east) : s;
      if (kPeriodic && px) {
        const int jc = min(max(j, 0), ny - 1);
        const T wx = static_cast<T>(*lds(&sh.wrap[jc]));
        s = (wrap_src && act) ? s - static_cast<T>(sh.lx[jc]) * wx : s;
      }
      const T cur = kIldlt ? s : staged_apply(s, c.f[k][1]);
      z[static_cast<unsigned>(act ? base[k] + i : sink)] = cur;
      if (kPeriodic) {
        first = j == ny - 1 ? cur : first;
        if (px && wrap_row && act) *lds(&sh.wrap[j]) = static_cast<double>(cur);
      }
      prev = cur;
      pub = write_lane(pub, lane_value(cur, 0), k);
    }
    if (wave > 0) publish_chunk<kPrefetch>(wave - 1, lane, ny - 1 - (d0 - wave * 64), ny, pub, sh);
    if (lane == 0) *lds(&sh.used[wave]) = t0 + kPrefetch;
  }
}

__device__ __forceinline__ void reset_stage(SweepShared &sh) {
  for (int k = threadIdx.x; k < kWaves * kMaxRing; k += kThreads) *lds(&sh.ready[k]) = 0;
  if (threadIdx.x < kWaves) *lds(&sh.used[threadIdx.x]) = 0;
}

// Operator fields the SpMV needs, read once per solve.
template <typename T>
struct Stencil {
  int nx, ny;
  bool px, py;
  global_ptr<const T> aP, aW, aS;
};

// Level-major SpMV: wave w takes levels w, w + kWaves, w + 2 kWaves, ...; its lanes walk the
// level's cells (contiguous positions base(d) + i), so every neighbour address is a scalar level
// base plus the lane's row and no index array is read. Row sums follow the CSR column order
// (ch_sparse_krylov.hpp:242-248): wrap (0, j), west, wrap (i, 0), south, centre, north,
// wrap (i, ny-1), east, wrap (nx-1, j). epi consumes the row's result and returns its
// contribution to the reduction; thread (w, lane) adds its contributions in (level, row) order,
// which the oracle replays for p^T A p (Reducer::sum_levels).
// One row, branch-free: every operand is loaded from a valid address and an absent entry gets a
// +0 coefficient by a select, so all loads of a row (and of neighbouring rows) can be in flight at
// once. The terms are added in the CSR column order; an absent term adds a signed zero to a sum
// that started at +0.0 and therefore never holds -0, which leaves it bit-identical to skipping it.
template <typename T>
__device__ __forceinline__ T spmv_cell(const Stencil<T> &A, const global_ptr<const T> p, T m_p, int i, int j, int bm,
                                       int b0, int bp, T &pq) {
  const int nx = A.nx, ny = A.ny;
  const int q = b0 + i;
  const bool has_w = i >= 1, has_s = j >= 1, has_n = j <= ny - 2, has_e = i <= nx - 2;
  const int qw = has_w ? bm + i - 1 : q, qs = has_s ? bm + i : q;
  const int qn = bp + i, qe = bp + i + 1;
  const T cw = has_w ? A.aW[q] : T(0), cs = has_s ? A.aS[q] : T(0);
  const T cn = has_n ? A.aS[qn] : T(0), ce = has_e ? A.aW[qe] : T(0);
  const T pw = p[qw], ps = p[qs], pc = p[q], pn = p[qn], pe = p[qe], cp = A.aP[q];
  T s = T(0);
  // wrap terms only for cells inside the grid (the lane may hold a cell past its level)
  const bool in_i = static_cast<unsigned>(i) < static_cast<unsigned>(nx);
  const bool in_j = static_cast<unsigned>(j) < static_cast<unsigned>(ny);
  if (A.px) {  // wrap (0, j) of row (nx-1, j)
    const bool on = i == nx - 1 && in_j;
    const int qx = on ? level_base(j, nx, ny) : q;
    s += (on ? A.aW[qx] : T(0)) * (p[qx] - m_p);
  }
  s += cw * (pw - m_p);
  if (A.py) {  // wrap (i, 0) of row (i, ny-1)
    const bool on = j == ny - 1 && in_i;
    const int qy = on ? level_base(i, nx, ny) + i : q;
    s += (on ? A.aS[qy] : T(0)) * (p[qy] - m_p);
  }
  s += cs * (ps - m_p);
  s += cp * (pc - m_p);
  s += cn * (pn - m_p);
  if (A.py) {  // wrap (i, ny-1) of row (i, 0)
    const bool on = j == 0 && in_i;
    const int qy = on ? level_base(i + ny - 1, nx, ny) + i : q;
    s += (on ? A.aS[q] : T(0)) * (p[qy] - m_p);
  }
  s += ce * (pe - m_p);
  pq = pc - m_p;
  if (A.px) {  // wrap (nx-1, j) of row (0, j)
    const bool on = i == 0 && in_j;
    const int qx = on ? level_base(nx - 1 + j, nx, ny) + nx - 1 : q;
    s += (on ? A.aW[q] : T(0)) * (p[qx] - m_p);
  }
  return s;
}

constexpr int kSpmvChunks = 4;  // 64-cell chunks per level computed before any result is stored

// epi(base, i, s, pq, ex) gets the result s of the row at position q = base + i (base: the level
// base, i: the row), pq = p[q] - m_p and ex = extra[q] (0 when extra is null);
// every load of a pass happens before its epilogues store anything.
template <typename T, typename F>
__device__ __forceinline__ double spmv_levels(const SysArgs &a, const Stencil<T> &A, const global_ptr<const T> p, T m_p,
                                              const global_ptr<const T> extra, F &&epi) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int nx = A.nx, ny = A.ny, nlev = nx + ny - 1;
  const __attribute__((address_space(4))) int *lb = (const __attribute__((address_space(4))) int *)a.lev_base;
  const int max_cells = nx < ny ? nx : ny;
  const int chunks = (max_cells + 63) >> 6;  // chunks per level that can hold cells (uniform)
  double acc = 0.0;
  if (chunks <= kSpmvChunks) {
    // two levels per pass, every chunk of both loaded first; contributions added in level order
    // (an absent cell adds +0.0, exact: the running sum is never -0)
    for (int d0 = wave; d0 < nlev; d0 += 2 * kWaves) {
      T sv[2][kSpmvChunks], pv[2][kSpmvChunks], xv[2][kSpmvChunks];
      int qv[2][kSpmvChunks], bv[2];
      bool ok[2][kSpmvChunks];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int d = d0 + u * kWaves;
        const int lo = level_lo(d, ny), hi = d < nx - 1 ? d : nx - 1;
        const int bm = lb[d - 1], b0 = lb[d], bp = lb[d + 1];
        bv[u] = b0;
#pragma unroll
        for (int c = 0; c < kSpmvChunks; ++c) {
          if (c < chunks) {
            const int i = lo + 64 * c + lane;
            ok[u][c] = d < nlev && i <= hi;
            qv[u][c] = i;
            sv[u][c] = spmv_cell<T>(A, p, m_p, i, d - i, bm, b0, bp, pv[u][c]);  // in-bounds loads even when !ok
            xv[u][c] = extra ? extra[b0 + i] : T(0);
          } else {
            ok[u][c] = false;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int c = 0; c < kSpmvChunks; ++c)
          if (c < chunks) acc = acc + (ok[u][c] ? epi(bv[u], qv[u][c], sv[u][c], pv[u][c], xv[u][c]) : 0.0);
    }
  } else {
    // wide levels: one level at a time, its chunks in groups
    for (int d = wave; d < nlev; d += kWaves) {
      const int lo = level_lo(d, ny), hi = d < nx - 1 ? d : nx - 1;
      const int bm = lb[d - 1], b0 = lb[d], bp = lb[d + 1];
      for (int i0 = lo; i0 <= hi; i0 += 64 * kSpmvChunks) {
        T sv[kSpmvChunks], pv[kSpmvChunks], xv[kSpmvChunks];
        bool ok[kSpmvChunks];
#pragma unroll
        for (int c = 0; c < kSpmvChunks; ++c) {
          const int i = i0 + 64 * c + lane;
          ok[c] = i <= hi;
          sv[c] = spmv_cell<T>(A, p, m_p, i, d - i, bm, b0, bp, pv[c]);  // in-bounds loads even when !ok
          xv[c] = extra ? extra[b0 + i] : T(0);
        }
#pragma unroll
        for (int c = 0; c < kSpmvChunks; ++c)
          acc = acc + (ok[c] ? epi(b0, i0 + 64 * c + lane, sv[c], pv[c], xv[c]) : 0.0);
      }
    }
  }
  return acc;
}

// A work array addressed as (level base, row): raw buffer accesses with the level base as the
// scalar offset and the row as the vector offset, so a lane's address costs one 32-bit value and
// no 64-bit arithmetic. Accesses past n + kPad are discarded (stores) or read 0 (loads) by the
// buffer range check.
template <typename T>
struct LevelArray {
  __amdgpu_buffer_rsrc_t rsrc;
  __device__ __forceinline__ LevelArray(const T *ptr, int n)
      : rsrc(__builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(ptr), 0, static_cast<int>((n + kPad) * sizeof(T)),
                                               0x00020000)) {}
  __device__ __forceinline__ LevelArray(const __attribute__((address_space(1))) T *ptr, int n)
      : LevelArray((const T *)ptr, n) {}
  __device__ __forceinline__ T get(int base, unsigned voff) const {
    if constexpr (sizeof(T) == 8) {
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, base * 8, 0));
    } else {
      return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, base * 4, 0));
    }
  }
  __device__ __forceinline__ void put(int base, int row, T v) const {
    const unsigned voff = static_cast<unsigned>(row) * sizeof(T);
    if constexpr (sizeof(T) == 8) {
      typedef unsigned int u2 __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rsrc, voff, base * 8, 0);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rsrc, voff, base * 4, 0);
    }
  }
};

// ---- windowed SpMV (nx <= 64 kWinChunks) ----------------------------------------------------------
//
// Wave w walks the levels [spmv_split[w], spmv_split[w+1]) in ascending order (contiguous ranges
// balanced by cell count, set by the host); lane l of row chunk c holds row i = 64 c + l of every
// level, as in the sweeps. A level's p, aW, aS and aP are loaded once: a cell's south and north
// neighbours are the same lane in the previous and next level, west and east the neighbouring
// lanes (a DPP shift; the chunk-boundary lane takes the adjacent chunk's value), so a cell reads
// four values where the level-major pass reads eleven. Loads run two levels ahead of the
// arithmetic. Row sums are the ones spmv_cell forms, term for term; thread (w, l) adds its
// contributions in (level, chunk) order.
constexpr int kWinChunks = 4;
#ifndef PP_WIN_BUFFERS
#define PP_WIN_BUFFERS 4
#endif
constexpr int kWinBuffers = PP_WIN_BUFFERS;  // rotating level buffers of the windowed SpMV
static_assert(sizeof(SysArgs::spmv_split) == (kWaves + 1) * sizeof(int), "one SpMV range per wave");

template <typename T>
struct WinLevel {
  T p[kWinChunks], w[kWinChunks], s[kWinChunks], c[kWinChunks];
};

// Position of row 64 c + lane in level d, clamped into the level's rows so every lane reads a real
// cell. Branch-free: for the levels just outside the grid (d = -1 and up to six past the last),
// the level-base table's margin holds 0 and the clamp lands on a position below n + kPad.
__device__ __forceinline__ int win_pos(const __attribute__((address_space(4))) int *lb, int d, int c, int lane,
                                       int nx, int ny) {
  const int lo = level_lo(d, ny), hi = d < nx - 1 ? d : nx - 1;
  const int top = hi > lo ? hi : lo;
  const int i = 64 * c + lane;
  return lb[d] + (i < lo ? lo : (i > top ? top : i));
}

template <typename T, int S>
__device__ __forceinline__ void win_shift(WinLevel<T> &L, T m_p) {
#pragma unroll
  for (int c = 0; c < S; ++c) L.p[c] = L.p[c] - m_p;
}

template <typename T>
struct WinArrays {
  LevelArray<T> p, w, s, c;
};

// One level's operands: positions lb[d] + clamp(row, lo, top), the clamp keeping every lane on a
// real cell of the level (levels just outside the grid read position 0 or a pad position).
template <typename T, int S>
__device__ __forceinline__ void win_load(const WinArrays<T> &V, const __attribute__((address_space(4))) int *lb, int d,
                                         int lane, int nx, int ny, WinLevel<T> &L) {
  const int lo = level_lo(d, ny), hi = d < nx - 1 ? d : nx - 1;
  const int top = hi > lo ? hi : lo;
  const int base = lb[d];
#pragma unroll
  for (int c = 0; c < S; ++c) {
    const int i = 64 * c + lane;
    const unsigned voff = static_cast<unsigned>(i < lo ? lo : (i > top ? top : i)) * sizeof(T);
    L.p[c] = V.p.get(base, voff);
    L.w[c] = V.w.get(base, voff);
    L.s[c] = V.s.get(base, voff);
    L.c[c] = V.c.get(base, voff);
  }
}

// Level d of the window: rows of every chunk from (prev, cur, next) = levels (d-1, d, d+1), whose p
// values already hold p - m_p (win_shift: each level's value serves three rows' sums, and the
// subtraction is the one a row-by-row SpMV repeats per use).
template <typename T, int S, bool kPeriodic, bool kExtra, typename F>
__device__ __forceinline__ double win_level(const Stencil<T> &A, const global_ptr<const T> p, T m_p,
                                            const LevelArray<T> &extra,
                                            const __attribute__((address_space(4))) int *lb, int d, int lane,
                                            const WinLevel<T> &prev, const WinLevel<T> &cur, const WinLevel<T> &next,
                                            double acc, F &epi) {
  const int nx = A.nx, ny = A.ny;
  const int b0 = lb[d];
  const int row_lo = level_lo(d, ny), row_hi = d < nx - 1 ? d : nx - 1;  // the level's rows
#pragma unroll
  for (int c = 0; c < S; ++c) {
    // a chunk with no row on this level contributes nothing (its loads were issued anyway, clamped
    // to real cells, so the prefetch stays unconditional); the branch is wave-uniform
    if (64 * c > row_hi || 64 * c + 63 < row_lo) continue;
    const int i = 64 * c + lane, j = d - i;
    const bool in_i = i < nx;
    const bool in_j = static_cast<unsigned>(j) < static_cast<unsigned>(ny);
    const bool act = in_i && in_j;
    const bool has_w = i >= 1, has_s = j >= 1, has_n = j <= ny - 2, has_e = i <= nx - 2;
    // p - m_p of the west neighbour: lane l-1, and lane 0 takes lane 63 of the previous chunk inside
    // the DPP shift itself (the lane without a source keeps the fill); likewise east
    const T pw = shr1_fill(prev.p[c], c > 0 ? lane_value(prev.p[c - 1], 63) : T(0));
    const T pe = shl1_fill(next.p[c], c + 1 < S ? lane_value(next.p[c + 1], 0) : T(0));
    const T ew = shl1_fill(next.w[c], c + 1 < S ? lane_value(next.w[c + 1], 0) : T(0));
    // Without periodicity the assembly stores aW = 0 at i = 0 and aS = 0 at j = 0 (no such entry),
    // and the values they multiply there are finite (the DPP fill, a clamped real cell), so the
    // unmasked product is the same signed zero the masked one gives. Periodic rows hold their wrap
    // coefficient in those slots, so there the masks stay.
    const T cw = (!kPeriodic || has_w) ? cur.w[c] : T(0), cs = (!kPeriodic || has_s) ? cur.s[c] : T(0);
    const T cn = has_n ? next.s[c] : T(0), ce = has_e ? ew : T(0);
    const T pc = cur.p[c];
    T s = T(0);
    const int q = kPeriodic ? (act ? b0 + i : win_pos(lb, d, c, lane, nx, ny)) : 0;
    if (kPeriodic && A.px) {  // wrap (0, j) of row (nx-1, j)
      const bool on = i == nx - 1 && in_j;
      const int qx = on ? level_base(j, nx, ny) : q;
      s += (on ? A.aW[qx] : T(0)) * (p[qx] - m_p);
    }
    s += cw * pw;
    if (kPeriodic && A.py) {  // wrap (i, 0) of row (i, ny-1)
      const bool on = j == ny - 1 && in_i;
      const int qy = on ? level_base(i, nx, ny) + i : q;
      s += (on ? A.aS[qy] : T(0)) * (p[qy] - m_p);
    }
    s += cs * prev.p[c];
    s += cur.c[c] * pc;
    s += cn * next.p[c];
    if (kPeriodic && A.py) {  // wrap (i, ny-1) of row (i, 0)
      const bool on = j == 0 && in_i;
      const int qy = on ? level_base(i + ny - 1, nx, ny) + i : q;
      s += (on ? A.aS[q] : T(0)) * (p[qy] - m_p);
    }
    s += ce * pe;
    if (kPeriodic && A.px) {  // wrap (nx-1, j) of row (0, j)
      const bool on = i == 0 && in_j;
      const int qx = on ? level_base(nx - 1 + j, nx, ny) + nx - 1 : q;
      s += (on ? A.aW[q] : T(0)) * (p[qx] - m_p);
    }
    const T ex = kExtra ? extra.get(b0, static_cast<unsigned>(i) * sizeof(T)) : T(0);
    acc = acc + (act ? epi(b0, i, s, pc, ex) : 0.0);
  }
  return acc;
}

// K rotating level buffers: the buffer of level L is (L - d_begin + 1) mod K. Step d computes level
// d from (d-1, d, d+1) after starting the loads of level d + K - 2, so a level's loads are issued
// K - 3 steps before it is first read (as "next"). The loop is unrolled K times, so every buffer
// index is a compile-time constant and no buffer is copied.
template <typename T, int S, int K, bool kPeriodic, bool kExtra, typename F>
__device__ __forceinline__ double spmv_window(const SysArgs &a, const Stencil<T> &A, const global_ptr<const T> p, T m_p,
                                              const global_ptr<const T> extra_ptr, F &epi) {
  static_assert(K >= 4, "prev, cur, next and one buffer in flight");
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const __attribute__((address_space(4))) int *lb = (const __attribute__((address_space(4))) int *)a.lev_base;
  const int d_begin = a.spmv_split[wave], d_end = a.spmv_split[wave + 1];
  double acc = 0.0;
  if (d_begin >= d_end) return acc;
  const int nx = A.nx, ny = A.ny, n = a.n;
  const WinArrays<T> V{LevelArray<T>(p, n), LevelArray<T>(A.aW, n), LevelArray<T>(A.aS, n), LevelArray<T>(A.aP, n)};
  const LevelArray<T> extra(kExtra ? extra_ptr : p, n);
  WinLevel<T> B[K];
#pragma unroll
  for (int b = 0; b < K - 1; ++b) win_load<T, S>(V, lb, d_begin - 1 + b, lane, nx, ny, B[b]);
  win_shift<T, S>(B[0], m_p);
  win_shift<T, S>(B[1], m_p);
  for (int d = d_begin; d < d_end; d += K) {
#pragma unroll
    for (int r = 0; r < K; ++r) {
      if (d + r >= d_end) return acc;
      win_load<T, S>(V, lb, d + r + K - 2, lane, nx, ny, B[(r + K - 1) % K]);
      win_shift<T, S>(B[(r + 2) % K], m_p);  // level d + r + 1, first read now (as next)
      acc = win_level<T, S, kPeriodic, kExtra>(A, p, m_p, extra, lb, d + r, lane, B[r % K], B[(r + 1) % K],
                                                B[(r + 2) % K], acc, epi);
    }
  }
  return acc;
}

template <typename T, bool kPeriodic, bool kExtra, typename F>
__device__ __forceinline__ double spmv_window_rows(const SysArgs &a, const Stencil<T> &A, const global_ptr<const T> p,
                                                   T m_p, const global_ptr<const T> extra, F &epi) {
  switch ((A.nx + 63) >> 6) {
    case 1: return spmv_window<T, 1, kWinBuffers, kPeriodic, kExtra>(a, A, p, m_p, extra, epi);
    case 2: return spmv_window<T, 2, kWinBuffers, kPeriodic, kExtra>(a, A, p, m_p, extra, epi);
    case 3: return spmv_window<T, 3, kWinBuffers, kPeriodic, kExtra>(a, A, p, m_p, extra, epi);
    default: return spmv_window<T, 4, kWinBuffers, kPeriodic, kExtra>(a, A, p, m_p, extra, epi);
  }
}

// A p (plus the epilogue) with the windowed pass where the rows fit its registers, else level-major.
template <typename T, typename F>
__device__ __forceinline__ double spmv_pass(const SysArgs &a, const Stencil<T> &A, const global_ptr<const T> p, T m_p,
                                            const global_ptr<const T> extra, F &&epi) {
  if (A.nx > 64 * kWinChunks) return spmv_levels<T>(a, A, p, m_p, extra, epi);
  const bool periodic = A.px || A.py;
  if (extra) {
    return periodic ? spmv_window_rows<T, true, true>(a, A, p, m_p, extra, epi)
                    : spmv_window_rows<T, false, true>(a, A, p, m_p, extra, epi);
  }
  return periodic ? spmv_window_rows<T, true, false>(a, A, p, m_p, extra, epi)
                  : spmv_window_rows<T, false, false>(a, A, p, m_p, extra, epi);
}

// z = M^{-1} (r - m_r): forward then backward sweep, staged through LDS when S <= kMaxStagedWaves.
template <typename T, bool kIldlt>
__device__ __forceinline__ void precondition(const SysArgs &a, const SweepArrays<T> &fwd, const SweepArrays<T> &bwd,
                                             SweepShared &sh, T m_r, T floor, unsigned long long *ticks,
                                             unsigned long long &t) {
  const int S = (a.nx + 63) >> 6;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const bool staged = S <= kMaxStagedWaves && a.ring != kDirectSweeps;
  reset_handoff(sh);
  reset_stage(sh);
  __syncthreads();
  // sweep waves carry the dependency chain: they win instruction issue over loader waves on the
  // same SIMD (wave priority, restored after the sweep)
  if (!staged)
    forward_solve<T, kIldlt>(a, fwd, sh, m_r, floor);
  else if (wave < S) {
    __builtin_amdgcn_s_setprio(3);
    if (a.periodic_x || a.periodic_y)
      staged_forward<T, kIldlt, true>(a, fwd, sh, m_r, floor, S);
    else
      staged_forward<T, kIldlt, false>(a, fwd, sh, m_r, floor, S);
    __builtin_amdgcn_s_setprio(0);
  }
  else
    stage_loader<T, false>(a, fwd, sh, S, m_r, floor);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long now = device_ticks();
    ticks[4] += now - t;
    t = now;
  }
  reset_handoff(sh);
  reset_stage(sh);
  __syncthreads();
  if (!staged)
    backward_solve<T, kIldlt>(a, bwd, sh, floor);
  else if (wave < S) {
    __builtin_amdgcn_s_setprio(3);
    if (a.periodic_x || a.periodic_y)
      staged_backward<T, kIldlt, true>(a, bwd, sh, floor, S);
    else
      staged_backward<T, kIldlt, false>(a, bwd, sh, floor, S);
    __builtin_amdgcn_s_setprio(0);
  }
  else
    stage_loader<T, true>(a, bwd, sh, S, T(0), floor);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long now = device_ticks();
    ticks[5] += now - t;
    t = now;
  }
}

// Streaming passes visit positions tid, tid + T, tid + 2T, ... in ascending order; kUnroll of
// them are loaded before any is used so each thread keeps several loads in flight.
constexpr int kUnroll = 8;

// Per-phase device timers (thread 0 keeps them; reported in pp_report.phase_ms).
struct PhaseClock {
  unsigned long long ticks[kPhases];
  unsigned long long t;
  __device__ __forceinline__ PhaseClock() : t(device_ticks()) {
    for (int k = 0; k < kPhases; ++k) ticks[k] = 0;
  }
  __device__ __forceinline__ void mark(int phase) {
    if (threadIdx.x == 0) {
      const unsigned long long now = device_ticks();
      ticks[phase] += now - t;
      t = now;
    }
  }
  __device__ __forceinline__ void finish(DevReport *rep) const {
    if (threadIdx.x == 0)
      for (int k = 0; k < kPhases; ++k) rep->phase_ticks[k] = ticks[k];
  }
};

struct Setup {
  double nd;      // max(1, n)
  double b_mean;  // mean of the assembled right-hand side (0 when the system is not projected)
  double floor;   // pivot floor of the factorization
};

// Assembly (+ missing-diagonal check), the right-hand-side mean and the factorization, shared by
// every solve. Returns false, with the status in LDS, when the system cannot be solved.
template <bool kIldlt>
__device__ __forceinline__ bool setup_system(const SysArgs &a, SweepShared &sh, int &parity, Setup &out) {
  const int tid = threadIdx.x;
  const int n = a.n, ny = a.ny;
  if (tid == 0) {
    *lds(&sh.spin_error) = 0;
    *lds(&sh.spin_limit) = a.spin_limit > 0 ? a.spin_limit : kSpinLimit;
    *lds(&sh.status) = kOk;
  }
  __syncthreads();

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
  if (nodiag) *lds(&sh.status) = kNoDiag;
  const double total = block_sum1(bsum, sh, parity);
  const double max_diag = block_max1(dmax, sh, parity);
  if (*lds(&sh.status) != kOk) return false;
  out.nd = static_cast<double>(n) < 1.0 ? 1.0 : static_cast<double>(n);
  out.b_mean = a.projected ? total / out.nd : 0.0;

  const double base = 1.0 < max_diag ? max_diag : 1.0;
  out.floor = kIldlt ? base * 1.0e-12 : sqrt(base * 1.0e-12);
  reset_handoff(sh);
  __syncthreads();
  factor_sweep<kIldlt>(a, sh, out.floor);
  __syncthreads();
  if (a.periodic_x)
    for (int j = tid; j < ny; j += kThreads) sh.lx[j] = as_global<const double>(a.LX)[j];
  if (*lds(&sh.spin_error)) *lds(&sh.status) = kSweepFail;
  return *lds(&sh.status) == kOk;
}

template <bool kIldlt, bool kMonitor>
__device__ __forceinline__ void solve_system(const SysArgs &a, SweepShared &sh) {
  const int tid = threadIdx.x;
  const int n = a.n, nx = a.nx, ny = a.ny;
  int parity = 0;
  PhaseClock clk;
  auto mark = [&](int phase) { clk.mark(phase); };
  // PCG monitor: one record per call of the reference's CGMonitor (ch_sparse_krylov.hpp:547-648)
  int monitor_calls = 0;
  auto monitor = [&](double alpha, double beta, double abs_r, double rel_r, double iter_r, double prec_r) {
    if (!kMonitor) return;
    if (tid == 0 && monitor_calls < a.monitor_cap) {
      MonitorRecord m;
      m.iteration = monitor_calls;
      m.pad = 0;
      m.alpha = alpha;
      m.beta = beta;
      m.abs_res = abs_r;
      m.rel_res = rel_r;
      m.iter_res = iter_r;
      m.prec_res = prec_r;
      a.monitor[monitor_calls] = m;
    }
    ++monitor_calls;
  };
  auto finish = [&]() {
    if (kMonitor && tid == 0) a.rep->monitor_calls = monitor_calls;
    clk.finish(a.rep);
  };
  Setup su;
  if (!setup_system<kIldlt>(a, sh, parity, su)) {
    finish();
    return;
  }
  const double nd = su.nd, b_mean = su.b_mean, floor = su.floor;

  const global_ptr<double> xv = as_global(a.x);
  const global_ptr<double> rv = as_global(a.r);
  const global_ptr<double> zv = as_global(a.z);
  const global_ptr<double> bv = as_global(a.b);
  const global_ptr<double> apv = as_global(a.y);  // Ap shares storage with the sweep intermediate
  const SweepArrays<double> fwd{a.r, a.Lw, a.Ls, a.dg, a.y, a.LY};
  const SweepArrays<double> bwd{a.y, a.Lw, a.Ls, a.dg, a.z, a.LY};
  const Stencil<double> stencil{nx, ny, a.periodic_x != 0, a.periodic_y != 0, as_global<const double>(a.aP),
                        as_global<const double>(a.aW), as_global<const double>(a.aS)};
  const bool projected = a.projected != 0;

  // r = b - A*0 (= b), x = 0; projections and norms
  double acc[2] = {0.0, 0.0};
  for (int q = tid; q < n; q += kThreads) {
    const double bq = projected ? bv[q] - b_mean : bv[q];
    const double rq = bq - 0.0;
    bv[q] = bq;
    rv[q] = rq;
    xv[q] = 0.0;
    acc[0] = acc[0] + bq * bq;
    acc[1] = acc[1] + rq;
  }
  block_sum<2>(acc, sh, parity);
  const double bnorm = sqrt(acc[0] < 0.0 ? 0.0 : acc[0]);
  double m_r = projected ? acc[1] / nd : 0.0;
  double rr = 0.0;
  for (int q = tid; q < n; q += kThreads) {
    const double tq = rv[q] - m_r;
    rr = rr + tq * tq;
  }
  rr = block_sum1(rr, sh, parity);
  double abs_res = sqrt(rr < 0.0 ? 0.0 : rr);
  double rel_res = abs_res / (bnorm < 1.0e-30 ? 1.0e-30 : bnorm);
  double iter_res = 0.0;
  int iterations = 0;
  double m_x = 0.0, m_p = 0.0;
  const double tol = a.tolerance;
  const global_ptr<double> p_raw = as_global(a.pa);
  const LevelArray<double> apv_rows(a.y, n);
  mark(0);

  if (rel_res < tol || abs_res < tol) monitor(0.0, 0.0, abs_res, rel_res, 0.0, 0.0);
  if (!(rel_res < tol || abs_res < tol)) {
    precondition<double, kIldlt>(a, fwd, bwd, sh, m_r, floor, clk.ticks, clk.t);
    double zs = 0.0;
    for (int q = tid; q < n; q += kThreads) zs = zs + zv[q];
    zs = block_sum1(zs, sh, parity);
    double m_z = projected ? zs / nd : 0.0;
    double rz = 0.0;
    for (int q = tid; q < n; q += kThreads) {
      const double zq = zv[q] - m_z;
      rz = rz + (rv[q] - m_r) * zq;
      p_raw[q] = zq;
    }
    rz = block_sum1(rz, sh, parity);
    if (!isfinite(rz) || rz <= 0.0) {
      if (tid == 0) *lds(&sh.status) = *lds(&sh.spin_error) ? kSweepFail : kRz0Fail;  // as for r.z below
      finish();
      return;
    }
    monitor(0.0, 0.0, abs_res, rel_res, 0.0, sqrt(rz < 0.0 ? 0.0 : rz));
    mark(0);
    const int limit = a.max_iterations < 1 ? 1 : a.max_iterations;
    // the loop head reads spin_error uniformly: the last sweeps ended before a barrier since
    for (int it = 0; it < limit && !*lds(&sh.spin_error); ++it) {
      // Ap and p^T A p (level-major order)
      double pap = spmv_pass<double>(a, stencil, p_raw, m_p, nullptr, [&](int base, int i, double apq, double pq, double) {
        apv_rows.put(base, i, apq);
        return pq * apq;
      });
      pap = block_sum1(pap, sh, parity);
      mark(1);
      if (!isfinite(pap) || fabs(pap) < 1.0e-30) {
        if (tid == 0) *lds(&sh.status) = kPapFail;
        finish();
        return;
      }
      const double alpha = rz / pap;
      // x += alpha p, r -= alpha Ap (on projected x, r), with the update / iterate norms
      double up[4] = {0.0, 0.0, 0.0, 0.0};
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        double xq[kUnroll], rq[kUnroll], pq[kUnroll], aq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          const int qq = q < n ? q : 0;
          xq[u] = xv[qq];
          rq[u] = rv[qq];
          pq[u] = p_raw[qq] - m_p;  // the projected p the SpMV used
          aq[u] = apv[qq];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          if (q < n) {
            const double delta = alpha * pq[u];
            const double xn = (xq[u] - m_x) + delta;
            const double rn = (rq[u] - m_r) - alpha * aq[u];
            xv[q] = xn;
            rv[q] = rn;
            up[0] = up[0] + xn;
            up[1] = up[1] + rn;
            up[2] = up[2] + delta * delta;
            up[3] = up[3] + xn * xn;
          }
        }
      }
      block_sum<4>(up, sh, parity);
      mark(2);
      m_x = projected ? up[0] / nd : 0.0;
      m_r = projected ? up[1] / nd : 0.0;
      rr = 0.0;
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        double rq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          rq[u] = rv[q < n ? q : 0];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          if (base + u * kThreads < n) {
            const double tq = rq[u] - m_r;
            rr = rr + tq * tq;
          }
        }
      }
      rr = block_sum1(rr, sh, parity);
      mark(3);
      iterations = it + 1;
      iter_res = sqrt(up[2] / (up[3] < 1.0e-30 ? 1.0e-30 : up[3]));
      abs_res = sqrt(rr < 0.0 ? 0.0 : rr);
      rel_res = abs_res / (bnorm < 1.0e-30 ? 1.0e-30 : bnorm);
      if (rel_res < tol || abs_res < tol) {
        monitor(alpha, 0.0, abs_res, rel_res, iter_res, 0.0);
        break;
      }

      precondition<double, kIldlt>(a, fwd, bwd, sh, m_r, floor, clk.ticks, clk.t);
      zs = 0.0;
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        double zq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          zq[u] = zv[q < n ? q : 0];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
          if (base + u * kThreads < n) zs = zs + zq[u];
      }
      zs = block_sum1(zs, sh, parity);
      mark(6);
      m_z = projected ? zs / nd : 0.0;
      double rzn = 0.0;
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        double zq[kUnroll], rq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          const int qq = q < n ? q : 0;
          zq[u] = zv[qq];
          rq[u] = rv[qq];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
          if (base + u * kThreads < n) rzn = rzn + (rq[u] - m_r) * (zq[u] - m_z);
      }
      rzn = block_sum1(rzn, sh, parity);
      mark(7);
      if (!isfinite(rzn) || rzn <= 0.0) {  // after a failed hand-off r.z is garbage: report the hand-off
        if (tid == 0) *lds(&sh.status) = *lds(&sh.spin_error) ? kSweepFail : kRzFail;
        finish();
        return;
      }
      const double beta = rzn / rz;
      monitor(alpha, beta, abs_res, rel_res, iter_res, sqrt(rzn < 0.0 ? 0.0 : rzn));
      double ps = 0.0;
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        double zq[kUnroll], pq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          const int qq = q < n ? q : 0;
          zq[u] = zv[qq];
          pq[u] = p_raw[qq] - m_p;  // the projected p the SpMV used
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          if (q < n) {
            const double pn = (zq[u] - m_z) + beta * pq[u];
            p_raw[q] = pn;
            ps = ps + pn;
          }
        }
      }
      ps = block_sum1(ps, sh, parity);
      mark(8);
      m_p = projected ? ps / nd : 0.0;
      rz = rzn;
    }
  }
  if (*lds(&sh.spin_error)) {
    if (tid == 0) *lds(&sh.status) = kSweepFail;
    finish();
    return;
  }

  // scatter the projected x, subtract the interior mean when the system is singular
  double xs = 0.0;
  for (int q = tid; q < n; q += kThreads) xs = xs + (xv[q] - m_x);
  xs = block_sum1(xs, sh, parity);
  const double mean = projected ? xs / static_cast<double>(nx * ny) : 0.0;
  for (int q = tid; q < n; q += kThreads) {
    const uint32_t c = a.cell[q];
    const int i = static_cast<int>(c >> 16), j = static_cast<int>(c & 0xffffu);
    const double xq = xv[q] - m_x;
    a.p_out[i * ny + j] = projected ? xq - mean : xq;
  }
  mark(9);
  if (tid == 0) {
    a.rep->iterations = iterations;
    a.rep->outer_cycles = 1;
    a.rep->rel = rel_res;
    a.rep->abs = abs_res;
    a.rep->iter_res = iter_res;
  }
  finish();
}

// ---- mixed precision: FP32 inner PCG inside FP64 iterative refinement ---------------------------
//
// No counterpart in the reference. Setup (assembly, regularization, ICC/ILDLT factorization) is the
// FP64 path's. Each refinement cycle computes the true residual r = b - A x in FP64 (projected like
// the FP64 solve), stops when |r| / |b| or |r| is below the tolerance, and otherwise solves
// A d = r / |r| approximately with PCG in FP32: FP32 copies of the operator and factor (reciprocal of
// the floored pivot instead of division), FP32 vectors, dot products and means accumulated in FP64.
// The inner solve stops at relative reduction inner_tolerance; x += |r| d in FP64.
constexpr int kMaxOuterCycles = 64;

template <bool kIldlt>
__device__ __forceinline__ void mixed_solve(const SysArgs &a, SweepShared &sh) {
  const int tid = threadIdx.x;
  const int n = a.n, nx = a.nx, ny = a.ny;
  int parity = 0;
  PhaseClock clk;
  Setup su;
  if (!setup_system<kIldlt>(a, sh, parity, su)) {
    clk.finish(a.rep);
    return;
  }
  const double nd = su.nd, floor = su.floor;
  const bool projected = a.projected != 0;
  const global_ptr<double> xv = as_global(a.x), rv = as_global(a.r), bv = as_global(a.b);
  const global_ptr<float> fP = as_global(a.fP), fW = as_global(a.fW), fS = as_global(a.fS);
  const global_ptr<float> fLw = as_global(a.fLw), fLs = as_global(a.fLs), fRd = as_global(a.fRd);
  const global_ptr<float> dv = as_global(a.fx), r32 = as_global(a.fr), p32 = as_global(a.fpa);
  const global_ptr<float> z32 = as_global(a.fz), ap32 = as_global(a.fy);
  const LevelArray<double> r_rows(a.r, n);
  const LevelArray<float> ap_rows(a.fy, n);

  // FP32 operator and factor, regularized FP64 right-hand side, x = 0
  double bb = 0.0;
  for (int q = tid; q < n; q += kThreads) {
    fP[q] = static_cast<float>(a.aP[q]);
    fW[q] = static_cast<float>(a.aW[q]);
    fS[q] = static_cast<float>(a.aS[q]);
    fLw[q] = static_cast<float>(a.Lw[q]);
    fLs[q] = static_cast<float>(a.Ls[q]);
    fRd[q] = static_cast<float>(1.0 / floor_max(a.dg[q], floor));
    const double bq = projected ? bv[q] - su.b_mean : bv[q];
    bv[q] = bq;
    xv[q] = 0.0;
    bb = bb + bq * bq;
  }
  bb = block_sum1(bb, sh, parity);
  const double bnorm = sqrt(bb < 0.0 ? 0.0 : bb);
  const Stencil<double> A64{nx, ny, a.periodic_x != 0, a.periodic_y != 0, as_global<const double>(a.aP),
                            as_global<const double>(a.aW), as_global<const double>(a.aS)};
  const Stencil<float> A32{nx, ny, a.periodic_x != 0, a.periodic_y != 0, fP, fW, fS};
  const SweepArrays<float> fwd{a.fr, a.fLw, a.fLs, a.fRd, a.fy, a.LY};
  const SweepArrays<float> bwd{a.fy, a.fLw, a.fLs, a.fRd, a.fz, a.LY};
  const double tol = a.tolerance;
  // default inner reduction 3 sqrt(tol), within [1e-5, 1e-2] (measured best at tol 1e-8: about three
  // cycles, fewest inner iterations); each cycle asks no more than the remaining 0.5 tol |b| / |r|
  const double inner_guess = 3.0 * sqrt(a.tolerance);
  const double inner_tol = a.inner_tolerance > 0.0 ? a.inner_tolerance
                                                    : (inner_guess < 1.0e-5 ? 1.0e-5 : (inner_guess > 1.0e-2 ? 1.0e-2 : inner_guess));
  const int limit = a.max_iterations < 1 ? 1 : a.max_iterations;
  const int inner_cap = a.inner_max_iterations > 0 ? a.inner_max_iterations : limit;
  clk.mark(0);

  double m_x = 0.0, rel_res = 0.0, abs_res = 0.0, iter_res = 0.0;
  int total = 0, cycles = 0;
  for (;;) {
    // true residual r = b - A (x - m_x) in FP64, projected
    double rs = spmv_pass<double>(a, A64, xv, m_x, bv, [&](int base, int i, double axq, double, double bq) {
      const double rq = bq - axq;
      r_rows.put(base, i, rq);
      return rq;
    });
    rs = block_sum1(rs, sh, parity);
    const double m_r64 = projected ? rs / nd : 0.0;
    double rr = 0.0;
    for (int q = tid; q < n; q += kThreads) {
      const double tq = rv[q] - m_r64;
      rr = rr + tq * tq;
    }
    rr = block_sum1(rr, sh, parity);
    clk.mark(1);
    abs_res = sqrt(rr < 0.0 ? 0.0 : rr);
    rel_res = abs_res / (bnorm < 1.0e-30 ? 1.0e-30 : bnorm);
    if (rel_res < tol || abs_res < tol || total >= limit || cycles >= kMaxOuterCycles) break;

    // inner FP32 PCG on A d = (r - m_r) / |r|
    const double scale = 1.0 / abs_res;
    const double needed = 0.5 * tol * (bnorm < 1.0e-30 ? 1.0e-30 : bnorm) / abs_res;
    const double cycle_tol = needed > inner_tol ? needed : inner_tol;
    double r0s = 0.0;
    for (int q = tid; q < n; q += kThreads) {
      const float rq = static_cast<float>((rv[q] - m_r64) * scale);
      r32[q] = rq;
      dv[q] = 0.0f;
      r0s = r0s + static_cast<double>(rq);
    }
    r0s = block_sum1(r0s, sh, parity);
    float m_r = projected ? static_cast<float>(r0s / nd) : 0.0f;
    float m_d = 0.0f;
    double rr0 = 0.0;
    for (int q = tid; q < n; q += kThreads) {
      const double tq = static_cast<double>(r32[q] - m_r);
      rr0 = rr0 + tq * tq;
    }
    rr0 = block_sum1(rr0, sh, parity);
    const double rn0 = sqrt(rr0 < 0.0 ? 0.0 : rr0);
    clk.mark(2);
    int inner = 0;
    precondition<float, kIldlt>(a, fwd, bwd, sh, m_r, 0.0f, clk.ticks, clk.t);
    double zs = 0.0;
    for (int q = tid; q < n; q += kThreads) zs = zs + static_cast<double>(z32[q]);
    zs = block_sum1(zs, sh, parity);
    float m_z = projected ? static_cast<float>(zs / nd) : 0.0f;
    double rz = 0.0;
    for (int q = tid; q < n; q += kThreads) {
      const float zq = z32[q] - m_z;
      rz = rz + static_cast<double>(r32[q] - m_r) * static_cast<double>(zq);
      p32[q] = zq;
    }
    rz = block_sum1(rz, sh, parity);
    float m_p = 0.0f;
    clk.mark(7);
    while (isfinite(rz) && rz > 0.0 && inner < inner_cap && total < limit && !*lds(&sh.spin_error)) {
      double pap = spmv_pass<float>(a, A32, p32, m_p, nullptr, [&](int base, int i, float apq, float pq, float) {
        ap_rows.put(base, i, apq);
        return static_cast<double>(pq) * static_cast<double>(apq);
      });
      pap = block_sum1(pap, sh, parity);
      clk.mark(1);
      if (!isfinite(pap) || fabs(pap) < 1.0e-30) break;
      const float alpha = static_cast<float>(rz / pap);
      // sums of d, r and r^2 in one pass: |r - m_r|^2 = sum r^2 - (sum r)^2 / n. r is projected
      // every iteration, so its mean is at rounding level and the subtraction cancels nothing.
      double up[3] = {0.0, 0.0, 0.0};
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        float dq[kUnroll], rq[kUnroll], pq[kUnroll], aq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          const int qq = q < n ? q : 0;
          dq[u] = dv[qq];
          rq[u] = r32[qq];
          pq[u] = p32[qq];
          aq[u] = ap32[qq];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          if (q < n) {
            const float dn = (dq[u] - m_d) + alpha * (pq[u] - m_p);
            const float rn = (rq[u] - m_r) - alpha * aq[u];
            dv[q] = dn;
            r32[q] = rn;
            up[0] = up[0] + static_cast<double>(dn);
            up[1] = up[1] + static_cast<double>(rn);
            up[2] = up[2] + static_cast<double>(rn) * static_cast<double>(rn);
          }
        }
      }
      block_sum<3>(up, sh, parity);
      m_d = projected ? static_cast<float>(up[0] / nd) : 0.0f;
      m_r = projected ? static_cast<float>(up[1] / nd) : 0.0f;
      const double rri = projected ? up[2] - up[1] * (up[1] / nd) : up[2];
      clk.mark(3);
      ++inner;
      ++total;
      if (sqrt(rri < 0.0 ? 0.0 : rri) < cycle_tol * rn0 || inner >= inner_cap || total >= limit) break;
      precondition<float, kIldlt>(a, fwd, bwd, sh, m_r, 0.0f, clk.ticks, clk.t);
      // mean(z) and (r - m_r).(z - m_z) in one pass: sum z, sum (r - m_r) z and sum (r - m_r), then
      // (r - m_r).(z - m_z) = sum (r - m_r) z - m_z sum (r - m_r)
      double zr[3] = {0.0, 0.0, 0.0};
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        float zq[kUnroll], rq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          const int qq = q < n ? q : 0;
          zq[u] = z32[qq];
          rq[u] = r32[qq];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
          if (base + u * kThreads < n) {
            const double rc = static_cast<double>(rq[u] - m_r), zc = static_cast<double>(zq[u]);
            zr[0] = zr[0] + zc;
            zr[1] = zr[1] + rc * zc;
            zr[2] = zr[2] + rc;
          }
      }
      block_sum<3>(zr, sh, parity);
      zs = zr[0];
      const double mz = projected ? zs / nd : 0.0;
      m_z = static_cast<float>(mz);
      const double rzn = zr[1] - mz * zr[2];
      clk.mark(7);
      if (!isfinite(rzn) || rzn <= 0.0) break;
      const float beta = static_cast<float>(rzn / rz);
      double ps = 0.0;
      for (int base = tid; base < n; base += kThreads * kUnroll) {
        float zq[kUnroll], pq[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          const int qq = q < n ? q : 0;
          zq[u] = z32[qq];
          pq[u] = p32[qq];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int q = base + u * kThreads;
          if (q < n) {
            const float pn = (zq[u] - m_z) + beta * (pq[u] - m_p);
            p32[q] = pn;
            ps = ps + static_cast<double>(pn);
          }
        }
      }
      ps = block_sum1(ps, sh, parity);
      m_p = projected ? static_cast<float>(ps / nd) : 0.0f;
      rz = rzn;
      clk.mark(8);
    }
    if (inner == 0 || *lds(&sh.spin_error)) break;  // no progress possible: stop like a silent non-convergence
    // x += |r| d in FP64 (x kept unprojected, its mean tracked for the lazy projection)
    double xs = 0.0, dd = 0.0, xx = 0.0;
    for (int q = tid; q < n; q += kThreads) {
      const double delta = abs_res * static_cast<double>(dv[q] - m_d);
      const double xn = (xv[q] - m_x) + delta;
      xv[q] = xn;
      xs = xs + xn;
      dd = dd + delta * delta;
      xx = xx + xn * xn;
    }
    double acc[3] = {xs, dd, xx};
    block_sum<3>(acc, sh, parity);
    m_x = projected ? acc[0] / nd : 0.0;
    iter_res = sqrt(acc[1] / (acc[2] < 1.0e-30 ? 1.0e-30 : acc[2]));
    ++cycles;
    clk.mark(2);
  }
  if (*lds(&sh.spin_error)) {
    if (tid == 0) *lds(&sh.status) = kSweepFail;
    clk.finish(a.rep);
    return;
  }
  double xs = 0.0;
  for (int q = tid; q < n; q += kThreads) xs = xs + (xv[q] - m_x);
  xs = block_sum1(xs, sh, parity);
  const double mean = projected ? xs / static_cast<double>(nx * ny) : 0.0;
  for (int q = tid; q < n; q += kThreads) {
    const uint32_t c = a.cell[q];
    const int i = static_cast<int>(c >> 16), j = static_cast<int>(c & 0xffffu);
    const double xq = xv[q] - m_x;
    a.p_out[i * ny + j] = projected ? xq - mean : xq;
  }
  clk.mark(9);
  if (tid == 0) {
    a.rep->iterations = total;
    a.rep->outer_cycles = cycles;
    a.rep->rel = rel_res;
    a.rep->abs = abs_res;
    a.rep->iter_res = iter_res;
  }
  clk.finish(a.rep);
}

}  // namespace

template <bool kIldlt, bool kMonitor>
__global__ void __launch_bounds__(kThreads) fp64_solve_kernel(const SysArgs *__restrict__ systems) {
  __shared__ SweepShared sh;
  const SysArgs &a = systems[blockIdx.x];
  solve_system<kIldlt, kMonitor>(a, sh);
  __syncthreads();
  if (threadIdx.x == 0) a.rep->status = *lds(&sh.status);
}

template <bool kIldlt>
__global__ void __launch_bounds__(kThreads) mixed_solve_kernel(const SysArgs *__restrict__ systems) {
  __shared__ SweepShared sh;
  const SysArgs &a = systems[blockIdx.x];
  mixed_solve<kIldlt>(a, sh);
  __syncthreads();
  if (threadIdx.x == 0) a.rep->status = *lds(&sh.status);
}

hipError_t launch_mixed_solve(const SysArgs *d_systems, int count, bool ildlt, hipStream_t stream) {
  if (ildlt)
    hipLaunchKernelGGL(mixed_solve_kernel<true>, dim3(count), dim3(kThreads), 0, stream, d_systems);
  else
    hipLaunchKernelGGL(mixed_solve_kernel<false>, dim3(count), dim3(kThreads), 0, stream, d_systems);
  return hipGetLastError();
}

hipError_t launch_fp64_solve(const SysArgs *d_systems, int count, bool ildlt, hipStream_t stream, bool monitor) {
  if (monitor) {
    if (ildlt)
      hipLaunchKernelGGL((fp64_solve_kernel<true, true>), dim3(count), dim3(kThreads), 0, stream, d_systems);
    else
      hipLaunchKernelGGL((fp64_solve_kernel<false, true>), dim3(count), dim3(kThreads), 0, stream, d_systems);
  } else {
    if (ildlt)
      hipLaunchKernelGGL((fp64_solve_kernel<true, false>), dim3(count), dim3(kThreads), 0, stream, d_systems);
    else
      hipLaunchKernelGGL((fp64_solve_kernel<false, false>), dim3(count), dim3(kThreads), 0, stream, d_systems);
  }
  return hipGetLastError();
}

}  // namespace pp
