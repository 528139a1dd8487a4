// Momentum predictor on the device: Solver::compute_momentum_fluxes and the assembly and scatter of
// Solver::solve_momentum_predictor (src/core/coupled/momentum.cpp:55-85, 87-155, 157-502), around
// the BiCGSTAB of pp_bicgstab.hip.
//
// Every expression is the reference's, term for term (compiled with -ffp-contract=off):
//   convection     H(u), H(v) with central differences                      momentum.cpp:55-85
//   stress         2 eta du/dx, eta (du/dy + dv/dx) with eta_previous_step_  momentum.cpp:87-155
//   right-hand side rho_f u + dt (-rho_f conv + rho_f g + f_s) + alpha visc   momentum.cpp:166-186
//   matrix rows    diagonal, 4 own-component and 4 cross entries with the
//                  wall / Neumann handling of add_*_entry                    momentum.cpp:287-446
//   face averages  operators.cpp:330-385
// A row is the ordered list of the contributions the reference inserts into its std::map row: a
// column's value is their sum in insertion order from 0.0 (the map's value-initialised entry), the
// row comes out in ascending columns and entries below 1e-18 in magnitude are dropped
// (ch_sparse_krylov.hpp:105-133), so the CSR is the reference's bit for bit. One thread per row;
// the 9 contributions stay in registers (all indices compile-time), each first occurrence of a
// column folds the later ones and lands at its rank among the row's kept columns.
#include <hip/hip_runtime.h>

#include "pp_device.hpp"
#include "pp_fields.hpp"
#include "pp_kernels.hpp"
#include "pp_solver.h"

namespace pp {

namespace {

constexpr int kBlock = 256;

struct Fields {
  GhostedConst u, v, rho, eta, eta_prev;
  __device__ explicit Fields(const MomArgs &a)
      : u(a.u, a.ny, a.G), v(a.v, a.ny + 1, a.G), rho(a.rho, a.ny, a.G), eta(a.eta, a.ny, a.G),
        eta_prev(a.eta_prev, a.ny, a.G) {}
};

// operators.cpp:330-385
__device__ __forceinline__ double face_x(const MomArgs &a, const GhostedConst &f, int i, int j) {
  const int nx = a.nx;
  if (i <= 0) return a.periodic_x ? 0.5 * (f(nx - 1, j) + f(0, j)) : f(0, j);
  if (i >= nx) return a.periodic_x ? 0.5 * (f(nx - 1, j) + f(0, j)) : f(nx - 1, j);
  return 0.5 * (f(i - 1, j) + f(i, j));
}

__device__ __forceinline__ double face_y(const MomArgs &a, const GhostedConst &f, int i, int j) {
  const int ny = a.ny;
  if (j <= 0) return a.periodic_y ? 0.5 * (f(i, ny - 1) + f(i, 0)) : f(i, 0);
  if (j >= ny) return a.periodic_y ? 0.5 * (f(i, ny - 1) + f(i, 0)) : f(i, ny - 1);
  return 0.5 * (f(i, j - 1) + f(i, j));
}

__device__ __forceinline__ double corner(const GhostedConst &f, int i, int j) {
  return 0.25 * (f(i, j) + f(i - 1, j) + f(i, j - 1) + f(i - 1, j - 1));
}

// H(u) on u face (i, j) and H(v) on v face (i, j) (momentum.cpp:61-82)
__device__ __forceinline__ double conv_u(const MomArgs &a, const Fields &F, int i, int j) {
  const double du_dx = (F.u(i + 1, j) - F.u(i - 1, j)) / (2.0 * a.dx);
  const double du_dy = (F.u(i, j + 1) - F.u(i, j - 1)) / (2.0 * a.dy);
  const double vf = 0.25 * (F.v(i - 1, j) + F.v(i, j) + F.v(i - 1, j + 1) + F.v(i, j + 1));
  return F.u(i, j) * du_dx + vf * du_dy;
}

__device__ __forceinline__ double conv_v(const MomArgs &a, const Fields &F, int i, int j) {
  const double dv_dx = (F.v(i + 1, j) - F.v(i - 1, j)) / (2.0 * a.dx);
  const double dv_dy = (F.v(i, j + 1) - F.v(i, j - 1)) / (2.0 * a.dy);
  const double uf = 0.25 * (F.u(i, j - 1) + F.u(i, j) + F.u(i + 1, j - 1) + F.u(i + 1, j));
  return uf * dv_dx + F.v(i, j) * dv_dy;
}

// stress_divergence_u / _v with eta_previous_step_ (momentum.cpp:91-155)
__device__ __forceinline__ double stress_u(const MomArgs &a, const Fields &F, int i, int j) {
  const GhostedConst &e = F.eta_prev;
  const double ee = face_x(a, e, i + 1, j), ew = face_x(a, e, i, j);
  const double normal = (2.0 * ee * ((F.u(i + 1, j) - F.u(i, j)) / a.dx) - 2.0 * ew * ((F.u(i, j) - F.u(i - 1, j)) / a.dx)) / a.dx;
  const double en = corner(e, i, j + 1), es = corner(e, i, j);
  const double shear = (en * ((F.u(i, j + 1) - F.u(i, j)) / a.dy) - es * ((F.u(i, j) - F.u(i, j - 1)) / a.dy)) / a.dy;
  const double cross = (en * ((F.v(i, j + 1) - F.v(i - 1, j + 1)) / a.dx) - es * ((F.v(i, j) - F.v(i - 1, j)) / a.dx)) / a.dy;
  return (normal + shear) + cross;
}

__device__ __forceinline__ double stress_v(const MomArgs &a, const Fields &F, int i, int j) {
  const GhostedConst &e = F.eta_prev;
  const double ee = corner(e, i + 1, j), ew = corner(e, i, j);
  const double shear = (ee * ((F.v(i + 1, j) - F.v(i, j)) / a.dx) - ew * ((F.v(i, j) - F.v(i - 1, j)) / a.dx)) / a.dx;
  const double en = face_y(a, e, i, j + 1), es = face_y(a, e, i, j);
  const double normal = (2.0 * en * ((F.v(i, j + 1) - F.v(i, j)) / a.dy) - 2.0 * es * ((F.v(i, j) - F.v(i, j - 1)) / a.dy)) / a.dy;
  const double cross = (ee * ((F.u(i + 1, j) - F.u(i + 1, j - 1)) / a.dy) - ew * ((F.u(i, j) - F.u(i, j - 1)) / a.dy)) / a.dx;
  return (shear + normal) + cross;
}

// One row's contributions in the reference's insertion order. col < 0: no matrix contribution.
struct Row {
  int col[kMomRow];
  double val[kMomRow];
  bool sub[kMomRow];  // entry -= value (a Dirichlet wall's reflected neighbour)
};

__device__ __forceinline__ int wrap(int k, int m, bool periodic) {
  return periodic ? (k < 0 ? m - 1 : (k >= m ? 0 : k)) : k;
}

__device__ __forceinline__ bool u_active(const MomArgs &a, int i, int j) {
  return j >= 0 && j < a.ny && i >= a.iu0 && i <= a.nx - 1;
}
__device__ __forceinline__ bool v_active(const MomArgs &a, int i, int j) {
  return i >= 0 && i < a.nx && j >= a.jv0 && j <= a.ny - 1;
}
__device__ __forceinline__ int u_row(const MomArgs &a, int i, int j) { return (i - a.iu0) * a.ny + j; }
__device__ __forceinline__ int v_row(const MomArgs &a, int i, int j) { return a.nu + i * a.njv + (j - a.jv0); }

// add_u_x_entry / add_u_y_entry / add_v_y_entry / add_v_x_entry (momentum.cpp:287-381) into slot k.
// The inactive neighbours that reach these are wall faces (periodic indices are wrapped first).
__device__ __forceinline__ void add_u_x(const MomArgs &a, Row &r, int k, double &b, int ui, int uj, double value) {
  r.col[k] = -1;
  r.val[k] = value;
  r.sub[k] = false;
  if (u_active(a, ui, uj)) {
    r.col[k] = u_row(a, ui, uj);
    return;
  }
  const int side = ui <= 0 ? 0 : 1;
  if (a.u_type[side] == PP_BC_DIRICHLET) {
    b -= value * a.u_value[side];
  } else {
    r.col[k] = u_row(a, side == 0 ? 1 : a.nx - 1, uj);
    b -= value * a.u_value[side] * a.dx;
  }
}

__device__ __forceinline__ void add_u_y(const MomArgs &a, Row &r, int k, double &b, int ui, int uj, double value) {
  r.val[k] = value;
  r.sub[k] = false;
  if (u_active(a, ui, uj)) {
    r.col[k] = u_row(a, ui, uj);
    return;
  }
  const int side = uj < 0 ? 2 : 3;
  r.col[k] = u_row(a, ui, side == 2 ? 0 : a.ny - 1);
  if (a.u_type[side] == PP_BC_DIRICHLET) {
    r.sub[k] = true;
    b -= value * (2.0 * a.u_value[side]);
  } else {
    b -= value * a.u_value[side] * a.dy;
  }
}

__device__ __forceinline__ void add_v_y(const MomArgs &a, Row &r, int k, double &b, int vi, int vj, double value) {
  r.col[k] = -1;
  r.val[k] = value;
  r.sub[k] = false;
  if (v_active(a, vi, vj)) {
    r.col[k] = v_row(a, vi, vj);
    return;
  }
  const int side = vj <= 0 ? 2 : 3;
  if (a.v_type[side] == PP_BC_DIRICHLET) {
    b -= value * a.v_value[side];
  } else {
    r.col[k] = v_row(a, vi, side == 2 ? 1 : a.ny - 1);
    b -= value * a.v_value[side] * a.dy;
  }
}

__device__ __forceinline__ void add_v_x(const MomArgs &a, Row &r, int k, double &b, int vi, int vj, double value) {
  r.val[k] = value;
  r.sub[k] = false;
  if (v_active(a, vi, vj)) {
    r.col[k] = v_row(a, vi, vj);
    return;
  }
  const int side = vi < 0 ? 0 : 1;
  r.col[k] = v_row(a, side == 0 ? 0 : a.nx - 1, vj);
  if (a.v_type[side] == PP_BC_DIRICHLET) {
    r.sub[k] = true;
    b -= value * (2.0 * a.v_value[side]);
  } else {
    b -= value * a.v_value[side] * a.dx;
  }
}

// Merge, sort and emit: entry k is kept when it is the first occurrence of its column and its folded
// value is not below 1e-18; it goes to its rank among the kept columns.
__device__ __forceinline__ void emit_row(const MomArgs &a, int row, const Row &r) {
  double acc[kMomRow];
  bool keep[kMomRow];
#pragma unroll
  for (int k = 0; k < kMomRow; ++k) {
    bool first = r.col[k] >= 0;
#pragma unroll
    for (int q = 0; q < k; ++q) first = first && r.col[q] != r.col[k];
    double s = 0.0;
#pragma unroll
    for (int q = k; q < kMomRow; ++q)
      if (r.col[q] == r.col[k]) s = r.sub[q] ? s - r.val[q] : s + r.val[q];
    acc[k] = s;
    keep[k] = first && !(fabs(s) < 1.0e-18);
  }
  int kept = 0;
#pragma unroll
  for (int k = 0; k < kMomRow; ++k) {
    int rank = 0;
#pragma unroll
    for (int q = 0; q < kMomRow; ++q) rank += (keep[q] && r.col[q] < r.col[k]) ? 1 : 0;
    if (keep[k]) {
      a.slot_col[static_cast<size_t>(row) * kMomRow + rank] = r.col[k];
      a.slot_val[static_cast<size_t>(row) * kMomRow + rank] = acc[k];
    }
    kept += keep[k] ? 1 : 0;
  }
  a.count[row] = kept;
}

__global__ void __launch_bounds__(kBlock) momentum_assemble_kernel(MomArgs a) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= a.n) return;
  const Fields F(a);
  const bool px = a.periodic_x != 0, py = a.periodic_y != 0;
  const int nx = a.nx, ny = a.ny;
  const double alpha = a.alpha, dx = a.dx, dy = a.dy, dt = a.dt;
  Row r;
  double b;
  if (row < a.nu) {  // u face (i, j), momentum.cpp:166-175, 383-413
    const int i = a.iu0 + row / ny, j = row - (row / ny) * ny;
    const GhostedConst prev(a.u_prev, ny, a.G), sfx(a.sfx, ny, a.G);
    const double rf = face_x(a, F.rho, i, j);
    const double hu = conv_u(a, F, i, j);
    const double conv = a.step_index == 0 ? hu : 1.5 * hu - 0.5 * prev(i, j);
    b = rf * F.u(i, j) + dt * (-rf * conv + rf * a.body_force_x + sfx(i, j)) + alpha * stress_u(a, F, i, j);
    a.x[row] = F.u(i, j);
    const double ee = face_x(a, F.eta, i + 1, j), ew = face_x(a, F.eta, i, j);
    const double en = corner(F.eta, i, j + 1), es = corner(F.eta, i, j);
    const double ce = 2.0 * ee / (dx * dx), cw = 2.0 * ew / (dx * dx), cn = en / (dy * dy), cs = es / (dy * dy);
    const double xn = en / (dx * dy), xs = es / (dx * dy);
    r.col[0] = row;
    r.val[0] = rf + alpha * (ce + cw + cn + cs);
    r.sub[0] = false;
    add_u_x(a, r, 1, b, wrap(i + 1, nx, px), j, -alpha * ce);
    add_u_x(a, r, 2, b, wrap(i - 1, nx, px), j, -alpha * cw);
    add_u_y(a, r, 3, b, i, wrap(j + 1, ny, py), -alpha * cn);
    add_u_y(a, r, 4, b, i, wrap(j - 1, ny, py), -alpha * cs);
    add_v_y(a, r, 5, b, wrap(i, nx, px), wrap(j + 1, ny, py), -alpha * xn);
    add_v_y(a, r, 6, b, wrap(i - 1, nx, px), wrap(j + 1, ny, py), alpha * xn);
    add_v_y(a, r, 7, b, wrap(i, nx, px), wrap(j, ny, py), alpha * xs);
    add_v_y(a, r, 8, b, wrap(i - 1, nx, px), wrap(j, ny, py), -alpha * xs);
  } else {  // v face (i, j), momentum.cpp:177-186, 415-446
    const int q = row - a.nu;
    const int i = q / a.njv, j = a.jv0 + (q - (q / a.njv) * a.njv);
    const GhostedConst prev(a.v_prev, ny + 1, a.G), sfy(a.sfy, ny + 1, a.G);
    const double rf = face_y(a, F.rho, i, j);
    const double hv = conv_v(a, F, i, j);
    const double conv = a.step_index == 0 ? hv : 1.5 * hv - 0.5 * prev(i, j);
    b = rf * F.v(i, j) + dt * (-rf * conv + rf * a.body_force_y + sfy(i, j)) + alpha * stress_v(a, F, i, j);
    a.x[row] = F.v(i, j);
    const double ee = corner(F.eta, i + 1, j), ew = corner(F.eta, i, j);
    const double en = face_y(a, F.eta, i, j + 1), es = face_y(a, F.eta, i, j);
    const double ce = ee / (dx * dx), cw = ew / (dx * dx), cn = 2.0 * en / (dy * dy), cs = 2.0 * es / (dy * dy);
    const double xe = ee / (dx * dy), xw = ew / (dx * dy);
    r.col[0] = row;
    r.val[0] = rf + alpha * (ce + cw + cn + cs);
    r.sub[0] = false;
    add_v_x(a, r, 1, b, wrap(i + 1, nx, px), j, -alpha * ce);
    add_v_x(a, r, 2, b, wrap(i - 1, nx, px), j, -alpha * cw);
    add_v_y(a, r, 3, b, i, wrap(j + 1, ny, py), -alpha * cn);
    add_v_y(a, r, 4, b, i, wrap(j - 1, ny, py), -alpha * cs);
    add_u_x(a, r, 5, b, wrap(i + 1, nx, px), wrap(j, ny, py), -alpha * xe);
    add_u_x(a, r, 6, b, wrap(i + 1, nx, px), wrap(j - 1, ny, py), alpha * xe);
    add_u_x(a, r, 7, b, wrap(i, nx, px), wrap(j, ny, py), alpha * xw);
    add_u_x(a, r, 8, b, wrap(i, nx, px), wrap(j - 1, ny, py), -alpha * xw);
  }
  a.rhs[row] = b;
  emit_row(a, row, r);
}

// Block-wide exclusive prefix of one int per thread (kMomScanBlock threads); returns the block total.
__device__ __forceinline__ int block_exclusive_scan(int v, int *lds_words, int &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(incl, off, 64);
    incl += lane >= off ? o : 0;
  }
  if (lane == 63) lds_words[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < kMomScanBlock / 64; ++w) {
      const int t = lds_words[w];
      lds_words[w] = run;
      run += t;
    }
    lds_words[kMomScanBlock / 64] = run;
  }
  __syncthreads();
  total = lds_words[kMomScanBlock / 64];
  return lds_words[wave] + incl - v;
}

// pass 1: entries per scan block
__global__ void __launch_bounds__(kMomScanBlock) momentum_count_kernel(MomArgs a) {
  __shared__ int words[kMomScanBlock / 64 + 1];
  const int row = blockIdx.x * kMomScanBlock + threadIdx.x;
  int total = 0;
  block_exclusive_scan(row < a.n ? a.count[row] : 0, words, total);
  if (threadIdx.x == 0) a.block_sum[blockIdx.x] = total;
}

// pass 2: exclusive prefix over the block sums (one block, any number of scan blocks)
__global__ void __launch_bounds__(kMomScanBlock) momentum_block_scan_kernel(MomArgs a, int blocks) {
  __shared__ int words[kMomScanBlock / 64 + 1];
  int carry = 0;
  for (int base = 0; base < blocks; base += kMomScanBlock) {
    const int k = base + threadIdx.x;
    const int v = k < blocks ? a.block_sum[k] : 0;
    int total = 0;
    const int ex = block_exclusive_scan(v, words, total);
    __syncthreads();
    if (k < blocks) a.block_sum[k] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) a.row_ptr[a.n] = carry;
}

// pass 3: row pointers and the compacted entries
__global__ void __launch_bounds__(kMomScanBlock) momentum_compact_kernel(MomArgs a) {
  __shared__ int words[kMomScanBlock / 64 + 1];
  const int row = blockIdx.x * kMomScanBlock + threadIdx.x;
  const int cnt = row < a.n ? a.count[row] : 0;
  int total = 0;
  const int start = a.block_sum[blockIdx.x] + block_exclusive_scan(cnt, words, total);
  if (row < a.n) {
    a.row_ptr[row] = start;
    for (int k = 0; k < cnt; ++k) {
      a.col[start + k] = a.slot_col[static_cast<size_t>(row) * kMomRow + k];
      a.val[start + k] = a.slot_val[static_cast<size_t>(row) * kMomRow + k];
    }
  }
}

// This step's convection fields into the u_prev / v_prev buffers: the interiors, zero ghosts (the
// reference fills fresh fields with 0.0 before apply_u_bc / apply_v_bc). Runs after the assembly,
// which read the previous step's values.
__global__ void __launch_bounds__(kBlock) momentum_convection_kernel(MomArgs a) {
  const int G = a.G, nx = a.nx, ny = a.ny;
  const size_t su = static_cast<size_t>(nx + 1 + 2 * G) * (ny + 2 * G);
  const size_t sv = static_cast<size_t>(nx + 2 * G) * (ny + 1 + 2 * G);
  const size_t t = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  const Fields F(a);
  if (t < su) {
    const int ld = ny + 2 * G;
    const int i = static_cast<int>(t / ld) - G, j = static_cast<int>(t % ld) - G;
    const bool in = i >= 0 && i <= nx && j >= 0 && j < ny;
    as_global(a.u_prev)[t] = in ? conv_u(a, F, i, j) : 0.0;
  } else if (t < su + sv) {
    const size_t e = t - su;
    const int ld = ny + 1 + 2 * G;
    const int i = static_cast<int>(e / ld) - G, j = static_cast<int>(e % ld) - G;
    const bool in = i >= 0 && i < nx && j >= 0 && j <= ny;
    as_global(a.v_prev)[e] = in ? conv_v(a, F, i, j) : 0.0;
  }
}

// u_star_ = u_, v_star_ = v_ with the active faces from the solution (momentum.cpp:474-489).
__global__ void __launch_bounds__(kBlock) momentum_scatter_kernel(MomArgs a) {
  const int G = a.G, nx = a.nx, ny = a.ny;
  const size_t su = static_cast<size_t>(nx + 1 + 2 * G) * (ny + 2 * G);
  const size_t sv = static_cast<size_t>(nx + 2 * G) * (ny + 1 + 2 * G);
  const size_t t = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t < su) {
    const int ld = ny + 2 * G;
    const int i = static_cast<int>(t / ld) - G, j = static_cast<int>(t % ld) - G;
    as_global(a.u_star)[t] = u_active(a, i, j) ? a.x[u_row(a, i, j)] : as_global(a.u)[t];
  } else if (t < su + sv) {
    const size_t e = t - su;
    const int ld = ny + 1 + 2 * G;
    const int i = static_cast<int>(e / ld) - G, j = static_cast<int>(e % ld) - G;
    as_global(a.v_star)[e] = v_active(a, i, j) ? a.x[v_row(a, i, j)] : as_global(a.v)[e];
  }
}

// Ghost fills: block 0 apply_u_bc and block 1 apply_v_bc on (u_prev, v_prev) or (u_star, v_star).
__global__ void __launch_bounds__(kThreads) momentum_ghost_kernel(MomArgs a, int star) {
  if (blockIdx.x == 0)
    u_bc(a, Ghosted{star ? a.u_star : a.u_prev, a.ny, a.G});
  else
    v_bc(a, Ghosted{star ? a.v_star : a.v_prev, a.ny + 1, a.G});
}

}  // namespace

hipError_t launch_momentum_system(const MomArgs &a, hipStream_t stream) {
  const int rows_blocks = (a.n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(momentum_assemble_kernel, dim3(rows_blocks), dim3(kBlock), 0, stream, a);
  const int scan_blocks = (a.n + kMomScanBlock - 1) / kMomScanBlock;
  hipLaunchKernelGGL(momentum_count_kernel, dim3(scan_blocks), dim3(kMomScanBlock), 0, stream, a);
  hipLaunchKernelGGL(momentum_block_scan_kernel, dim3(1), dim3(kMomScanBlock), 0, stream, a, scan_blocks);
  hipLaunchKernelGGL(momentum_compact_kernel, dim3(scan_blocks), dim3(kMomScanBlock), 0, stream, a);
  const size_t cells = static_cast<size_t>(a.nx + 1 + 2 * a.G) * (a.ny + 2 * a.G) +
                       static_cast<size_t>(a.nx + 2 * a.G) * (a.ny + 1 + 2 * a.G);
  hipLaunchKernelGGL(momentum_convection_kernel, dim3(static_cast<unsigned>((cells + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(momentum_ghost_kernel, dim3(2), dim3(kThreads), 0, stream, a, 0);
  return hipGetLastError();
}

hipError_t launch_momentum_scatter(const MomArgs &a, hipStream_t stream) {
  const size_t cells = static_cast<size_t>(a.nx + 1 + 2 * a.G) * (a.ny + 2 * a.G) +
                       static_cast<size_t>(a.nx + 2 * a.G) * (a.ny + 1 + 2 * a.G);
  hipLaunchKernelGGL(momentum_scatter_kernel, dim3(static_cast<unsigned>((cells + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, stream, a);
  hipLaunchKernelGGL(momentum_ghost_kernel, dim3(2), dim3(kThreads), 0, stream, a, 1);
  return hipGetLastError();
}

}  // namespace pp
