// Cahn-Hilliard semi-implicit linear solve on the device (Solver::solve_phase_linear_system_eq25,
// src/core/coupled/phase_field.cpp:437-503): ILDLT(0) of the 13-point operator
// A = alpha0 I - beta1 Lap + beta2 Lap^2 (ch_sparse_krylov.hpp:252-335) and PCG without projection
// from x0 = c (ch_sparse_krylov.hpp:515-660) with that preconditioner or the diagonal one
// (421-431).
//
// The factorization and both triangular sweeps run level by level: a row's lower entries all lie
// on earlier levels (level = 1 + the largest level of its lower columns, set on the host from the
// pattern), so the rows of one level are independent and one workgroup processes them in parallel
// with a barrier between levels. Within a row every operation is the reference's, in its order:
// the factor's cross terms (find_lower_entry), the forward sum in CSR order, the division by
// max(D, floor) after the whole forward sweep, the backward sum over the children in ascending row
// order. For the 13-point stencil in natural order the levels are the anti-diagonals d = 2i + j
// (2 nx + ny - 2 of them); a periodic y direction chains (i, 0) to (i - 1, ny - 1) and the levels
// become the rows themselves.
//
// Reductions (dot products, norms) use the 512-thread canonical tree over the natural row order:
// thread t sums rows t, t + 512, ..., then the 64-lane tree (offsets 32..1) and the 8-wave tree
// (offsets 4..1). The CSR row sums of the SpMV follow apply_matrix (235-250).
#include <hip/hip_runtime.h>

#include "pp_device.hpp"
#include "pp_kernels.hpp"

namespace pp {

namespace {

struct ChShared {
  double red[2][kWaves * 4];
  volatile int flag;
};

template <int K>
__device__ __forceinline__ void ch_block_sum(double (&v)[K], ChShared &sh, int &parity) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = v[k] + __shfl_down(v[k], off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *slot = sh.red[parity];
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) slot[k * kWaves + wave] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double *src = slot + k * kWaves;
    double w[kWaves / 2];
#pragma unroll
    for (int t = 0; t < kWaves / 2; ++t) w[t] = src[t] + src[t + kWaves / 2];
#pragma unroll
    for (int off = kWaves / 4; off >= 1; off >>= 1) {
#pragma unroll
      for (int t = 0; t < off; ++t) w[t] = w[t] + w[t + off];
    }
    v[k] = w[0];
  }
  parity ^= 1;
}

__device__ __forceinline__ double ch_sum1(double v, ChShared &sh, int &parity) {
  double a[1] = {v};
  ch_block_sum<1>(a, sh, parity);
  return a[0];
}

// L(row, target) for target < row, 0 when absent (find_lower_entry, ch_sparse_krylov.hpp:265-276)
__device__ __forceinline__ double lower_of(const ChArgs &a, int row, int target) {
  for (int p = a.row_ptr[row]; p < a.row_ptr[row + 1]; ++p) {
    const int c = a.col[p];
    if (c >= row || c > target) break;
    if (c == target) return a.lower[p];
  }
  return 0.0;
}

__global__ void __launch_bounds__(kThreads) ch_factor_kernel(ChArgs a) {
  __shared__ ChShared sh;
  const int tid = threadIdx.x;
  // diagonal preconditioner and the pivot floor from max |a_ii| (ch_sparse_krylov.hpp:258-263, 421-431)
  double m = 0.0;
  for (int r = tid; r < a.n; r += kThreads) {
    const double d = a.val[a.diag_pos[r]];
    const double ad = fabs(d);
    a.dinv[r] = 1.0 / (ad < 1.0e-30 ? 1.0e-30 : ad);
    m = m < ad ? ad : m;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_down(m, off, 64);
    m = m < o ? o : m;
  }
  if (tid == 0) sh.flag = 0;
  if ((tid & 63) == 0) sh.red[0][tid >> 6] = m;
  __syncthreads();
  double max_diag = sh.red[0][0];
  for (int w = 1; w < kWaves; ++w) max_diag = max_diag < sh.red[0][w] ? sh.red[0][w] : max_diag;
  const double floor = (1.0 < max_diag ? max_diag : 1.0) * 1.0e-12;
  if (tid == 0) a.state->floor = floor;
  // ILDLT(0), level by level (ch_sparse_krylov.hpp:278-320)
  bool failed = false;
  for (int lev = 0; lev < a.nlev; ++lev) {
    for (int e = a.lev_ptr[lev] + tid; e < a.lev_ptr[lev + 1]; e += kThreads) {
      const int row = a.lev_rows[e];
      const int beg = a.row_ptr[row], end = a.row_ptr[row + 1];
      for (int p = beg; p < end; ++p) {
        const int c = a.col[p];
        if (c >= row) break;
        double s = a.val[p];
        for (int q = beg; q < p; ++q) {
          const int k = a.col[q];
          if (k >= c) break;
          const double l_ck = lower_of(a, c, k);
          if (l_ck != 0.0) s -= a.lower[q] * a.D[k] * l_ck;
        }
        const double dc = a.D[c];
        a.lower[p] = s / (dc < floor ? floor : dc);
      }
      double d = a.val[a.diag_pos[row]];
      for (int p = beg; p < end; ++p) {
        const int c = a.col[p];
        if (c >= row) break;
        d -= (a.lower[p] * a.lower[p]) * a.D[c];
      }
      failed = failed || !isfinite(d) || d <= floor;
      a.D[row] = d;
    }
    __syncthreads();
  }
  if (failed) sh.flag = 1;
  __syncthreads();
  if (tid == 0) a.state->ildlt_ok = sh.flag ? 0 : 1;
}

// z = M^{-1} r with the ILDLT factor (ch_sparse_krylov.hpp:486-512) or the diagonal (446-452).
__device__ void ch_precondition(const ChArgs &a, bool ildlt, double floor) {
  const int tid = threadIdx.x;
  if (!ildlt) {
    for (int r = tid; r < a.n; r += kThreads) a.z[r] = a.dinv[r] * a.r[r];
    __syncthreads();
    return;
  }
  for (int lev = 0; lev < a.nlev; ++lev) {
    for (int e = a.lev_ptr[lev] + tid; e < a.lev_ptr[lev + 1]; e += kThreads) {
      const int row = a.lev_rows[e];
      double s = a.r[row];
      for (int p = a.row_ptr[row]; p < a.row_ptr[row + 1]; ++p) {
        const int c = a.col[p];
        if (c >= row) break;
        s -= a.lower[p] * a.y[c];
      }
      a.y[row] = s;
    }
    __syncthreads();
  }
  for (int r = tid; r < a.n; r += kThreads) {
    const double d = a.D[r];
    a.z[r] = a.y[r] / (d < floor ? floor : d);
  }
  __syncthreads();
  for (int lev = a.nlev - 1; lev >= 0; --lev) {
    for (int e = a.lev_ptr[lev] + tid; e < a.lev_ptr[lev + 1]; e += kThreads) {
      const int row = a.lev_rows[e];
      double s = a.z[row];
      for (int t = a.lt_ptr[row]; t < a.lt_ptr[row + 1]; ++t) s -= a.lower[a.lt_pos[t]] * a.z[a.lt_row[t]];
      a.z[row] = s;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ double row_times(const ChArgs &a, const double *x, int row) {
  double s = 0.0;
  for (int p = a.row_ptr[row]; p < a.row_ptr[row + 1]; ++p) s += a.val[p] * x[a.col[p]];
  return s;
}

// PCG from x (ch_sparse_krylov.hpp:515-660, no projection); the stop and breakdown tests are the
// reference's. ildlt selects the preconditioner.
__global__ void __launch_bounds__(kThreads) ch_pcg_kernel(ChArgs a, int ildlt) {
  __shared__ ChShared sh;
  const int tid = threadIdx.x, n = a.n;
  int parity = 0;
  const double floor = a.state->floor;
  const bool use_ildlt = ildlt != 0;
  double acc[2] = {0.0, 0.0};
  for (int r = tid; r < n; r += kThreads) {
    const double rr = a.rhs[r] - row_times(a, a.x, r);
    a.r[r] = rr;
    acc[0] = acc[0] + a.rhs[r] * a.rhs[r];
    acc[1] = acc[1] + rr * rr;
  }
  ch_block_sum<2>(acc, sh, parity);
  const double bnorm = sqrt(acc[0] < 0.0 ? 0.0 : acc[0]);
  const double bden = bnorm < 1.0e-30 ? 1.0e-30 : bnorm;
  double abs_res = sqrt(acc[1] < 0.0 ? 0.0 : acc[1]);
  double rel_res = abs_res / bden, iter_res = 0.0;
  const double tol = a.tolerance;
  int iterations = 0, status = kChOk;
  if (!(rel_res < tol || abs_res < tol)) {
    ch_precondition(a, use_ildlt, floor);
    double rz = 0.0;
    for (int r = tid; r < n; r += kThreads) {
      rz = rz + a.r[r] * a.z[r];
      a.p[r] = a.z[r];
    }
    rz = ch_sum1(rz, sh, parity);
    if (!isfinite(rz) || rz <= 0.0) {
      status = kChRz0Fail;
    } else {
      const int limit = a.max_iterations < 1 ? 1 : a.max_iterations;
      for (int it = 0; it < limit; ++it) {
        double pap = 0.0;
        for (int r = tid; r < n; r += kThreads) {
          const double v = row_times(a, a.p, r);
          a.ap[r] = v;
          pap = pap + a.p[r] * v;
        }
        pap = ch_sum1(pap, sh, parity);
        if (!isfinite(pap) || fabs(pap) < 1.0e-30) {
          status = kChPapFail;
          break;
        }
        const double alpha = rz / pap;
        double up[3] = {0.0, 0.0, 0.0};
        for (int r = tid; r < n; r += kThreads) {
          const double delta = alpha * a.p[r];
          const double xn = a.x[r] + delta;
          const double rn = a.r[r] - alpha * a.ap[r];
          a.x[r] = xn;
          a.r[r] = rn;
          up[0] = up[0] + delta * delta;
          up[1] = up[1] + xn * xn;
          up[2] = up[2] + rn * rn;
        }
        ch_block_sum<3>(up, sh, parity);
        iterations = it + 1;
        iter_res = sqrt(up[0] / (up[1] < 1.0e-30 ? 1.0e-30 : up[1]));
        abs_res = sqrt(up[2] < 0.0 ? 0.0 : up[2]);
        rel_res = abs_res / bden;
        if (rel_res < tol || abs_res < tol) break;
        ch_precondition(a, use_ildlt, floor);
        double rzn = 0.0;
        for (int r = tid; r < n; r += kThreads) rzn = rzn + a.r[r] * a.z[r];
        rzn = ch_sum1(rzn, sh, parity);
        if (!isfinite(rzn) || rzn <= 0.0) {
          status = kChRzFail;
          break;
        }
        const double beta = rzn / rz;
        for (int r = tid; r < n; r += kThreads) a.p[r] = a.z[r] + beta * a.p[r];
        rz = rzn;
        __syncthreads();
      }
    }
  }
  // the equation residual |rhs - A x| / |rhs| of the returned x (phase_field.cpp:481-502)
  __syncthreads();
  double eq = 0.0;
  for (int r = tid; r < n; r += kThreads) {
    const double d = a.rhs[r] - row_times(a, a.x, r);
    eq = eq + d * d;
  }
  eq = ch_sum1(eq, sh, parity);
  if (tid == 0) {
    ChState &s = *a.state;
    s.status = status;
    s.iterations = iterations;
    s.abs_res = abs_res;
    s.rel_res = rel_res;
    s.iter_res = iter_res;
    s.eq_res = sqrt(eq / (acc[0] < 1.0e-30 ? 1.0e-30 : acc[0]));
  }
}

}  // namespace

hipError_t launch_ch_factor(const ChArgs &a, hipStream_t stream) {
  hipLaunchKernelGGL(ch_factor_kernel, dim3(1), dim3(kThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_ch_pcg(const ChArgs &a, bool ildlt, hipStream_t stream) {
  hipLaunchKernelGGL(ch_pcg_kernel, dim3(1), dim3(kThreads), 0, stream, a, ildlt ? 1 : 0);
  return hipGetLastError();
}

}  // namespace pp
