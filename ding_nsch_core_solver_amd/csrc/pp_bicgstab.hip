// BiCGSTAB with the diagonal preconditioner for the momentum predictor's CSR system.
//
// The reference solves it on the host with build_diagonal_preconditioner (ch_sparse_krylov.hpp:
// 421-431) and solve_preconditioned_bicgstab (839-963), called at momentum.cpp:469-472. Here every
// vector pass is a full-chip kernel: 512-thread workgroups, workgroup c owning rows [512 c, 512 c
// + 512), one thread per row. SpMV row sums run over the row's CSR entries in order from 0.0
// (apply_matrix, 235-250), and every vector update is the reference's expression term for term,
// so only the dot products differ from the reference, by summation order: each kernel leaves one
// partial per workgroup (the canonical 64-lane and 8-wave trees over its 512 rows) and the host
// adds the partials in workgroup order. The scalar recurrence (rho, alpha, omega, beta, the
// breakdown checks and the stopping tests) runs on the host between launches.
#include <hip/hip_runtime.h>

#include "pp_kernels.hpp"

namespace pp {
namespace {

constexpr int kRows = kBicgRows;  // rows per workgroup, one per thread

// K per-thread values summed over the workgroup: lanes with offsets 32..1, then the 8 wave sums
// with offsets 4, 2, 1. Thread 0 stores sum k at partials[k * chunks + blockIdx.x].
template <int K>
__device__ __forceinline__ void chunk_partials(double (&v)[K], double *partials, int chunks) {
  __shared__ double red[K * 8];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = v[k] + __shfl_down(v[k], off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k * 8 + wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double w[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) w[t] = red[k * 8 + t] + red[k * 8 + t + 4];
      w[0] = w[0] + w[2];
      w[1] = w[1] + w[3];
      partials[k * chunks + blockIdx.x] = w[0] + w[1];
    }
  }
}

__device__ __forceinline__ double csr_row(const int *__restrict__ rp, const int *__restrict__ ci,
                                          const double *__restrict__ val, const double *__restrict__ x, int row) {
  double sum = 0.0;
  const int end = rp[row + 1];
  for (int pos = rp[row]; pos < end; ++pos) sum += val[pos] * x[ci[pos]];
  return sum;
}

// Diagonal preconditioner (1 / max(|a_ii|, 1e-30), the diagonal entry located as
// finalize_row_maps records it), r = b - A x, r_hat = r, p = v = 0; partials b.b, r.r.
__global__ void __launch_bounds__(kRows) bicg_setup_kernel(BicgArgs a) {
  const int q = blockIdx.x * kRows + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < a.n) {
    int diag = -1;
    for (int pos = a.row_ptr[q]; pos < a.row_ptr[q + 1]; ++pos)
      if (a.col_idx[pos] == q) diag = pos;
    double d = 1.0;
    if (diag < 0) {
      *a.flag = 1;
    } else {
      const double m = fabs(a.values[diag]);
      d = 1.0 / (m < 1.0e-30 ? 1.0e-30 : m);
    }
    a.dinv[q] = d;
    const double ax = csr_row(a.row_ptr, a.col_idx, a.values, a.x, q);
    const double b = a.rhs[q];
    const double r = b - ax;
    a.r[q] = r;
    a.r_hat[q] = r;
    a.p[q] = 0.0;
    a.v[q] = 0.0;
    v[0] = v[0] + b * b;
    v[1] = v[1] + r * r;
  }
  chunk_partials<2>(v, a.partials, a.chunks);
}

// p = r + beta (p - omega v); p_hat = M^-1 p
__global__ void __launch_bounds__(kRows) bicg_direction_kernel(BicgArgs a, double beta, double omega) {
  const int q = blockIdx.x * kRows + threadIdx.x;
  if (q >= a.n) return;
  const double p = a.r[q] + beta * (a.p[q] - omega * a.v[q]);
  a.p[q] = p;
  a.p_hat[q] = a.dinv[q] * p;
}

// v = A p_hat; partial r_hat.v
__global__ void __launch_bounds__(kRows) bicg_spmv_v_kernel(BicgArgs a) {
  const int q = blockIdx.x * kRows + threadIdx.x;
  double v[1] = {0.0};
  if (q < a.n) {
    const double y = csr_row(a.row_ptr, a.col_idx, a.values, a.p_hat, q);
    a.v[q] = y;
    v[0] = v[0] + a.r_hat[q] * y;
  }
  chunk_partials<1>(v, a.partials, a.chunks);
}

// s = r - alpha v; update = alpha p_hat; s_hat = M^-1 s; partial s.s
__global__ void __launch_bounds__(kRows) bicg_half_step_kernel(BicgArgs a, double alpha) {
  const int q = blockIdx.x * kRows + threadIdx.x;
  double v[1] = {0.0};
  if (q < a.n) {
    const double s = a.r[q] - alpha * a.v[q];
    a.s[q] = s;
    a.update[q] = alpha * a.p_hat[q];
    a.s_hat[q] = a.dinv[q] * s;
    v[0] = v[0] + s * s;
  }
  chunk_partials<1>(v, a.partials, a.chunks);
}

// converged at the half step: x += update; partials update.update, x.x
__global__ void __launch_bounds__(kRows) bicg_half_exit_kernel(BicgArgs a) {
  const int q = blockIdx.x * kRows + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < a.n) {
    const double u = a.update[q];
    const double x = a.x[q] + u;
    a.x[q] = x;
    v[0] = v[0] + u * u;
    v[1] = v[1] + x * x;
  }
  chunk_partials<2>(v, a.partials, a.chunks);
}

// t = A s_hat; partials t.t, t.s
__global__ void __launch_bounds__(kRows) bicg_spmv_t_kernel(BicgArgs a) {
  const int q = blockIdx.x * kRows + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < a.n) {
    const double t = csr_row(a.row_ptr, a.col_idx, a.values, a.s_hat, q);
    a.t[q] = t;
    v[0] = v[0] + t * t;
    v[1] = v[1] + t * a.s[q];
  }
  chunk_partials<2>(v, a.partials, a.chunks);
}

// update += omega s_hat; x += update; r = s - omega t; partials update.update, x.x, r.r, r_hat.r
__global__ void __launch_bounds__(kRows) bicg_full_step_kernel(BicgArgs a, double omega) {
  const int q = blockIdx.x * kRows + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (q < a.n) {
    const double u = a.update[q] + omega * a.s_hat[q];
    a.update[q] = u;
    const double x = a.x[q] + u;
    a.x[q] = x;
    const double r = a.s[q] - omega * a.t[q];
    a.r[q] = r;
    v[0] = v[0] + u * u;
    v[1] = v[1] + x * x;
    v[2] = v[2] + r * r;
    v[3] = v[3] + a.r_hat[q] * r;
  }
  chunk_partials<4>(v, a.partials, a.chunks);
}

}  // namespace

hipError_t launch_bicg(BicgStep step, const BicgArgs &a, double s0, double s1, hipStream_t stream) {
  const dim3 grid(a.chunks), block(kRows);
  switch (step) {
    case BicgStep::kSetup: bicg_setup_kernel<<<grid, block, 0, stream>>>(a); break;
    case BicgStep::kDirection: bicg_direction_kernel<<<grid, block, 0, stream>>>(a, s0, s1); break;
    case BicgStep::kSpmvV: bicg_spmv_v_kernel<<<grid, block, 0, stream>>>(a); break;
    case BicgStep::kHalfStep: bicg_half_step_kernel<<<grid, block, 0, stream>>>(a, s0); break;
    case BicgStep::kHalfExit: bicg_half_exit_kernel<<<grid, block, 0, stream>>>(a); break;
    case BicgStep::kSpmvT: bicg_spmv_t_kernel<<<grid, block, 0, stream>>>(a); break;
    case BicgStep::kFullStep: bicg_full_step_kernel<<<grid, block, 0, stream>>>(a, s0); break;
  }
  return hipGetLastError();
}

}  // namespace pp
