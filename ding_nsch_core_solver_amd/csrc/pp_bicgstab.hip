// BiCGSTAB with the diagonal preconditioner for the momentum predictor's CSR system.
//
// The reference solves it on the host with build_diagonal_preconditioner (ch_sparse_krylov.hpp:
// 421-431) and solve_preconditioned_bicgstab (839-963), called at momentum.cpp:469-472. Here every
// vector pass is a full-chip kernel: 512-thread workgroups, each owning one chunk of 512 rows
// (paired_chunk: each XCD takes interleaved u and v chunks of the same cells), one thread per row. SpMV row sums run over the row's
// CSR entries in order from 0.0 (apply_matrix, 235-250), and every vector update is the
// reference's expression term for term, so only the dot products differ from the reference, by
// summation order: each pass leaves one partial per chunk (the canonical 64-lane and 8-wave trees
// over its 512 rows), and the partials are reduced the same way (thread t adds chunks t, t + 512,
// ... in order, then the same trees).
//
// The scalar recurrence lives on the device. Each pass starts by forming the scalars it needs
// from the previous pass's partials (every workgroup computes the same values in the same order)
// and runs the reference's test at that point of the iteration: a stop (convergence or a
// breakdown) is decided identically by every workgroup, which then write nothing, and workgroup 0
// records it in the state block. Every later pass sees the state and returns at once, so the host
// can queue several iterations between its checks.
#include <hip/hip_runtime.h>

#include <math.h>

#include "pp_chunked.hpp"
#include "pp_kernels.hpp"

namespace pp {
namespace {

constexpr int kRows = kBicgRows;  // rows per workgroup, one per thread
static_assert(kRows == chunked::kRows, "BiCGSTAB chunks are the shared 512-row chunks");
using chunked::paired_chunk;
using chunked::chunk_partials;
using chunked::dot_sums;

// The state's status and half-exit flag as one value per workgroup: thread 0 reads them and the
// barrier hands them to every wave, so a workgroup never splits on a stop another one records.
struct StateFlags {
  int status, half_exit;
};
__device__ __forceinline__ StateFlags block_flags(const BicgArgs &a) {
  __shared__ int flags[2];
  if (threadIdx.x == 0) {
    flags[0] = a.state->status;
    flags[1] = a.state->half_exit;
  }
  __syncthreads();
  return {flags[0], flags[1]};
}

// The K sums of partial slots [slot, slot + K): summed here by every workgroup (small systems),
// or read from the state where the sum pass left them (a.presummed: systems of more than
// kBicgPresumChunks workgroups, where every workgroup re-reading every partial would cost
// chunks^2 loads). Both add in the same order, so the bits do not depend on the choice.
template <int K>
__device__ __forceinline__ void slot_sums(const BicgArgs &a, int slot, double (&out)[K]) {
  if (a.presummed) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = a.state->sums[slot + k];
  } else {
    dot_sums<K>(a.partials + static_cast<size_t>(slot) * a.chunks, a.chunks, out);
  }
}

__device__ __forceinline__ double csr_row(const int *__restrict__ rp, const int *__restrict__ ci,
                                          const double *__restrict__ val, const double *__restrict__ x, int row) {
  double sum = 0.0;
  const int end = rp[row + 1];
  for (int pos = rp[row]; pos < end; ++pos) sum += val[pos] * x[ci[pos]];
  return sum;
}

__device__ __forceinline__ double max0(double v) { return v < 0.0 ? 0.0 : v; }      // std::max(v, 0.0)
__device__ __forceinline__ double floor30(double v) { return v < 1.0e-30 ? 1.0e-30 : v; }  // std::max(v, 1e-30)

__device__ __forceinline__ bool converged(const BicgArgs &a, double rel, double abs_res) {
  return (a.use_rel && rel < a.tolerance) || abs_res < a.tolerance;
}


// Workgroup 0 records a stop (status != 0) with the report fields.
__device__ __forceinline__ void record(const BicgArgs &a, int status, int iterations, double abs_res, double rel,
                                       double iter_res, bool with_iter_res) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.state->iterations = iterations;
    a.state->abs_res = abs_res;
    a.state->rel_res = rel;
    if (with_iter_res) a.state->iter_res = iter_res;
    a.state->status = status;
  }
}

// Diagonal preconditioner (1 / max(|a_ii|, 1e-30), the diagonal entry located as
// finalize_row_maps records it), r = b - A x, r_hat = r, p = v = 0 (both parities); partials b.b,
// r.r.
__global__ void __launch_bounds__(kRows) bicg_setup_kernel(BicgArgs a) {
  const int chunk = paired_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < a.n) {
    int diag = -1;
    for (int pos = a.row_ptr[q]; pos < a.row_ptr[q + 1]; ++pos)
      if (a.col_idx[pos] == q) diag = pos;
    double d = 1.0;
    if (diag < 0) {
      a.state->status = kBicgNoDiag;
    } else {
      d = 1.0 / floor30(fabs(a.values[diag]));
    }
    a.dinv[q] = d;
    const double ax = csr_row(a.row_ptr, a.col_idx, a.values, a.x, q);
    const double b = a.rhs[q];
    const double r = b - ax;
    a.r[q] = r;
    a.r_hat[q] = r;
    a.p[0][q] = 0.0;
    a.p[1][q] = 0.0;
    a.v[0][q] = 0.0;
    a.v[1][q] = 0.0;
    v[0] = v[0] + b * b;
    v[1] = v[1] + r * r;
  }
  chunk_partials<2>(v, a.p_setup, a.chunks, chunk);
}

// Pass A of iteration k (or the report after the last iteration, final): the stop test of the
// previous iteration (half-step exit, full-step convergence, or the initial residual at k = 0),
// the rho check, then p = r + beta (p_old - omega v_old), p_hat = M^-1 p and v = A p_hat; partial
// r_hat.v. A row's neighbours' p_hat are recomputed from r, p_old, v_old and M^-1 with the owner's
// expression (same bits), so p and v are double-buffered by iteration parity.
__global__ void __launch_bounds__(kRows) bicg_pass_a_kernel(BicgArgs a, int k, int final) {
  // loads that need no scalar are issued first, so their latency overlaps the partial sums
  const double *__restrict__ p_old = a.p[(k + 1) & 1];
  const double *__restrict__ v_old = a.v[(k + 1) & 1];
  const double *__restrict__ r = a.r;
  const double *__restrict__ dinv = a.dinv;
  const int chunk = paired_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  int beg = 0, end = 0;
  double rq = 0.0, pq_old = 0.0, vq_old = 0.0, dq = 0.0, rhq = 0.0;
  if (q < a.n) {
    beg = a.row_ptr[q];
    end = a.row_ptr[q + 1];
    rq = r[q];
    pq_old = p_old[q];
    vq_old = v_old[q];
    dq = dinv[q];
    rhq = a.r_hat[q];
  }
  const StateFlags fl = block_flags(a);
  if (fl.status != kBicgRunning) return;
  double bn, rho_new;
  if (k == 0) {
    double rr[2];
    slot_sums<2>(a, kSlotSetup, rr);  // b.b, r.r
    bn = sqrt(max0(rr[0]));
    if (blockIdx.x == 0 && threadIdx.x == 0) a.state->rhs_norm = bn;
    const double abs_res = sqrt(max0(rr[1]));
    const double rel = abs_res / floor30(bn);
    if (converged(a, rel, abs_res) || final) {
      record(a, kBicgDone, 0, abs_res, rel, 0.0, false);
      return;
    }
    rho_new = rr[1];  // r_hat = r: r_hat . r is r . r, summed in the same order
  } else if (fl.half_exit) {
    double d[2];
    slot_sums<2>(a, kSlotExit, d);  // update.update, x.x after x += update
    record(a, kBicgDone, k, a.state->abs_res, a.state->rel_res, sqrt(d[0] / floor30(d[1])), true);
    return;
  } else {
    bn = a.state->rhs_norm;
    double d[4];
    slot_sums<4>(a, kSlotX, d);  // update.update, x.x, r.r, r_hat.r of iteration k - 1
    const double iter_res = sqrt(d[0] / floor30(d[1]));
    const double abs_res = sqrt(max0(d[2]));
    const double rel = abs_res / floor30(bn);
    if (converged(a, rel, abs_res) || final) {
      record(a, kBicgDone, k, abs_res, rel, iter_res, true);
      return;
    }
    rho_new = d[3];
  }
  if (!isfinite(rho_new) || fabs(rho_new) < 1.0e-30) {
    record(a, kBicgRho, k, 0.0, 0.0, 0.0, false);
    return;
  }
  const double rho_old = k == 0 ? 1.0 : a.state->rho[(k - 1) & 1];
  const double alpha = k == 0 ? 1.0 : a.state->alpha;
  const double omega = k == 0 ? 1.0 : a.state->omega;
  const double beta = (k == 0) ? 0.0 : (rho_new / rho_old) * (alpha / omega);
  if (blockIdx.x == 0 && threadIdx.x == 0) a.state->rho[k & 1] = rho_new;
  double v[1] = {0.0};
  if (q < a.n) {
    const double pq = rq + beta * (pq_old - omega * vq_old);
    a.p[k & 1][q] = pq;
    a.p_hat[q] = dq * pq;
    double sum = 0.0;
    for (int pos = beg; pos < end; ++pos) {
      const int j = a.col_idx[pos];
      const double pj = r[j] + beta * (p_old[j] - omega * v_old[j]);
      sum += a.values[pos] * (dinv[j] * pj);
    }
    a.v[k & 1][q] = sum;
    v[0] = v[0] + rhq * sum;
  }
  chunk_partials<1>(v, a.p_v, a.chunks, chunk);
}

// Pass B: alpha = rho / r_hat.v (with its check); s = r - alpha v; update = alpha p_hat;
// s_hat = M^-1 s; t = A s_hat (neighbours' s_hat recomputed from r, v, M^-1); partials s.s, t.t,
// t.s. t is formed before the half-step test, which the next pass makes; on a half exit it goes
// unused.
__global__ void __launch_bounds__(kRows) bicg_pass_b_kernel(BicgArgs a, int k) {
  const double *__restrict__ vk = a.v[k & 1];
  const double *__restrict__ r = a.r;
  const double *__restrict__ dinv = a.dinv;
  const int chunk = paired_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  int beg = 0, end = 0;
  double rq = 0.0, vq = 0.0, dq = 0.0, phq = 0.0;
  if (q < a.n) {  // issued before the partial sums, as in pass A
    beg = a.row_ptr[q];
    end = a.row_ptr[q + 1];
    rq = r[q];
    vq = vk[q];
    dq = dinv[q];
    phq = a.p_hat[q];
  }
  const StateFlags fl = block_flags(a);
  if (fl.status != kBicgRunning) return;
  double d[1];
  slot_sums<1>(a, kSlotV, d);
  const double rhat_v = d[0];
  if (!isfinite(rhat_v) || fabs(rhat_v) < 1.0e-30) {
    record(a, kBicgRhatV, k, 0.0, 0.0, 0.0, false);
    return;
  }
  const double alpha = a.state->rho[k & 1] / rhat_v;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.state->alpha = alpha;
  double v[3] = {0.0, 0.0, 0.0};
  if (q < a.n) {
    const double sq = rq - alpha * vq;
    a.s[q] = sq;
    a.update[q] = alpha * phq;
    a.s_hat[q] = dq * sq;
    double sum = 0.0;
    for (int pos = beg; pos < end; ++pos) {
      const int j = a.col_idx[pos];
      const double sj = r[j] - alpha * vk[j];
      sum += a.values[pos] * (dinv[j] * sj);
    }
    a.t[q] = sum;
    v[0] = v[0] + sq * sq;
    v[1] = v[1] + sum * sum;
    v[2] = v[2] + sum * sq;
  }
  chunk_partials<3>(v, a.p_b, a.chunks, chunk);
}

// Pass C: the half-step test on |s| (converged: x += update, partials update.update, x.x, and
// workgroup 0 marks the half exit for the next pass A); else omega = t.s / t.t with both checks,
// update += omega s_hat; x += update; r = s - omega t; partials update.update, x.x, r.r, r_hat.r.
__global__ void __launch_bounds__(kRows) bicg_pass_c_kernel(BicgArgs a, int k) {
  const int chunk = paired_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  double uq = 0.0, xq = 0.0, shq = 0.0, sq = 0.0, tq = 0.0, rhq = 0.0;
  if (q < a.n) {  // issued before the partial sums, as in pass A
    uq = a.update[q];
    xq = a.x[q];
    shq = a.s_hat[q];
    sq = a.s[q];
    tq = a.t[q];
    rhq = a.r_hat[q];
  }
  const StateFlags fl = block_flags(a);
  if (fl.status != kBicgRunning) return;
  const double bn = a.state->rhs_norm;
  double d[3];
  slot_sums<3>(a, kSlotB, d);  // s.s, t.t, t.s
  const double s_norm = sqrt(max0(d[0]));
  if (converged(a, s_norm / floor30(bn), s_norm)) {
    double v[2] = {0.0, 0.0};
    if (q < a.n) {
      const double u = uq;
      const double x = xq + u;
      a.x[q] = x;
      v[0] = v[0] + u * u;
      v[1] = v[1] + x * x;
    }
    chunk_partials<2>(v, a.p_exit, a.chunks, chunk);
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // status stays running: every workgroup does its rows
      a.state->abs_res = s_norm;
      a.state->rel_res = s_norm / floor30(bn);
      a.state->half_exit = 1;
    }
    return;
  }
  const double tt = d[1];
  if (!isfinite(tt) || fabs(tt) < 1.0e-30) {
    record(a, kBicgTT, k, 0.0, 0.0, 0.0, false);
    return;
  }
  const double omega = d[2] / tt;
  if (!isfinite(omega) || fabs(omega) < 1.0e-30) {
    record(a, kBicgOmega, k, 0.0, 0.0, 0.0, false);
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.state->omega = omega;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (q < a.n) {
    const double u = uq + omega * shq;
    a.update[q] = u;
    const double x = xq + u;
    a.x[q] = x;
    const double r = sq - omega * tq;
    a.r[q] = r;
    v[0] = v[0] + u * u;
    v[1] = v[1] + x * x;
    v[2] = v[2] + r * r;
    v[3] = v[3] + rhq * r;
  }
  chunk_partials<4>(v, a.p_x, a.chunks, chunk);
}

// Sum pass for large systems: one workgroup reduces partial slots [first, first + count) as
// dot_sums does and leaves the totals in state->sums.
template <int K>
__global__ void __launch_bounds__(kRows) bicg_sum_kernel(BicgArgs a, int first) {
  const StateFlags fl = block_flags(a);
  if (fl.status != kBicgRunning) return;
  double d[K];
  dot_sums<K>(a.partials + static_cast<size_t>(first) * a.chunks, a.chunks, d);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) a.state->sums[first + k] = d[k];
  }
}

}  // namespace

hipError_t launch_bicg_setup(const BicgArgs &a, hipStream_t stream) {
  bicg_setup_kernel<<<dim3(a.chunks), dim3(kRows), 0, stream>>>(a);
  return hipGetLastError();
}

// Before pass A of iteration k: the sums it may read (the setup's at k = 0; else the full step's
// and the half exit's, whichever the state selects).
void presum_a(const BicgArgs &a, int k, hipStream_t stream) {
  if (!a.presummed) return;
  if (k == 0) {
    bicg_sum_kernel<2><<<dim3(1), dim3(kRows), 0, stream>>>(a, kSlotSetup);
  } else {
    bicg_sum_kernel<4><<<dim3(1), dim3(kRows), 0, stream>>>(a, kSlotX);
    bicg_sum_kernel<2><<<dim3(1), dim3(kRows), 0, stream>>>(a, kSlotExit);
  }
}

hipError_t launch_bicg_iteration(const BicgArgs &a, int iter, hipStream_t stream) {
  const dim3 grid(a.chunks), block(kRows);
  presum_a(a, iter, stream);
  bicg_pass_a_kernel<<<grid, block, 0, stream>>>(a, iter, 0);
  if (a.presummed) bicg_sum_kernel<1><<<dim3(1), block, 0, stream>>>(a, kSlotV);
  bicg_pass_b_kernel<<<grid, block, 0, stream>>>(a, iter);
  if (a.presummed) bicg_sum_kernel<3><<<dim3(1), block, 0, stream>>>(a, kSlotB);
  bicg_pass_c_kernel<<<grid, block, 0, stream>>>(a, iter);
  return hipGetLastError();
}

hipError_t launch_bicg_report(const BicgArgs &a, int iterations, hipStream_t stream) {
  presum_a(a, iterations, stream);
  bicg_pass_a_kernel<<<dim3(a.chunks), dim3(kRows), 0, stream>>>(a, iterations, 1);
  return hipGetLastError();
}

}  // namespace pp
