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
// become the rows themselves. Such chains are factored and swept by batches of rows in one wavefront
// instead (ch_factor_seq_kernel, ch_precondition_seq).
//
// Reductions (dot products, norms) of this one-workgroup PCG use the 512-thread canonical tree
// over the natural row order: thread t sums rows t, t + 512, ..., then the 64-lane tree (offsets
// 32..1) and the 8-wave tree (offsets 4..1). The CSR row sums of the SpMV follow apply_matrix
// (235-250).
//
// With the diagonal preconditioner (the one the reference falls back to whenever ILDLT(0) fails,
// which is the case at the production grid sizes) nothing is sequential, and the CG runs as
// full-chip passes instead (ch_diag_* below): one 512-row chunk per workgroup, the chunked
// reductions of pp_chunked.hpp, the scalar recurrence and the reference's stop and breakdown tests
// on the device.
#include <hip/hip_runtime.h>

#include "pp_chunked.hpp"
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

using lds_double = __attribute__((address_space(3))) double;

// ILDLT(0) by batches of B rows, for factors whose level schedule is a chain (the batch structure of
// the sequential sweeps below: a row's lower entries inside its batch are L(r, r-2) and L(r, r-1)
// only). Wave 0 does the work, lane k holding row b0 + k of the batch; D of the finished rows is kept
// in LDS (Dl, n doubles). Per batch:
//  1. every lane factors its row's entries before the batch, in CSR order: their cross terms
//     L(c, k) (c and k before the batch, final; positions fx_pos) and D are final, and the row's
//     own earlier entries are its own work (kept in the lane's LDS record Lr);
//  2. the numerators of the in-batch entries, over the row's entries before the batch: L(r-2, k) and
//     L(r-1, k) are entries of rows r-2, r-1 before the batch, found in their lanes' records (fab);
//  3. the rows one after the other, every lane holding the running values: L(r, r-2), then
//     L(r, r-1) with its last cross term (L(r, r-2) D(r-2)) L(r-1, r-2), then D(r), whose
//     subtractions end with the two in-batch entries (the far ones were subtracted in step 1).
// Every operation is the reference's (ch_sparse_krylov.hpp:278-320), in its order; the skip of a zero
// L(c, k) included.
__global__ void __launch_bounds__(kThreads) ch_factor_seq_kernel(ChArgs a) {
  __shared__ ChShared sh;
  extern __shared__ double ch_fac_lds[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int W = kChSeqWidth;
  lds_double *Dl = (lds_double *)ch_fac_lds;
  lds_double *Lr = Dl + a.n;       // the batch's entries before the batch: lane k at k * W
  lds_double *rec = Lr + 64 * W;   // chain records: numerator of L(r, r-2), of L(r, r-1), D so far, flags
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
  if (wave == 0) {
    const int n = a.n, B = a.seq_batch;
    bool failed = false;
    for (int b0 = 0; b0 < n; b0 += B) {
      const int cnt = min(B, n - b0);
      const bool own = lane < cnt;
      const int r = own ? b0 + lane : 0;
      const size_t base = static_cast<size_t>(r) * W;
      const int fc = own ? a.fo_cnt[r] : 0;
      lds_double *mine = Lr + lane * W;
      // 1. entries before the batch, and D's subtractions for them
      double d = a.val[a.diag_pos[r]];
      for (int e = 0; e < fc; ++e) {
        const int p = a.fo_pos[base + e];
        double s = a.val[p];
        for (int e2 = 0; e2 < e; ++e2) {
          const int pos = a.fx_pos[(base + e) * W + e2];
          const double l_ck = pos >= 0 ? a.lower[pos] : 0.0;
          if (l_ck != 0.0) s -= mine[e2] * Dl[a.fo_col[base + e2]] * l_ck;
        }
        const double dc = Dl[a.fo_col[base + e]];
        const double l = s / (dc < floor ? floor : dc);
        mine[e] = l;
        a.lower[p] = l;
        d -= (l * l) * dc;
      }
      __asm__ volatile("" ::: "memory");  // the lanes' records are complete (LDS runs a wave's accesses in order)
      // 2. numerators of L(r, r-2) and L(r, r-1) over the entries before the batch
      const int p2 = own ? a.fin_pos[2 * r] : -1, p1 = own ? a.fin_pos[2 * r + 1] : -1;
      const lds_double *row2 = Lr + (lane >= 2 ? lane - 2 : 0) * W, *row1 = Lr + (lane >= 1 ? lane - 1 : 0) * W;
      double s2 = p2 >= 0 ? a.val[p2] : 0.0, s1 = p1 >= 0 ? a.val[p1] : 0.0;
      for (int e2 = 0; e2 < fc; ++e2) {
        const int code = a.fab[base + e2];
        const int i2 = (code & 0xff) - 1, i1 = (code >> 8) - 1;
        const double t = mine[e2] * Dl[a.fo_col[base + e2]];
        if (p2 >= 0 && i2 >= 0) {
          const double l_ck = row2[i2];
          if (l_ck != 0.0) s2 -= t * l_ck;
        }
        if (p1 >= 0 && i1 >= 0) {
          const double l_ck = row1[i1];
          if (l_ck != 0.0) s1 -= t * l_ck;
        }
      }
      if (own) {
        lds_double *q = rec + lane * 4;
        q[0] = s2;
        q[1] = s1;
        q[2] = d;
        q[3] = static_cast<double>((p2 >= 0 ? 2 : 0) | (p1 >= 0 ? 1 : 0));
      }
      __asm__ volatile("" ::: "memory");
      // 3. the rows of the batch in order
      double dm1 = 0.0, dm2 = 0.0, lm = 0.0;  // D(r-1), D(r-2), L(r-1, r-2) of the running row
      double my_l2 = 0.0, my_l1 = 0.0, my_d = 0.0;
      for (int k = 0; k < cnt; ++k) {
        const lds_double *q = rec + k * 4;
        const double n2 = q[0], n1 = q[1], dk = q[2];
        const int fk = static_cast<int>(q[3]);
        double l2 = 0.0, l1 = 0.0, t = n1, dd = dk;
        if (fk & 2) {
          l2 = n2 / (dm2 < floor ? floor : dm2);
          if (lm != 0.0) t -= l2 * dm2 * lm;
        }
        if (fk & 1) l1 = t / (dm1 < floor ? floor : dm1);
        if (fk & 2) dd -= (l2 * l2) * dm2;
        if (fk & 1) dd -= (l1 * l1) * dm1;
        dm2 = dm1;
        dm1 = dd;
        lm = l1;
        if (lane == k) {
          my_l2 = l2;
          my_l1 = l1;
          my_d = dd;
        }
      }
      if (own) {
        if (p2 >= 0) a.lower[p2] = my_l2;
        if (p1 >= 0) a.lower[p1] = my_l1;
        a.D[b0 + lane] = my_d;
        Dl[b0 + lane] = my_d;
        failed = failed || !isfinite(my_d) || my_d <= floor;
      }
      // the batch's factor entries are read from memory by later batches' lanes
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    if (failed) sh.flag = 1;
  }
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

// z = M^{-1} r with the ILDLT factor by batches of B rows (apply_preconditioner, 486-512), for
// factors whose level schedule is a chain. Wave 0 does the sweeps; ys (LDS, n doubles) holds y,
// then z. Forward, per batch: lane k subtracts row b0+k's lower entries that lie before the batch
// from r, in CSR order (their y are final), then the rows run one after the other, each subtracting
// its in-batch entries L(r, r-2) y_{r-2}, L(r, r-1) y_{r-1} (the last ones in CSR order) with every
// lane holding the running values. Backward, batches and rows descending: each row subtracts the
// in-batch children r+1, r+2 first (they lead the ascending child list), then the products of its
// children outside the batch, formed beforehand by lane k in list order. Every operation is the
// reference's, in its order; only who computes it changes. A batch's operands are packed per row,
// so its loads do not depend on the sweep's values: wave 1 stages them in LDS during the batch
// before.
struct SeqFwd {
  double r, l2, l1;
  int cnt, flags;
  int col[kChSeqWidth];
  double val[kChSeqWidth];
};

struct SeqBwd {
  double c1, c2;
  int cnt, flags;
  int row[kChSeqWidth];
  double val[kChSeqWidth];
};

__device__ __forceinline__ void load_fwd(const ChArgs &a, int r, bool own, SeqFwd &o) {
  const int rr = own ? r : 0;
  const size_t base = static_cast<size_t>(rr) * kChSeqWidth;
  o.r = a.r[rr];
  o.cnt = own ? a.fo_cnt[rr] : 0;
  o.l2 = a.fin_val[2 * rr];
  o.l1 = a.fin_val[2 * rr + 1];
  o.flags = (a.fin_pos[2 * rr] >= 0 ? 2 : 0) | (a.fin_pos[2 * rr + 1] >= 0 ? 1 : 0);
#pragma unroll
  for (int e = 0; e < kChSeqWidth; ++e) {
    o.col[e] = a.fo_col[base + e];
    o.val[e] = a.fo_val[base + e];
  }
}

__device__ __forceinline__ void load_bwd(const ChArgs &a, int r, bool own, SeqBwd &o) {
  const int rr = own ? r : 0;
  const size_t base = static_cast<size_t>(rr) * kChSeqWidth;
  o.cnt = own ? a.bf_cnt[rr] : 0;
  o.c1 = a.bin_val[2 * rr];
  o.c2 = a.bin_val[2 * rr + 1];
  o.flags = (a.bin_pos[2 * rr] >= 0 ? 1 : 0) | (a.bin_pos[2 * rr + 1] >= 0 ? 2 : 0);
#pragma unroll
  for (int e = 0; e < kChSeqWidth; ++e) {
    o.row[e] = a.bf_row[base + e];
    o.val[e] = a.bf_val[base + e];
  }
}

// Row records of a batch in LDS for the sequential part: lane k writes its row's operands, then
// every lane reads row k's record with broadcast reads (no lane-to-lane moves). Which in-batch
// entries a row has is a bit per row in a wave-uniform mask (scalar registers).
constexpr int kRecF = 4;                  // forward: s, L(r, r-2), L(r, r-1), unused
constexpr int kRecB = 4 + kChSeqWidth;    // backward: z, L(r+1, r), L(r+2, r), unused, far products

// The chains: one wave steps through the batch's rows, every lane computing each row. A wave issues
// at most one instruction every four cycles, so the chains are cut to few instructions per row:
// groups of rows whose in-batch entries are all present take a path without flag tests or selects
// (most rows of a 13-point operator), the others test their bits (has1, has2: bit k for row k of
// the batch) and select. Every lane stores the rows' values (the same in all lanes) to their slots of
// the LDS vector after each group: of the ways to hand the values on, this one costs the chain least
// (in a one-wave timing of the forward recurrence, cycles per row: 72, against 82 storing each row as
// it is computed, 107 for a lane-0 store beside a sink, 151 for an exec-masked store, 82 for a
// per-lane select, 75 for a DPP chain kept in registers; 42 with no values handed on at all).

// The backward batch's chain: row k of the batch (from the top, row top - k) subtracts its in-batch
// children (bit k of has1: r+1, of has2: r+2) and then its NF far products (+0.0 past its own list),
// and lane 0 stores its z at ys[top - k].
template <int NF>
__device__ __forceinline__ void bwd_chain(const lds_double *__restrict__ rec, int cnt, unsigned long long has1,
                                          unsigned long long has2, lds_double *__restrict__ ys, int top) {
  constexpr int G = 4;
  const unsigned long long both = has1 & has2;
  double z1 = 0.0, z2 = 0.0;
  int k = 0;
  for (; k + G <= cnt; k += G) {
    double v[G][3 + NF];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const lds_double *q = rec + (k + j) * kRecB;
      v[j][0] = q[0];
      v[j][1] = q[1];
      v[j][2] = q[2];
#pragma unroll
      for (int e = 0; e < NF; ++e) v[j][3 + e] = q[4 + e];
    }
    double res[G];
    if (((both >> k) & 15) == 15) {
#pragma unroll
      for (int j = 0; j < G; ++j) {
        double t = v[j][0] - v[j][1] * z1;
        t = t - v[j][2] * z2;
#pragma unroll
        for (int e = 0; e < NF; ++e) t -= v[j][3 + e];
        z2 = z1;
        z1 = t;
        res[j] = t;
      }
    } else {
#pragma unroll
      for (int j = 0; j < G; ++j) {
        double t = v[j][0];
        const double t1 = t - v[j][1] * z1;
        t = (has1 >> (k + j)) & 1 ? t1 : t;
        const double t2 = t - v[j][2] * z2;
        t = (has2 >> (k + j)) & 1 ? t2 : t;
#pragma unroll
        for (int e = 0; e < NF; ++e) t -= v[j][3 + e];
        z2 = z1;
        z1 = t;
        res[j] = t;
      }
    }
    // stored after the group, so no store waits on a result just computed
    lds_double *base = ys + (top - k - (G - 1));  // the group's rows top-k .. top-k-G+1
#pragma unroll
    for (int j = 0; j < G; ++j) base[G - 1 - j] = res[j];
  }
  for (; k < cnt; ++k) {
    const lds_double *q = rec + k * kRecB;
    double t = q[0];
    const double t1 = t - q[1] * z1;
    t = (has1 >> k) & 1 ? t1 : t;
    const double t2 = t - q[2] * z2;
    t = (has2 >> k) & 1 ? t2 : t;
#pragma unroll
    for (int e = 0; e < NF; ++e) t -= q[4 + e];
    z2 = z1;
    z1 = t;
    ys[top - k] = t;
  }
}

// The forward batch's chain: row k of the batch subtracts its in-batch entries L(r, r-2) y(r-2)
// (bit k of has2) and L(r, r-1) y(r-1) (bit k of has1), and lane 0 stores its y at out[k].
__device__ __forceinline__ void fwd_chain(const lds_double *__restrict__ rec, int cnt, unsigned long long has1,
                                          unsigned long long has2, lds_double *__restrict__ out) {
  constexpr int G = 4;
  const unsigned long long both = has1 & has2;
  double y1 = 0.0, y2 = 0.0;
  int k = 0;
  for (; k + G <= cnt; k += G) {
    double v[G][3];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const lds_double *q = rec + (k + j) * kRecF;
      v[j][0] = q[0];
      v[j][1] = q[1];
      v[j][2] = q[2];
    }
    double res[G];
    if (((both >> k) & 15) == 15) {
#pragma unroll
      for (int j = 0; j < G; ++j) {
        double t = v[j][0] - v[j][1] * y2;
        t = t - v[j][2] * y1;
        y2 = y1;
        y1 = t;
        res[j] = t;
      }
    } else {
#pragma unroll
      for (int j = 0; j < G; ++j) {
        double t = v[j][0];
        const double t2 = t - v[j][1] * y2;
        t = (has2 >> (k + j)) & 1 ? t2 : t;
        const double t1 = t - v[j][2] * y1;
        t = (has1 >> (k + j)) & 1 ? t1 : t;
        y2 = y1;
        y1 = t;
        res[j] = t;
      }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) out[k + j] = res[j];  // after the group, as in the backward chain
  }
  for (; k < cnt; ++k) {
    const lds_double *q = rec + k * kRecF;
    double t = q[0];
    const double t2 = t - q[1] * y2;
    t = (has2 >> k) & 1 ? t2 : t;
    const double t1 = t - q[2] * y1;
    t = (has1 >> k) & 1 ? t1 : t;
    y2 = y1;
    y1 = t;
    out[k] = t;
  }
}

// A batch's operands are staged in LDS by wave 1 while wave 0 sweeps the batch before (two buffers,
// one workgroup barrier per batch), so wave 0 reads them at LDS latency. One buffer: per row the
// doubles (forward: r, L(r, r-2), L(r, r-1); backward: L(r+1, r), L(r+2, r), unused; then the 12
// packed values) and the ints (count, flags, the 12 columns or rows), field-major so each lane's
// reads and writes are stride 1.
using lds_int = __attribute__((address_space(3))) int;
constexpr int kStageDbl = 3 + kChSeqWidth, kStageInt = 2 + kChSeqWidth;
constexpr int kChStageDoubles = 64 * kStageDbl + 64 * kStageInt / 2;

__device__ __forceinline__ void put_stage(lds_double *buf, int lane, double d0, double d1, double d2,
                                          const double (&val)[kChSeqWidth], int cnt, int flags,
                                          const int (&idx)[kChSeqWidth]) {
  lds_int *bi = (lds_int *)(buf + 64 * kStageDbl);
  buf[lane] = d0;
  buf[64 + lane] = d1;
  buf[128 + lane] = d2;
#pragma unroll
  for (int e = 0; e < kChSeqWidth; ++e) buf[(3 + e) * 64 + lane] = val[e];
  bi[lane] = cnt;
  bi[64 + lane] = flags;
#pragma unroll
  for (int e = 0; e < kChSeqWidth; ++e) bi[(2 + e) * 64 + lane] = idx[e];
}

__device__ __forceinline__ void get_stage(const lds_double *buf, int lane, double &d0, double &d1, double &d2,
                                          double (&val)[kChSeqWidth], int &cnt, int &flags,
                                          int (&idx)[kChSeqWidth]) {
  const lds_int *bi = (const lds_int *)(buf + 64 * kStageDbl);
  d0 = buf[lane];
  d1 = buf[64 + lane];
  d2 = buf[128 + lane];
#pragma unroll
  for (int e = 0; e < kChSeqWidth; ++e) val[e] = buf[(3 + e) * 64 + lane];
  cnt = bi[lane];
  flags = bi[64 + lane];
#pragma unroll
  for (int e = 0; e < kChSeqWidth; ++e) idx[e] = bi[(2 + e) * 64 + lane];
}

__device__ __forceinline__ void stage_fwd(const ChArgs &a, int r, bool own, lds_double *buf, int lane) {
  SeqFwd o;
  load_fwd(a, r, own, o);
  put_stage(buf, lane, o.r, o.l2, o.l1, o.val, o.cnt, o.flags, o.col);
}

__device__ __forceinline__ void stage_bwd(const ChArgs &a, int r, bool own, lds_double *buf, int lane) {
  SeqBwd o;
  load_bwd(a, r, own, o);
  put_stage(buf, lane, o.c1, o.c2, 0.0, o.val, o.cnt, o.flags, o.row);
}

__device__ void ch_precondition_seq(const ChArgs &a, double floor, lds_double *ys, lds_double *rec) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  lds_double *stage = rec + 64 * kRecB;  // two staging buffers after the records
  const int n = a.n, B = a.seq_batch, nb = (n + B - 1) / B;
  if (wave == 1) stage_fwd(a, lane, lane < min(B, n), stage, lane);
  __syncthreads();
  for (int bi = 0; bi < nb; ++bi) {  // the same barriers in every wave
    const int b0 = bi * B;
    if (wave == 0) {
      SeqFwd cur;
      get_stage(stage + (bi & 1) * kChStageDoubles, lane, cur.r, cur.l2, cur.l1, cur.val, cur.cnt, cur.flags, cur.col);
      const int cnt = min(B, n - b0);
      const bool own = lane < cnt;
      double s = cur.r, g[kChSeqWidth];
#pragma unroll
      for (int e = 0; e < kChSeqWidth; ++e) g[e] = ys[cur.col[e]];  // column 0 past the row's list
#pragma unroll
      for (int e = 0; e < kChSeqWidth; ++e) s = e < cur.cnt ? s - cur.val[e] * g[e] : s;
      const unsigned long long has2 = __ballot(own && (cur.flags & 2)), has1 = __ballot(own && (cur.flags & 1));
      if (own) {
        lds_double *q = rec + lane * kRecF;
        q[0] = s;
        q[1] = cur.l2;
        q[2] = cur.l1;
      }
      fwd_chain(rec, cnt, has1, has2, ys + b0);
    } else if (wave == 1 && bi + 1 < nb) {
      stage_fwd(a, b0 + B + lane, lane < min(B, n - b0 - B), stage + ((bi + 1) & 1) * kChStageDoubles, lane);
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += kThreads) {
    const double d = a.D[r];
    ys[r] = ys[r] / (d < floor ? floor : d);
  }
  if (wave == 1) {
    const int top = (nb - 1) * B;
    stage_bwd(a, n - 1 - lane, lane < n - top, stage, lane);
  }
  __syncthreads();
  for (int t = 0; t < nb; ++t) {  // batches from the top
    const int b0 = (nb - 1 - t) * B;
    if (wave == 0) {
      SeqBwd cur;
      double unused;
      get_stage(stage + (t & 1) * kChStageDoubles, lane, cur.c1, cur.c2, unused, cur.val, cur.cnt, cur.flags, cur.row);
      const int hi = min(b0 + B, n) - 1, cnt = hi - b0 + 1;
      const int r = hi - lane;  // lane k: the k-th row of the batch from the top
      const bool own = lane < cnt;
      int nfar = own ? cur.cnt : 0;  // the batch's longest list of far children
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) nfar = max(nfar, __shfl_xor(nfar, off, 64));
      nfar = __builtin_amdgcn_readfirstlane(nfar);
      const unsigned long long has1 = __ballot(own && (cur.flags & 1)), has2 = __ballot(own && (cur.flags & 2));
      if (own) {
        lds_double *q = rec + lane * kRecB;
        q[0] = ys[r];
        q[1] = cur.c1;
        q[2] = cur.c2;
        // products of the children outside the batch in list order; +0.0 past the row's list
        // (t - (+0.0) leaves t unchanged bit for bit, -0.0 included)
        double g[kChSeqWidth];
#pragma unroll
        for (int e = 0; e < kChSeqWidth; ++e) g[e] = ys[cur.row[e]];  // row 0 past the row's list
#pragma unroll
        for (int e = 0; e < kChSeqWidth; ++e) q[4 + e] = e < cur.cnt ? cur.val[e] * g[e] : 0.0;
      }
      switch (nfar) {  // the chain with the batch's far count fixed: reads issued together, no branches
        case 0: bwd_chain<0>(rec, cnt, has1, has2, ys, hi); break;
        case 1: bwd_chain<1>(rec, cnt, has1, has2, ys, hi); break;
        case 2: bwd_chain<2>(rec, cnt, has1, has2, ys, hi); break;
        case 3: bwd_chain<3>(rec, cnt, has1, has2, ys, hi); break;
        case 4: bwd_chain<4>(rec, cnt, has1, has2, ys, hi); break;
        case 5: bwd_chain<5>(rec, cnt, has1, has2, ys, hi); break;
        case 6: bwd_chain<6>(rec, cnt, has1, has2, ys, hi); break;
        case 7: bwd_chain<7>(rec, cnt, has1, has2, ys, hi); break;
        case 8: bwd_chain<8>(rec, cnt, has1, has2, ys, hi); break;
        case 9: bwd_chain<9>(rec, cnt, has1, has2, ys, hi); break;
        case 10: bwd_chain<10>(rec, cnt, has1, has2, ys, hi); break;
        case 11: bwd_chain<11>(rec, cnt, has1, has2, ys, hi); break;
        default: bwd_chain<kChSeqWidth>(rec, cnt, has1, has2, ys, hi); break;
      }
    } else if (wave == 1 && t + 1 < nb) {
      stage_bwd(a, b0 - 1 - lane, lane < B, stage + ((t + 1) & 1) * kChStageDoubles, lane);
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += kThreads) a.z[r] = ys[r];
  __syncthreads();
}

// Gathers the factor's values into the sequential sweeps' packed per-row layout (once per operator).
__global__ void __launch_bounds__(kThreads) ch_pack_kernel(ChArgs a) {
  for (int r = blockIdx.x * kThreads + threadIdx.x; r < a.n; r += gridDim.x * kThreads) {
    const size_t base = static_cast<size_t>(r) * kChSeqWidth;
    for (int e = 0; e < kChSeqWidth; ++e) {
      const int pf = a.fo_pos[base + e], pb = a.bf_pos[base + e];
      a.fo_val[base + e] = pf >= 0 ? a.lower[pf] : 0.0;
      a.bf_val[base + e] = pb >= 0 ? a.lower[pb] : 0.0;
    }
    for (int e = 0; e < 2; ++e) {
      const int pi = a.fin_pos[2 * r + e], pc = a.bin_pos[2 * r + e];
      a.fin_val[2 * r + e] = pi >= 0 ? a.lower[pi] : 0.0;
      a.bin_val[2 * r + e] = pc >= 0 ? a.lower[pc] : 0.0;
    }
  }
}

// (A x)_row in CSR order (apply_matrix, 235-250). The first 13 entries (the whole 13-point row) are
// loaded together, positions past the row clamped to its last entry and their products left out,
// so a row costs two load round trips rather than two per entry.
__device__ __forceinline__ double row_times(const ChArgs &a, const double *x, int row) {
  constexpr int kW = 13;
  const int beg = a.row_ptr[row], end = a.row_ptr[row + 1];  // every row holds its diagonal
  double v[kW], xv[kW];
#pragma unroll
  for (int e = 0; e < kW; ++e) {
    const int p = min(beg + e, end - 1);
    v[e] = a.val[p];
    xv[e] = x[a.col[p]];
  }
  double s = 0.0;
#pragma unroll
  for (int e = 0; e < kW; ++e) s = beg + e < end ? s + v[e] * xv[e] : s;
  for (int p = beg + kW; p < end; ++p) s += a.val[p] * x[a.col[p]];
  return s;
}

// PCG from x (ch_sparse_krylov.hpp:515-660, no projection); the stop and breakdown tests are the
// reference's. ildlt selects the preconditioner.
__global__ void __launch_bounds__(kThreads) ch_pcg_kernel(ChArgs a, int ildlt) {
  __shared__ ChShared sh;
  extern __shared__ double ch_ys[];  // the sequential sweeps' vector, row records and staging (a.seq_batch > 0)
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
    if (use_ildlt && a.seq_batch > 0)
      ch_precondition_seq(a, floor, (lds_double *)ch_ys, (lds_double *)ch_ys + a.n);
    else
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
        if (use_ildlt && a.seq_batch > 0)
          ch_precondition_seq(a, floor, (lds_double *)ch_ys, (lds_double *)ch_ys + a.n);
        else
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

// ---- diagonal-preconditioned CG as full-chip passes ------------------------------------------

using chunked::row_chunk;
using chunked::chunk_partials;
using chunked::dot_sums;
constexpr int kRows = chunked::kRows;
static_assert(kRows == kChRows, "the host sizes the CH chunks with kChRows");

__device__ __forceinline__ double max0(double v) { return v < 0.0 ? 0.0 : v; }            // std::max(v, 0.0)
__device__ __forceinline__ double floor30(double v) { return v < 1.0e-30 ? 1.0e-30 : v; }  // std::max(v, 1e-30)

// The state's stop flag as one value per workgroup (thread 0 reads it, the barrier hands it on).
__device__ __forceinline__ bool block_stopped(const ChDiagArgs &a) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = a.state->stopped;
  __syncthreads();
  return flag != 0;
}

// The K sums of partial slots [slot, slot + K), summed here or taken from the sum pass.
template <int K>
__device__ __forceinline__ void ch_slot_sums(const ChDiagArgs &a, int slot, double (&out)[K]) {
  if (a.presummed) {
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = a.state->sums[slot + k];
  } else {
    dot_sums<K>(a.partials + static_cast<size_t>(slot) * a.chunks, a.chunks, out);
  }
}

__device__ __forceinline__ double csr_row(const ChDiagArgs &a, const double *x, int row) {
  double s = 0.0;
  const int end = a.row_ptr[row + 1];
  for (int p = a.row_ptr[row]; p < end; ++p) s += a.val[p] * x[a.col[p]];
  return s;
}

// Workgroup 0 records a stop with the report fields.
__device__ __forceinline__ void ch_record(const ChDiagArgs &a, int status, int iterations, double abs_res,
                                          double rel_res, double iter_res) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ChDiagState &s = *a.state;
    s.iterations = iterations;
    s.abs_res = abs_res;
    s.rel_res = rel_res;
    s.iter_res = iter_res;
    s.status = status;
    s.stopped = 1;
  }
}

// r = b - A x, z = M^-1 r, p = z (ch_sparse_krylov.hpp:530-583); partials b.b, r.r, r.z. z and r.z
// are formed before the initial convergence test, which pass A of iteration 0 makes; they go unused
// when it stops.
__global__ void __launch_bounds__(kRows) ch_diag_setup_kernel(ChDiagArgs a) {
  const int chunk = row_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (q < a.n) {
    const double b = a.rhs[q];
    const double r = b - csr_row(a, a.x, q);
    const double z = a.dinv[q] * r;
    a.r[q] = r;
    a.z[q] = z;
    a.p[0][q] = z;
    v[0] = v[0] + b * b;
    v[1] = v[1] + r * r;
    v[2] = v[2] + r * z;
  }
  chunk_partials<3>(v, a.partials + kChSlotSetup * static_cast<size_t>(a.chunks), a.chunks, chunk);
}

// Pass A of iteration k (final: the report after the limit's last iteration). The tests of the
// previous step: at k = 0 the initial residual and r.z > 0 (540-583); after an iteration its
// residuals (603-612), then the new r.z > 0 and beta (640-652). Then p_k = z + beta p_{k-1} (p_0 = z),
// A p_k with the neighbours' p_k formed from their z and p_{k-1} by the same expression, and the
// partial p.Ap.
__global__ void __launch_bounds__(kRows) ch_diag_pass_a_kernel(ChDiagArgs a, int k, int final) {
  if (block_stopped(a)) return;
  const double tol = a.tolerance;
  double beta = 0.0;
  if (k == 0) {
    double s[3];
    ch_slot_sums<3>(a, kChSlotSetup, s);
    const double bnorm = sqrt(max0(s[0]));
    const double abs_res = sqrt(max0(s[1]));
    const double rel = abs_res / floor30(bnorm);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.state->bnorm = bnorm;
      a.state->bsq = s[0];
      a.state->rz[0] = s[2];
    }
    if (rel < tol || abs_res < tol) {
      ch_record(a, kChOk, 0, abs_res, rel, 0.0);
      return;
    }
    if (!isfinite(s[2]) || s[2] <= 0.0) {
      ch_record(a, kChRz0Fail, 0, abs_res, rel, 0.0);
      return;
    }
  } else {
    double u[4];
    ch_slot_sums<4>(a, kChSlotUpd, u);
    const double iter_res = sqrt(u[0] / floor30(u[1]));
    const double abs_res = sqrt(max0(u[2]));
    const double rel = abs_res / floor30(a.state->bnorm);
    if (rel < tol || abs_res < tol || final) {
      ch_record(a, kChOk, k, abs_res, rel, iter_res);
      return;
    }
    if (!isfinite(u[3]) || u[3] <= 0.0) {
      ch_record(a, kChRzFail, k, abs_res, rel, iter_res);
      return;
    }
    beta = u[3] / a.state->rz[(k - 1) & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.state->rz[k & 1] = u[3];
  }
  const int chunk = row_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  const double *__restrict__ z = a.z;
  const double *__restrict__ po = a.p[(k - 1) & 1];
  double v[1] = {0.0};
  if (q < a.n) {
    double s = 0.0;
    const int end = a.row_ptr[q + 1];
    for (int e = a.row_ptr[q]; e < end; ++e) {
      const int c = a.col[e];
      const double pc = k == 0 ? z[c] : z[c] + beta * po[c];
      s += a.val[e] * pc;
    }
    const double pq = k == 0 ? z[q] : z[q] + beta * po[q];
    a.p[k & 1][q] = pq;
    a.ap[q] = s;
    v[0] = v[0] + pq * s;
  }
  chunk_partials<1>(v, a.partials + kChSlotPap * static_cast<size_t>(a.chunks), a.chunks, chunk);
}

// Pass B of iteration k: p^T A p valid (588-594), alpha, x += alpha p, r -= alpha A p, z = M^-1 r
// (595-601, 614-621); partials delta.delta, x.x, r.r, r.z.
__global__ void __launch_bounds__(kRows) ch_diag_pass_b_kernel(ChDiagArgs a, int k) {
  if (block_stopped(a)) return;
  double pap[1];
  ch_slot_sums<1>(a, kChSlotPap, pap);
  if (!isfinite(pap[0]) || fabs(pap[0]) < 1.0e-30) {
    ch_record(a, kChPapFail, k, 0.0, 0.0, 0.0);
    return;
  }
  const double alpha = a.state->rz[k & 1] / pap[0];
  const int chunk = row_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (q < a.n) {
    const double delta = alpha * a.p[k & 1][q];
    const double xn = a.x[q] + delta;
    const double rn = a.r[q] - alpha * a.ap[q];
    const double zn = a.dinv[q] * rn;
    a.x[q] = xn;
    a.r[q] = rn;
    a.z[q] = zn;
    v[0] = v[0] + delta * delta;
    v[1] = v[1] + xn * xn;
    v[2] = v[2] + rn * rn;
    v[3] = v[3] + rn * zn;
  }
  chunk_partials<4>(v, a.partials + kChSlotUpd * static_cast<size_t>(a.chunks), a.chunks, chunk);
}

// |rhs - A x|^2 of the returned x (phase_field.cpp:481-502), per chunk.
__global__ void __launch_bounds__(kRows) ch_diag_eq_kernel(ChDiagArgs a) {
  const int chunk = row_chunk(a.chunks), q = chunk * kRows + threadIdx.x;
  double v[1] = {0.0};
  if (q < a.n) {
    const double d = a.rhs[q] - csr_row(a, a.x, q);
    v[0] = v[0] + d * d;
  }
  chunk_partials<1>(v, a.partials + kChSlotEq * static_cast<size_t>(a.chunks), a.chunks, chunk);
}

__global__ void __launch_bounds__(kRows) ch_diag_finish_kernel(ChDiagArgs a) {
  double eq[1];
  dot_sums<1>(a.partials + kChSlotEq * static_cast<size_t>(a.chunks), a.chunks, eq);
  if (threadIdx.x == 0) a.state->eq_res = sqrt(eq[0] / floor30(a.state->bsq));
}

// Large systems: the K sums of slots [first, first + K) once, into the state.
template <int K>
__global__ void __launch_bounds__(kRows) ch_diag_sum_kernel(ChDiagArgs a, int first) {
  if (block_stopped(a)) return;
  double d[K];
  dot_sums<K>(a.partials + static_cast<size_t>(first) * a.chunks, a.chunks, d);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) a.state->sums[first + k] = d[k];
  }
}

}  // namespace

hipError_t launch_ch_factor(const ChArgs &a, hipStream_t stream) {
  if (a.seq_batch > 0) {
    const size_t lds = (static_cast<size_t>(a.n) + 64 * (kChSeqWidth + 4)) * sizeof(double);
    hipLaunchKernelGGL(ch_factor_seq_kernel, dim3(1), dim3(kThreads), lds, stream, a);
  } else {
    hipLaunchKernelGGL(ch_factor_kernel, dim3(1), dim3(kThreads), 0, stream, a);
  }
  return hipGetLastError();
}

hipError_t launch_ch_pcg(const ChArgs &a, bool ildlt, hipStream_t stream) {
  const size_t lds =
      ildlt && a.seq_batch > 0 ? (static_cast<size_t>(a.n) + 64 * kRecB + 2 * kChStageDoubles) * sizeof(double) : 0;
  hipLaunchKernelGGL(ch_pcg_kernel, dim3(1), dim3(kThreads), lds, stream, a, ildlt ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_ch_pack(const ChArgs &a, hipStream_t stream) {
  const int blocks = (a.n + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(ch_pack_kernel, dim3(blocks < 64 ? blocks : 64), dim3(kThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_ch_diag_setup(const ChDiagArgs &a, hipStream_t stream) {
  ch_diag_setup_kernel<<<dim3(a.chunks), dim3(kRows), 0, stream>>>(a);
  return hipGetLastError();
}

// Before pass A of iteration k: the sums it reads (the setup's at k = 0, else pass B's).
static void ch_presum_a(const ChDiagArgs &a, int k, hipStream_t stream) {
  if (!a.presummed) return;
  if (k == 0)
    ch_diag_sum_kernel<3><<<dim3(1), dim3(kRows), 0, stream>>>(a, kChSlotSetup);
  else
    ch_diag_sum_kernel<4><<<dim3(1), dim3(kRows), 0, stream>>>(a, kChSlotUpd);
}

hipError_t launch_ch_diag_iteration(const ChDiagArgs &a, int k, hipStream_t stream) {
  const dim3 grid(a.chunks), block(kRows);
  ch_presum_a(a, k, stream);
  ch_diag_pass_a_kernel<<<grid, block, 0, stream>>>(a, k, 0);
  if (a.presummed) ch_diag_sum_kernel<1><<<dim3(1), block, 0, stream>>>(a, kChSlotPap);
  ch_diag_pass_b_kernel<<<grid, block, 0, stream>>>(a, k);
  return hipGetLastError();
}

hipError_t launch_ch_diag_report(const ChDiagArgs &a, int iterations, hipStream_t stream) {
  ch_presum_a(a, iterations, stream);
  ch_diag_pass_a_kernel<<<dim3(a.chunks), dim3(kRows), 0, stream>>>(a, iterations, 1);
  return hipGetLastError();
}

hipError_t launch_ch_diag_finish(const ChDiagArgs &a, hipStream_t stream) {
  ch_diag_eq_kernel<<<dim3(a.chunks), dim3(kRows), 0, stream>>>(a);
  ch_diag_finish_kernel<<<dim3(1), dim3(kRows), 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace pp
