// Device building blocks of the pressure-projection engine (gfx950, wave64).
//
// Data layout: every per-cell array of a system is stored in anti-diagonal ("wavefront") order.
// Level d holds the cells (i, d - i), i ascending, and levels are concatenated:
//     pos(i, j) = off[i + j] + i - max(0, i + j - ny + 1)
// The natural-ordering ICC/ILDLT recurrences (ch_sparse_krylov.hpp:337-419, 444-513) make cell
// (i, j) depend only on level i + j - 1 (plus the periodic wrap cells of level j or i, which are
// long finished), so a sweep advances one level per step and every load of a step is one
// contiguous run of memory.
//
// Sweeps run on S = ceil(nx / 64) "sweep waves": lane l of wave w owns row i = 64 w + l and walks
// the levels. The west neighbour (row i-1, previous level) arrives by a one-lane shuffle; lane 0
// takes it from wave w-1 through an LDS ring with a published/consumed counter pair, so the waves
// form a pipeline without block-wide barriers. The south neighbour is the lane's own previous value.
//
// Reductions use one fixed order (the "canonical tree"): thread t of the block accumulates
// positions t, t + T, t + 2T, ... sequentially from 0.0; lanes of a wave combine with offsets
// 32, 16, ..., 1 (v[l] += v[l + off]); waves combine the same way with offsets W/2, ..., 1.
// oracle/pp_oracle.cpp reproduces this order to check whole solves bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

constexpr int kThreads = 1024;            // threads per system block
constexpr int kWaves = kThreads / 64;     // 16
constexpr int kMaxSide = 1024;            // nx, ny <= 1024 (sweep waves <= 16)
constexpr int kMaxLevels = 2 * kMaxSide - 1;
constexpr int kRing = 64;                 // hand-off ring depth between sweep waves
constexpr int kSpinLimit = 1 << 24;       // bounded waits: a broken hand-off ends as an error

enum DevStatus : int {
  kOk = 0, kFactorFail = 1, kPapFail = 2, kRz0Fail = 3, kRzFail = 4, kNoDiag = 5, kSweepFail = 6
};

struct DevReport {
  int status;
  int iterations;
  int outer_cycles;
  int pad;
  double rel, abs, iter_res;
};

__device__ __forceinline__ int level_lo(int d, int ny) { return d - ny + 1 > 0 ? d - ny + 1 : 0; }

// std::max(value, floor) as the reference writes it: floor only when value < floor.
__device__ __forceinline__ double floor_max(double value, double floor) { return value < floor ? floor : value; }
__device__ __forceinline__ float floor_max(float value, float floor) { return value < floor ? floor : value; }

// LDS state shared by the phases of one system block.
struct SweepShared {
  int off[kMaxLevels + 1];              // level offsets of the wavefront layout
  double red[2][kWaves * 4];            // reduction slots, double-buffered
  volatile double hand[kWaves][kRing];  // hand[w]: boundary values flowing into wave w
  volatile int cnt[kWaves];             // values published into hand[w]
  volatile int ack[kWaves];             // values consumed from hand[w]
  volatile double wrap[kMaxSide];       // periodic-x wrap line (row 0 -> row nx-1 and back)
  volatile int spin_error;
  volatile int status;                  // DevStatus of the system, uniform after each barrier
};

__device__ __forceinline__ int cell_pos(const SweepShared &sh, int i, int j, int ny) {
  const int d = i + j;
  return sh.off[d] + i - level_lo(d, ny);
}

// Sum of K per-thread partials over the block in the canonical tree; every thread gets the result.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], SweepShared &sh, int &parity) {
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
  // wave tree with offsets 8, 4, 2, 1; the first level is folded into the loads
#pragma unroll 1
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

__device__ __forceinline__ double block_sum1(double v, SweepShared &sh, int &parity) {
  double a[1] = {v};
  block_sum<1>(a, sh, parity);
  return a[0];
}

__device__ __forceinline__ double block_max1(double v, SweepShared &sh, int &parity) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    v = v < o ? o : v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *slot = sh.red[parity];
  if (lane == 0) slot[wave] = v;
  __syncthreads();
  double m = slot[0];
#pragma unroll
  for (int t = 1; t < kWaves; ++t) m = m < slot[t] ? slot[t] : m;
  parity ^= 1;
  return m;
}

// Bounded wait on an LDS counter; on timeout records an error and returns (never hangs).
__device__ __forceinline__ void wait_at_least(volatile int *counter, int target, SweepShared &sh) {
  int guard = 0;
  while (*counter < target) {
    __builtin_amdgcn_s_sleep(1);
    if (++guard > kSpinLimit) {
      sh.spin_error = 1;
      return;
    }
  }
}

__device__ __forceinline__ void reset_handoff(SweepShared &sh) {
  if (threadIdx.x < kWaves) {
    sh.cnt[threadIdx.x] = 0;
    sh.ack[threadIdx.x] = 0;
  }
}

// ---- forward wavefront (factorization and forward substitution share the pattern) -----------
//
// Helper that moves lane 63's value of wave w to lane 0 of wave w+1 for rows in ascending order.
// `j_in` is the column lane 0 of this wave works on at this level, `j_out` that of lane 63.
template <typename T>
__device__ __forceinline__ T forward_receive(T west_from_shuffle, int wave, int lane, int j_in, int ny,
                                             SweepShared &sh) {
  T west = west_from_shuffle;
  if (wave > 0 && j_in >= 0 && j_in < ny) {
    wait_at_least(&sh.cnt[wave], j_in + 1, sh);
    const T h = static_cast<T>(sh.hand[wave][j_in % kRing]);
    if (lane == 0) west = h;
    if (lane == 0) sh.ack[wave] = j_in + 1;
  }
  return west;
}

template <typename T>
__device__ __forceinline__ void forward_send(T value, int wave, int lane, int S, int j_out, int ny, SweepShared &sh) {
  if (wave + 1 < S && j_out >= 0 && j_out < ny) {
    wait_at_least(&sh.ack[wave + 1], j_out + 1 - kRing, sh);
    if (lane == 63) {
      sh.hand[wave + 1][j_out % kRing] = static_cast<double>(value);
      sh.cnt[wave + 1] = j_out + 1;
    }
  }
}

// Backward direction: lane 0 of wave w sends to lane 63 of wave w-1; columns descend, so the
// k-th hand-off carries column ny-1-k.
template <typename T>
__device__ __forceinline__ T backward_receive(T east_from_shuffle, int wave, int lane, int S, int j63, int ny,
                                              SweepShared &sh) {
  T east = east_from_shuffle;
  if (wave + 1 < S && j63 >= 0 && j63 < ny) {
    const int k = ny - 1 - j63;
    wait_at_least(&sh.cnt[wave], k + 1, sh);
    const T h = static_cast<T>(sh.hand[wave][k % kRing]);
    if (lane == 63) east = h;
    if (lane == 63) sh.ack[wave] = k + 1;
  }
  return east;
}

template <typename T>
__device__ __forceinline__ void backward_send(T value, int wave, int lane, int j0, int ny, SweepShared &sh) {
  if (wave > 0 && j0 >= 0 && j0 < ny) {
    const int k = ny - 1 - j0;
    wait_at_least(&sh.ack[wave - 1], k + 1 - kRing, sh);
    if (lane == 0) {
      sh.hand[wave - 1][k % kRing] = static_cast<double>(value);
      sh.cnt[wave - 1] = k + 1;
    }
  }
}

}  // namespace pp
