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
// Reductions use one fixed order (the "canonical tree"): thread t of the T = 512 threads accumulates
// positions t, t + T, t + 2T, ... sequentially from 0.0; lanes of a wave combine with offsets
// 32, 16, ..., 1 (v[l] += v[l + off]); waves combine the same way with offsets W/2, ..., 1.
// oracle/pp_oracle.cpp reproduces this order to check whole solves bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

constexpr int kThreads = 512;             // threads per system block (8 waves, 256 VGPRs per lane)
constexpr int kWaves = kThreads / 64;     // 8
constexpr int kMaxRows = 64 * kWaves;     // nx <= 512: one sweep wave per 64 rows
constexpr int kMaxSide = 1024;            // ny <= 1024 where the periodic-x wrap line sits in LDS
constexpr int kMaxColumns = 65535;        // ny <= 65535 otherwise (the (i << 16) | j cell code)
constexpr int kRing = 64;                 // hand-off ring depth between sweep waves
constexpr int kPad = 1024;                // padding after every work array (lanes outside the grid)
constexpr int kStageDoubles = 16384;      // 128 KB of LDS for the sweep operand ring
constexpr int kMaxRing = 64;              // deepest operand ring in levels (per sweep wave)
constexpr int kStageUnit = 8;             // levels a loader moves per step
constexpr int kMaxStagedWaves = 6;        // more sweep waves leave too few loader waves
constexpr int kSpinLimit = 1 << 24;       // bounded waits: a broken hand-off ends as an error
constexpr int kLevPad = 80;               // level-base table margin: sweeps index up to 71 levels past the grid

enum DevStatus : int {
  kOk = 0, kFactorFail = 1, kPapFail = 2, kRz0Fail = 3, kRzFail = 4, kNoDiag = 5, kSweepFail = 6
};

struct DevReport {
  int status;
  int iterations;
  int outer_cycles;
  int monitor_calls;  // PCG monitor calls made (records written up to the log's capacity)
  double rel, abs, iter_res;
  unsigned long long phase_ticks[10];
};

__device__ __forceinline__ int level_lo(int d, int ny) { return d - ny + 1 > 0 ? d - ny + 1 : 0; }

// Explicit address spaces. Pointers that arrive through a struct are generic ("flat"); flat loads
// count against both the vector-memory and the LDS wait counters, and volatile flat accesses are
// followed by a full vmcnt(0) drain. Casting to the global (1) and LDS (3) spaces gives global_* and
// ds_* instructions, so LDS polling never waits on in-flight prefetch loads.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T *;

template <typename T>
__device__ __forceinline__ global_ptr<T> as_global(T *p) {
  return (global_ptr<T>)p;
}

template <typename T>
__device__ __forceinline__ volatile __attribute__((address_space(3))) T *lds(volatile T *p) {
  return (volatile __attribute__((address_space(3))) T *)(p);
}

// std::max(value, floor) as the reference writes it: floor only when value < floor.
__device__ __forceinline__ double floor_max(double value, double floor) { return value < floor ? floor : value; }
__device__ __forceinline__ float floor_max(float value, float floor) { return value < floor ? floor : value; }

// LDS state shared by the phases of one system block.
struct SweepShared {
  double red[2][kWaves * 4];            // reduction slots, double-buffered
  volatile double hand[kWaves][kRing];  // hand[w]: boundary values flowing into wave w
  volatile int cnt[kWaves];             // values published into hand[w]
  volatile int ack[kWaves];             // values consumed from hand[w]
  volatile double wrap[kMaxSide];       // periodic-x wrap line (row 0 -> row nx-1 and back)
  double lx[kMaxSide];                  // periodic-x wrap lower entries LX[j], copied after factoring
  // Operand staging ring for the sweeps: loader waves copy the four per-level operands of every
  // sweep wave into stage[] ahead of use; ready[w][slot] holds step + 1 once a slot is filled and
  // used[w] counts the steps sweep wave w has finished with.
  double stage[kStageDoubles];
  volatile int ready[kWaves * kMaxRing];
  volatile int used[kWaves];
  // Write targets for lanes that have nothing to publish, so hand-off stores need no branch.
  volatile double dummy_d[64];
  volatile int dummy_i[64];
  volatile int spin_error;              // set by the first wait that gives up; every later wait returns at once
  volatile int spin_limit;              // sleeps a wait may spend before it gives up (kSpinLimit unless lowered)
  volatile int status;                  // DevStatus of the system, uniform after each barrier
};

// Device clock for per-phase timing (s_memrealtime, 100 MHz).
__device__ __forceinline__ unsigned long long device_ticks() { return __builtin_amdgcn_s_memrealtime(); }

constexpr int kPhases = 10;  // setup, spmv, update, rr, forward, backward, zsum, rz, p update, output

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

// Bounded wait on an LDS counter. A wait that spins past spin_limit sleeps records spin_error and
// returns; once spin_error is set every wait returns at once, so one broken hand-off costs a single
// wait budget and the sweeps run to their end without further waiting. The solve loops read
// spin_error at a point where no wave can be inside a wait (after a block-wide barrier that follows
// the sweeps and precedes the next ones) and leave with kSweepFail, so every wave takes the same
// path and the kernel exits.
__device__ __forceinline__ void wait_at_least(volatile int *counter, int target, SweepShared &sh) {
  int guard = 0;
  while (*lds(counter) < target) {
    if (*lds(&sh.spin_error)) return;
    if (++guard > *lds(&sh.spin_limit)) {
      *lds(&sh.spin_error) = 1;
      return;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__device__ __forceinline__ void reset_handoff(SweepShared &sh) {
  if (threadIdx.x < kWaves) {
    *lds(&sh.cnt[threadIdx.x]) = 0;
    *lds(&sh.ack[threadIdx.x]) = 0;
  }
}

// ---- lane shifts and hand-offs between sweep waves ---------------------------------------------

// Lane l receives lane l-1's value (lane 0 receives +0): one DPP wave_shr per 32-bit half, with
// bound_ctrl so the lane without a source reads zero.
__device__ __forceinline__ double wave_shr1(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Lane l receives lane l+1's value (lane 63 receives +0).
__device__ __forceinline__ double wave_shl1(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float wave_shr1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}

__device__ __forceinline__ float wave_shl1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}

// Lane l receives lane l-1's value and lane 0 receives `fill` (wave_shr with bound_ctrl off: the
// lane without a source keeps the old value, which is fill).
__device__ __forceinline__ double shr1_fill(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Lane l receives lane l+1's value and lane 63 receives `fill`.
__device__ __forceinline__ double shl1_fill(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x130, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float shr1_fill(float v, float fill) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138, 0xF, 0xF, false));
}

__device__ __forceinline__ float shl1_fill(float v, float fill) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x130, 0xF, 0xF, false));
}

// Value held by lane k (compile-time k), broadcast through SGPRs.
__device__ __forceinline__ double lane_value(double v, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float lane_value(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Lane k of the result holds v (wave-uniform, e.g. from lane_value); other lanes keep `into`.
// The lane comparison is loop-invariant, so this is one select per 32-bit half.
template <typename T>
__device__ __forceinline__ T write_lane(T into, T v, int k) {
  return (threadIdx.x & 63) == static_cast<unsigned>(k) ? v : into;
}

// Hand-off protocol. Ring hand[w] carries boundary values into wave w; cnt[w] counts values
// published, ack[w] values consumed. The k-th value of a stream sits in slot k % kRing.
// Consumers read a batch of P values at once (lane k of the wave holds batch entry k);
// producers publish one value per level and check ring space once per batch of P levels.
// With P <= 31 and kRing = 64 a producer never waits on a consumer that waits on it.

template <typename T, int P>
__device__ __forceinline__ T receive_batch(int stream_wave, int lane, int k0, int count, SweepShared &sh) {
  T hv = T(0);
  const int klo = k0 > 0 ? k0 : 0;
  const int khi = (k0 + P - 1) < (count - 1) ? (k0 + P - 1) : (count - 1);
  if (klo <= khi) {
    wait_at_least(&sh.cnt[stream_wave], khi + 1, sh);
    const int kl = k0 + lane;
    if (lane < P && kl >= klo && kl <= khi) hv = static_cast<T>(*lds(&sh.hand[stream_wave][kl % kRing]));
    if (lane == 0) *lds(&sh.ack[stream_wave]) = khi + 1;
  }
  return hv;
}

// The same batch with every lane holding all P values: hk[k] = stream value k0 + k (a broadcast LDS
// read per value), for the lane that takes it through shr1_fill / shl1_fill. Slots outside the
// published range are read too; their values only reach lanes whose cell is outside the grid at that
// level (the stream index is the boundary lane's column), which compute on padding anyway.
template <typename T, int P>
__device__ __forceinline__ void receive_chunk(int stream_wave, int k0, int count, SweepShared &sh, T (&hk)[P]) {
  const int klo = k0 > 0 ? k0 : 0;
  const int khi = (k0 + P - 1) < (count - 1) ? (k0 + P - 1) : (count - 1);
  if (klo <= khi) {
    wait_at_least(&sh.cnt[stream_wave], khi + 1, sh);
#pragma unroll
    for (int k = 0; k < P; ++k) hk[k] = static_cast<T>(*lds(&sh.hand[stream_wave][(k0 + k) & (kRing - 1)]));
    if ((threadIdx.x & 63) == 0) *lds(&sh.ack[stream_wave]) = khi + 1;
  } else {
#pragma unroll
    for (int k = 0; k < P; ++k) hk[k] = T(0);
  }
}

// Before a batch whose last published index is k_last: make sure its slot has been consumed.
__device__ __forceinline__ void reserve_batch(int stream_wave, int k_last, SweepShared &sh) {
  if (k_last >= 0) wait_at_least(&sh.ack[stream_wave], k_last + 1 - kRing, sh);
}

template <typename T>
__device__ __forceinline__ void publish(int stream_wave, int lane, int src_lane, int k, int count, T value,
                                        SweepShared &sh) {
  if (k >= 0 && k < count && lane == src_lane) {
    *lds(&sh.hand[stream_wave][k % kRing]) = static_cast<double>(value);
    *lds(&sh.cnt[stream_wave]) = k + 1;
  }
}

// Branch-free publish: every lane executes the two LDS stores, only src_lane with a valid index
// targets the ring and counter, the others write their dummy slots.
template <typename T>
__device__ __forceinline__ void publish_nb(int stream_wave, int lane, int src_lane, int k, int count, T value,
                                           SweepShared &sh) {
  const bool real = lane == src_lane && k >= 0 && k < count;
  volatile double *slot = real ? &sh.hand[stream_wave][k & (kRing - 1)] : &sh.dummy_d[lane];
  volatile int *ctr = real ? &sh.cnt[stream_wave] : &sh.dummy_i[lane];
  *lds(slot) = static_cast<double>(value);
  *lds(ctr) = k + 1;
}

// Chunk publish: lane kk (< kChunk) holds stream value k0 + kk; the values in range go into the ring,
// then lane 0 advances the counter past the last of them (one LDS write per lane, one counter write).
template <int kChunk, typename T>
__device__ __forceinline__ void publish_chunk(int stream_wave, int lane, int k0, int count, T pub, SweepShared &sh) {
  const int k = k0 + lane;
  if (lane < kChunk && k >= 0 && k < count) *lds(&sh.hand[stream_wave][k & (kRing - 1)]) = static_cast<double>(pub);
  const int last = (k0 + kChunk - 1) < (count - 1) ? (k0 + kChunk - 1) : (count - 1);
  if (lane == 0 && last >= 0) *lds(&sh.cnt[stream_wave]) = last + 1;
}

// Number of cells on levels < d of an nx x ny grid (0 <= d <= nx + ny - 1).
__device__ __forceinline__ int level_offset(int d, int nx, int ny) {
  const int a = nx < ny ? nx : ny, b = nx < ny ? ny : nx;
  if (d <= a) return d * (d + 1) / 2;
  if (d <= b) return a * (a + 1) / 2 + (d - a) * a;
  const int m = nx + ny - 1 - d;
  return nx * ny - m * (m + 1) / 2;
}

// off(d) - lo(d): cell (i, d - i) sits at level_base(d) + i. Branch-free closed form, so with a
// wave-uniform d the whole computation stays on the scalar unit.
__device__ __forceinline__ int level_base(int d, int nx, int ny) {
  const int a = nx < ny ? nx : ny, b = nx < ny ? ny : nx;
  const int d1 = min(d, a), d2 = min(max(d - a, 0), b - a), d3 = max(d - b, 0);
  const int off = d1 * (d1 + 1) / 2 + d2 * a + d3 * (a - 1) - d3 * (d3 - 1) / 2;
  return off - max(0, d - ny + 1);
}

__device__ __forceinline__ int grid_pos(int i, int j, int nx, int ny) {
  const int d = i + j;
  return level_offset(d, nx, ny) + i - level_lo(d, ny);
}

}  // namespace pp
