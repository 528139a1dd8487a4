// Chunked reductions shared by the full-chip Krylov passes (BiCGSTAB for the momentum system,
// the diagonal-preconditioned CG of the Cahn-Hilliard solve): 512-thread workgroups, one row per
// thread, each workgroup owning one chunk of 512 rows. A pass leaves one partial per chunk (the
// 64-lane and 8-wave trees over its rows); the next pass sums the partials the same way (thread t
// adds chunks t, t + 512, ... in order, then the same trees), so every workgroup derives the same
// scalars in the same order.
#pragma once

#include <hip/hip_runtime.h>

namespace pp {
namespace chunked {

constexpr int kRows = 512;  // rows per workgroup, one per thread

// Row chunk of this workgroup. Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD
// b mod 8), and each XCD has its own L2, so XCD x is given one contiguous run of chunks: a row's
// neighbours one grid line away are then mostly in chunks of the same XCD and their lines are
// shared in its L2. The map is a bijection on [0, chunks) for any dispatch order; partials are
// indexed by chunk, so the summation order does not depend on it.
__device__ __forceinline__ int row_chunk(int chunks) {
  const int b = blockIdx.x, x = b & 7, i = b >> 3;
  const int per = chunks >> 3, rem = chunks & 7;
  return x * per + (x < rem ? x : rem) + i;
}

// Row chunk of this workgroup for a system made of two blocks of nearly equal size (the u and v
// unknowns of the momentum system, the v row of a face ~n/2 after the u rows it couples to): the
// contiguous run of positions row_chunk gives XCD x is mapped to interleaved chunks, position 2m to
// chunk m and 2m + 1 to chunk ceil(chunks/2) + m, so a u chunk and the v chunk it couples to are
// worked on by the same XCD at about the same time and the coupling gathers hit its L2. A
// bijection on [0, chunks) like row_chunk, so the partials (indexed by chunk) sum the same way.
__device__ __forceinline__ int paired_chunk(int chunks) {
  const int pos = row_chunk(chunks), half = (chunks + 1) >> 1;
  return (pos & 1) ? half + (pos >> 1) : (pos >> 1);
}

// K per-thread values summed over the workgroup: lanes with offsets 32..1, then the 8 wave sums
// with offsets 4, 2, 1. Thread 0 stores sum k at partials[k * chunks + chunk], chunk being the row
// chunk the workgroup worked on.
template <int K>
__device__ __forceinline__ void chunk_partials(double (&v)[K], double *partials, int chunks, int chunk) {
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
      partials[k * chunks + chunk] = w[0] + w[1];
    }
  }
}

// The K dot products of a previous pass, reduced like a pass's own rows: thread t adds the
// partials of workgroups t, t + 512, ... in order from 0.0, then the 64-lane and 8-wave trees.
// Every thread receives all K.
template <int K>
__device__ __forceinline__ void dot_sums(const double *partials, int chunks, double (&out)[K]) {
  __shared__ double red[K * 8];
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double *p = partials + static_cast<size_t>(k) * chunks;
    double total = 0.0;
    for (int c = threadIdx.x; c < chunks; c += kRows) total = total + p[c];
    v[k] = total;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = v[k] + __shfl_down(v[k], off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();  // an earlier call's red has been read
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k * 8 + wave] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double w[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) w[t] = red[k * 8 + t] + red[k * 8 + t + 4];
    w[0] = w[0] + w[2];
    w[1] = w[1] + w[3];
    out[k] = w[0] + w[1];
  }
}

}  // namespace chunked
}  // namespace pp
