// Velocity correction and pressure-history shift after a pressure solve
// (Solver::apply_pressure_velocity_correction, operators.cpp:480-534), on the device, on ghosted
// Field2D buffers.
//
// Three launches on the caller's stream:
//   faces    one thread per u and v face (the split extrapolation evaluated in place), and, in
//            block 0, the mean of the correction summed by one lane in the reference's i-outer,
//            j-inner order (terms staged through LDS by the block's other waves), so it is the
//            reference's value bit for bit;
//   history  one thread per cell: pressure_prev <- pressure, new pressure interior;
//   finish   one block: boundary speeds, boundary faces and ghost layers in the reference's pass
//            order (layer by layer: x ghosts, then y ghosts over the full ghosted x range).
// The y pass of a layer reads x-ghost cells of deeper layers that the reference has not rewritten
// yet, i.e. the caller's old values: the buffers are updated in place, as the reference updates
// its fields.
#include <hip/hip_runtime.h>

#include "pp_device.hpp"
#include "pp_fields.hpp"
#include "pp_kernels.hpp"
#include "pp_solver.h"

namespace pp {

namespace {

constexpr int kSumChunk = 1024;  // terms per LDS stage of the sequential mean (two stages, 16 KB)
constexpr int kSumGroup = 32;   // terms the summing lane loads per step (kSumChunk: a multiple of 2x)
constexpr int kSumStage = kSumChunk + kSumGroup;  // one stage plus the group the last step preloads
constexpr int kWaveSize = 64;
constexpr int kFaceBlock = 256;  // threads per block of the face and history kernels
constexpr int kMeanSlot = 2;     // work[0], work[1]: boundary speeds; work[2]: the pressure mean

__device__ __forceinline__ double block_max(double v, double *red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    v = v > o ? v : o;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double m = red[0];
  for (int w = 1; w < kWaves; ++w) m = m > red[w] ? m : red[w];
  __syncthreads();
  return m;
}

// periodic_boundary_speed_stat (operators.cpp:463-478); a maximum, so any order gives its value
__device__ double boundary_speed(const CorrArgs &a, const GhostedConst &u, const GhostedConst &v, double *red) {
  double m = 0.0;
  if (a.periodic_x)
    for (int j = threadIdx.x; j < a.ny; j += kThreads) {
      m = fmax(m, fabs(u(0, j)));
      m = fmax(m, fabs(u(a.nx, j)));
    }
  if (a.periodic_y)
    for (int i = threadIdx.x; i < a.nx; i += kThreads) {
      m = fmax(m, fabs(v(i, 0)));
      m = fmax(m, fabs(v(i, a.ny)));
    }
  return block_max(m, red);
}

// apply_pressure_bc (operators.cpp:143-181) on a ghosted nx x ny field
__device__ void pressure_ghosts(const CorrArgs &a, const Ghosted &f) {
  const int nx = a.nx, ny = a.ny, G = a.G;
  for (int k = 1; k <= G; ++k) {
    for (int j = threadIdx.x; j < ny; j += kThreads) {
      if (a.periodic_x) {
        f(-k, j) = f(nx - k, j);
        f(nx - 1 + k, j) = f(k - 1, j);
      } else {
        const double lo = 2.0 * (static_cast<double>(k) - 0.5) * a.p_value[0] * a.dx;
        const double ro = 2.0 * (static_cast<double>(k) - 0.5) * a.p_value[1] * a.dx;
        f(-k, j) = a.p_type[0] == PP_BC_DIRICHLET ? 2.0 * a.p_value[0] - f(k - 1, j) : f(k - 1, j) + lo;
        f(nx - 1 + k, j) = a.p_type[1] == PP_BC_DIRICHLET ? 2.0 * a.p_value[1] - f(nx - k, j) : f(nx - k, j) + ro;
      }
    }
    __syncthreads();
    for (int i = -G + static_cast<int>(threadIdx.x); i < nx + G; i += kThreads) {
      if (a.periodic_y) {
        f(i, -k) = f(i, ny - k);
        f(i, ny - 1 + k) = f(i, k - 1);
      } else {
        const double bo = 2.0 * (static_cast<double>(k) - 0.5) * a.p_value[2] * a.dy;
        const double to = 2.0 * (static_cast<double>(k) - 0.5) * a.p_value[3] * a.dy;
        f(i, -k) = a.p_type[2] == PP_BC_DIRICHLET ? 2.0 * a.p_value[2] - f(i, k - 1) : f(i, k - 1) + bo;
        f(i, ny - 1 + k) = a.p_type[3] == PP_BC_DIRICHLET ? 2.0 * a.p_value[3] - f(i, ny - k) : f(i, ny - k) + to;
      }
    }
    __syncthreads();
  }
}

// rho_u_face / rho_v_face (operators.cpp:330-350)
__device__ __forceinline__ double rho_u_face(const CorrArgs &a, const GhostedConst &r, int i, int j) {
  const int nx = a.nx;
  if (i <= 0) return a.periodic_x ? 0.5 * (r(nx - 1, j) + r(0, j)) : r(0, j);
  if (i >= nx) return a.periodic_x ? 0.5 * (r(nx - 1, j) + r(0, j)) : r(nx - 1, j);
  return 0.5 * (r(i - 1, j) + r(i, j));
}

__device__ __forceinline__ double rho_v_face(const CorrArgs &a, const GhostedConst &r, int i, int j) {
  const int ny = a.ny;
  if (j <= 0) return a.periodic_y ? 0.5 * (r(i, ny - 1) + r(i, 0)) : r(i, 0);
  if (j >= ny) return a.periodic_y ? 0.5 * (r(i, ny - 1) + r(i, 0)) : r(i, ny - 1);
  return 0.5 * (r(i, j - 1) + r(i, j));
}

// The split extrapolation p_ext (operators.cpp:413-422) at (i, j), read where the face updates
// need it: interior cells (p^n on the first step, else 2 p^n - p^{n-1}) and the first ghost layer
// off the corners, which apply_pressure_bc fills from the interior (periodic copy, Dirichlet
// 2v - f, Neumann f + v h), evaluated here with the same expressions as pressure_ghosts.
__device__ __forceinline__ double p_ext_interior(const CorrArgs &a, const GhostedConst &p, const GhostedConst &pm,
                                                 int i, int j) {
  return a.step_index == 0 ? p(i, j) : 2.0 * p(i, j) - pm(i, j);
}

__device__ double p_ext_at(const CorrArgs &a, const GhostedConst &p, const GhostedConst &pm, int i, int j) {
  const int nx = a.nx, ny = a.ny;
  const double layer = 2.0 * (1.0 - 0.5);
  if (i < 0 || i >= nx) {
    const int side = i < 0 ? 0 : 1;
    if (a.periodic_x) return p_ext_interior(a, p, pm, i < 0 ? nx - 1 : 0, j);
    const double f = p_ext_interior(a, p, pm, i < 0 ? 0 : nx - 1, j);
    return a.p_type[side] == PP_BC_DIRICHLET ? 2.0 * a.p_value[side] - f : f + layer * a.p_value[side] * a.dx;
  }
  if (j < 0 || j >= ny) {
    const int side = j < 0 ? 2 : 3;
    if (a.periodic_y) return p_ext_interior(a, p, pm, i, j < 0 ? ny - 1 : 0);
    const double f = p_ext_interior(a, p, pm, i, j < 0 ? 0 : ny - 1);
    return a.p_type[side] == PP_BC_DIRICHLET ? 2.0 * a.p_value[side] - f : f + layer * a.p_value[side] * a.dy;
  }
  return p_ext_interior(a, p, pm, i, j);
}

// Sum of the nx x ny interior of f in i-outer, j-inner order with one rounding per term, as a
// plain sequential loop computes it (the reference's mean, operators.cpp:526-532). The additions
// form one dependent chain, so the chain runs in one lane; waves 1..7 stage the next chunk of
// terms into LDS while lane 0 adds the current one. The tail of the last chunk is padded with
// +0.0, which leaves the running sum unchanged bit for bit (it starts at +0.0 and round-to-nearest
// never produces -0.0 from a sum of a nonzero and its negation). Runs in a kFaceBlock-thread
// block; returns the sum in thread 0.
__device__ double sequential_sum(const GhostedConst &f, int nx, int ny, double (*stage)[kSumStage]) {
  const int tid = threadIdx.x, n = nx * ny;
  const int chunks = (n + kSumChunk - 1) / kSumChunk;
  constexpr int kStagers = kFaceBlock - kWaveSize;
  constexpr int kPerStager = (kSumChunk + kStagers - 1) / kStagers;
  // every load of a chunk is issued before the first LDS store (term indices clamped into the grid)
  auto fill = [&](int c) {
    const int base = c * kSumChunk, m = min(kSumChunk, n - base);
    const int padded = (m + 2 * kSumGroup - 1) / (2 * kSumGroup) * (2 * kSumGroup);
    double *dst = stage[c & 1];
    double val[kPerStager];
#pragma unroll
    for (int k = 0; k < kPerStager; ++k) {
      const int e = min(base + tid - kWaveSize + k * kStagers, n - 1), i = e / ny, j = e - i * ny;
      val[k] = f(i, j);
    }
#pragma unroll
    for (int k = 0; k < kPerStager; ++k) {
      const int q = tid - kWaveSize + k * kStagers;
      if (q < padded) dst[q] = q < m ? val[k] : 0.0;
    }
  };
  if (tid >= kWaveSize) fill(0);
  __syncthreads();
  double acc = 0.0;
  for (int c = 0; c < chunks; ++c) {
    if (tid >= kWaveSize) {
      if (c + 1 < chunks) fill(c + 1);
    } else if (tid == 0) {
      // groups A and B of kSumGroup terms alternate: B is loaded while A is added, then the next
      // A while B is added. Volatile loads and scheduling barriers keep the loads one group ahead
      // of the additions (plain loads get folded into a load at their use).
      typedef double pair_t __attribute__((ext_vector_type(2)));
      const volatile __attribute__((address_space(3))) pair_t *src =
          reinterpret_cast<const volatile __attribute__((address_space(3))) pair_t *>(lds<double>(stage[c & 1]));
      constexpr int kVec = kSumGroup / 2;
      const int pairs = (min(kSumChunk, n - c * kSumChunk) + 2 * kSumGroup - 1) / (2 * kSumGroup);
      pair_t A[kVec], B[kVec];
#pragma unroll
      for (int r = 0; r < kVec; ++r) A[r] = src[r];
      for (int p = 0; p < pairs; ++p) {
#pragma unroll
        for (int r = 0; r < kVec; ++r) B[r] = src[2 * kVec * p + kVec + r];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < kVec; ++r) {
          acc += A[r].x;
          acc += A[r].y;
        }
        __builtin_amdgcn_sched_barrier(0);
        // the last pair's "next" group is the stage's padding: read, never added
#pragma unroll
        for (int r = 0; r < kVec; ++r) A[r] = src[2 * kVec * (p + 1) + r];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < kVec; ++r) {
          acc += B[r].x;
          acc += B[r].y;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __syncthreads();
  }
  return acc;
}

// Sum of the nx x ny interior of f in the tree order: thread t of the kFaceBlock threads adds the
// cells of i-outer index e = t, t + kFaceBlock, ... from 0.0, eight loads in flight at a time; the
// 64-lane tree (offsets 32..1) and the wave tree (offsets 2, 1) combine the threads. Returns the
// sum in thread 0.
__device__ double tree_sum(const GhostedConst &f, int nx, int ny, double *red) {
  const int tid = threadIdx.x, n = nx * ny;
  constexpr int kAhead = 8;
  double acc = 0.0;
  for (int e0 = tid; e0 < n; e0 += kAhead * kFaceBlock) {
    double val[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      const int e = min(e0 + k * kFaceBlock, n - 1), i = e / ny, j = e - i * ny;
      val[k] = f(i, j);
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
      if (e0 + k * kFaceBlock < n) acc = acc + val[k];
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_down(acc, off, 64);
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  double total = 0.0;
  if (tid == 0) {
    static_assert(kFaceBlock == 256, "the wave tree below is written for four waves");
    const double w0 = red[0] + red[2], w1 = red[1] + red[3];
    total = w0 + w1;
  }
  return total;
}

}  // namespace

// Block 0: the mean of the correction (when no pressure edge is Dirichlet), in the reference's
// sequential order or the tree order (CorrArgs::exact_mean), into work[kMeanSlot].
// Other blocks: one face each, u = u* - dt (implicit + explicit)
// (operators.cpp:490-519), every face of the (nx+1) x ny and nx x (ny+1) fields including the
// boundary faces, which the finish kernel then overwrites as the reference does.
__global__ void __launch_bounds__(kFaceBlock) correction_faces_kernel(const CorrArgs *__restrict__ args) {
  __shared__ __align__(16) double stage[2][kSumStage];
  const CorrArgs &a = *args;
  const int nx = a.nx, ny = a.ny, G = a.G;
  const GhostedConst rho{a.rho_mid, ny, G}, p_corr{a.p_corr, ny, G};
  if (blockIdx.x == 0) {
    if (!a.has_dirichlet) {
      const double sum = a.exact_mean ? sequential_sum(p_corr, nx, ny, stage) : tree_sum(p_corr, nx, ny, stage[0]);
      if (threadIdx.x == 0) a.work[kMeanSlot] = sum / static_cast<double>(nx * ny);
    }
    return;
  }
  const GhostedConst pres{a.pressure, ny, G}, prev{a.pressure_prev, ny, G};
  const int q = (blockIdx.x - 1) * kFaceBlock + static_cast<int>(threadIdx.x);
  const int n_u = (nx + 1) * ny, n_v = nx * (ny + 1);
  if (q < n_u) {
    const GhostedConst u_star{a.u_star, ny, G};
    const int i = q / ny, j = q - i * ny;
    const double ex = a.split ? (1.0 / rho_u_face(a, rho, i, j) - 1.0 / a.rho_ref) *
                                    ((p_ext_at(a, pres, prev, i, j) - p_ext_at(a, pres, prev, i - 1, j)) / a.dx)
                              : 0.0;
    const double gp = (p_corr(i, j) - p_corr(i - 1, j)) / a.dx;
    const double im = a.split ? gp / a.rho_ref : gp / rho_u_face(a, rho, i, j);
    Ghosted{a.u, ny, G}(i, j) = u_star(i, j) - a.dt * (im + ex);
  } else if (q < n_u + n_v) {
    const GhostedConst v_star{a.v_star, ny + 1, G};
    const int r = q - n_u, i = r / (ny + 1), j = r - i * (ny + 1);
    const double ex = a.split ? (1.0 / rho_v_face(a, rho, i, j) - 1.0 / a.rho_ref) *
                                    ((p_ext_at(a, pres, prev, i, j) - p_ext_at(a, pres, prev, i, j - 1)) / a.dy)
                              : 0.0;
    const double gp = (p_corr(i, j) - p_corr(i, j - 1)) / a.dy;
    const double im = a.split ? gp / a.rho_ref : gp / rho_v_face(a, rho, i, j);
    Ghosted{a.v, ny + 1, G}(i, j) = v_star(i, j) - a.dt * (im + ex);
  }
}

// Pressure history (operators.cpp:524-533): previous <- pressure over the whole ghosted buffer,
// the new pressure interior <- the correction minus the mean (when no edge is Dirichlet). Each
// thread owns one cell of both buffers, so the old value is read before it is replaced; the
// pressure ghost cells keep their old values for the finish kernel's ghost passes.
__global__ void __launch_bounds__(kFaceBlock) correction_history_kernel(const CorrArgs *__restrict__ args) {
  const CorrArgs &a = *args;
  const int ld = a.ny + 2 * a.G;
  const int cells = (a.nx + 2 * a.G) * ld;
  const int q = blockIdx.x * kFaceBlock + static_cast<int>(threadIdx.x);
  if (q >= cells) return;
  const global_ptr<double> pressure = as_global(a.pressure);
  const double old = pressure[q];
  as_global(a.pressure_prev)[q] = old;
  const int i = q / ld - a.G, j = q % ld - a.G;
  if (i >= 0 && i < a.nx && j >= 0 && j < a.ny) {
    const double c = as_global(a.p_corr)[q];
    pressure[q] = a.has_dirichlet ? c : c - as_global(a.work)[kMeanSlot];
  }
}

// The ordered part in one block: boundary speeds before and after (operators.cpp:463-478),
// apply_u_bc / apply_v_bc on the corrected faces and apply_pressure_bc on the new pressure.
__global__ void __launch_bounds__(kThreads) correction_finish_kernel(const CorrArgs *__restrict__ args) {
  __shared__ double red[kWaves];
  const CorrArgs &a = *args;
  const int ny = a.ny, G = a.G;
  const double speed_pre =
      boundary_speed(a, GhostedConst{a.u_star, ny, G}, GhostedConst{a.v_star, ny + 1, G}, red);
  u_bc(a, Ghosted{a.u, ny, G});
  v_bc(a, Ghosted{a.v, ny + 1, G});
  const double speed_post = boundary_speed(a, GhostedConst{a.u, ny, G}, GhostedConst{a.v, ny + 1, G}, red);
  pressure_ghosts(a, Ghosted{a.pressure, ny, G});
  if (threadIdx.x == 0) {
    a.work[0] = speed_pre;
    a.work[1] = speed_post;
  }
}

// apply_pressure_bc (operators.cpp:143-181) on one ghosted nx x ny field in place, one block.
__global__ void __launch_bounds__(kThreads) pressure_ghost_kernel(const CorrArgs *__restrict__ args, double *field) {
  const CorrArgs &a = *args;
  pressure_ghosts(a, Ghosted{field, a.ny, a.G});
}

hipError_t launch_pressure_ghosts(const CorrArgs *d_args, double *d_field, hipStream_t stream) {
  hipLaunchKernelGGL(pressure_ghost_kernel, dim3(1), dim3(kThreads), 0, stream, d_args, d_field);
  return hipGetLastError();
}

hipError_t launch_velocity_correction(const CorrArgs *d_args, int nx, int ny, int ghost, hipStream_t stream) {
  const int faces = (nx + 1) * ny + nx * (ny + 1);
  const int cells = (nx + 2 * ghost) * (ny + 2 * ghost);
  hipLaunchKernelGGL(correction_faces_kernel, dim3(1 + (faces + kFaceBlock - 1) / kFaceBlock), dim3(kFaceBlock), 0,
                     stream, d_args);
  hipLaunchKernelGGL(correction_history_kernel, dim3((cells + kFaceBlock - 1) / kFaceBlock), dim3(kFaceBlock), 0,
                     stream, d_args);
  hipLaunchKernelGGL(correction_finish_kernel, dim3(1), dim3(kThreads), 0, stream, d_args);
  return hipGetLastError();
}

}  // namespace pp
