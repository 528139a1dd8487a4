// Velocity correction and pressure-history shift after a pressure solve
// (Solver::apply_pressure_velocity_correction, operators.cpp:480-534), on the device, on ghosted
// Field2D buffers.
//
// One workgroup runs the whole step: the face updates are parallel over faces, the boundary-face
// rules and the ghost layers follow the reference's pass order (layer by layer: x ghosts, then y
// ghosts over the full ghosted x range). The y pass of a layer reads x-ghost cells of deeper layers
// that the reference has not rewritten yet, i.e. the caller's old values: the buffers are updated
// in place, as the reference updates its fields. The mean removed from the new pressure is summed
// by one lane in the reference's i-outer, j-inner order, so the result is the reference's bit for
// bit.
#include <hip/hip_runtime.h>

#include "pp_device.hpp"
#include "pp_kernels.hpp"
#include "pp_solver.h"

namespace pp {

namespace {

// Field2D element (i, j) of a field with `cols` interior columns and G ghost layers.
struct Ghosted {
  double *f;
  int cols, G;
  __device__ __forceinline__ double &operator()(int i, int j) const {
    return f[static_cast<size_t>(i + G) * (cols + 2 * G) + (j + G)];
  }
};

struct GhostedConst {
  const double *f;
  int cols, G;
  __device__ __forceinline__ double operator()(int i, int j) const {
    return f[static_cast<size_t>(i + G) * (cols + 2 * G) + (j + G)];
  }
};

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

// apply_u_bc (operators.cpp:182-233) on the (nx+1) x ny face field
__device__ void u_bc(const CorrArgs &a, const Ghosted &u) {
  const int fnx = a.nx + 1, ny = a.ny, G = a.G;
  for (int j = threadIdx.x; j < ny; j += kThreads) {
    if (a.periodic_x) {
      u(fnx - 1, j) = u(0, j);
    } else {
      u(0, j) = a.u_type[0] == PP_BC_DIRICHLET ? a.u_value[0] : u(1, j) + a.u_value[0] * a.dx;
      u(fnx - 1, j) = a.u_type[1] == PP_BC_DIRICHLET ? a.u_value[1] : u(fnx - 2, j) + a.u_value[1] * a.dx;
    }
  }
  __syncthreads();
  for (int k = 1; k <= G; ++k) {
    for (int j = threadIdx.x; j < ny; j += kThreads) {
      if (a.periodic_x) {
        u(-k, j) = u(fnx - k - 1, j);
        u(fnx - 1 + k, j) = u(k, j);
      } else {
        const double lo = 2.0 * static_cast<double>(k) * a.u_value[0] * a.dx;
        const double ro = 2.0 * static_cast<double>(k) * a.u_value[1] * a.dx;
        u(-k, j) = a.u_type[0] == PP_BC_DIRICHLET ? 2.0 * a.u_value[0] - u(k, j) : u(k, j) + lo;
        u(fnx - 1 + k, j) = a.u_type[1] == PP_BC_DIRICHLET ? 2.0 * a.u_value[1] - u(fnx - 1 - k, j)
                                                            : u(fnx - 1 - k, j) + ro;
      }
    }
    __syncthreads();
    for (int i = -G + static_cast<int>(threadIdx.x); i < fnx + G; i += kThreads) {
      if (a.periodic_y) {
        u(i, -k) = u(i, ny - k);
        u(i, ny - 1 + k) = u(i, k - 1);
      } else {
        const double bo = 2.0 * (static_cast<double>(k) - 0.5) * a.u_value[2] * a.dy;
        const double to = 2.0 * (static_cast<double>(k) - 0.5) * a.u_value[3] * a.dy;
        u(i, -k) = a.u_type[2] == PP_BC_DIRICHLET ? 2.0 * a.u_value[2] - u(i, k - 1) : u(i, k - 1) + bo;
        u(i, ny - 1 + k) = a.u_type[3] == PP_BC_DIRICHLET ? 2.0 * a.u_value[3] - u(i, ny - k) : u(i, ny - k) + to;
      }
    }
    __syncthreads();
  }
}

// apply_v_bc (operators.cpp:239-283) on the nx x (ny+1) face field
__device__ void v_bc(const CorrArgs &a, const Ghosted &v) {
  const int nx = a.nx, fny = a.ny + 1, G = a.G;
  for (int i = threadIdx.x; i < nx; i += kThreads) {
    if (a.periodic_y) {
      v(i, fny - 1) = v(i, 0);
    } else {
      v(i, 0) = a.v_type[2] == PP_BC_DIRICHLET ? a.v_value[2] : v(i, 1) + a.v_value[2] * a.dy;
      v(i, fny - 1) = a.v_type[3] == PP_BC_DIRICHLET ? a.v_value[3] : v(i, fny - 2) + a.v_value[3] * a.dy;
    }
  }
  __syncthreads();
  for (int k = 1; k <= G; ++k) {
    for (int j = threadIdx.x; j < fny; j += kThreads) {
      if (a.periodic_x) {
        v(-k, j) = v(nx - k, j);
        v(nx - 1 + k, j) = v(k - 1, j);
      } else {
        const double lo = 2.0 * (static_cast<double>(k) - 0.5) * a.v_value[0] * a.dx;
        const double ro = 2.0 * (static_cast<double>(k) - 0.5) * a.v_value[1] * a.dx;
        v(-k, j) = a.v_type[0] == PP_BC_DIRICHLET ? 2.0 * a.v_value[0] - v(k - 1, j) : v(k - 1, j) + lo;
        v(nx - 1 + k, j) = a.v_type[1] == PP_BC_DIRICHLET ? 2.0 * a.v_value[1] - v(nx - k, j) : v(nx - k, j) + ro;
      }
    }
    __syncthreads();
    for (int i = -G + static_cast<int>(threadIdx.x); i < nx + G; i += kThreads) {
      if (a.periodic_y) {
        v(i, -k) = v(i, fny - k - 1);
        v(i, fny - 1 + k) = v(i, k);
      } else {
        const double bo = 2.0 * static_cast<double>(k) * a.v_value[2] * a.dy;
        const double to = 2.0 * static_cast<double>(k) * a.v_value[3] * a.dy;
        v(i, -k) = a.v_type[2] == PP_BC_DIRICHLET ? 2.0 * a.v_value[2] - v(i, k) : v(i, k) + bo;
        v(i, fny - 1 + k) = a.v_type[3] == PP_BC_DIRICHLET ? 2.0 * a.v_value[3] - v(i, fny - 1 - k)
                                                            : v(i, fny - 1 - k) + to;
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

}  // namespace

__global__ void __launch_bounds__(kThreads) velocity_correction_kernel(const CorrArgs *__restrict__ args) {
  __shared__ double red[kWaves];
  __shared__ double mean_sh;
  const CorrArgs &a = *args;
  const int nx = a.nx, ny = a.ny, G = a.G, tid = threadIdx.x;
  const GhostedConst rho{a.rho_mid, ny, G}, p_corr{a.p_corr, ny, G};
  const GhostedConst u_star{a.u_star, ny, G}, v_star{a.v_star, ny + 1, G};
  const Ghosted u{a.u, ny, G}, v{a.v, ny + 1, G}, pres{a.pressure, ny, G}, prev{a.pressure_prev, ny, G};
  const Ghosted p_ext{a.p_ext, ny, G};

  const double speed_pre = boundary_speed(a, u_star, v_star, red);

  // split extrapolation p_ext = p^n (first step) or 2 p^n - p^{n-1}, ghost-filled on a zeroed field
  // (operators.cpp:413-422)
  if (a.split) {
    const size_t cells = static_cast<size_t>(nx + 2 * G) * (ny + 2 * G);
    for (size_t q = tid; q < cells; q += kThreads) a.p_ext[q] = 0.0;
    __syncthreads();
    for (int q = tid; q < nx * ny; q += kThreads) {
      const int i = q / ny, j = q - i * ny;
      p_ext(i, j) = a.step_index == 0 ? pres(i, j) : 2.0 * pres(i, j) - prev(i, j);
    }
    __syncthreads();
    pressure_ghosts(a, p_ext);
  }
  const GhostedConst pe{a.p_ext, ny, G};

  // u = u* - dt (implicit + explicit) (operators.cpp:490-519)
  const int n_u = (nx + 1) * ny;
  for (int q = tid; q < n_u; q += kThreads) {
    const int i = q / ny, j = q - i * ny;
    const double ex = a.split ? (1.0 / rho_u_face(a, rho, i, j) - 1.0 / a.rho_ref) * ((pe(i, j) - pe(i - 1, j)) / a.dx)
                              : 0.0;
    const double gp = (p_corr(i, j) - p_corr(i - 1, j)) / a.dx;
    const double im = a.split ? gp / a.rho_ref : gp / rho_u_face(a, rho, i, j);
    u(i, j) = u_star(i, j) - a.dt * (im + ex);
  }
  const int n_v = nx * (ny + 1);
  for (int q = tid; q < n_v; q += kThreads) {
    const int i = q / (ny + 1), j = q - i * (ny + 1);
    const double ex = a.split ? (1.0 / rho_v_face(a, rho, i, j) - 1.0 / a.rho_ref) * ((pe(i, j) - pe(i, j - 1)) / a.dy)
                              : 0.0;
    const double gp = (p_corr(i, j) - p_corr(i, j - 1)) / a.dy;
    const double im = a.split ? gp / a.rho_ref : gp / rho_v_face(a, rho, i, j);
    v(i, j) = v_star(i, j) - a.dt * (im + ex);
  }
  __syncthreads();
  u_bc(a, u);
  v_bc(a, v);
  const double speed_post = boundary_speed(a, GhostedConst{a.u, ny, G}, GhostedConst{a.v, ny + 1, G}, red);

  // pressure history (operators.cpp:524-533): previous <- pressure (whole buffer), pressure <-
  // the correction's interior, its mean removed when no edge is Dirichlet, ghosts refilled
  const size_t cells = static_cast<size_t>(nx + 2 * G) * (ny + 2 * G);
  for (size_t q = tid; q < cells; q += kThreads) a.pressure_prev[q] = a.pressure[q];
  __syncthreads();
  for (int q = tid; q < nx * ny; q += kThreads) {
    const int i = q / ny, j = q - i * ny;
    pres(i, j) = p_corr(i, j);
  }
  __syncthreads();
  if (!a.has_dirichlet) {
    if (tid == 0) {
      double mean = 0.0;
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) mean += pres(i, j);
      mean_sh = mean / static_cast<double>(nx * ny);
    }
    __syncthreads();
    const double mean = mean_sh;
    for (int q = tid; q < nx * ny; q += kThreads) {
      const int i = q / ny, j = q - i * ny;
      pres(i, j) -= mean;
    }
    __syncthreads();
  }
  pressure_ghosts(a, pres);
  if (tid == 0) {
    a.speeds[0] = speed_pre;
    a.speeds[1] = speed_post;
  }
}

hipError_t launch_velocity_correction(const CorrArgs *d_args, hipStream_t stream) {
  hipLaunchKernelGGL(velocity_correction_kernel, dim3(1), dim3(kThreads), 0, stream, d_args);
  return hipGetLastError();
}

}  // namespace pp
