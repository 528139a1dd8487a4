// Ghosted Field2D access and the velocity ghost fills shared by the correction and momentum
// kernels. A Field2D of a x b interior values and G ghost layers stores element (i, j) at
// (i + G) * (b + 2G) + (j + G) (src/core/base/field.hpp:23-33).
#pragma once

#include <hip/hip_runtime.h>

#include "pp_device.hpp"
#include "pp_solver.h"

namespace pp {

// Field2D element (i, j) of a field with `cols` interior columns and G ghost layers (global
// address space, so the accesses are global_load / global_store rather than flat ones).
struct Ghosted {
  global_ptr<double> f;
  int cols, G;
  __device__ Ghosted(double *p, int c, int g) : f(as_global(p)), cols(c), G(g) {}
  __device__ __forceinline__ __attribute__((address_space(1))) double &operator()(int i, int j) const {
    return f[static_cast<size_t>(i + G) * (cols + 2 * G) + (j + G)];
  }
};

struct GhostedConst {
  global_ptr<const double> f;
  int cols, G;
  __device__ GhostedConst(const double *p, int c, int g) : f(as_global(p)), cols(c), G(g) {}
  __device__ __forceinline__ double operator()(int i, int j) const {
    return f[static_cast<size_t>(i + G) * (cols + 2 * G) + (j + G)];
  }
};

// The ghost fills run in one kThreads-thread block; Args carries nx, ny, G, dx, dy, periodic_x,
// periodic_y and the resolved velocity BCs u_type / u_value / v_type / v_value (left, right,
// bottom, top; unset edges already replaced by their no-slip / wall-speed Dirichlet defaults).

// apply_u_bc (operators.cpp:182-233) on the (nx+1) x ny face field
template <class Args>
__device__ void u_bc(const Args &a, const Ghosted &u) {
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
template <class Args>
__device__ void v_bc(const Args &a, const Ghosted &v) {
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

}  // namespace pp
