// Face gradients of the pressure solve's adjoint (no counterpart in the reference).
//
// The coefficient-mode system is A(k) p = src + b_bc(k) with A = sum over faces f of
// c_f (e_a - e_b)(e_a - e_b)^T, c_f = k_f / h^2, and, on a wall face of cell c, a Dirichlet
// contribution 2 c_f e_c e_c^T to A and 2 c_f value e_c to b, or a Neumann contribution
// c_f value h e_c to b. For a loss L(p) with g = dL/dp the adjoint solution lam = A^+ P g (P the
// zero-mean projection when the system is singular) gives
//   dL/dsrc = lam,  dL/dc_f = -(lam_a - lam_b)(p_a - p_b)            (interior / periodic face)
//   dL/dc_f = -2 lam_c (p_c - value)  (Dirichlet wall),  lam_c value h  (Neumann wall)
// and dL/dk_f = dL/dc_f / h^2. Along a periodic direction face nx (ny) is face 0: its gradient is
// reported on face 0 and face nx (ny) gets 0.
#include "pp_kernels.hpp"

namespace pp {

namespace {

__device__ __forceinline__ double wall_gradient(const GradArgs &g, int side, int c, double h) {
  int type = g.bc_type[side];
  double value = g.bc_value[side];
  if (type == 0) {  // unset edges are homogeneous Neumann
    type = 2;
    value = 0.0;
  }
  if (type == 1) return (-2.0 * g.lam[c] * (g.p[c] - value)) / (h * h);
  return (g.lam[c] * value * h) / (h * h);
}

__device__ __forceinline__ double pair_gradient(const GradArgs &g, int a, int b, double h) {
  return -((g.lam[b] - g.lam[a]) * (g.p[b] - g.p[a])) / (h * h);
}

__global__ void __launch_bounds__(256) face_gradient_kernel(GradArgs g) {
  const int nx = g.nx, ny = g.ny;
  const int n_u = (nx + 1) * ny, n_v = nx * (ny + 1);
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n_u + n_v; f += gridDim.x * blockDim.x) {
    if (f < n_u) {
      const int i = f / ny, j = f - i * ny;
      double out;
      if (i >= 1 && i <= nx - 1)
        out = pair_gradient(g, (i - 1) * ny + j, i * ny + j, g.dx);
      else if (g.periodic_x)
        out = i == 0 ? pair_gradient(g, (nx - 1) * ny + j, j, g.dx) : 0.0;
      else
        out = i == 0 ? wall_gradient(g, 0, j, g.dx) : wall_gradient(g, 1, (nx - 1) * ny + j, g.dx);
      g.dku[f] = out;
    } else {
      const int e = f - n_u;
      const int i = e / (ny + 1), j = e - i * (ny + 1);
      double out;
      if (j >= 1 && j <= ny - 1)
        out = pair_gradient(g, i * ny + j - 1, i * ny + j, g.dy);
      else if (g.periodic_y)
        out = j == 0 ? pair_gradient(g, i * ny + ny - 1, i * ny, g.dy) : 0.0;
      else
        out = j == 0 ? wall_gradient(g, 2, i * ny, g.dy) : wall_gradient(g, 3, i * ny + ny - 1, g.dy);
      g.dkv[e] = out;
    }
  }
}

}  // namespace

hipError_t launch_face_gradient(const GradArgs &g, hipStream_t stream) {
  const int faces = (g.nx + 1) * g.ny + g.nx * (g.ny + 1);
  const int blocks = (faces + 255) / 256;
  hipLaunchKernelGGL(face_gradient_kernel, dim3(blocks), dim3(256), 0, stream, g);
  return hipGetLastError();
}

}  // namespace pp
