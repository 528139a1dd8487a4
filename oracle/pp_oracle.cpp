// CPU oracle for the pressure-projection path. TEST INFRASTRUCTURE ONLY.
//
// A single-threaded restatement of the reference's variable-coefficient pressure solve, written
// for checking the HIP engine: only tests/, the smoke check and bench.py's CPU-baseline leg may
// load it. It follows, operation for operation, the reference functions cited at each routine
// (paths relative to the reference repository):
//   assembly        src/core/coupled/pressure.cpp:123-196, operators.cpp:330-350, 388-390, 411-444
//   regularization  pressure.cpp:198-212
//   CSR             src/core/linear_algebra/ch_sparse_krylov.hpp:105-133
//   ICC(0)/ILDLT(0) ch_sparse_krylov.hpp:337-419 / 252-335
//   sweeps, SpMV    ch_sparse_krylov.hpp:444-513 / 235-250
//   PCG             ch_sparse_krylov.hpp:515-660 with pressure.cpp:225-239 projections
//   scatter         pressure.cpp:241-248, operators.cpp:447-460, ghost fill operators.cpp:143-181
// Arithmetic is unfused double precision (-ffp-contract=off) in the reference's order, so with
// sequential reductions it reproduces the reference bit for bit. The reduction mode can be
// switched to a pairwise tree to measure how strongly the result depends on summation order.
//
// Pinning: bitwise against the compiled reference (oracle/_ref) and the committed golden vectors
// in tests/golden/, plus closed-form invariants (symmetry, A*1 = 0, IC(0) pattern identity, CG
// residual identity) in tests/test_oracle.py.
//
// Extras with no reference counterpart (mixed-precision refinement and the adjoint) are defined
// in this file too, in FP64, as the checkers of the engine's new functions.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {

struct OrProblem {
  int nx, ny;
  double lx, ly, dt;
  int periodic_x, periodic_y;
  int bc_type[4];  // left, right, bottom, top; 0 unset, 1 dirichlet, 2 neumann
  double bc_value[4];
  int max_iterations;
  double tolerance;
  double density_ratio;
  int step_index;
};

// Velocity boundary conditions of the correction step (config.hpp:105-123): u and v per edge
// (left, right, bottom, top; 0 unset, 1 dirichlet, 2 neumann) and the wall speeds that unset u
// bottom/top edges take.
struct OrVelocityBc {
  int u_type[4];
  double u_value[4];
  int v_type[4];
  double v_value[4];
  double bottom_wall_velocity_x, top_wall_velocity_x;
};

// The scalar fields of one CGMonitor call (CGMonitorData, ch_sparse_krylov.hpp:47-58); record k
// belongs to monitor call k (iteration k).
struct OrMonitorRecord {
  int iteration, pad;
  double alpha, beta, absolute_residual, relative_residual, iterate_residual, preconditioned_residual;
};

}  // extern "C"

namespace {

enum Side { kLeft = 0, kRight = 1, kBottom = 2, kTop = 3 };
enum BcType { kUnset = 0, kDirichlet = 1, kNeumann = 2 };
enum Reduction { kSequential = 0, kPairwise = 1, kCanonical = 2 };

// Summation orders. kSequential is std::accumulate / dot_product (the reference). kPairwise is a
// recursive halving tree over natural order. kCanonical is the HIP engine's block reduction:
// positions of the wavefront layout (level d = i + j ascending, i ascending within a level) are
// dealt round-robin to 512 threads, each thread sums its positions in order from 0.0, lanes of a
// 64-wide wave combine with offsets 32..1 (v[l] += v[l + off]) and the 8 wave sums combine with
// offsets 4..1.
struct Reducer {
  int mode = kSequential;
  std::vector<int> perm;  // kCanonical: natural index of each wavefront position
  int nx = 0, ny = 0;

  Reducer(int m, int nx_, int ny_) : mode(m), nx(nx_), ny(ny_) {
    const int nx = nx_, ny = ny_;
    if (mode == kCanonical) {
      perm.reserve(static_cast<std::size_t>(nx) * ny);
      for (int d = 0; d < nx + ny - 1; ++d)
        for (int i = std::max(0, d - ny + 1); i <= std::min(d, nx - 1); ++i) perm.push_back(i * ny + (d - i));
    }
  }

  template <class F>
  double sum(std::size_t n, F f) const {
    if (mode == kSequential) {
      double s = 0.0;
      for (std::size_t k = 0; k < n; ++k) s += f(k);
      return s;
    }
    if (mode == kPairwise) {
      std::vector<double> t(n);
      for (std::size_t k = 0; k < n; ++k) t[k] = f(k);
      std::size_t len = n;
      while (len > 1) {
        const std::size_t half = (len + 1) / 2;
        for (std::size_t k = 0; k + half < len; ++k) t[k] = t[k] + t[k + half];
        len = half;
      }
      return n ? t[0] : 0.0;
    }
    constexpr std::size_t T = 512, W = 8;
    double part[T];
    for (std::size_t t = 0; t < T; ++t) {
      double s = 0.0;
      for (std::size_t q = t; q < n; q += T) s = s + f(static_cast<std::size_t>(perm[q]));
      part[t] = s;
    }
    double wsum[W];
    for (std::size_t w = 0; w < W; ++w) {
      double *v = part + 64 * w;
      for (std::size_t off = 32; off >= 1; off >>= 1)
        for (std::size_t l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
      wsum[w] = v[0];
    }
    for (std::size_t off = W / 2; off >= 1; off >>= 1)
      for (std::size_t t = 0; t < off; ++t) wsum[t] = wsum[t] + wsum[t + off];
    return wsum[0];
  }

  // The engine's SpMV order for p^T A p. Grids with nx <= 256 use its windowed pass: wave w of 8
  // takes the levels [split[w], split[w+1]) (window_split below), lane l of chunk c holds row
  // 64 c + l, and each of the 512 threads sums its cells in (level, chunk) order from 0.0. Wider
  // grids use the level-major pass: wave w of 8 takes levels
  // d = w, w + 8, ...; within a level, cell i goes to lane (i - lo(d)) mod 64; each of the 512
  // threads sums its cells in (level, row) order from 0.0; then the same lane and wave trees.
  // Other modes: the same as sum().
  template <class F>
  double sum_levels(std::size_t n, F f) const {
    if (mode != kCanonical) return sum(n, f);
    constexpr std::size_t T = 512, W = 8;
    double part[T];
    for (std::size_t t = 0; t < T; ++t) part[t] = 0.0;
    if (nx <= 256) {
      int split[W + 1];
      window_split(nx, ny, split);
      for (std::size_t w = 0; w < W; ++w)
        for (int d = split[w]; d < split[w + 1]; ++d) {
          const int lo = std::max(0, d - ny + 1), hi = std::min(d, nx - 1);
          for (int c = 0; 64 * c < nx; ++c)
            for (int l = 0; l < 64; ++l) {
              const int i = 64 * c + l;
              if (i < lo || i > hi) continue;
              const std::size_t t = 64 * w + static_cast<std::size_t>(l);
              part[t] = part[t] + f(static_cast<std::size_t>(i) * static_cast<std::size_t>(ny) +
                                    static_cast<std::size_t>(d - i));
            }
        }
      return tree(part);
    }
    for (int d = 0; d < nx + ny - 1; ++d) {
      const int lo = std::max(0, d - ny + 1), hi = std::min(d, nx - 1);
      const std::size_t w = static_cast<std::size_t>(d) % W;
      for (int i = lo; i <= hi; ++i) {
        const std::size_t t = 64 * w + static_cast<std::size_t>(i - lo) % 64;
        part[t] = part[t] + f(static_cast<std::size_t>(i) * static_cast<std::size_t>(ny) + static_cast<std::size_t>(d - i));
      }
    }
    return tree(part);
  }

  // Lane tree (offsets 32 .. 1) inside each wave, then the wave tree (offsets 4, 2, 1).
  static double tree(double *part) {
    constexpr std::size_t W = 8;
    double wsum[W];
    for (std::size_t w = 0; w < W; ++w) {
      double *v = part + 64 * w;
      for (std::size_t off = 32; off >= 1; off >>= 1)
        for (std::size_t l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
      wsum[w] = v[0];
    }
    for (std::size_t off = W / 2; off >= 1; off >>= 1)
      for (std::size_t t = 0; t < off; ++t) wsum[t] = wsum[t] + wsum[t + off];
    return wsum[0];
  }

  // The engine's SpMV level ranges: boundary w at the first level where the running cell count
  // reaches w n / 8 (empty ranges on small grids).
  static void window_split(int nx, int ny, int *split) {
    constexpr int W = 8;
    const int nlev = nx + ny - 1;
    // a level costs its 64-row chunks that hold a row of it, plus 160 (the engine's kSpmvLevelCost)
    auto cost = [&](int d) { return 64LL * (std::min(d, nx - 1) / 64 - std::max(0, d - ny + 1) / 64 + 1) + 160; };
    long long total = 0;
    for (int d = 0; d < nlev; ++d) total += cost(d);
    long long cum = 0;
    int w = 1;
    split[0] = 0;
    for (int d = 0; d < nlev && w < W; ++d) {
      cum += cost(d);
      while (w < W && cum * W >= w * total) split[w++] = d + 1;
    }
    while (w <= W) split[w++] = nlev;
  }
};

struct Bc {
  int type;
  double value;
};

struct Grid {
  const OrProblem &pb;
  double dx, dy;
  explicit Grid(const OrProblem &p)
      : pb(p), dx(p.lx / std::max(static_cast<double>(p.nx), 1.0)), dy(p.ly / std::max(static_cast<double>(p.ny), 1.0)) {}

  // Unset edges act as homogeneous Neumann (operators.cpp:36-60).
  Bc bc(int side) const {
    if (pb.bc_type[side] != kUnset) return {pb.bc_type[side], pb.bc_value[side]};
    return {kNeumann, 0.0};
  }
  // operators.cpp:114-128
  bool has_dirichlet() const {
    if (!pb.periodic_x && (bc(kLeft).type == kDirichlet || bc(kRight).type == kDirichlet)) return true;
    if (!pb.periodic_y && (bc(kBottom).type == kDirichlet || bc(kTop).type == kDirichlet)) return true;
    return false;
  }
};

// Cell-centred field accessor over an interior-only i-major array.
struct CellView {
  const double *a;
  int ny;
  double operator()(int i, int j) const { return a[static_cast<std::size_t>(i) * ny + j]; }
};

// Face densities: arithmetic mean of the two cells, wall faces take the adjacent cell, periodic
// faces average across the wrap (operators.cpp:330-350).
double rho_xface(const Grid &g, const CellView &rho, int i, int j) {
  const int nx = g.pb.nx;
  if (i <= 0) return g.pb.periodic_x ? 0.5 * (rho(nx - 1, j) + rho(0, j)) : rho(0, j);
  if (i >= nx) return g.pb.periodic_x ? 0.5 * (rho(nx - 1, j) + rho(0, j)) : rho(nx - 1, j);
  return 0.5 * (rho(i - 1, j) + rho(i, j));
}

double rho_yface(const Grid &g, const CellView &rho, int i, int j) {
  const int ny = g.pb.ny;
  if (j <= 0) return g.pb.periodic_y ? 0.5 * (rho(i, ny - 1) + rho(i, 0)) : rho(i, 0);
  if (j >= ny) return g.pb.periodic_y ? 0.5 * (rho(i, ny - 1) + rho(i, 0)) : rho(i, ny - 1);
  return 0.5 * (rho(i, j - 1) + rho(i, j));
}

// (u(i+1,j)-u(i,j))/dx + (v(i,j+1)-v(i,j))/dy on the staggered faces (operators.cpp:388-390).
double cell_divergence(const Grid &g, const double *u, const double *v, int i, int j) {
  const int ny = g.pb.ny;
  const double ue = u[static_cast<std::size_t>(i + 1) * ny + j], uw = u[static_cast<std::size_t>(i) * ny + j];
  const double vn = v[static_cast<std::size_t>(i) * (ny + 1) + j + 1], vs = v[static_cast<std::size_t>(i) * (ny + 1) + j];
  return (ue - uw) / g.dx + (vn - vs) / g.dy;
}

// Pressure value with periodic wrap; only wrapped neighbours are ever read outside the interior
// (the non-periodic wall fluxes are zero), so ghost layers reduce to the wrap (operators.cpp:413-444).
double wrap_value(const Grid &g, const std::vector<double> &p, int i, int j) {
  const int nx = g.pb.nx, ny = g.pb.ny;
  if (i < 0) i += nx;
  if (i >= nx) i -= nx;
  if (j < 0) j += ny;
  if (j >= ny) j -= ny;
  return p[static_cast<std::size_t>(i) * ny + j];
}

// Split explicit term sum over faces of (1 - rho_ref/rho_f) * grad(p_ext) (operators.cpp:424-444).
double split_explicit_divergence(const Grid &g, const CellView &rho, const std::vector<double> &pext, double rho_ref,
                                 int i, int j) {
  const OrProblem &pb = g.pb;
  const bool east_open = pb.periodic_x || i < pb.nx - 1;
  const bool west_open = pb.periodic_x || i > 0;
  const bool north_open = pb.periodic_y || j < pb.ny - 1;
  const bool south_open = pb.periodic_y || j > 0;
  const double pc = pext[static_cast<std::size_t>(i) * pb.ny + j];
  const double fe = east_open ? (1.0 - rho_ref / rho_xface(g, rho, i + 1, j)) * ((wrap_value(g, pext, i + 1, j) - pc) / g.dx) : 0.0;
  const double fw = west_open ? (1.0 - rho_ref / rho_xface(g, rho, i, j)) * ((pc - wrap_value(g, pext, i - 1, j)) / g.dx) : 0.0;
  const double fn = north_open ? (1.0 - rho_ref / rho_yface(g, rho, i, j + 1)) * ((wrap_value(g, pext, i, j + 1) - pc) / g.dy) : 0.0;
  const double fs = south_open ? (1.0 - rho_ref / rho_yface(g, rho, i, j)) * ((pc - wrap_value(g, pext, i, j - 1)) / g.dy) : 0.0;
  return (fe - fw) / g.dx + (fn - fs) / g.dy;
}

struct Csr {
  int n = 0;
  std::vector<int> row_ptr, col, diag_pos;
  std::vector<double> val;
};

struct System {
  Csr A;
  std::vector<double> rhs;
  bool has_dirichlet = false;
  double assembly_max_diag = 0.0;
};

// Coefficient mode (no reference counterpart): face inverse densities k_u ((nx+1) x ny), k_v
// (nx x (ny+1)) give c_f = k_f / h^2, src gives the right-hand side before boundary terms, and the
// boundary right-hand-side terms are added only when rhs_bc is set. Along a periodic direction the
// last face is face 0.
struct Coeffs {
  const double *ku, *kv, *src;
  bool rhs_bc;
  double kx(const OrProblem &pb, int i, int j) const {
    if (pb.periodic_x && i == pb.nx) i = 0;
    return ku[static_cast<std::size_t>(i) * pb.ny + j];
  }
  double ky(const OrProblem &pb, int i, int j) const {
    if (pb.periodic_y && j == pb.ny) j = 0;
    return kv[static_cast<std::size_t>(i) * (pb.ny + 1) + j];
  }
};

// Assembly + regularization + CSR conversion (pressure.cpp:123-212, ch_sparse_krylov.hpp:105-133).
// Each row gathers its entries in the reference's insertion order (east, west, north, south,
// centre), accumulating repeated columns in that order, then sorts by column as std::map would.
System assemble(const OrProblem &pb, const double *rho_in, const double *u, const double *v, const double *p_n,
                const double *p_nm1, bool use_split, const Reducer &R, const Coeffs *cf = nullptr) {
  Grid g(pb);
  const int nx = pb.nx, ny = pb.ny, n = nx * ny;
  const CellView rho{rho_in, ny};
  System sys;
  sys.has_dirichlet = g.has_dirichlet();
  sys.rhs.assign(static_cast<std::size_t>(n), 0.0);
  const double rho_ref = std::min(1.0, pb.density_ratio);

  std::vector<double> pext;
  if (use_split) {
    pext.assign(static_cast<std::size_t>(n), 0.0);
    for (int k = 0; k < n; ++k) {
      const double pn = p_n ? p_n[k] : 0.0, pm = p_nm1 ? p_nm1[k] : 0.0;
      pext[static_cast<std::size_t>(k)] = (pb.step_index == 0) ? pn : (2.0 * pn - pm);
    }
  }

  Csr &A = sys.A;
  A.n = n;
  A.row_ptr.assign(1, 0);
  A.diag_pos.assign(static_cast<std::size_t>(n), -1);
  double max_diag = 0.0;
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) {
      const int row = i * ny + j;
      const bool has_e = pb.periodic_x || i < nx - 1;
      const bool has_w = pb.periodic_x || i > 0;
      const bool has_n = pb.periodic_y || j < ny - 1;
      const bool has_s = pb.periodic_y || j > 0;
      double ce, cw, cn, cs, b;
      if (cf) {
        ce = cf->kx(pb, i + 1, j) / (g.dx * g.dx);
        cw = cf->kx(pb, i, j) / (g.dx * g.dx);
        cn = cf->ky(pb, i, j + 1) / (g.dy * g.dy);
        cs = cf->ky(pb, i, j) / (g.dy * g.dy);
        b = cf->src[static_cast<std::size_t>(row)];
      } else {
        ce = use_split ? 1.0 / (g.dx * g.dx) : 1.0 / (rho_xface(g, rho, i + 1, j) * g.dx * g.dx);
        cw = use_split ? 1.0 / (g.dx * g.dx) : 1.0 / (rho_xface(g, rho, i, j) * g.dx * g.dx);
        cn = use_split ? 1.0 / (g.dy * g.dy) : 1.0 / (rho_yface(g, rho, i, j + 1) * g.dy * g.dy);
        cs = use_split ? 1.0 / (g.dy * g.dy) : 1.0 / (rho_yface(g, rho, i, j) * g.dy * g.dy);
        b = use_split ? -(split_explicit_divergence(g, rho, pext, rho_ref, i, j) +
                          rho_ref * cell_divergence(g, u, v, i, j) / pb.dt)
                      : -cell_divergence(g, u, v, i, j) / pb.dt;
      }
      const bool rhs_bc = !cf || cf->rhs_bc;
      double diag = 0.0;

      int cols[5];
      double vals[5];
      int m = 0;
      auto put = [&](int c, double value) {
        for (int k = 0; k < m; ++k) {
          if (cols[k] == c) {
            vals[k] += value;
            return;
          }
        }
        cols[m] = c;
        vals[m] = 0.0 + value;
        ++m;
      };
      auto boundary = [&](int side, double coeff, double spacing) {
        const Bc bc = g.bc(side);
        if (bc.type == kDirichlet) {
          diag += 2.0 * coeff;
          if (rhs_bc) b += 2.0 * coeff * bc.value;
        } else if (rhs_bc) {
          b += coeff * bc.value * spacing;
        }
      };
      if (has_e) {
        diag += ce;
        put((pb.periodic_x && i == nx - 1 ? 0 : i + 1) * ny + j, -ce);
      } else {
        boundary(kRight, ce, g.dx);
      }
      if (has_w) {
        diag += cw;
        put((pb.periodic_x && i == 0 ? nx - 1 : i - 1) * ny + j, -cw);
      } else {
        boundary(kLeft, cw, g.dx);
      }
      if (has_n) {
        diag += cn;
        put(i * ny + (pb.periodic_y && j == ny - 1 ? 0 : j + 1), -cn);
      } else {
        boundary(kTop, cn, g.dy);
      }
      if (has_s) {
        diag += cs;
        put(i * ny + (pb.periodic_y && j == 0 ? ny - 1 : j - 1), -cs);
      } else {
        boundary(kBottom, cs, g.dy);
      }
      put(row, diag);
      sys.rhs[static_cast<std::size_t>(row)] = b;
      max_diag = std::max(max_diag, diag);

      // sort by column (insertion sort on <= 5 entries), then drop |v| < 1e-18 like finalize_row_maps
      for (int a = 1; a < m; ++a)
        for (int k = a; k > 0 && cols[k - 1] > cols[k]; --k) {
          std::swap(cols[k - 1], cols[k]);
          std::swap(vals[k - 1], vals[k]);
        }
      for (int k = 0; k < m; ++k) {
        if (std::abs(vals[k]) < 1.0e-18) continue;
        A.col.push_back(cols[k]);
        A.val.push_back(vals[k]);
        if (cols[k] == row) A.diag_pos[static_cast<std::size_t>(row)] = static_cast<int>(A.val.size()) - 1;
      }
      A.row_ptr.push_back(static_cast<int>(A.val.size()));
    }
  }
  for (int r = 0; r < n; ++r)
    if (A.diag_pos[static_cast<std::size_t>(r)] < 0)
      throw std::runtime_error("sparse matrix assembly failed to produce a diagonal entry");
  sys.assembly_max_diag = max_diag;

  if (!sys.has_dirichlet) {
    const double total = R.sum(sys.rhs.size(), [&](std::size_t k) { return sys.rhs[k]; });
    const double mean = total / std::max(static_cast<double>(n), 1.0);
    for (double &x : sys.rhs) x -= mean;
  }
  return sys;
}

struct Factor {
  int kind = 1;  // 1 ICC(0), 2 ILDLT(0)
  std::vector<double> lower, diag;
  double floor = 1.0e-30;
  // transpose lists: for column c, the (row, L value) pairs of the rows below it, ascending rows
  std::vector<int> t_ptr, t_row;
  std::vector<double> t_val;
};

// ICC(0) (ch_sparse_krylov.hpp:337-419) and ILDLT(0) (252-335) in natural order on the CSR
// pattern. Returns false when a pivot is non-finite or at/below the floor.
bool factorize(const Csr &A, int kind, Factor &F) {
  const int n = A.n;
  F.kind = kind;
  F.lower.assign(A.val.size(), 0.0);
  F.diag.assign(static_cast<std::size_t>(n), 0.0);
  double max_diag = 0.0;
  for (int r = 0; r < n; ++r) max_diag = std::max(max_diag, std::abs(A.val[static_cast<std::size_t>(A.diag_pos[static_cast<std::size_t>(r)])]));
  F.floor = (kind == 2) ? std::max(1.0, max_diag) * 1.0e-12 : std::sqrt(std::max(1.0, max_diag) * 1.0e-12);
  const double floor_sq = F.floor * F.floor;

  // L(row, target) if that strictly-lower entry exists, else 0 (scan stops past the target).
  auto lookup = [&](int row, int target) {
    for (int q = A.row_ptr[static_cast<std::size_t>(row)]; q < A.row_ptr[static_cast<std::size_t>(row) + 1]; ++q) {
      const int c = A.col[static_cast<std::size_t>(q)];
      if (c >= row || c > target) break;
      if (c == target) return F.lower[static_cast<std::size_t>(q)];
    }
    return 0.0;
  };

  for (int r = 0; r < n; ++r) {
    const int beg = A.row_ptr[static_cast<std::size_t>(r)], end = A.row_ptr[static_cast<std::size_t>(r) + 1];
    for (int q = beg; q < end; ++q) {
      const int c = A.col[static_cast<std::size_t>(q)];
      if (c >= r) break;
      double s = A.val[static_cast<std::size_t>(q)];
      for (int q2 = beg; q2 < q; ++q2) {
        const int k = A.col[static_cast<std::size_t>(q2)];
        if (k >= c) break;
        const double ljk = lookup(c, k);
        if (ljk != 0.0) {
          if (kind == 2)
            s -= F.lower[static_cast<std::size_t>(q2)] * F.diag[static_cast<std::size_t>(k)] * ljk;
          else
            s -= F.lower[static_cast<std::size_t>(q2)] * ljk;
        }
      }
      F.lower[static_cast<std::size_t>(q)] = s / std::max(F.diag[static_cast<std::size_t>(c)], F.floor);
    }
    double d = A.val[static_cast<std::size_t>(A.diag_pos[static_cast<std::size_t>(r)])];
    for (int q = beg; q < end; ++q) {
      const int c = A.col[static_cast<std::size_t>(q)];
      if (c >= r) break;
      const double l = F.lower[static_cast<std::size_t>(q)];
      if (kind == 2)
        d -= (l * l) * F.diag[static_cast<std::size_t>(c)];
      else
        d -= l * l;
    }
    if (kind == 2) {
      if (!std::isfinite(d) || d <= F.floor) return false;
      F.diag[static_cast<std::size_t>(r)] = d;
    } else {
      if (!std::isfinite(d) || d <= floor_sq) return false;
      F.diag[static_cast<std::size_t>(r)] = std::sqrt(d);
    }
  }

  // column-wise transpose of the strictly lower part, rows ascending within each column
  F.t_ptr.assign(static_cast<std::size_t>(n) + 1, 0);
  for (int r = 0; r < n; ++r)
    for (int q = A.row_ptr[static_cast<std::size_t>(r)]; q < A.row_ptr[static_cast<std::size_t>(r) + 1]; ++q)
      if (A.col[static_cast<std::size_t>(q)] < r) ++F.t_ptr[static_cast<std::size_t>(A.col[static_cast<std::size_t>(q)]) + 1];
  for (int c = 0; c < n; ++c) F.t_ptr[static_cast<std::size_t>(c) + 1] += F.t_ptr[static_cast<std::size_t>(c)];
  F.t_row.assign(static_cast<std::size_t>(F.t_ptr[static_cast<std::size_t>(n)]), 0);
  F.t_val.assign(F.t_row.size(), 0.0);
  std::vector<int> fill(F.t_ptr.begin(), F.t_ptr.end() - 1);
  for (int r = 0; r < n; ++r)
    for (int q = A.row_ptr[static_cast<std::size_t>(r)]; q < A.row_ptr[static_cast<std::size_t>(r) + 1]; ++q) {
      const int c = A.col[static_cast<std::size_t>(q)];
      if (c < r) {
        const int at = fill[static_cast<std::size_t>(c)]++;
        F.t_row[static_cast<std::size_t>(at)] = r;
        F.t_val[static_cast<std::size_t>(at)] = F.lower[static_cast<std::size_t>(q)];
      }
    }
  return true;
}

// y = A x, row sums in column order (ch_sparse_krylov.hpp:235-250).
void spmv(const Csr &A, const double *x, double *y) {
  for (int r = 0; r < A.n; ++r) {
    double s = 0.0;
    for (int q = A.row_ptr[static_cast<std::size_t>(r)]; q < A.row_ptr[static_cast<std::size_t>(r) + 1]; ++q)
      s += A.val[static_cast<std::size_t>(q)] * x[A.col[static_cast<std::size_t>(q)]];
    y[r] = s;
  }
}

// z = M^{-1} r by forward and backward sweeps (ch_sparse_krylov.hpp:444-513).
void precondition(const Csr &A, const Factor &F, const double *r, double *z) {
  const int n = A.n;
  std::vector<double> y(static_cast<std::size_t>(n), 0.0);
  for (int row = 0; row < n; ++row) {
    double s = r[row];
    for (int q = A.row_ptr[static_cast<std::size_t>(row)]; q < A.row_ptr[static_cast<std::size_t>(row) + 1]; ++q) {
      const int c = A.col[static_cast<std::size_t>(q)];
      if (c >= row) break;
      s -= F.lower[static_cast<std::size_t>(q)] * y[static_cast<std::size_t>(c)];
    }
    y[static_cast<std::size_t>(row)] = (F.kind == 2) ? s : s / std::max(F.diag[static_cast<std::size_t>(row)], F.floor);
  }
  if (F.kind == 2)
    for (int row = 0; row < n; ++row) y[static_cast<std::size_t>(row)] /= std::max(F.diag[static_cast<std::size_t>(row)], F.floor);
  for (int row = 0; row < n; ++row) z[row] = y[static_cast<std::size_t>(row)];
  for (int row = n - 1; row >= 0; --row) {
    double s = z[row];
    for (int t = F.t_ptr[static_cast<std::size_t>(row)]; t < F.t_ptr[static_cast<std::size_t>(row) + 1]; ++t)
      s -= F.t_val[static_cast<std::size_t>(t)] * z[F.t_row[static_cast<std::size_t>(t)]];
    z[row] = (F.kind == 2) ? s : s / std::max(F.diag[static_cast<std::size_t>(row)], F.floor);
  }
}

struct Report {
  int iterations = 0;
  double abs_res = 0.0, rel_res = 0.0, iterate_res = 0.0;
};

// Zero-mean projection (pressure.cpp:225-239).
void project(std::vector<double> &w, const Reducer &R) {
  if (w.empty()) return;
  const double mean = R.sum(w.size(), [&](std::size_t k) { return w[k]; }) /
                      std::max(static_cast<double>(w.size()), 1.0);
  for (double &x : w) x -= mean;
}

// PCG from x0 = 0 (ch_sparse_krylov.hpp:515-660, pressure.cpp:50-57).
// log (optional) receives one record per CGMonitor call at the reference's call sites
// (ch_sparse_krylov.hpp:547-555, 569-579, 612-621, 636-648).
Report pcg(const Csr &A, const Factor &F, const std::vector<double> &b, int max_iterations, double tol, bool projected,
           const Reducer &R, std::vector<double> &x, std::vector<OrMonitorRecord> *log = nullptr) {
  auto monitor = [&](double alpha, double beta, double abs_r, double rel_r, double iter_r, double prec_r) {
    if (!log) return;
    OrMonitorRecord m{static_cast<int>(log->size()), 0, alpha, beta, abs_r, rel_r, iter_r, prec_r};
    log->push_back(m);
  };
  const int n = A.n;
  if (n <= 0) throw std::runtime_error("cannot solve an empty linear system");
  auto dot = [&](const std::vector<double> &a, const std::vector<double> &c) {
    return R.sum(static_cast<std::size_t>(n), [&](std::size_t k) { return a[k] * c[k]; });
  };
  x.assign(static_cast<std::size_t>(n), 0.0);
  std::vector<double> ax(static_cast<std::size_t>(n));
  spmv(A, x.data(), ax.data());
  std::vector<double> r(static_cast<std::size_t>(n));
  for (int k = 0; k < n; ++k) r[static_cast<std::size_t>(k)] = b[static_cast<std::size_t>(k)] - ax[static_cast<std::size_t>(k)];
  if (projected) project(r, R);
  const double bnorm = std::sqrt(std::max(dot(b, b), 0.0));
  Report rep;
  rep.abs_res = std::sqrt(std::max(dot(r, r), 0.0));
  rep.rel_res = rep.abs_res / std::max(bnorm, 1.0e-30);
  if (rep.rel_res < tol || rep.abs_res < tol) {
    monitor(0.0, 0.0, rep.abs_res, rep.rel_res, rep.iterate_res, 0.0);
    return rep;
  }

  std::vector<double> z(static_cast<std::size_t>(n));
  precondition(A, F, r.data(), z.data());
  if (projected) project(z, R);
  double rz = dot(r, z);
  if (!std::isfinite(rz) || rz <= 0.0) throw std::runtime_error("pcg preconditioner breakdown: non-positive r^T M^{-1} r");
  monitor(0.0, 0.0, rep.abs_res, rep.rel_res, rep.iterate_res, std::sqrt(std::max(rz, 0.0)));
  std::vector<double> p(z), ap(static_cast<std::size_t>(n));
  const int limit = std::max(1, max_iterations);
  for (int it = 0; it < limit; ++it) {
    spmv(A, p.data(), ap.data());
    const double pap = R.sum_levels(static_cast<std::size_t>(n), [&](std::size_t k) { return p[k] * ap[k]; });
    if (!std::isfinite(pap) || std::abs(pap) < 1.0e-30) throw std::runtime_error("pcg breakdown: p^T A p is invalid");
    const double alpha = rz / pap;
    double upd = 0.0, xsq = 0.0;
    for (int k = 0; k < n; ++k) {
      const std::size_t s = static_cast<std::size_t>(k);
      const double delta = alpha * p[s];
      x[s] += delta;
      r[s] -= alpha * ap[s];
      upd += delta * delta;
      xsq += x[s] * x[s];
    }
    if (R.mode != kSequential) {
      upd = R.sum(static_cast<std::size_t>(n), [&](std::size_t k) {
        const double delta = alpha * p[k];
        return delta * delta;
      });
      xsq = R.sum(static_cast<std::size_t>(n), [&](std::size_t k) { return x[k] * x[k]; });
    }
    if (projected) {
      project(x, R);
      project(r, R);
    }
    rep.iterations = it + 1;
    rep.iterate_res = std::sqrt(upd / std::max(xsq, 1.0e-30));
    rep.abs_res = std::sqrt(std::max(dot(r, r), 0.0));
    rep.rel_res = rep.abs_res / std::max(bnorm, 1.0e-30);
    if (rep.rel_res < tol || rep.abs_res < tol) {
      monitor(alpha, 0.0, rep.abs_res, rep.rel_res, rep.iterate_res, 0.0);
      break;
    }
    precondition(A, F, r.data(), z.data());
    if (projected) project(z, R);
    const double rz_new = dot(r, z);
    if (!std::isfinite(rz_new) || rz_new <= 0.0)
      throw std::runtime_error("pcg preconditioner breakdown: non-positive new r^T M^{-1} r");
    const double beta = rz_new / rz;
    monitor(alpha, beta, rep.abs_res, rep.rel_res, rep.iterate_res, std::sqrt(std::max(rz_new, 0.0)));
    for (int k = 0; k < n; ++k) {
      const std::size_t s = static_cast<std::size_t>(k);
      p[s] = z[s] + beta * p[s];
    }
    if (projected) project(p, R);
    rz = rz_new;
  }
  return rep;
}

void copy_log(const std::vector<OrMonitorRecord> &log, OrMonitorRecord *records, int capacity, int *count) {
  const std::size_t m = std::min(log.size(), static_cast<std::size_t>(std::max(capacity, 0)));
  if (records) std::copy(log.begin(), log.begin() + static_cast<std::ptrdiff_t>(m), records);
  if (count) *count = static_cast<int>(log.size());
}

void set_err(char *err, int len, const char *msg) {
  if (err && len > 0) std::snprintf(err, static_cast<std::size_t>(len), "%s", msg);
}


int solve_common(int scheme, const OrProblem *pb, const double *rho, const double *u, const double *v,
                 const double *p_n, const double *p_nm1, const Coeffs *cf, int reduction, double *p_out,
                 int *iterations, double *residual, char *err, int errlen,
                 std::vector<OrMonitorRecord> *log = nullptr) {
  try {
    const bool split = scheme == 3 || scheme == 4;
    const int kind = (scheme == 2 || scheme == 4) ? 2 : 1;
    const Reducer R(reduction, pb->nx, pb->ny);
    System sys = assemble(*pb, rho, u, v, p_n, p_nm1, split, R, cf);
    Factor F;
    if (!factorize(sys.A, kind, F))
      throw std::runtime_error(kind == 2 ? "ILDLT preconditioner construction failed"
                                         : "ICC preconditioner construction failed");
    std::vector<double> x;
    Report rep = pcg(sys.A, F, sys.rhs, pb->max_iterations, pb->tolerance, !sys.has_dirichlet, R, x, log);
    const int n = pb->nx * pb->ny;
    for (int k = 0; k < n; ++k) p_out[k] = x[static_cast<std::size_t>(k)];
    if (!sys.has_dirichlet) {
      // subtract_mean: i-outer, j-inner accumulation (operators.cpp:447-460)
      double mean = R.sum(static_cast<std::size_t>(n), [&](std::size_t k) { return p_out[k]; });
      mean /= static_cast<double>(n);
      for (int k = 0; k < n; ++k) p_out[k] -= mean;
    }
    *iterations = rep.iterations;
    *residual = rep.rel_res;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

double wall_gradient(const Grid &g, int side, double lam, double p, double h) {
  const Bc bc = g.bc(side);
  if (bc.type == kDirichlet) return (-2.0 * lam * (p - bc.value)) / (h * h);
  return (lam * bc.value * h) / (h * h);
}

double pair_gradient(const double *lam, const double *p, std::size_t a, std::size_t b, double h) {
  return -((lam[b] - lam[a]) * (p[b] - p[a])) / (h * h);
}

// apply_pressure_bc (operators.cpp:143-181) on a ghosted nx x ny field with G ghost layers.
void fill_pressure_ghosts(const OrProblem &pb, int G, double *f) {
  const Grid g(pb);
  const int nx = pb.nx, ny = pb.ny, ld = ny + 2 * G;
  auto at = [&](int i, int j) -> double & { return f[static_cast<std::size_t>(i + G) * ld + (j + G)]; };
  const Bc L = g.bc(kLeft), Rb = g.bc(kRight), B = g.bc(kBottom), T = g.bc(kTop);
  for (int k = 1; k <= G; ++k) {
    for (int j = 0; j < ny; ++j) {
      if (pb.periodic_x) {
        at(-k, j) = at(nx - k, j);
        at(nx - 1 + k, j) = at(k - 1, j);
      } else {
        const double lo = 2.0 * (static_cast<double>(k) - 0.5) * L.value * g.dx;
        const double ro = 2.0 * (static_cast<double>(k) - 0.5) * Rb.value * g.dx;
        at(-k, j) = (L.type == kDirichlet) ? 2.0 * L.value - at(k - 1, j) : at(k - 1, j) + lo;
        at(nx - 1 + k, j) = (Rb.type == kDirichlet) ? 2.0 * Rb.value - at(nx - k, j) : at(nx - k, j) + ro;
      }
    }
    for (int i = -G; i < nx + G; ++i) {
      if (pb.periodic_y) {
        at(i, -k) = at(i, ny - k);
        at(i, ny - 1 + k) = at(i, k - 1);
      } else {
        const double bo = 2.0 * (static_cast<double>(k) - 0.5) * B.value * g.dy;
        const double to = 2.0 * (static_cast<double>(k) - 0.5) * T.value * g.dy;
        at(i, -k) = (B.type == kDirichlet) ? 2.0 * B.value - at(i, k - 1) : at(i, k - 1) + bo;
        at(i, ny - 1 + k) = (T.type == kDirichlet) ? 2.0 * T.value - at(i, ny - k) : at(i, ny - k) + to;
      }
    }
  }
}

}  // namespace

extern "C" {

// Scheme codes shared with the engine: 1 icpcg, 2 ildlt_pcg, 3 split_icpcg, 4 split_ildlt_pcg.
// Full pressure correction: assemble, factor, PCG, scatter, subtract the interior mean (when no
// Dirichlet edge). p_out receives the interior (i*ny+j). Returns 0 or 1 (message in err).
int or_pressure_solve(int scheme, const OrProblem *pb, const double *rho, const double *u, const double *v,
                      const double *p_n, const double *p_nm1, int reduction, double *p_out, int *iterations,
                      double *residual, char *err, int errlen) {
  return solve_common(scheme, pb, rho, u, v, p_n, p_nm1, nullptr, reduction, p_out, iterations, residual, err, errlen);
}

// or_pressure_solve with the PCG monitor log: up to `capacity` records into `records`, *count =
// the number of monitor calls (also on failure: the calls made before the breakdown).
int or_pressure_solve_monitored(int scheme, const OrProblem *pb, const double *rho, const double *u, const double *v,
                                const double *p_n, const double *p_nm1, int reduction, double *p_out, int *iterations,
                                double *residual, OrMonitorRecord *records, int capacity, int *count, char *err,
                                int errlen) {
  std::vector<OrMonitorRecord> log;
  const int rc = solve_common(scheme, pb, rho, u, v, p_n, p_nm1, nullptr, reduction, p_out, iterations, residual, err,
                              errlen, &log);
  copy_log(log, records, capacity, count);
  return rc;
}

// Coefficient-level solve (no reference counterpart): scheme 1 or 2 on c_f = k_f / h^2 with
// right-hand side src plus the reference's boundary terms.
int or_solve_coeffs(int scheme, const OrProblem *pb, const double *ku, const double *kv, const double *src,
                    int reduction, double *p_out, int *iterations, double *residual, char *err, int errlen) {
  if (scheme != 1 && scheme != 2) {
    set_err(err, errlen, "coefficient-level solves and the adjoint take PP_ICPCG or PP_ILDLT_PCG");
    return 1;
  }
  const Coeffs cf{ku, kv, src, true};
  return solve_common(scheme, pb, nullptr, nullptr, nullptr, nullptr, nullptr, &cf, reduction, p_out, iterations,
                      residual, err, errlen);
}

// FP64 adjoint of or_solve_coeffs for L(p) with gradient dLdp: lam = A^+ P dLdp by the same PCG
// (source dLdp, no boundary right-hand-side terms), then per face
//   dL/dk = -(lam_a - lam_b)(p_a - p_b) / h^2 (interior and periodic faces, reported on face 0),
//           -2 lam_c (p_c - value) / h^2 (Dirichlet wall), lam_c value h / h^2 (Neumann wall).
int or_adjoint(int scheme, const OrProblem *pb, const double *ku, const double *kv, const double *p,
               const double *dLdp, int reduction, double *dLdsrc, double *dLdku, double *dLdkv, int *iterations,
               double *residual, char *err, int errlen) {
  if (scheme != 1 && scheme != 2) {
    set_err(err, errlen, "coefficient-level solves and the adjoint take PP_ICPCG or PP_ILDLT_PCG");
    return 1;
  }
  const Coeffs cf{ku, kv, dLdp, false};
  if (int rc = solve_common(scheme, pb, nullptr, nullptr, nullptr, nullptr, nullptr, &cf, reduction, dLdsrc,
                            iterations, residual, err, errlen))
    return rc;
  const Grid g(*pb);
  const int nx = pb->nx, ny = pb->ny;
  const double *lam = dLdsrc;
  auto cell = [&](int i, int j) { return static_cast<std::size_t>(i) * ny + j; };
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j) {
      double out;
      if (i >= 1 && i <= nx - 1)
        out = pair_gradient(lam, p, cell(i - 1, j), cell(i, j), g.dx);
      else if (pb->periodic_x)
        out = i == 0 ? pair_gradient(lam, p, cell(nx - 1, j), cell(0, j), g.dx) : 0.0;
      else if (i == 0)
        out = wall_gradient(g, kLeft, lam[cell(0, j)], p[cell(0, j)], g.dx);
      else
        out = wall_gradient(g, kRight, lam[cell(nx - 1, j)], p[cell(nx - 1, j)], g.dx);
      dLdku[static_cast<std::size_t>(i) * ny + j] = out;
    }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j) {
      double out;
      if (j >= 1 && j <= ny - 1)
        out = pair_gradient(lam, p, cell(i, j - 1), cell(i, j), g.dy);
      else if (pb->periodic_y)
        out = j == 0 ? pair_gradient(lam, p, cell(i, ny - 1), cell(i, 0), g.dy) : 0.0;
      else if (j == 0)
        out = wall_gradient(g, kBottom, lam[cell(i, 0)], p[cell(i, 0)], g.dy);
      else
        out = wall_gradient(g, kTop, lam[cell(i, ny - 1)], p[cell(i, ny - 1)], g.dy);
      dLdkv[static_cast<std::size_t>(i) * (ny + 1) + j] = out;
    }
  return 0;
}

// ---- system-level access for tests ------------------------------------------------------------
struct OrSystem {
  System sys;
  Factor F;
  Reducer R{kSequential, 1, 1};
};

void *or_sys_build(const OrProblem *pb, int use_split, const double *rho, const double *u, const double *v,
                   const double *p_n, const double *p_nm1, int reduction, char *err, int errlen) {
  try {
    auto *h = new OrSystem();
    h->R = Reducer(reduction, pb->nx, pb->ny);
    h->sys = assemble(*pb, rho, u, v, p_n, p_nm1, use_split != 0, h->R);
    return h;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

void or_sys_free(void *h) { delete static_cast<OrSystem *>(h); }

void or_sys_sizes(void *h, int *n, int *nnz, int *has_dirichlet) {
  auto *o = static_cast<OrSystem *>(h);
  *n = o->sys.A.n;
  *nnz = static_cast<int>(o->sys.A.val.size());
  *has_dirichlet = o->sys.has_dirichlet ? 1 : 0;
}

void or_sys_csr(void *h, int *row_ptr, int *col, double *val, double *rhs) {
  auto *o = static_cast<OrSystem *>(h);
  std::copy(o->sys.A.row_ptr.begin(), o->sys.A.row_ptr.end(), row_ptr);
  std::copy(o->sys.A.col.begin(), o->sys.A.col.end(), col);
  std::copy(o->sys.A.val.begin(), o->sys.A.val.end(), val);
  std::copy(o->sys.rhs.begin(), o->sys.rhs.end(), rhs);
}

int or_sys_factor(void *h, int kind, double *lower, double *diag, double *floor_out) {
  auto *o = static_cast<OrSystem *>(h);
  const bool ok = factorize(o->sys.A, kind, o->F);
  if (lower) std::copy(o->F.lower.begin(), o->F.lower.end(), lower);
  if (diag) std::copy(o->F.diag.begin(), o->F.diag.end(), diag);
  if (floor_out) *floor_out = o->F.floor;
  return ok ? 1 : 0;
}

void or_sys_apply_matrix(void *h, const double *x, double *y) { spmv(static_cast<OrSystem *>(h)->sys.A, x, y); }

void or_sys_apply_precond(void *h, const double *r, double *z) {
  auto *o = static_cast<OrSystem *>(h);
  precondition(o->sys.A, o->F, r, z);
}

int or_sys_pcg_monitored(void *h, int max_iterations, double tolerance, double *x, int *iterations, double *rel,
                         double *abs_res, OrMonitorRecord *records, int capacity, int *count, char *err, int errlen);

int or_sys_pcg(void *h, int max_iterations, double tolerance, double *x, int *iterations, double *rel, double *abs_res,
               char *err, int errlen) {
  return or_sys_pcg_monitored(h, max_iterations, tolerance, x, iterations, rel, abs_res, nullptr, 0, nullptr, err,
                              errlen);
}

// or_sys_pcg with the PCG monitor log (see or_pressure_solve_monitored).
int or_sys_pcg_monitored(void *h, int max_iterations, double tolerance, double *x, int *iterations, double *rel,
                         double *abs_res, OrMonitorRecord *records, int capacity, int *count, char *err, int errlen) {
  auto *o = static_cast<OrSystem *>(h);
  std::vector<OrMonitorRecord> log;
  struct Flush {
    std::vector<OrMonitorRecord> &log;
    OrMonitorRecord *records;
    int capacity;
    int *count;
    ~Flush() { copy_log(log, records, capacity, count); }
  } flush{log, records, capacity, count};
  try {
    std::vector<double> xv;
    const Report rep = pcg(o->sys.A, o->F, o->sys.rhs, max_iterations, tolerance, !o->sys.has_dirichlet, o->R, xv,
                           &log);
    std::copy(xv.begin(), xv.end(), x);
    *iterations = rep.iterations;
    *rel = rep.rel_res;
    *abs_res = rep.abs_res;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

}  // extern "C"

// Velocity ghost fills on ghosted face fields of G layers (Field2D layout).
namespace {

// apply_u_bc (operators.cpp:182-233) on an (nx+1) x ny face field; unset edges: no-slip, bottom/top
// at the wall speed (62-86). Layer by layer: the boundary faces, then x ghosts, then y ghosts over the
// whole ghosted x range, which reads x-ghost cells of deeper layers before they are rewritten.
void fill_u_ghosts(const OrProblem &pbr, const OrVelocityBc &vbr, int G, double *u) {
  const OrProblem *pb = &pbr;
  const OrVelocityBc *vb = &vbr;
  const Grid g(pbr);
  const int nx = pb->nx, ny = pb->ny;
  const int ld_u = ny + 2 * G;
  auto uref = [&](double *f, int i, int j) -> double & { return f[static_cast<std::size_t>(i + G) * ld_u + (j + G)]; };
  auto uval = [&](const double *f, int i, int j) -> double { return f[static_cast<std::size_t>(i + G) * ld_u + (j + G)]; };
  auto ubc = [&](int s) -> Bc {
    if (vb->u_type[s] != kUnset) return {vb->u_type[s], vb->u_value[s]};
    if (s == kBottom) return {kDirichlet, vb->bottom_wall_velocity_x};
    if (s == kTop) return {kDirichlet, vb->top_wall_velocity_x};
    return {kDirichlet, 0.0};
  };
  {
    const Bc L = ubc(kLeft), Rb = ubc(kRight), B = ubc(kBottom), T = ubc(kTop);
    const int fnx = nx + 1;
    for (int j = 0; j < ny; ++j) {
      if (pb->periodic_x) {
        uref(u, fnx - 1, j) = uval(u, 0, j);
      } else {
        uref(u, 0, j) = (L.type == kDirichlet) ? L.value : uval(u, 1, j) + L.value * g.dx;
        uref(u, fnx - 1, j) = (Rb.type == kDirichlet) ? Rb.value : uval(u, fnx - 2, j) + Rb.value * g.dx;
      }
    }
    for (int k = 1; k <= G; ++k) {
      for (int j = 0; j < ny; ++j) {
        if (pb->periodic_x) {
          uref(u, -k, j) = uval(u, fnx - k - 1, j);
          uref(u, fnx - 1 + k, j) = uval(u, k, j);
        } else {
          const double lo = 2.0 * static_cast<double>(k) * L.value * g.dx;
          const double ro = 2.0 * static_cast<double>(k) * Rb.value * g.dx;
          uref(u, -k, j) = (L.type == kDirichlet) ? 2.0 * L.value - uval(u, k, j) : uval(u, k, j) + lo;
          uref(u, fnx - 1 + k, j) = (Rb.type == kDirichlet) ? 2.0 * Rb.value - uval(u, fnx - 1 - k, j)
                                                            : uval(u, fnx - 1 - k, j) + ro;
        }
      }
      for (int i = -G; i < fnx + G; ++i) {
        if (pb->periodic_y) {
          uref(u, i, -k) = uval(u, i, ny - k);
          uref(u, i, ny - 1 + k) = uval(u, i, k - 1);
        } else {
          const double bo = 2.0 * (static_cast<double>(k) - 0.5) * B.value * g.dy;
          const double to = 2.0 * (static_cast<double>(k) - 0.5) * T.value * g.dy;
          uref(u, i, -k) = (B.type == kDirichlet) ? 2.0 * B.value - uval(u, i, k - 1) : uval(u, i, k - 1) + bo;
          uref(u, i, ny - 1 + k) = (T.type == kDirichlet) ? 2.0 * T.value - uval(u, i, ny - k) : uval(u, i, ny - k) + to;
        }
      }
    }
  }
}

// apply_v_bc (operators.cpp:239-283) on an nx x (ny+1) face field; unset edges: no-slip (88-112).
void fill_v_ghosts(const OrProblem &pbr, const OrVelocityBc &vbr, int G, double *v) {
  const OrProblem *pb = &pbr;
  const OrVelocityBc *vb = &vbr;
  const Grid g(pbr);
  const int nx = pb->nx, ny = pb->ny;
  const int ld_v = ny + 1 + 2 * G;
  auto vref = [&](double *f, int i, int j) -> double & { return f[static_cast<std::size_t>(i + G) * ld_v + (j + G)]; };
  auto vval = [&](const double *f, int i, int j) -> double { return f[static_cast<std::size_t>(i + G) * ld_v + (j + G)]; };
  auto vbc = [&](int s) -> Bc {
    if (vb->v_type[s] != kUnset) return {vb->v_type[s], vb->v_value[s]};
    return {kDirichlet, 0.0};
  };
  {
    const Bc L = vbc(kLeft), Rb = vbc(kRight), B = vbc(kBottom), T = vbc(kTop);
    const int fny = ny + 1;
    for (int i = 0; i < nx; ++i) {
      if (pb->periodic_y) {
        vref(v, i, fny - 1) = vval(v, i, 0);
      } else {
        vref(v, i, 0) = (B.type == kDirichlet) ? B.value : vval(v, i, 1) + B.value * g.dy;
        vref(v, i, fny - 1) = (T.type == kDirichlet) ? T.value : vval(v, i, fny - 2) + T.value * g.dy;
      }
    }
    for (int k = 1; k <= G; ++k) {
      for (int j = 0; j < fny; ++j) {
        if (pb->periodic_x) {
          vref(v, -k, j) = vval(v, nx - k, j);
          vref(v, nx - 1 + k, j) = vval(v, k - 1, j);
        } else {
          const double lo = 2.0 * (static_cast<double>(k) - 0.5) * L.value * g.dx;
          const double ro = 2.0 * (static_cast<double>(k) - 0.5) * Rb.value * g.dx;
          vref(v, -k, j) = (L.type == kDirichlet) ? 2.0 * L.value - vval(v, k - 1, j) : vval(v, k - 1, j) + lo;
          vref(v, nx - 1 + k, j) = (Rb.type == kDirichlet) ? 2.0 * Rb.value - vval(v, nx - k, j) : vval(v, nx - k, j) + ro;
        }
      }
      for (int i = -G; i < nx + G; ++i) {
        if (pb->periodic_y) {
          vref(v, i, -k) = vval(v, i, fny - k - 1);
          vref(v, i, fny - 1 + k) = vval(v, i, k);
        } else {
          const double bo = 2.0 * static_cast<double>(k) * B.value * g.dy;
          const double to = 2.0 * static_cast<double>(k) * T.value * g.dy;
          vref(v, i, -k) = (B.type == kDirichlet) ? 2.0 * B.value - vval(v, i, k) : vval(v, i, k) + bo;
          vref(v, i, fny - 1 + k) = (T.type == kDirichlet) ? 2.0 * T.value - vval(v, i, fny - 1 - k)
                                                           : vval(v, i, fny - 1 - k) + to;
        }
      }
    }
  }
}

}  // namespace

extern "C" {

// Velocity correction and pressure-history shift after a pressure solve
// (Solver::apply_pressure_velocity_correction, operators.cpp:480-534), on ghosted fields of G
// ghost layers in Field2D layout (element (i, j) of an a x b field at (i + G) * (b + 2G) + j + G):
// rho_mid, p_corr, pressure, pressure_prev are nx x ny; u_star, u are (nx+1) x ny; v_star, v are
// nx x (ny+1). u, v and pressure are updated in place (their ghost cells take part, as in the
// reference, where the y pass of a layer reads x-ghost cells of deeper layers not yet rewritten);
// pressure_prev holds the previous step's pressure on entry (the split extrapolation reads it) and
// the pressure before this correction, ghosts included, on return. speeds[0] / speeds[1] get the
// periodic boundary speed before / after (operators.cpp:463-478).
void or_velocity_correction_order(int split, const OrProblem *pb, const OrVelocityBc *vb, int G, const double *rho_mid,
                                  const double *u_star, const double *v_star, const double *p_corr, double *u,
                                  double *v, double *pressure, double *pressure_prev, double *speeds, int tree_mean);

void or_velocity_correction(int split, const OrProblem *pb, const OrVelocityBc *vb, int G, const double *rho_mid,
                            const double *u_star, const double *v_star, const double *p_corr, double *u, double *v,
                            double *pressure, double *pressure_prev, double *speeds) {
  or_velocity_correction_order(split, pb, vb, G, rho_mid, u_star, v_star, p_corr, u, v, pressure, pressure_prev, speeds,
                               0);
}

// The same with the pressure mean's summation order chosen: tree_mean = 0 the reference's
// sequential sum (subtract_mean, operators.cpp:447-460), 1 the engine's tree order (256 threads:
// thread t adds the cells of i-outer index t, t + 256, ... from 0.0; lanes of each 64-lane wave
// combine with offsets 32..1, the four wave sums as (w0 + w2) + (w1 + w3)).
void or_velocity_correction_order(int split, const OrProblem *pb, const OrVelocityBc *vb, int G, const double *rho_mid,
                                  const double *u_star, const double *v_star, const double *p_corr, double *u,
                                  double *v, double *pressure, double *pressure_prev, double *speeds, int tree_mean) {
  const Grid g(*pb);
  const int nx = pb->nx, ny = pb->ny;
  const int ld_c = ny + 2 * G, ld_u = ny + 2 * G, ld_v = ny + 1 + 2 * G;
  auto cell = [&](const double *f, int i, int j) -> double { return f[static_cast<std::size_t>(i + G) * ld_c + (j + G)]; };
  auto cref = [&](double *f, int i, int j) -> double & { return f[static_cast<std::size_t>(i + G) * ld_c + (j + G)]; };
  auto uref = [&](double *f, int i, int j) -> double & { return f[static_cast<std::size_t>(i + G) * ld_u + (j + G)]; };
  auto uval = [&](const double *f, int i, int j) -> double { return f[static_cast<std::size_t>(i + G) * ld_u + (j + G)]; };
  auto vref = [&](double *f, int i, int j) -> double & { return f[static_cast<std::size_t>(i + G) * ld_v + (j + G)]; };
  auto vval = [&](const double *f, int i, int j) -> double { return f[static_cast<std::size_t>(i + G) * ld_v + (j + G)]; };
  const std::size_t n_cell = static_cast<std::size_t>(nx + 2 * G) * ld_c;

  // periodic_boundary_speed_stat (operators.cpp:463-478)
  auto speed = [&](const double *uf, const double *vf) {
    double m = 0.0;
    if (pb->periodic_x)
      for (int j = 0; j < ny; ++j) {
        m = std::max(m, std::abs(uval(uf, 0, j)));
        m = std::max(m, std::abs(uval(uf, nx, j)));
      }
    if (pb->periodic_y)
      for (int i = 0; i < nx; ++i) {
        m = std::max(m, std::abs(vval(vf, i, 0)));
        m = std::max(m, std::abs(vval(vf, i, ny)));
      }
    return m;
  };
  speeds[0] = speed(u_star, v_star);

  // rho_u_face / rho_v_face (operators.cpp:330-350)
  auto rho_u = [&](int i, int j) {
    if (i <= 0) return pb->periodic_x ? 0.5 * (cell(rho_mid, nx - 1, j) + cell(rho_mid, 0, j)) : cell(rho_mid, 0, j);
    if (i >= nx) return pb->periodic_x ? 0.5 * (cell(rho_mid, nx - 1, j) + cell(rho_mid, 0, j)) : cell(rho_mid, nx - 1, j);
    return 0.5 * (cell(rho_mid, i - 1, j) + cell(rho_mid, i, j));
  };
  auto rho_v = [&](int i, int j) {
    if (j <= 0) return pb->periodic_y ? 0.5 * (cell(rho_mid, i, ny - 1) + cell(rho_mid, i, 0)) : cell(rho_mid, i, 0);
    if (j >= ny) return pb->periodic_y ? 0.5 * (cell(rho_mid, i, ny - 1) + cell(rho_mid, i, 0)) : cell(rho_mid, i, ny - 1);
    return 0.5 * (cell(rho_mid, i, j - 1) + cell(rho_mid, i, j));
  };
  // split extrapolation, ghost-filled on a zeroed field (operators.cpp:413-422)
  const double rho_ref = split ? std::min(1.0, pb->density_ratio) : 1.0;
  std::vector<double> p_ext(n_cell, 0.0);
  if (split) {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        cref(p_ext.data(), i, j) = (pb->step_index == 0) ? cell(pressure, i, j)
                                                          : (2.0 * cell(pressure, i, j) - cell(pressure_prev, i, j));
    fill_pressure_ghosts(*pb, G, p_ext.data());
  }
  // u = u* - dt (implicit + explicit), implicit = grad p / rho_f (non-split) or grad p / rho_ref,
  // explicit = (1/rho_f - 1/rho_ref) grad p_ext (split) or 0 (operators.cpp:490-519)
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const double ex = split ? (1.0 / rho_u(i, j) - 1.0 / rho_ref) *
                                    ((cell(p_ext.data(), i, j) - cell(p_ext.data(), i - 1, j)) / g.dx)
                              : 0.0;
      const double gp = (cell(p_corr, i, j) - cell(p_corr, i - 1, j)) / g.dx;
      const double im = split ? gp / rho_ref : gp / rho_u(i, j);
      uref(u, i, j) = uval(u_star, i, j) - pb->dt * (im + ex);
    }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j) {
      const double ex = split ? (1.0 / rho_v(i, j) - 1.0 / rho_ref) *
                                    ((cell(p_ext.data(), i, j) - cell(p_ext.data(), i, j - 1)) / g.dy)
                              : 0.0;
      const double gp = (cell(p_corr, i, j) - cell(p_corr, i, j - 1)) / g.dy;
      const double im = split ? gp / rho_ref : gp / rho_v(i, j);
      vref(v, i, j) = vval(v_star, i, j) - pb->dt * (im + ex);
    }

  fill_u_ghosts(*pb, *vb, G, u);
  fill_v_ghosts(*pb, *vb, G, v);
  speeds[1] = speed(u, v);

  // pressure history (operators.cpp:524-533): previous <- pressure; pressure <- p_corr (interior),
  // mean removed when no edge is Dirichlet (subtract_mean, 447-460), ghosts refilled
  std::copy(pressure, pressure + n_cell, pressure_prev);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) cref(pressure, i, j) = cell(p_corr, i, j);
  if (!g.has_dirichlet()) {
    double mean = 0.0;
    if (tree_mean) {
      constexpr int T = 256;
      const int n = nx * ny;
      double part[T];
      for (int t = 0; t < T; ++t) {
        double acc = 0.0;
        for (int e = t; e < n; e += T) acc = acc + cell(pressure, e / ny, e % ny);
        part[t] = acc;
      }
      double wsum[T / 64];
      for (int w = 0; w < T / 64; ++w) {
        double *v = part + 64 * w;
        for (int off = 32; off >= 1; off >>= 1)
          for (int l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
        wsum[w] = v[0];
      }
      mean = (wsum[0] + wsum[2]) + (wsum[1] + wsum[3]);
    } else {
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) mean += cell(pressure, i, j);
    }
    mean /= static_cast<double>(nx * ny);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) cref(pressure, i, j) -= mean;
  }
  fill_pressure_ghosts(*pb, G, pressure);
}

// Ghost fill of a pressure-like field with 3 ghost layers (operators.cpp:143-181).
void or_apply_pressure_bc(const OrProblem *pb, double *f) { fill_pressure_ghosts(*pb, 3, f); }

// apply_pressure_bc (operators.cpp:143-181) on a ghosted nx x ny field with G ghost layers.
void or_pressure_ghosts(const OrProblem *pb, int G, double *f) { fill_pressure_ghosts(*pb, G, f); }

}  // extern "C"

// ---- momentum predictor's linear solve: BiCGSTAB with the diagonal preconditioner ---------------
//
// build_diagonal_preconditioner (ch_sparse_krylov.hpp:421-431) and solve_preconditioned_bicgstab
// (ch_sparse_krylov.hpp:839-963), as momentum.cpp:469-472 calls them, on a CSR matrix in the form
// finalize_row_maps produces (ch_sparse_krylov.hpp:105-133: ascending columns, one diagonal entry
// per row). SpMV row sums follow apply_matrix (235-250); dot products are either dot_product's
// sequential sum (84-90, kSequential) or the HIP engine's order (kChunked, bicg_sum below).
namespace {

constexpr int kChunked = 3;

// The engine's BiCGSTAB dot-product order. Rows are cut into chunks of 512: in chunk c, thread t
// holds the term of row 512 c + t added to 0.0 (0.0 past the last row), and the 64-lane trees
// (offsets 32..1) and the 8-wave tree (offsets 4..1) give the chunk sum. The chunk sums are then
// reduced the same way: thread t adds chunk sums t, t + 512, ... in order from 0.0, and the same
// lane and wave trees give the total.
double tree512(double *part) {
  constexpr std::size_t W = 8;
  double wsum[W];
  for (std::size_t w = 0; w < W; ++w) {
    double *v = part + 64 * w;
    for (std::size_t off = 32; off >= 1; off >>= 1)
      for (std::size_t l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
    wsum[w] = v[0];
  }
  for (std::size_t off = W / 2; off >= 1; off >>= 1)
    for (std::size_t t = 0; t < off; ++t) wsum[t] = wsum[t] + wsum[t + off];
  return wsum[0];
}

template <class F>
double bicg_sum(int mode, std::size_t n, F f) {
  if (mode == kSequential) {
    double s = 0.0;
    for (std::size_t k = 0; k < n; ++k) s += f(k);
    return s;
  }
  constexpr std::size_t T = 512;
  std::vector<double> chunk;
  for (std::size_t c0 = 0; c0 < n; c0 += T) {
    double part[T];
    for (std::size_t t = 0; t < T; ++t) {
      double s = 0.0;
      if (c0 + t < n) s = s + f(c0 + t);
      part[t] = s;
    }
    chunk.push_back(tree512(part));
  }
  double part[T];
  for (std::size_t t = 0; t < T; ++t) {
    double s = 0.0;
    for (std::size_t c = t; c < chunk.size(); c += T) s = s + chunk[c];
    part[t] = s;
  }
  return tree512(part);
}

struct BicgReport {
  int iterations = 0;
  double abs_res = 0.0, rel_res = 0.0, iter_res = 0.0;
};

void csr_apply(int n, const int *rp, const int *ci, const double *val, const std::vector<double> &x,
               std::vector<double> &y) {
  y.assign(static_cast<std::size_t>(n), 0.0);
  for (int row = 0; row < n; ++row) {
    double sum = 0.0;
    for (int pos = rp[row]; pos < rp[row + 1]; ++pos) sum += val[pos] * x[static_cast<std::size_t>(ci[pos])];
    y[static_cast<std::size_t>(row)] = sum;
  }
}

BicgReport bicgstab(int mode, int n, const int *rp, const int *ci, const double *val, const double *rhs,
                    std::vector<double> &x, int max_iterations, double tolerance, bool use_rel) {
  if (n <= 0) throw std::runtime_error("cannot solve an empty linear system");
  const std::size_t N = static_cast<std::size_t>(n);
  // diagonal preconditioner: 1 / max(|a_ii|, 1e-30), the diagonal located as finalize_row_maps does
  std::vector<double> dinv(N, 1.0);
  for (int row = 0; row < n; ++row) {
    int diag = -1;
    for (int pos = rp[row]; pos < rp[row + 1]; ++pos)
      if (ci[pos] == row) diag = pos;
    if (diag < 0) throw std::runtime_error("sparse matrix assembly failed to produce a diagonal entry");
    dinv[static_cast<std::size_t>(row)] = 1.0 / std::max(std::abs(val[diag]), 1.0e-30);
  }
  auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
    return bicg_sum(mode, N, [&](std::size_t k) { return a[k] * b[k]; });
  };
  std::vector<double> ax;
  csr_apply(n, rp, ci, val, x, ax);
  std::vector<double> r(N), b(rhs, rhs + N);
  for (std::size_t k = 0; k < N; ++k) r[k] = rhs[k] - ax[k];
  const double rhs_norm = std::sqrt(std::max(dot(b, b), 0.0));
  BicgReport rep;
  rep.abs_res = std::sqrt(std::max(dot(r, r), 0.0));
  rep.rel_res = rep.abs_res / std::max(rhs_norm, 1.0e-30);
  if ((use_rel && rep.rel_res < tolerance) || rep.abs_res < tolerance) return rep;

  std::vector<double> r_hat = r, p(N, 0.0), v(N, 0.0), phat(N), s(N, 0.0), shat(N), t(N, 0.0), update(N, 0.0);
  double rho_old = 1.0, alpha = 1.0, omega = 1.0;
  const int limit = std::max(1, max_iterations);
  for (int iter = 0; iter < limit; ++iter) {
    const double rho_new = dot(r_hat, r);
    if (!std::isfinite(rho_new) || std::abs(rho_new) < 1.0e-30) throw std::runtime_error("bicgstab breakdown: invalid rho");
    const double beta = (iter == 0) ? 0.0 : (rho_new / rho_old) * (alpha / omega);
    for (std::size_t k = 0; k < N; ++k) p[k] = r[k] + beta * (p[k] - omega * v[k]);
    for (std::size_t k = 0; k < N; ++k) phat[k] = dinv[k] * p[k];
    csr_apply(n, rp, ci, val, phat, v);
    const double rhat_v = dot(r_hat, v);
    if (!std::isfinite(rhat_v) || std::abs(rhat_v) < 1.0e-30)
      throw std::runtime_error("bicgstab breakdown: invalid r_hat^T A M^{-1} p");
    alpha = rho_new / rhat_v;
    for (std::size_t k = 0; k < N; ++k) {
      s[k] = r[k] - alpha * v[k];
      update[k] = alpha * phat[k];
    }
    const double s_norm = std::sqrt(std::max(dot(s, s), 0.0));
    if ((use_rel && s_norm / std::max(rhs_norm, 1.0e-30) < tolerance) || s_norm < tolerance) {
      for (std::size_t k = 0; k < N; ++k) x[k] += update[k];
      const double update_sq = dot(update, update), x_norm_sq = dot(x, x);
      rep.iterations = iter + 1;
      rep.iter_res = std::sqrt(update_sq / std::max(x_norm_sq, 1.0e-30));
      rep.abs_res = s_norm;
      rep.rel_res = s_norm / std::max(rhs_norm, 1.0e-30);
      return rep;
    }
    for (std::size_t k = 0; k < N; ++k) shat[k] = dinv[k] * s[k];
    csr_apply(n, rp, ci, val, shat, t);
    const double tt = dot(t, t);
    if (!std::isfinite(tt) || std::abs(tt) < 1.0e-30) throw std::runtime_error("bicgstab breakdown: invalid t^T t");
    omega = dot(t, s) / tt;
    if (!std::isfinite(omega) || std::abs(omega) < 1.0e-30) throw std::runtime_error("bicgstab breakdown: invalid omega");
    for (std::size_t k = 0; k < N; ++k) {
      update[k] += omega * shat[k];
      x[k] += update[k];
      r[k] = s[k] - omega * t[k];
    }
    const double update_sq = dot(update, update), x_norm_sq = dot(x, x);
    rep.iterations = iter + 1;
    rep.iter_res = std::sqrt(update_sq / std::max(x_norm_sq, 1.0e-30));
    rep.abs_res = std::sqrt(std::max(dot(r, r), 0.0));
    rep.rel_res = rep.abs_res / std::max(rhs_norm, 1.0e-30);
    if ((use_rel && rep.rel_res < tolerance) || rep.abs_res < tolerance) return rep;
    rho_old = rho_new;
  }
  return rep;
}

}  // namespace

extern "C" {

// BiCGSTAB (diagonal preconditioner) on a CSR system; x holds the initial guess and receives the
// solution. reduction: 0 the reference's sequential dot products, 3 the engine's chunked order.
// Returns 0, or 1 with the reference's error text in err.
int or_bicgstab(int n, const int *row_ptr, const int *col_idx, const double *values, const double *rhs, double *x,
                int max_iterations, double tolerance, int use_relative, int reduction, int *iterations,
                double *abs_res, double *rel_res, double *iter_res, char *err, int errlen) {
  try {
    std::vector<double> xv(x, x + std::max(n, 0));
    const BicgReport rep = bicgstab(reduction == kChunked ? kChunked : kSequential, n, row_ptr, col_idx, values, rhs,
                                    xv, max_iterations, tolerance, use_relative != 0);
    std::copy(xv.begin(), xv.end(), x);
    *iterations = rep.iterations;
    *abs_res = rep.abs_res;
    *rel_res = rep.rel_res;
    *iter_res = rep.iter_res;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

}  // extern "C"

// ---- momentum predictor (Solver::compute_momentum_fluxes + solve_momentum_predictor) -------------
//
// momentum.cpp:55-85 (central convection H(u), H(v)), 87-155 (stress divergence), 157-502 (the
// implicit full-viscous-tensor u-v system with its right-hand side, BiCGSTAB with the diagonal
// preconditioner, scatter into u_star_ / v_star_ and their ghost fills). Face densities and
// viscosities as operators.cpp:330-385. Fields are ghosted (G layers, Field2D layout): u (nx+1) x ny,
// v nx x (ny+1), rho_mid / eta / eta_prev nx x ny (their ghosts filled by the caller, as
// apply_scalar_bc leaves them), surface forces on the u and v faces.
//
// Unknowns: the active u faces (i-major, j inner; i = 1..nx-1 with walls, 0..nx-1 periodic) then the
// active v faces (j = 1..ny-1 with walls, 0..ny-1 periodic). Each matrix row is the list of
// contributions the reference inserts into its row map, in its order; a column's value is the sum of
// its contributions in that order starting from 0.0, and the row is emitted in ascending columns
// without entries below 1e-18 in magnitude (ch_sparse_krylov.hpp:105-133).

extern "C" {

struct OrMomentumProblem {
  int nx, ny;
  double lx, ly, dt, re;
  double body_force_x, body_force_y;
  int periodic_x, periodic_y;
  int step_index;
  int max_iterations;
  double tolerance;
};

}  // extern "C"

namespace {

struct MomentumSystem {
  int nu = 0, nv = 0;
  std::vector<int> row_ptr, col;
  std::vector<double> val, rhs, x0;
  std::vector<double> u_adv, v_adv;  // ghosted, after their ghost fills
};

MomentumSystem momentum_system(const OrMomentumProblem &mp, const OrVelocityBc &vb, int G, const double *u,
                               const double *v, const double *rho, const double *eta, const double *eta_prev,
                               const double *sfx, const double *sfy, const double *u_prev, const double *v_prev) {
  const int nx = mp.nx, ny = mp.ny;
  const bool px = mp.periodic_x != 0, py = mp.periodic_y != 0;
  const double dx = mp.lx / nx, dy = mp.ly / ny, dt = mp.dt;
  const int ldc = ny + 2 * G, ldu = ny + 2 * G, ldv = ny + 1 + 2 * G;
  auto C = [&](const double *f, int i, int j) { return f[static_cast<std::size_t>(i + G) * ldc + (j + G)]; };
  auto U = [&](const double *f, int i, int j) { return f[static_cast<std::size_t>(i + G) * ldu + (j + G)]; };
  auto V = [&](const double *f, int i, int j) { return f[static_cast<std::size_t>(i + G) * ldv + (j + G)]; };
  // face averages (operators.cpp:330-385)
  auto rho_u = [&](int i, int j) {
    if (i <= 0) return px ? 0.5 * (C(rho, nx - 1, j) + C(rho, 0, j)) : C(rho, 0, j);
    if (i >= nx) return px ? 0.5 * (C(rho, nx - 1, j) + C(rho, 0, j)) : C(rho, nx - 1, j);
    return 0.5 * (C(rho, i - 1, j) + C(rho, i, j));
  };
  auto rho_v = [&](int i, int j) {
    if (j <= 0) return py ? 0.5 * (C(rho, i, ny - 1) + C(rho, i, 0)) : C(rho, i, 0);
    if (j >= ny) return py ? 0.5 * (C(rho, i, ny - 1) + C(rho, i, 0)) : C(rho, i, ny - 1);
    return 0.5 * (C(rho, i, j - 1) + C(rho, i, j));
  };
  auto eta_u = [&](const double *e, int i, int j) {
    if (i <= 0) return px ? 0.5 * (C(e, nx - 1, j) + C(e, 0, j)) : C(e, 0, j);
    if (i >= nx) return px ? 0.5 * (C(e, nx - 1, j) + C(e, 0, j)) : C(e, nx - 1, j);
    return 0.5 * (C(e, i - 1, j) + C(e, i, j));
  };
  auto eta_v = [&](const double *e, int i, int j) {
    if (j <= 0) return py ? 0.5 * (C(e, i, ny - 1) + C(e, i, 0)) : C(e, i, 0);
    if (j >= ny) return py ? 0.5 * (C(e, i, ny - 1) + C(e, i, 0)) : C(e, i, ny - 1);
    return 0.5 * (C(e, i, j - 1) + C(e, i, j));
  };
  auto eta_c = [&](const double *e, int i, int j) {
    return 0.25 * (C(e, i, j) + C(e, i - 1, j) + C(e, i, j - 1) + C(e, i - 1, j - 1));
  };
  // convection (momentum.cpp:55-85)
  auto hu = [&](int i, int j) {
    const double du_dx = (U(u, i + 1, j) - U(u, i - 1, j)) / (2.0 * dx);
    const double du_dy = (U(u, i, j + 1) - U(u, i, j - 1)) / (2.0 * dy);
    const double vf = 0.25 * (V(v, i - 1, j) + V(v, i, j) + V(v, i - 1, j + 1) + V(v, i, j + 1));
    return U(u, i, j) * du_dx + vf * du_dy;
  };
  auto hv = [&](int i, int j) {
    const double dv_dx = (V(v, i + 1, j) - V(v, i - 1, j)) / (2.0 * dx);
    const double dv_dy = (V(v, i, j + 1) - V(v, i, j - 1)) / (2.0 * dy);
    const double uf = 0.25 * (U(u, i, j - 1) + U(u, i, j) + U(u, i + 1, j - 1) + U(u, i + 1, j));
    return uf * dv_dx + V(v, i, j) * dv_dy;
  };
  // stress divergence with the previous step's viscosity (momentum.cpp:91-155)
  auto visc_u = [&](int i, int j) {
    const double *e = eta_prev;
    const double ee = eta_u(e, i + 1, j), ew = eta_u(e, i, j);
    const double normal = (2.0 * ee * ((U(u, i + 1, j) - U(u, i, j)) / dx) -
                           2.0 * ew * ((U(u, i, j) - U(u, i - 1, j)) / dx)) / dx;
    const double en = eta_c(e, i, j + 1), es = eta_c(e, i, j);
    const double shear = (en * ((U(u, i, j + 1) - U(u, i, j)) / dy) - es * ((U(u, i, j) - U(u, i, j - 1)) / dy)) / dy;
    const double cross = (en * ((V(v, i, j + 1) - V(v, i - 1, j + 1)) / dx) - es * ((V(v, i, j) - V(v, i - 1, j)) / dx)) / dy;
    return (normal + shear) + cross;
  };
  auto visc_v = [&](int i, int j) {
    const double *e = eta_prev;
    const double ee = eta_c(e, i + 1, j), ew = eta_c(e, i, j);
    const double shear = (ee * ((V(v, i + 1, j) - V(v, i, j)) / dx) - ew * ((V(v, i, j) - V(v, i - 1, j)) / dx)) / dx;
    const double en = eta_v(e, i, j + 1), es = eta_v(e, i, j);
    const double normal = (2.0 * en * ((V(v, i, j + 1) - V(v, i, j)) / dy) -
                           2.0 * es * ((V(v, i, j) - V(v, i, j - 1)) / dy)) / dy;
    const double cross = (ee * ((U(u, i + 1, j) - U(u, i + 1, j - 1)) / dy) - ew * ((U(u, i, j) - U(u, i, j - 1)) / dy)) / dx;
    return (shear + normal) + cross;
  };

  MomentumSystem M;
  const int iu0 = px ? 0 : 1, iu1 = px ? nx - 1 : nx - 1;  // active u faces i in [iu0, iu1]
  const int jv0 = py ? 0 : 1, jv1 = py ? ny - 1 : ny - 1;  // active v faces j in [jv0, jv1]
  const int niu = iu1 - iu0 + 1, njv = jv1 - jv0 + 1;
  M.nu = niu * ny;
  M.nv = nx * njv;
  auto u_active = [&](int i, int j) { return j >= 0 && j < ny && i >= iu0 && i <= iu1; };
  auto v_active = [&](int i, int j) { return i >= 0 && i < nx && j >= jv0 && j <= jv1; };
  auto u_row = [&](int i, int j) { return (i - iu0) * ny + j; };
  auto v_row = [&](int i, int j) { return M.nu + i * njv + (j - jv0); };
  auto wrap_ui = [&](int i) { return px ? (i < 0 ? nx - 1 : (i >= nx ? 0 : i)) : i; };
  auto wrap_uj = [&](int j) { return py ? (j < 0 ? ny - 1 : (j >= ny ? 0 : j)) : j; };
  auto wrap_vi = [&](int i) { return px ? (i < 0 ? nx - 1 : (i >= nx ? 0 : i)) : i; };
  auto wrap_vj = [&](int j) { return py ? (j < 0 ? ny - 1 : (j >= ny ? 0 : j)) : j; };
  // effective velocity BCs (operators.cpp:62-112)
  auto ubc = [&](int s) -> Bc {
    if (vb.u_type[s] != kUnset) return {vb.u_type[s], vb.u_value[s]};
    if (s == kBottom) return {kDirichlet, vb.bottom_wall_velocity_x};
    if (s == kTop) return {kDirichlet, vb.top_wall_velocity_x};
    return {kDirichlet, 0.0};
  };
  auto vbcf = [&](int s) -> Bc {
    if (vb.v_type[s] != kUnset) return {vb.v_type[s], vb.v_value[s]};
    return {kDirichlet, 0.0};
  };
  const double alpha = 0.5 * dt / mp.re;
  const int n = M.nu + M.nv;
  M.rhs.assign(static_cast<std::size_t>(n), 0.0);
  M.x0.assign(static_cast<std::size_t>(n), 0.0);
  M.row_ptr.assign(1, 0);

  struct Contribution {
    int col;
    double value;
    bool subtract;  // entry -= value (a Dirichlet wall's reflected neighbour) instead of +=
  };
  auto emit = [&](const std::vector<Contribution> &list, int row) {
    std::vector<std::pair<int, double>> merged;  // columns in first-insertion order, values folded
    for (const Contribution &c : list) {
      auto it = std::find_if(merged.begin(), merged.end(), [&](const std::pair<int, double> &m) { return m.first == c.col; });
      if (it == merged.end()) {
        merged.push_back({c.col, 0.0});
        it = merged.end() - 1;
      }
      it->second = c.subtract ? it->second - c.value : it->second + c.value;
    }
    std::sort(merged.begin(), merged.end());
    bool diag = false;
    for (const auto &m : merged) {
      if (std::abs(m.second) < 1.0e-18) continue;
      M.col.push_back(m.first);
      M.val.push_back(m.second);
      diag = diag || m.first == row;
    }
    if (!diag) throw std::runtime_error("sparse matrix assembly failed to produce a diagonal entry");
    M.row_ptr.push_back(static_cast<int>(M.col.size()));
  };
  // neighbour contributions with the boundary handling of momentum.cpp:287-381
  auto add_u_x = [&](std::vector<Contribution> &list, double &b, int ui, int uj, double value) {
    if (u_active(ui, uj)) return list.push_back({u_row(ui, uj), value, false});
    if (!px && (ui <= 0 || ui >= nx)) {
      const bool left = ui <= 0;
      const Bc bc = ubc(left ? kLeft : kRight);
      if (bc.type == kDirichlet) {
        b -= value * bc.value;
      } else {
        list.push_back({u_row(left ? 1 : nx - 1, uj), value, false});
        b -= value * bc.value * dx;
      }
      return;
    }
    throw std::runtime_error("unexpected inactive u x-neighbor in momentum predictor");
  };
  auto add_u_y = [&](std::vector<Contribution> &list, double &b, int ui, int uj, double value) {
    if (u_active(ui, uj)) return list.push_back({u_row(ui, uj), value, false});
    if (!py && (uj < 0 || uj >= ny)) {
      const bool bottom = uj < 0;
      const Bc bc = ubc(bottom ? kBottom : kTop);
      const int jin = bottom ? 0 : ny - 1;
      if (bc.type == kDirichlet) {
        list.push_back({u_row(ui, jin), value, true});
        b -= value * (2.0 * bc.value);
      } else {
        list.push_back({u_row(ui, jin), value, false});
        b -= value * bc.value * dy;
      }
      return;
    }
    throw std::runtime_error("unexpected inactive u y-neighbor in momentum predictor");
  };
  auto add_v_y = [&](std::vector<Contribution> &list, double &b, int vi, int vj, double value) {
    if (v_active(vi, vj)) return list.push_back({v_row(vi, vj), value, false});
    if (!py && (vj <= 0 || vj >= ny)) {
      const bool bottom = vj <= 0;
      const Bc bc = vbcf(bottom ? kBottom : kTop);
      if (bc.type == kDirichlet) {
        b -= value * bc.value;
      } else {
        list.push_back({v_row(vi, bottom ? 1 : ny - 1), value, false});
        b -= value * bc.value * dy;
      }
      return;
    }
    throw std::runtime_error("unexpected inactive v y-neighbor in momentum predictor");
  };
  auto add_v_x = [&](std::vector<Contribution> &list, double &b, int vi, int vj, double value) {
    if (v_active(vi, vj)) return list.push_back({v_row(vi, vj), value, false});
    if (!px && (vi < 0 || vi >= nx)) {
      const bool left = vi < 0;
      const Bc bc = vbcf(left ? kLeft : kRight);
      const int iin = left ? 0 : nx - 1;
      if (bc.type == kDirichlet) {
        list.push_back({v_row(iin, vj), value, true});
        b -= value * (2.0 * bc.value);
      } else {
        list.push_back({v_row(iin, vj), value, false});
        b -= value * bc.value * dx;
      }
      return;
    }
    throw std::runtime_error("unexpected inactive v x-neighbor in momentum predictor");
  };

  std::vector<Contribution> list;
  for (int i = iu0; i <= iu1; ++i)
    for (int j = 0; j < ny; ++j) {
      const int row = u_row(i, j);
      const double rf = rho_u(i, j);
      const double conv = mp.step_index == 0 ? hu(i, j) : 1.5 * hu(i, j) - 0.5 * U(u_prev, i, j);
      double b = rf * U(u, i, j) + dt * (-rf * conv + rf * mp.body_force_x + U(sfx, i, j)) + alpha * visc_u(i, j);
      M.x0[static_cast<std::size_t>(row)] = U(u, i, j);
      const double ee = eta_u(eta, i + 1, j), ew = eta_u(eta, i, j), en = eta_c(eta, i, j + 1), es = eta_c(eta, i, j);
      const double ce = 2.0 * ee / (dx * dx), cw = 2.0 * ew / (dx * dx), cn = en / (dy * dy), cs = es / (dy * dy);
      const double xn = en / (dx * dy), xs = es / (dx * dy);
      list.clear();
      list.push_back({row, rf + alpha * (ce + cw + cn + cs), false});
      add_u_x(list, b, wrap_ui(i + 1), j, -alpha * ce);
      add_u_x(list, b, wrap_ui(i - 1), j, -alpha * cw);
      add_u_y(list, b, i, wrap_uj(j + 1), -alpha * cn);
      add_u_y(list, b, i, wrap_uj(j - 1), -alpha * cs);
      add_v_y(list, b, wrap_vi(i), wrap_vj(j + 1), -alpha * xn);
      add_v_y(list, b, wrap_vi(i - 1), wrap_vj(j + 1), alpha * xn);
      add_v_y(list, b, wrap_vi(i), wrap_vj(j), alpha * xs);
      add_v_y(list, b, wrap_vi(i - 1), wrap_vj(j), -alpha * xs);
      M.rhs[static_cast<std::size_t>(row)] = b;
      emit(list, row);
    }
  for (int i = 0; i < nx; ++i)
    for (int j = jv0; j <= jv1; ++j) {
      const int row = v_row(i, j);
      const double rf = rho_v(i, j);
      const double conv = mp.step_index == 0 ? hv(i, j) : 1.5 * hv(i, j) - 0.5 * V(v_prev, i, j);
      double b = rf * V(v, i, j) + dt * (-rf * conv + rf * mp.body_force_y + V(sfy, i, j)) + alpha * visc_v(i, j);
      M.x0[static_cast<std::size_t>(row)] = V(v, i, j);
      const double ee = eta_c(eta, i + 1, j), ew = eta_c(eta, i, j), en = eta_v(eta, i, j + 1), es = eta_v(eta, i, j);
      const double ce = ee / (dx * dx), cw = ew / (dx * dx), cn = 2.0 * en / (dy * dy), cs = 2.0 * es / (dy * dy);
      const double xe = ee / (dx * dy), xw = ew / (dx * dy);
      list.clear();
      list.push_back({row, rf + alpha * (ce + cw + cn + cs), false});
      add_v_x(list, b, wrap_vi(i + 1), j, -alpha * ce);
      add_v_x(list, b, wrap_vi(i - 1), j, -alpha * cw);
      add_v_y(list, b, i, wrap_vj(j + 1), -alpha * cn);
      add_v_y(list, b, i, wrap_vj(j - 1), -alpha * cs);
      add_u_x(list, b, wrap_ui(i + 1), wrap_uj(j), -alpha * xe);
      add_u_x(list, b, wrap_ui(i + 1), wrap_uj(j - 1), alpha * xe);
      add_u_x(list, b, wrap_ui(i), wrap_uj(j), alpha * xw);
      add_u_x(list, b, wrap_ui(i), wrap_uj(j - 1), -alpha * xw);
      M.rhs[static_cast<std::size_t>(row)] = b;
      emit(list, row);
    }

  // the convection fields as stored for the next step: interiors, then apply_u_bc / apply_v_bc
  // on buffers that start at zero (compute_momentum_fluxes, momentum.cpp:55-85)
  const OrProblem gp{nx, ny, mp.lx, mp.ly, dt, mp.periodic_x, mp.periodic_y, {0, 0, 0, 0}, {0, 0, 0, 0}, 0, 0.0, 1.0, 0};
  M.u_adv.assign(static_cast<std::size_t>(nx + 1 + 2 * G) * ldu, 0.0);
  M.v_adv.assign(static_cast<std::size_t>(nx + 2 * G) * ldv, 0.0);
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j) M.u_adv[static_cast<std::size_t>(i + G) * ldu + (j + G)] = hu(i, j);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j) M.v_adv[static_cast<std::size_t>(i + G) * ldv + (j + G)] = hv(i, j);
  fill_u_ghosts(gp, vb, G, M.u_adv.data());
  fill_v_ghosts(gp, vb, G, M.v_adv.data());
  return M;
}

}  // namespace

extern "C" {

// The momentum predictor's system: n = active u faces + active v faces; row_ptr (n + 1), col / val
// (up to 9 n), rhs and x0 (n) are written when not null (sizes from or_momentum_sizes).
int or_momentum_system(const OrMomentumProblem *mp, const OrVelocityBc *vb, int G, const double *u, const double *v,
                       const double *rho, const double *eta, const double *eta_prev, const double *sfx,
                       const double *sfy, const double *u_prev, const double *v_prev, int *n_out, int *nnz_out,
                       int *row_ptr, int *col, double *val, double *rhs, double *x0, char *err, int errlen) {
  try {
    const MomentumSystem M = momentum_system(*mp, *vb, G, u, v, rho, eta, eta_prev, sfx, sfy, u_prev, v_prev);
    *n_out = M.nu + M.nv;
    *nnz_out = static_cast<int>(M.col.size());
    if (row_ptr) std::copy(M.row_ptr.begin(), M.row_ptr.end(), row_ptr);
    if (col) std::copy(M.col.begin(), M.col.end(), col);
    if (val) std::copy(M.val.begin(), M.val.end(), val);
    if (rhs) std::copy(M.rhs.begin(), M.rhs.end(), rhs);
    if (x0) std::copy(M.x0.begin(), M.x0.end(), x0);
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// The whole predictor: system, BiCGSTAB from x0 = u / v (reduction 0: sequential dot products as the
// reference, 3: the engine's chunked order), u_star / v_star = u / v with the active faces replaced by
// the solution and their ghosts refilled (apply_u_velocity_bc / apply_v_bc), and u_prev / v_prev
// replaced by this step's convection fields (momentum_u_rhs_prev_ = u_adv). Returns 0, or 1 with the
// reference's text.
int or_momentum_predictor(const OrMomentumProblem *mp, const OrVelocityBc *vb, int G, const double *u, const double *v,
                          const double *rho, const double *eta, const double *eta_prev, const double *sfx,
                          const double *sfy, double *u_prev, double *v_prev, double *u_star, double *v_star,
                          int reduction, int *iterations, double *abs_res, char *err, int errlen) {
  try {
    const int nx = mp->nx, ny = mp->ny;
    MomentumSystem M = momentum_system(*mp, *vb, G, u, v, rho, eta, eta_prev, sfx, sfy, u_prev, v_prev);
    const int n = M.nu + M.nv;
    std::vector<double> x = M.x0;
    const BicgReport rep = bicgstab(reduction == kChunked ? kChunked : kSequential, n, M.row_ptr.data(), M.col.data(),
                                    M.val.data(), M.rhs.data(), x, mp->max_iterations, mp->tolerance, false);
    const int ldu = ny + 2 * G, ldv = ny + 1 + 2 * G;
    const std::size_t su = static_cast<std::size_t>(nx + 1 + 2 * G) * ldu, sv = static_cast<std::size_t>(nx + 2 * G) * ldv;
    std::copy(u, u + su, u_star);
    std::copy(v, v + sv, v_star);
    const bool px = mp->periodic_x != 0, py = mp->periodic_y != 0;
    const int iu0 = px ? 0 : 1, jv0 = py ? 0 : 1, njv = py ? ny : ny - 1;
    for (int i = iu0; i <= nx - 1; ++i)
      for (int j = 0; j < ny; ++j) u_star[static_cast<std::size_t>(i + G) * ldu + (j + G)] = x[static_cast<std::size_t>((i - iu0) * ny + j)];
    for (int i = 0; i < nx; ++i)
      for (int j = jv0; j <= ny - 1; ++j)
        v_star[static_cast<std::size_t>(i + G) * ldv + (j + G)] = x[static_cast<std::size_t>(M.nu + i * njv + (j - jv0))];
    const OrProblem gp{nx, ny, mp->lx, mp->ly, mp->dt, mp->periodic_x, mp->periodic_y, {0, 0, 0, 0}, {0, 0, 0, 0}, 0,
                       0.0, 1.0, 0};
    fill_u_ghosts(gp, *vb, G, u_star);
    fill_v_ghosts(gp, *vb, G, v_star);
    std::copy(M.u_adv.begin(), M.u_adv.end(), u_prev);
    std::copy(M.v_adv.begin(), M.v_adv.end(), v_prev);
    *iterations = rep.iterations;
    *abs_res = rep.abs_res;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

}  // extern "C"

// ---- Cahn-Hilliard semi-implicit solve (Solver::solve_phase_linear_system_eq25) ------------------
//
// phase_field.cpp:437-503 with the cached operator and preconditioner of 109-135: the operator
// A = alpha0 I - beta1 Lap + beta2 Lap^2 (ch_sparse_krylov.hpp:135-233: build_laplacian_matrix,
// multiply, build_cahn_hilliard_operator, each through std::map rows, so every entry is the sum of
// its contributions in map-insertion order from 0.0), the ILDLT(0) preconditioner when it builds
// (build_preconditioner, 252-335, 434-442) else the diagonal one, and PCG without projection from
// x0 = c (515-660). When that PCG throws, the preconditioner becomes the diagonal one (and stays so
// for the operator) and PCG runs again from the x the failed attempt left.

namespace {

// The engine's orders: the one-workgroup ILDLT PCG sums with the 512-thread canonical tree over
// natural rows (kNaturalTree); the full-chip diagonal CG with BiCGSTAB's chunked order (kChunked,
// bicg_sum). or_ch_solve's reduction 2 selects each attempt's order by its preconditioner.
constexpr int kNaturalTree = 2;

template <class F>
double ch_sum(int mode, std::size_t n, F f) {
  if (mode == kChunked) return bicg_sum(kChunked, n, f);
  if (mode != kNaturalTree) {
    double s = 0.0;
    for (std::size_t k = 0; k < n; ++k) s += f(k);
    return s;
  }
  constexpr std::size_t T = 512, W = 8;
  double part[T];
  for (std::size_t t = 0; t < T; ++t) {
    double s = 0.0;
    for (std::size_t q = t; q < n; q += T) s = s + f(q);
    part[t] = s;
  }
  double wsum[W];
  for (std::size_t w = 0; w < W; ++w) {
    double *v = part + 64 * w;
    for (std::size_t off = 32; off >= 1; off >>= 1)
      for (std::size_t l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
    wsum[w] = v[0];
  }
  for (std::size_t off = W / 2; off >= 1; off >>= 1)
    for (std::size_t t = 0; t < off; ++t) wsum[t] = wsum[t] + wsum[t + off];
  return wsum[0];
}

struct ChCsr {
  int n = 0;
  std::vector<int> rp{0}, ci, diag;
  std::vector<double> v;
};

// Rows given as (column, value) lists already merged and sorted: drop |v| < 1e-18, record diagonals.
void push_row(ChCsr &A, const std::vector<std::pair<int, double>> &row, int r) {
  int d = -1;
  for (const auto &e : row) {
    if (std::abs(e.second) < 1.0e-18) continue;
    if (e.first == r) d = static_cast<int>(A.v.size());
    A.ci.push_back(e.first);
    A.v.push_back(e.second);
  }
  if (d < 0) throw std::runtime_error("sparse matrix assembly failed to produce a diagonal entry");
  A.diag.push_back(d);
  A.rp.push_back(static_cast<int>(A.v.size()));
  ++A.n;
}

// Accumulates contributions into a row: the first contribution to a column starts it at 0.0.
struct RowAccum {
  std::vector<std::pair<int, double>> e;  // insertion order
  void add(int col, double value) {
    for (auto &x : e)
      if (x.first == col) {
        x.second += value;
        return;
      }
    e.push_back({col, 0.0 + value});
  }
  std::vector<std::pair<int, double>> sorted() const {
    auto s = e;
    std::sort(s.begin(), s.end());
    return s;
  }
};

ChCsr laplacian(int nx, int ny, double dx, double dy, bool px, bool py) {
  const double ix = 1.0 / (dx * dx), iy = 1.0 / (dy * dy);
  ChCsr L;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const int r = i * ny + j;
      RowAccum a;
      double diag = 0.0;
      if (px || i > 0) {
        a.add((px && i == 0 ? nx - 1 : i - 1) * ny + j, ix);
        diag -= ix;
      }
      if (px || i < nx - 1) {
        a.add((px && i == nx - 1 ? 0 : i + 1) * ny + j, ix);
        diag -= ix;
      }
      if (py || j > 0) {
        a.add(i * ny + (py && j == 0 ? ny - 1 : j - 1), iy);
        diag -= iy;
      }
      if (py || j < ny - 1) {
        a.add(i * ny + (py && j == ny - 1 ? 0 : j + 1), iy);
        diag -= iy;
      }
      a.add(r, diag);
      push_row(L, a.sorted(), r);
    }
  return L;
}

ChCsr ch_operator(int nx, int ny, double dx, double dy, bool px, bool py, double alpha0, double beta1, double beta2) {
  const ChCsr L = laplacian(nx, ny, dx, dy, px, py);
  ChCsr A;
  for (int r = 0; r < L.n; ++r) {
    RowAccum l2;  // (L L)(r, :) in the traversal order of multiply()
    for (int pa = L.rp[r]; pa < L.rp[r + 1]; ++pa) {
      const int mid = L.ci[pa];
      for (int pb = L.rp[mid]; pb < L.rp[mid + 1]; ++pb) l2.add(L.ci[pb], L.v[pa] * L.v[pb]);
    }
    std::vector<std::pair<int, double>> l2row;  // finalize_row_maps of the product
    for (const auto &e : l2.sorted())
      if (!(std::abs(e.second) < 1.0e-18)) l2row.push_back(e);
    RowAccum a;
    a.add(r, alpha0);
    for (int p = L.rp[r]; p < L.rp[r + 1]; ++p) a.add(L.ci[p], -beta1 * L.v[p]);
    for (const auto &e : l2row) a.add(e.first, beta2 * e.second);
    push_row(A, a.sorted(), r);
  }
  return A;
}

struct ChPrecond {
  int type = 0;  // 0 diagonal, 1 ILDLT
  std::vector<double> dinv, lower, D;
  double floor = 0.0;
};

// try_build_incomplete_ldlt (ch_sparse_krylov.hpp:252-335)
bool ildlt(const ChCsr &A, ChPrecond &P) {
  P.lower.assign(A.v.size(), 0.0);
  P.D.assign(static_cast<std::size_t>(A.n), 0.0);
  double max_diag = 0.0;
  for (int r = 0; r < A.n; ++r) max_diag = std::max(max_diag, std::abs(A.v[static_cast<std::size_t>(A.diag[r])]));
  P.floor = std::max(1.0, max_diag) * 1.0e-12;
  auto lower_of = [&](int row, int col) {  // L(row, col) for col < row, 0 when absent
    for (int p = A.rp[row]; p < A.rp[row + 1] && A.ci[p] < row && A.ci[p] <= col; ++p)
      if (A.ci[p] == col) return P.lower[static_cast<std::size_t>(p)];
    return 0.0;
  };
  for (int r = 0; r < A.n; ++r) {
    for (int p = A.rp[r]; p < A.rp[r + 1] && A.ci[p] < r; ++p) {
      const int c = A.ci[p];
      double s = A.v[static_cast<std::size_t>(p)];
      for (int q = A.rp[r]; q < p && A.ci[q] < c; ++q) {
        const int k = A.ci[q];
        const double l_ck = lower_of(c, k);
        if (l_ck != 0.0) s -= P.lower[static_cast<std::size_t>(q)] * P.D[static_cast<std::size_t>(k)] * l_ck;
      }
      P.lower[static_cast<std::size_t>(p)] = s / std::max(P.D[static_cast<std::size_t>(c)], P.floor);
    }
    double d = A.v[static_cast<std::size_t>(A.diag[r])];
    for (int p = A.rp[r]; p < A.rp[r + 1] && A.ci[p] < r; ++p)
      d -= (P.lower[static_cast<std::size_t>(p)] * P.lower[static_cast<std::size_t>(p)]) * P.D[static_cast<std::size_t>(A.ci[p])];
    if (!std::isfinite(d) || d <= P.floor) return false;
    P.D[static_cast<std::size_t>(r)] = d;
  }
  P.type = 1;
  return true;
}

ChPrecond diagonal_precond(const ChCsr &A) {
  ChPrecond P;
  P.dinv.assign(static_cast<std::size_t>(A.n), 1.0);
  for (int r = 0; r < A.n; ++r) P.dinv[static_cast<std::size_t>(r)] = 1.0 / std::max(std::abs(A.v[static_cast<std::size_t>(A.diag[r])]), 1.0e-30);
  return P;
}

// apply_preconditioner, ILDLT branch (ch_sparse_krylov.hpp:486-512): unit-lower forward in CSR
// order, division by max(D, floor), unit-upper backward over the children in ascending row order.
void ch_apply(const ChCsr &A, const ChPrecond &P, const std::vector<double> &r, std::vector<double> &z) {
  const std::size_t n = static_cast<std::size_t>(A.n);
  if (P.type == 0) {
    z.resize(n);
    for (std::size_t k = 0; k < n; ++k) z[k] = P.dinv[k] * r[k];
    return;
  }
  std::vector<double> y(n, 0.0);
  for (int row = 0; row < A.n; ++row) {
    double s = r[static_cast<std::size_t>(row)];
    for (int p = A.rp[row]; p < A.rp[row + 1] && A.ci[p] < row; ++p)
      s -= P.lower[static_cast<std::size_t>(p)] * y[static_cast<std::size_t>(A.ci[p])];
    y[static_cast<std::size_t>(row)] = s;
  }
  for (std::size_t k = 0; k < n; ++k) y[k] /= std::max(P.D[k], P.floor);
  // children lists in ascending child row order
  std::vector<std::vector<std::pair<int, double>>> lt(n);
  for (int row = 0; row < A.n; ++row)
    for (int p = A.rp[row]; p < A.rp[row + 1] && A.ci[p] < row; ++p)
      lt[static_cast<std::size_t>(A.ci[p])].push_back({row, P.lower[static_cast<std::size_t>(p)]});
  z = y;
  for (int row = A.n - 1; row >= 0; --row) {
    double s = z[static_cast<std::size_t>(row)];
    for (const auto &e : lt[static_cast<std::size_t>(row)]) s -= e.second * z[static_cast<std::size_t>(e.first)];
    z[static_cast<std::size_t>(row)] = s;
  }
}

struct ChReport {
  int iterations = 0;
  double abs_res = 0.0, rel_res = 0.0, iter_res = 0.0;
};

// solve_preconditioned_cg without projection (ch_sparse_krylov.hpp:515-660)
ChReport ch_pcg(int mode, const ChCsr &A, const ChPrecond &P, const std::vector<double> &b, int max_it, double tol,
                std::vector<double> &x) {
  const std::size_t n = static_cast<std::size_t>(A.n);
  if (A.n <= 0) throw std::runtime_error("cannot solve an empty linear system");
  auto spmv = [&](const std::vector<double> &in, std::vector<double> &out) {
    out.assign(n, 0.0);
    for (int row = 0; row < A.n; ++row) {
      double s = 0.0;
      for (int p = A.rp[row]; p < A.rp[row + 1]; ++p) s += A.v[static_cast<std::size_t>(p)] * in[static_cast<std::size_t>(A.ci[p])];
      out[static_cast<std::size_t>(row)] = s;
    }
  };
  std::vector<double> ax, r(n), z, p, ap;
  spmv(x, ax);
  for (std::size_t k = 0; k < n; ++k) r[k] = b[k] - ax[k];
  const double bn = std::sqrt(std::max(ch_sum(mode, n, [&](std::size_t k) { return b[k] * b[k]; }), 0.0));
  ChReport rep;
  rep.abs_res = std::sqrt(std::max(ch_sum(mode, n, [&](std::size_t k) { return r[k] * r[k]; }), 0.0));
  rep.rel_res = rep.abs_res / std::max(bn, 1.0e-30);
  if (rep.rel_res < tol || rep.abs_res < tol) return rep;
  ch_apply(A, P, r, z);
  double rz = ch_sum(mode, n, [&](std::size_t k) { return r[k] * z[k]; });
  if (!std::isfinite(rz) || rz <= 0.0) throw std::runtime_error("pcg preconditioner breakdown: non-positive r^T M^{-1} r");
  p = z;
  const int limit = std::max(1, max_it);
  for (int it = 0; it < limit; ++it) {
    spmv(p, ap);
    const double pap = ch_sum(mode, n, [&](std::size_t k) { return p[k] * ap[k]; });
    if (!std::isfinite(pap) || std::abs(pap) < 1.0e-30) throw std::runtime_error("pcg breakdown: p^T A p is invalid");
    const double alpha = rz / pap;
    std::vector<double> delta(n);
    for (std::size_t k = 0; k < n; ++k) {
      delta[k] = alpha * p[k];
      x[k] += delta[k];
      r[k] -= alpha * ap[k];
    }
    const double up = ch_sum(mode, n, [&](std::size_t k) { return delta[k] * delta[k]; });
    const double xx = ch_sum(mode, n, [&](std::size_t k) { return x[k] * x[k]; });
    rep.iterations = it + 1;
    rep.iter_res = std::sqrt(up / std::max(xx, 1.0e-30));
    rep.abs_res = std::sqrt(std::max(ch_sum(mode, n, [&](std::size_t k) { return r[k] * r[k]; }), 0.0));
    rep.rel_res = rep.abs_res / std::max(bn, 1.0e-30);
    if (rep.rel_res < tol || rep.abs_res < tol) break;
    ch_apply(A, P, r, z);
    const double rzn = ch_sum(mode, n, [&](std::size_t k) { return r[k] * z[k]; });
    if (!std::isfinite(rzn) || rzn <= 0.0)
      throw std::runtime_error("pcg preconditioner breakdown: non-positive new r^T M^{-1} r");
    const double beta = rzn / rz;
    for (std::size_t k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
    rz = rzn;
  }
  return rep;
}

}  // namespace

extern "C" {

struct OrChProblem {
  int nx, ny;
  double lx, ly;
  int periodic_x, periodic_y;
  double alpha0, beta1, beta2;
  int max_iterations;
  double tolerance;
};

// The CH operator's CSR (sizes first with null arrays).
int or_ch_operator(const OrChProblem *cp, int *n, int *nnz, int *row_ptr, int *col, double *val, char *err, int errlen) {
  try {
    const ChCsr A = ch_operator(cp->nx, cp->ny, cp->lx / cp->nx, cp->ly / cp->ny, cp->periodic_x != 0, cp->periodic_y != 0,
                              cp->alpha0, cp->beta1, cp->beta2);
    *n = A.n;
    *nnz = static_cast<int>(A.v.size());
    if (row_ptr) std::copy(A.rp.begin(), A.rp.end(), row_ptr);
    if (col) std::copy(A.ci.begin(), A.ci.end(), col);
    if (val) std::copy(A.v.begin(), A.v.end(), val);
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// ILDLT(0) of the CH operator: returns 1 and fills lower (nnz, aligned with the CSR) and D (n) when it
// builds, 0 when it fails (build_preconditioner then takes the diagonal one).
int or_ch_ildlt(const OrChProblem *cp, double *lower, double *D, double *floor_out) {
  const ChCsr A = ch_operator(cp->nx, cp->ny, cp->lx / cp->nx, cp->ly / cp->ny, cp->periodic_x != 0, cp->periodic_y != 0,
                            cp->alpha0, cp->beta1, cp->beta2);
  ChPrecond P;
  const bool ok = ildlt(A, P);
  std::copy(P.lower.begin(), P.lower.end(), lower);
  std::copy(P.D.begin(), P.D.end(), D);
  *floor_out = P.floor;
  return ok ? 1 : 0;
}

// The solve: rhs and c (in: the initial guess, out: the solution) are natural-order nx*ny arrays.
// precond_in: -1 build (ILDLT if it builds, else diagonal), 0 diagonal (the cached state after a
// fallback); precond_out: 1 ILDLT, 0 diagonal, 2 diagonal after a fallback. reduction: 0 sequential
// (the reference), 2 the engine's orders (the ILDLT PCG's natural-order tree, the diagonal CG's
// chunked order), 3 the chunked order for every attempt.
int or_ch_solve(const OrChProblem *cp, const double *rhs, double *c, int precond_in, int reduction, int *iterations,
                double *abs_res, double *rel_res, double *iter_res, int *precond_out, char *err, int errlen) {
  try {
    const ChCsr A = ch_operator(cp->nx, cp->ny, cp->lx / cp->nx, cp->ly / cp->ny, cp->periodic_x != 0, cp->periodic_y != 0,
                              cp->alpha0, cp->beta1, cp->beta2);
    ChPrecond P;
    if (precond_in != 0 && ildlt(A, P)) {
      P.type = 1;
    } else {
      P = diagonal_precond(A);
    }
    const std::vector<double> b(rhs, rhs + A.n);
    std::vector<double> x(c, c + A.n);
    ChReport rep;
    int used = P.type;
    // reduction 2: the engine's order of each attempt (ILDLT: natural tree, diagonal: chunked)
    auto order = [&](const ChPrecond &Q) { return reduction == kNaturalTree && Q.type == 0 ? kChunked : reduction; };
    try {
      rep = ch_pcg(order(P), A, P, b, cp->max_iterations, cp->tolerance, x);
    } catch (const std::runtime_error &) {
      P = diagonal_precond(A);
      used = 2;
      rep = ch_pcg(order(P), A, P, b, cp->max_iterations, cp->tolerance, x);
    }
    std::copy(x.begin(), x.end(), c);
    *iterations = rep.iterations;
    *abs_res = rep.abs_res;
    *rel_res = rep.rel_res;
    *iter_res = rep.iter_res;
    *precond_out = used;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

}  // extern "C"
