// C++ host mirror of the reference pressure backends over the engine's C ABI (see pp_backend.hpp).
#include "pp_backend.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace ppamd {

namespace {

struct SchemeName {
  const char *name;
  int code;
};

// hip_ + the reference's scheme names (pressure.cpp:1048-1064), split aliases included.
constexpr SchemeName kSchemes[] = {
    {"hip_icpcg", PP_ICPCG},
    {"hip_ildlt_pcg", PP_ILDLT_PCG},
    {"hip_split_icpcg", PP_SPLIT_ICPCG},
    {"hip_liu_split_icpcg", PP_SPLIT_ICPCG},
    {"hip_paper_split_icpcg", PP_SPLIT_ICPCG},
    {"hip_split_ildlt_pcg", PP_SPLIT_ILDLT_PCG},
    {"hip_liu_split_ildlt_pcg", PP_SPLIT_ILDLT_PCG},
    {"hip_paper_split_ildlt_pcg", PP_SPLIT_ILDLT_PCG},
};

pp_problem make_problem(const PressureSettings &s, int code, int nx, int ny, int step_index) {
  pp_problem pb;
  std::memset(&pb, 0, sizeof(pb));
  pb.nx = nx;
  pb.ny = ny;
  pb.dx = s.lx / nx;
  pb.dy = s.ly / ny;
  pb.dt = s.dt;
  pb.periodic_x = s.periodic_x ? 1 : 0;
  pb.periodic_y = s.periodic_y ? 1 : 0;
  const BoundarySpec *edges[4] = {&s.left, &s.right, &s.bottom, &s.top};
  for (int k = 0; k < 4; ++k) {
    pb.bc_type[k] = edges[k]->type;
    pb.bc_value[k] = edges[k]->value;
  }
  pb.scheme = code;
  pb.precision = PP_FP64;
  pb.max_iterations = s.poisson_iterations;
  pb.tolerance = s.pressure_tolerance;
  pb.density_ratio = s.density_ratio;
  pb.step_index = step_index;
  return pb;
}

pp_velocity_bc velocity_bc(const VelocitySettings &vel) {
  pp_velocity_bc vb;
  std::memset(&vb, 0, sizeof(vb));
  const BoundarySpec *ue[4] = {&vel.u_left, &vel.u_right, &vel.u_bottom, &vel.u_top};
  const BoundarySpec *ve[4] = {&vel.v_left, &vel.v_right, &vel.v_bottom, &vel.v_top};
  for (int k = 0; k < 4; ++k) {
    vb.u_type[k] = ue[k]->type;
    vb.u_value[k] = ue[k]->value;
    vb.v_type[k] = ve[k]->type;
    vb.v_value[k] = ve[k]->value;
  }
  vb.bottom_wall_velocity_x = vel.bottom_wall_velocity_x;
  vb.top_wall_velocity_x = vel.top_wall_velocity_x;
  return vb;
}

// All fields of a pressure step share the pressure grid's ghost count and the MAC shapes.
bool same_grid(const FieldView &pressure, const FieldView &rho_mid, const FieldView &u_star, const FieldView &v_star,
               const FieldView &p_corr, const FieldView &u, const FieldView &v, const FieldView &pressure_previous) {
  const int nx = pressure.nx, ny = pressure.ny, g = pressure.ghost;
  auto same = [&](const FieldView &f, int rows, int cols) {
    return f.data && f.nx == rows && f.ny == cols && f.ghost == g;
  };
  return same(pressure, nx, ny) && same(rho_mid, nx, ny) && same(p_corr, nx, ny) && same(pressure_previous, nx, ny) &&
         same(u_star, nx + 1, ny) && same(u, nx + 1, ny) && same(v_star, nx, ny + 1) && same(v, nx, ny + 1);
}

}  // namespace

int hip_scheme_code(const std::string &scheme) {
  for (const auto &s : kSchemes)
    if (scheme == s.name) return s.code;
  return 0;
}

bool hip_scheme_is_split(const std::string &scheme) {
  const int c = hip_scheme_code(scheme);
  return c == PP_SPLIT_ICPCG || c == PP_SPLIT_ILDLT_PCG;
}

BoundarySpec effective_bc(const BoundarySpec &b) {
  if (b.type == PP_BC_UNSET) return {PP_BC_NEUMANN, 0.0};
  return b;
}

bool has_dirichlet(const PressureSettings &s) {
  auto dir = [](const BoundarySpec &b) { return effective_bc(b).type == PP_BC_DIRICHLET; };
  if (!s.periodic_x && (dir(s.left) || dir(s.right))) return true;
  if (!s.periodic_y && (dir(s.bottom) || dir(s.top))) return true;
  return false;
}

void apply_pressure_bc(const PressureSettings &s, const FieldView &f) {
  const int nx = f.nx, ny = f.ny;
  const double dx = s.lx / nx, dy = s.ly / ny;
  const BoundarySpec L = effective_bc(s.left), R = effective_bc(s.right);
  const BoundarySpec B = effective_bc(s.bottom), T = effective_bc(s.top);
  // x ghosts on interior rows first, then y ghosts over the full (ghosted) x extent, so corners
  // take the y rule applied to x ghosts
  for (int g = 1; g <= f.ghost; ++g) {
    const double shift = 2.0 * (static_cast<double>(g) - 0.5);
    for (int j = 0; j < ny; ++j) {
      if (s.periodic_x) {
        f(-g, j) = f(nx - g, j);
        f(nx - 1 + g, j) = f(g - 1, j);
      } else {
        f(-g, j) = L.type == PP_BC_DIRICHLET ? 2.0 * L.value - f(g - 1, j) : f(g - 1, j) + shift * L.value * dx;
        f(nx - 1 + g, j) = R.type == PP_BC_DIRICHLET ? 2.0 * R.value - f(nx - g, j) : f(nx - g, j) + shift * R.value * dx;
      }
    }
    for (int i = -f.ghost; i < nx + f.ghost; ++i) {
      if (s.periodic_y) {
        f(i, -g) = f(i, ny - g);
        f(i, ny - 1 + g) = f(i, g - 1);
      } else {
        f(i, -g) = B.type == PP_BC_DIRICHLET ? 2.0 * B.value - f(i, g - 1) : f(i, g - 1) + shift * B.value * dy;
        f(i, ny - 1 + g) = T.type == PP_BC_DIRICHLET ? 2.0 * T.value - f(i, ny - g) : f(i, ny - g) + shift * T.value * dy;
      }
    }
  }
}

PressureBackend::PressureBackend(int device) {
  if (pp_create(&ctx_, device) != PP_OK) throw std::runtime_error("pressure engine: device initialisation failed");
}

PressureBackend::~PressureBackend() { pp_destroy(ctx_); }

double PressureBackend::solve(const std::string &scheme, const PressureSettings &s, const FieldView &rho_mid,
                              const FieldView &u_star, const FieldView &v_star, const FieldView &pressure,
                              const FieldView &pressure_previous, int step_index,
                              const FieldView &pressure_correction) {
  return solve(scheme, s, rho_mid, u_star, v_star, pressure, pressure_previous, step_index, pressure_correction,
               CGMonitor());
}

double PressureBackend::solve(const std::string &scheme, const PressureSettings &s, const FieldView &rho_mid,
                              const FieldView &u_star, const FieldView &v_star, const FieldView &pressure,
                              const FieldView &pressure_previous, int step_index,
                              const FieldView &pressure_correction, const CGMonitor &monitor) {
  const int code = hip_scheme_code(scheme);
  if (code == 0) throw std::runtime_error("unsupported pressure scheme: " + scheme);
  const int nx = pressure_correction.nx, ny = pressure_correction.ny;
  const pp_problem pb = make_problem(s, code, nx, ny, step_index);

  // zero-fill as the reference backends do before solving, so a failed solve leaves zeros
  const long total = static_cast<long>(nx + 2 * pressure_correction.ghost) * pressure_correction.ld();
  for (long k = 0; k < total; ++k) pressure_correction.data[k] = 0.0;
  last_iterations_ = 0;
  const bool split = code == PP_SPLIT_ICPCG || code == PP_SPLIT_ILDLT_PCG;
  const double *pn = split && pressure.data ? pressure.interior() : nullptr;
  const double *pm = split && pressure_previous.data ? pressure_previous.interior() : nullptr;
  int rc = PP_OK;
  if (monitor) {
    std::vector<pp_monitor_record> log(static_cast<size_t>(std::max(1, s.poisson_iterations)) + 1);
    int calls = 0;
    rc = pp_solve_monitored(ctx_, &pb, rho_mid.interior(), rho_mid.ld(), u_star.interior(), u_star.ld(),
                            v_star.interior(), v_star.ld(), pn, pressure.ld(), pm, pressure_previous.ld(),
                            pressure_correction.interior(), pressure_correction.ld(), log.data(),
                            static_cast<int>(log.size()), &calls, &report_);
    for (int k = 0; k < std::min(calls, static_cast<int>(log.size())); ++k) monitor(log[static_cast<size_t>(k)]);
  } else {
    rc = pp_solve(ctx_, &pb, rho_mid.interior(), rho_mid.ld(), u_star.interior(), u_star.ld(), v_star.interior(),
                  v_star.ld(), pn, pressure.ld(), pm, pressure_previous.ld(), pressure_correction.interior(),
                  pressure_correction.ld(), &report_);
  }
  if (rc != PP_OK) {
    apply_pressure_bc(s, pressure_correction);
    throw std::runtime_error(report_.message[0] ? report_.message : pp_last_error(ctx_));
  }
  last_iterations_ = report_.iterations;
  apply_pressure_bc(s, pressure_correction);
  return report_.relative_residual;
}

void PressureBackend::set_exact_mean(bool exact) { pp_set_exact_mean(ctx_, exact ? 1 : 0); }

void PressureBackend::correct_velocity(const std::string &scheme, const PressureSettings &s,
                                       const VelocitySettings &vel, const FieldView &rho_mid,
                                       const FieldView &u_star, const FieldView &v_star,
                                       const FieldView &pressure_correction, int step_index, const FieldView &u,
                                       const FieldView &v, const FieldView &pressure,
                                       const FieldView &pressure_previous) {
  const int code = hip_scheme_code(scheme);
  if (code == 0) throw std::runtime_error("unsupported pressure scheme: " + scheme);
  const int nx = pressure.nx, ny = pressure.ny, g = pressure.ghost;
  if (!same_grid(pressure, rho_mid, u_star, v_star, pressure_correction, u, v, pressure_previous))
    throw std::runtime_error("velocity correction: field shapes or ghost counts do not match the pressure grid");
  const pp_problem pb = make_problem(s, code, nx, ny, step_index);
  const pp_velocity_bc vb = velocity_bc(vel);
  const int rc = pp_velocity_correction(ctx_, &pb, &vb, g, rho_mid.data, u_star.data, v_star.data,
                                        pressure_correction.data, u.data, v.data, pressure.data,
                                        pressure_previous.data, speeds_);
  if (rc != PP_OK) throw std::runtime_error(pp_last_error(ctx_));
}

double PressureBackend::pressure_step(const std::string &scheme, const PressureSettings &s,
                                     const VelocitySettings &vel, const FieldView &rho_mid, const FieldView &u_star,
                                     const FieldView &v_star, int step_index, const FieldView &pressure_correction,
                                     const FieldView &u, const FieldView &v, const FieldView &pressure,
                                     const FieldView &pressure_previous) {
  const int code = hip_scheme_code(scheme);
  if (code == 0) throw std::runtime_error("unsupported pressure scheme: " + scheme);
  if (!same_grid(pressure, rho_mid, u_star, v_star, pressure_correction, u, v, pressure_previous))
    throw std::runtime_error("pressure step: field shapes or ghost counts do not match the pressure grid");
  const pp_problem pb = make_problem(s, code, pressure.nx, pressure.ny, step_index);
  const pp_velocity_bc vb = velocity_bc(vel);
  last_iterations_ = 0;
  const int rc = pp_pressure_step(ctx_, &pb, &vb, pressure.ghost, rho_mid.data, u_star.data, v_star.data,
                                  pressure_correction.data, u.data, v.data, pressure.data, pressure_previous.data,
                                  speeds_, &report_);
  if (rc != PP_OK) throw std::runtime_error(report_.message[0] ? report_.message : pp_last_error(ctx_));
  last_iterations_ = report_.iterations;
  return report_.relative_residual;
}

LinearSolveReport PressureBackend::solve_bicgstab(const SparseMatrixCSR &matrix, const std::vector<double> &rhs,
                                                  int max_iterations, double tolerance, std::vector<double> &x,
                                                  bool use_relative_tolerance) {
  if (matrix.n <= 0) throw std::runtime_error("cannot solve an empty linear system");
  if (static_cast<int>(rhs.size()) != matrix.n) throw std::runtime_error("bicgstab rhs size mismatch");
  if (static_cast<int>(matrix.row_ptr.size()) != matrix.n + 1 ||
      matrix.col_idx.size() < static_cast<size_t>(matrix.row_ptr.back()) ||
      matrix.values.size() < static_cast<size_t>(matrix.row_ptr.back()))
    throw std::runtime_error("bicgstab: CSR arrays do not match n");
  if (static_cast<int>(x.size()) != matrix.n) x.assign(static_cast<size_t>(matrix.n), 0.0);
  const int rc = pp_bicgstab(ctx_, matrix.n, matrix.row_ptr.data(), matrix.col_idx.data(), matrix.values.data(),
                             rhs.data(), x.data(), max_iterations, tolerance, use_relative_tolerance ? 1 : 0, &report_);
  if (rc != PP_OK) throw std::runtime_error(report_.message[0] ? report_.message : pp_last_error(ctx_));
  last_iterations_ = report_.iterations;
  LinearSolveReport out;
  out.iterations = report_.iterations;
  out.absolute_residual = report_.absolute_residual;
  out.relative_residual = report_.relative_residual;
  out.iterate_residual = report_.iterate_residual;
  return out;
}

}  // namespace ppamd
