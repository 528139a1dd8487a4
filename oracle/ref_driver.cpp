// Checker-only driver over the reference solver compiled from its own sources.
//
// This file is test infrastructure: it is linked against the reference's src/core/*.cpp by
// oracle/Makefile into oracle/_ref/libdingref.so, and only tests/, the smoke check and the
// CPU-baseline leg of bench.py load that library. It reaches the Solver's private members the
// same way the reference's own unit tests do (src/tests/core_solver_unit_tests.cpp:7-9).
//
// Entry points (extern "C", plain pointers):
//   ref_pressure_solve   one call of Solver::solve_pressure_correction() (pressure.cpp:1031-1089)
//   ref_sys_*            the assembled CSR system, the ICC/ILDLT factors, one SpMV, one
//                        preconditioner application and one PCG solve, taken from the reference's
//                        own functions (pressure.cpp:123-248, ch_sparse_krylov.hpp:105-660).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/types.h>
#include <vector>

#define private public
#include "core/coupled/solver.hpp"
#undef private

extern "C" {

struct RefProblem {
  int nx, ny;
  double lx, ly, dt;
  int periodic_x, periodic_y;
  int bc_type[4];     // left, right, bottom, top; 0 unset, 1 dirichlet, 2 neumann
  double bc_value[4];
  int max_iterations;
  double tolerance;
  double density_ratio;
  int step_index;
};

}  // extern "C"

namespace {

ding::Config make_config(const RefProblem &pb, const char *scheme) {
  ding::Config cfg;
  cfg.name = "oracle_ref";
  cfg.nx = pb.nx;
  cfg.ny = pb.ny;
  cfg.ghost = 3;
  cfg.lx = pb.lx;
  cfg.ly = pb.ly;
  cfg.dt = pb.dt;
  cfg.periodic_x = pb.periodic_x != 0;
  cfg.periodic_y = pb.periodic_y != 0;
  ding::BoundaryConditionSpec *bcs[4] = {&cfg.pressure_bc_left, &cfg.pressure_bc_right, &cfg.pressure_bc_bottom,
                                         &cfg.pressure_bc_top};
  for (int s = 0; s < 4; ++s) {
    bcs[s]->type = static_cast<ding::BoundaryConditionType>(pb.bc_type[s]);
    bcs[s]->value = pb.bc_value[s];
  }
  cfg.poisson_iterations = pb.max_iterations;
  cfg.pressure_tolerance = pb.tolerance;
  cfg.density_ratio = pb.density_ratio;
  cfg.pressure_scheme = scheme ? scheme : "icpcg";
  cfg.verbose = false;
  cfg.write_vtk = false;
  return cfg;
}

void load_interior(ding::Field2D &f, const double *src) {
  if (!src) return;
  for (int i = 0; i < f.nx; ++i)
    for (int j = 0; j < f.ny; ++j) f(i, j) = src[static_cast<std::size_t>(i) * f.ny + j];
}

std::unique_ptr<ding::Solver> make_solver(const RefProblem &pb, const char *scheme, const double *rho,
                                          const double *u, const double *v, const double *p_n,
                                          const double *p_nm1) {
  auto s = std::make_unique<ding::Solver>(make_config(pb, scheme));
  s->validate_boundary_configuration();
  load_interior(s->rho_mid_, rho);
  load_interior(s->u_star_, u);
  load_interior(s->v_star_, v);
  load_interior(s->pressure_, p_n);
  load_interior(s->pressure_previous_step_, p_nm1);
  s->apply_pressure_bc(s->pressure_);
  s->apply_pressure_bc(s->pressure_previous_step_);
  s->current_step_index_ = pb.step_index;
  return s;
}

void set_err(char *err, int errlen, const char *msg) {
  if (err && errlen > 0) {
    std::snprintf(err, static_cast<std::size_t>(errlen), "%s", msg);
  }
}

struct RefSystem {
  std::unique_ptr<ding::Solver> solver;
  ding::PressureLinearSystem system;
  ding::ch_sparse_krylov::SparseMatrixCSR matrix;
  ding::ch_sparse_krylov::KrylovPreconditioner precond;
  bool factored = false;
};

}  // namespace

extern "C" {

// One pressure solve through the reference dispatcher. p_out receives the interior of
// pressure_correction_ (row i*ny+j); p_out_ghosted (may be NULL) receives the whole Field2D
// buffer including its 3 ghost layers. Returns 0 on success, 1 on a thrown error (message in err).
int ref_pressure_solve(const char *scheme, const RefProblem *pb, const double *rho, const double *u,
                       const double *v, const double *p_n, const double *p_nm1, double *p_out,
                       double *p_out_ghosted, int *iterations, double *residual, char *err, int errlen) {
  try {
    auto s = make_solver(*pb, scheme, rho, u, v, p_n, p_nm1);
    const double res = s->solve_pressure_correction();
    if (residual) *residual = res;
    if (iterations) *iterations = s->last_pressure_iterations_;
    if (p_out) {
      for (int i = 0; i < pb->nx; ++i)
        for (int j = 0; j < pb->ny; ++j)
          p_out[static_cast<std::size_t>(i) * pb->ny + j] = s->pressure_correction_(i, j);
    }
    if (p_out_ghosted) {
      std::memcpy(p_out_ghosted, s->pressure_correction_.data.data(),
                  s->pressure_correction_.data.size() * sizeof(double));
    }
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

struct RefVelocityBc {
  int u_type[4];      // left, right, bottom, top; 0 unset, 1 dirichlet, 2 neumann
  double u_value[4];
  int v_type[4];
  double v_value[4];
  double bottom_wall_velocity_x, top_wall_velocity_x;
};

// Solver::apply_pressure_velocity_correction() (operators.cpp:480-534) on whole Field2D buffers
// with G ghost layers: rho_mid, u_star, v_star, p_corr are read; u, v, pressure and pressure_prev
// (pressure_previous_step_) are read and written. speeds[0] / speeds[1] receive the periodic
// boundary speed before / after the correction. With repeats > 1 the call runs that many times on
// the same solver state (for timing; the outputs are those of the last call) and *seconds (if not
// null) receives the median wall time of one call, solver set-up and copies excluded.
int ref_velocity_correction(const char *scheme, const RefProblem *pb, const RefVelocityBc *vb, int G,
                            const double *rho_mid, const double *u_star, const double *v_star, const double *p_corr,
                            double *u, double *v, double *pressure, double *pressure_prev, double *speeds,
                            int repeats, double *seconds, char *err, int errlen) {
  try {
    ding::Config cfg = make_config(*pb, scheme);
    cfg.ghost = G;
    ding::BoundaryConditionSpec *ub[4] = {&cfg.u_bc_left, &cfg.u_bc_right, &cfg.u_bc_bottom, &cfg.u_bc_top};
    ding::BoundaryConditionSpec *vbs[4] = {&cfg.v_bc_left, &cfg.v_bc_right, &cfg.v_bc_bottom, &cfg.v_bc_top};
    for (int k = 0; k < 4; ++k) {
      ub[k]->type = static_cast<ding::BoundaryConditionType>(vb->u_type[k]);
      ub[k]->value = vb->u_value[k];
      vbs[k]->type = static_cast<ding::BoundaryConditionType>(vb->v_type[k]);
      vbs[k]->value = vb->v_value[k];
    }
    cfg.bottom_wall_velocity_x = vb->bottom_wall_velocity_x;
    cfg.top_wall_velocity_x = vb->top_wall_velocity_x;
    auto s = std::make_unique<ding::Solver>(cfg);
    auto in = [](ding::Field2D &f, const double *src) {
      std::memcpy(f.data.data(), src, f.data.size() * sizeof(double));
    };
    auto out = [](const ding::Field2D &f, double *dst) {
      std::memcpy(dst, f.data.data(), f.data.size() * sizeof(double));
    };
    in(s->rho_mid_, rho_mid);
    in(s->u_star_, u_star);
    in(s->v_star_, v_star);
    in(s->pressure_correction_, p_corr);
    in(s->u_, u);
    in(s->v_, v);
    in(s->pressure_, pressure);
    in(s->pressure_previous_step_, pressure_prev);
    s->current_step_index_ = pb->step_index;
    std::vector<double> times;
    for (int r = 0; r < std::max(1, repeats); ++r) {
      const auto t0 = std::chrono::steady_clock::now();
      s->apply_pressure_velocity_correction();
      times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    if (seconds) {
      std::sort(times.begin(), times.end());
      *seconds = times[times.size() / 2];
    }
    out(s->u_, u);
    out(s->v_, v);
    out(s->pressure_, pressure);
    out(s->pressure_previous_step_, pressure_prev);
    speeds[0] = s->last_boundary_speed_pre_correction_;
    speeds[1] = s->last_boundary_speed_post_correction_;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// Ghost fill of a pressure-like field (Field2D buffer with 3 ghost layers) by apply_pressure_bc.
int ref_apply_pressure_bc(const RefProblem *pb, double *field_ghosted) {
  auto s = std::make_unique<ding::Solver>(make_config(*pb, "icpcg"));
  ding::Field2D f(pb->nx, pb->ny, 3, 0.0);
  std::memcpy(f.data.data(), field_ghosted, f.data.size() * sizeof(double));
  s->apply_pressure_bc(f);
  std::memcpy(field_ghosted, f.data.data(), f.data.size() * sizeof(double));
  return 0;
}

// apply_pressure_bc on a Field2D buffer with G ghost layers (the loop runs over field.ghost).
int ref_pressure_ghosts(const RefProblem *pb, int G, double *field_ghosted) {
  auto s = std::make_unique<ding::Solver>(make_config(*pb, "icpcg"));
  ding::Field2D f(pb->nx, pb->ny, G, 0.0);
  std::memcpy(f.data.data(), field_ghosted, f.data.size() * sizeof(double));
  s->apply_pressure_bc(f);
  std::memcpy(field_ghosted, f.data.data(), f.data.size() * sizeof(double));
  return 0;
}

// Assemble (split or not), regularize and convert to CSR with the reference's own functions.
void *ref_sys_build(const RefProblem *pb, int use_split, const double *rho, const double *u, const double *v,
                    const double *p_n, const double *p_nm1, char *err, int errlen) {
  try {
    auto sys = std::make_unique<RefSystem>();
    sys->solver = make_solver(*pb, use_split ? "split_icpcg" : "icpcg", rho, u, v, p_n, p_nm1);
    ding::Solver &s = *sys->solver;
    if (use_split) {
      ding::Field2D p_ext(pb->nx, pb->ny, 3, 0.0);
      s.build_liu_split_pressure_extrapolation(p_ext);
      sys->system = s.assemble_pressure_linear_system(true, s.liu_split_reference_density(), &p_ext);
    } else {
      sys->system = s.assemble_pressure_linear_system(false, 1.0, nullptr);
    }
    ding::ch_sparse_krylov::finalize_row_maps(sys->system.row_maps, sys->matrix);
    return sys.release();
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

void ref_sys_free(void *h) { delete static_cast<RefSystem *>(h); }

void ref_sys_sizes(void *h, int *n, int *nnz, int *has_dirichlet) {
  auto *sys = static_cast<RefSystem *>(h);
  *n = sys->matrix.n;
  *nnz = static_cast<int>(sys->matrix.values.size());
  *has_dirichlet = sys->system.has_dirichlet_boundary ? 1 : 0;
}

void ref_sys_csr(void *h, int *row_ptr, int *col_idx, double *values, double *rhs) {
  auto *sys = static_cast<RefSystem *>(h);
  std::copy(sys->matrix.row_ptr.begin(), sys->matrix.row_ptr.end(), row_ptr);
  std::copy(sys->matrix.col_idx.begin(), sys->matrix.col_idx.end(), col_idx);
  std::copy(sys->matrix.values.begin(), sys->matrix.values.end(), values);
  std::copy(sys->system.rhs.begin(), sys->system.rhs.end(), rhs);
}

// kind 1 = ICC(0), 2 = ILDLT(0). Returns 1 on success, 0 when the factorization fails.
int ref_sys_factor(void *h, int kind, double *lower, double *diag, double *floor_out) {
  auto *sys = static_cast<RefSystem *>(h);
  bool ok = (kind == 2) ? ding::ch_sparse_krylov::try_build_incomplete_ldlt(sys->matrix, sys->precond)
                        : ding::ch_sparse_krylov::try_build_incomplete_cholesky(sys->matrix, sys->precond);
  sys->factored = ok;
  if (lower) std::copy(sys->precond.lower.begin(), sys->precond.lower.end(), lower);
  if (diag) std::copy(sys->precond.diag.begin(), sys->precond.diag.end(), diag);
  if (floor_out) *floor_out = sys->precond.diagonal_floor;
  return ok ? 1 : 0;
}

void ref_sys_apply_matrix(void *h, const double *x, double *y) {
  auto *sys = static_cast<RefSystem *>(h);
  std::vector<double> xv(x, x + sys->matrix.n), yv;
  ding::ch_sparse_krylov::apply_matrix(sys->matrix, xv, yv);
  std::copy(yv.begin(), yv.end(), y);
}

void ref_sys_apply_precond(void *h, const double *r, double *z) {
  auto *sys = static_cast<RefSystem *>(h);
  std::vector<double> rv(r, r + sys->matrix.n), zv;
  ding::ch_sparse_krylov::apply_preconditioner(sys->matrix, sys->precond, rv, zv);
  std::copy(zv.begin(), zv.end(), z);
}

// PCG from x0 = 0 with the reference's nullspace projection (pressure.cpp:50-57, 225-239).
// The scalar fields of one CGMonitor call (CGMonitorData, ch_sparse_krylov.hpp:47-58).
struct RefMonitorRecord {
  int iteration, pad;
  double alpha, beta, absolute_residual, relative_residual, iterate_residual, preconditioned_residual;
};

int ref_sys_pcg_monitored(void *h, int max_iterations, double tolerance, double *x, int *iterations, double *rel,
                          double *abs_res, RefMonitorRecord *records, int capacity, int *count, char *err,
                          int errlen);

int ref_sys_pcg(void *h, int max_iterations, double tolerance, double *x, int *iterations, double *rel,
                double *abs_res, char *err, int errlen) {
  return ref_sys_pcg_monitored(h, max_iterations, tolerance, x, iterations, rel, abs_res, nullptr, 0, nullptr, err,
                               errlen);
}

// The reference's solve_preconditioned_cg with a CGMonitor that records the scalar fields of every
// call: up to `capacity` records, *count = the number of calls (also when the PCG throws).
int ref_sys_pcg_monitored(void *h, int max_iterations, double tolerance, double *x, int *iterations, double *rel,
                          double *abs_res, RefMonitorRecord *records, int capacity, int *count, char *err,
                          int errlen) {
  auto *sys = static_cast<RefSystem *>(h);
  int calls = 0;
  const ding::ch_sparse_krylov::CGMonitor monitor = [&](const ding::ch_sparse_krylov::CGMonitorData &d) {
    if (records && calls < capacity)
      records[calls] = RefMonitorRecord{d.iteration, 0, d.alpha, d.beta, d.absolute_residual, d.relative_residual,
                                        d.iterate_residual, d.preconditioned_residual};
    ++calls;
  };
  struct Count {
    int &calls;
    int *count;
    ~Count() {
      if (count) *count = calls;
    }
  } counted{calls, count};
  try {
    std::vector<double> xv(static_cast<std::size_t>(sys->matrix.n), 0.0);
    const auto report = ding::ch_sparse_krylov::solve_preconditioned_cg(
        sys->matrix, sys->precond, sys->system.rhs, max_iterations, tolerance, xv,
        sys->solver->pressure_nullspace_projection(), monitor);
    std::copy(xv.begin(), xv.end(), x);
    *iterations = report.iterations;
    *rel = report.relative_residual;
    *abs_res = report.absolute_residual;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

}  // extern "C"

// ---- the momentum predictor's linear solve -------------------------------------------------------

extern "C" {

// The reference's build_diagonal_preconditioner + solve_preconditioned_bicgstab
// (ch_sparse_krylov.hpp:421-431, 839-963) on the CSR system given, after the reference's own
// finalize_row_maps (105-133) rebuilt it from per-row maps, as momentum.cpp:469-472 calls them.
// repeats > 1 times the solve (from the same x) and returns the mean in *seconds.
int ref_bicgstab(int n, const int *row_ptr, const int *col_idx, const double *values, const double *rhs, double *x,
                 int max_iterations, double tolerance, int use_relative, int *iterations, double *abs_res,
                 double *rel_res, double *iter_res, int repeats, double *seconds, char *err, int errlen) {
  namespace K = ding::ch_sparse_krylov;
  try {
    std::vector<std::map<int, double>> rows(static_cast<std::size_t>(std::max(n, 0)));
    for (int r = 0; r < n; ++r)
      for (int pos = row_ptr[r]; pos < row_ptr[r + 1]; ++pos) rows[static_cast<std::size_t>(r)][col_idx[pos]] = values[pos];
    K::SparseMatrixCSR matrix;
    K::finalize_row_maps(rows, matrix);
    const std::vector<double> b(rhs, rhs + std::max(n, 0));
    const std::vector<double> x0(x, x + std::max(n, 0));
    std::vector<double> xv;
    K::LinearSolveReport rep;
    const int reps = std::max(1, repeats);
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < reps; ++k) {
      xv = x0;
      const K::KrylovPreconditioner pre = K::build_diagonal_preconditioner(matrix);
      rep = K::solve_preconditioned_bicgstab(matrix, pre, b, max_iterations, tolerance, xv, use_relative != 0);
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count() / reps;
    std::copy(xv.begin(), xv.end(), x);
    *iterations = rep.iterations;
    *abs_res = rep.absolute_residual;
    *rel_res = rep.relative_residual;
    *iter_res = rep.iterate_residual;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}


// ---- full runs (solver.cpp:224-283) split at the pressure step ---------------------------------------
//
// A reference Solver built from a RefRunConfig and stepped in three parts, so that a caller can put
// another pressure step between the reference's phase / momentum stages and the rest of the
// timestep: ref_run_phase_momentum (the history shifts, advance_phase_state, compute_momentum_fluxes
// and solve_momentum_predictor of advance_one_timestep, solver.cpp:224-270), ref_run_pressure (the
// reference's own solve_pressure_correction + apply_pressure_velocity_correction, 271-272) and
// ref_run_finish (refresh_chemical_potential_for_current_phase and compute_diagnostics, 274-282).
// ref_run_field exposes the Solver's Field2D buffers (ghost layers included) for reading and writing.

struct RefRunConfig {
  int nx, ny;
  double lx, ly, dt;
  int periodic_x, periodic_y;
  double re, ca, pe, cn, density_ratio, viscosity_ratio, surface_tension_multiplier;
  double body_force_x, body_force_y;
  double center_x, center_y, radius;
  double width;  // > 0: c = 0.5 - 0.5 tanh((r - radius) / width) replaces the initial phase
  int ch_inner_iterations;
  double ch_tolerance;
  int momentum_iterations;
  double momentum_tolerance;
  int poisson_iterations;
  double pressure_tolerance;
};

struct RefRun {
  std::unique_ptr<ding::Solver> solver;
  int step = 0;
  double ch_residual = 0.0, momentum_residual = 0.0, pressure_residual = 0.0;
};

void *ref_run_create(const RefRunConfig *rc, const char *scheme, const char *restart_file, char *err, int errlen) {
  try {
    ding::Config cfg;
    cfg.name = "oracle_ref_run";
    cfg.mode = "coupled";
    cfg.nx = rc->nx;
    cfg.ny = rc->ny;
    cfg.ghost = 3;
    cfg.lx = rc->lx;
    cfg.ly = rc->ly;
    cfg.dt = rc->dt;
    cfg.periodic_x = rc->periodic_x != 0;
    cfg.periodic_y = rc->periodic_y != 0;
    cfg.re = rc->re;
    cfg.ca = rc->ca;
    cfg.pe = rc->pe;
    cfg.cn = rc->cn;
    cfg.density_ratio = rc->density_ratio;
    cfg.viscosity_ratio = rc->viscosity_ratio;
    cfg.surface_tension_multiplier = rc->surface_tension_multiplier;
    cfg.body_force_x = rc->body_force_x;
    cfg.body_force_y = rc->body_force_y;
    cfg.interface_center_x = rc->center_x;
    cfg.interface_center_y = rc->center_y;
    cfg.interface_radius = rc->radius;
    cfg.ch_inner_iterations = rc->ch_inner_iterations;
    cfg.ch_tolerance = rc->ch_tolerance;
    cfg.momentum_iterations = rc->momentum_iterations;
    cfg.momentum_tolerance = rc->momentum_tolerance;
    cfg.poisson_iterations = rc->poisson_iterations;
    cfg.pressure_tolerance = rc->pressure_tolerance;
    cfg.pressure_scheme = scheme ? scheme : "icpcg";
    cfg.verbose = false;
    cfg.write_vtk = false;
    cfg.write_restart = false;
    if (restart_file && *restart_file) {
      cfg.restart = true;
      cfg.restart_file = restart_file;
    }
    auto run = std::make_unique<RefRun>();
    run->solver = std::make_unique<ding::Solver>(cfg);
    ding::Solver &s = *run->solver;
    s.initialize();
    if (s.restarted_from_snapshot_) {
      run->step = s.restart_step_;
    } else if (rc->width > 0.0) {
      // the initial phase with another interface width, then the tail of Solver::initialize()
      for (int i = 0; i < cfg.nx; ++i)
        for (int j = 0; j < cfg.ny; ++j) {
          const double x = (i + 0.5) * s.dx_, y = (j + 0.5) * s.dy_;
          const double r = std::sqrt((x - rc->center_x) * (x - rc->center_x) + (y - rc->center_y) * (y - rc->center_y));
          s.c_(i, j) = 0.5 - 0.5 * std::tanh((r - rc->radius) / rc->width);
        }
      s.apply_scalar_bc(s.c_);
      s.c_previous_step_ = s.c_;
      s.c_two_steps_back_ = s.c_;
      s.update_materials();
      s.rho_previous_step_ = s.rho_;
      s.eta_previous_step_ = s.eta_;
      s.rho_mid_ = s.rho_;
      s.eta_mid_ = s.eta_;
      s.refresh_chemical_potential_for_current_phase();
      s.clear_surface_tension_force();
      s.initial_mass_ = s.compute_mass(s.c_);
    }
    return run.release();
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

void ref_run_destroy(void *h) { delete static_cast<RefRun *>(h); }

int ref_run_step_index(void *h) { return static_cast<RefRun *>(h)->step; }

// advance_one_timestep(step) up to the momentum predictor: history shifts and advance_phase_state
// (solver.cpp:224-263).
int ref_run_phase(void *h, double *seconds, char *err, int errlen) {
  auto *run = static_cast<RefRun *>(h);
  ding::Solver &s = *run->solver;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    const int step = run->step;
    s.current_step_index_ = step;
    s.c_two_steps_back_ = s.c_previous_step_;
    s.c_previous_step_ = s.c_;
    s.rho_previous_step_ = s.rho_;
    s.eta_previous_step_ = s.eta_;
    s.u_previous_step_ = s.u_;
    s.v_previous_step_ = s.v_;
    s.current_coupling_iteration_ = 0;
    run->ch_residual = s.advance_phase_state(step).ch_residual;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// ref_run_phase with another CH linear solve: the CH step of solve_cahn_hilliard_semi_implicit
// (phase_field.cpp:667-708) and the rest of advance_phase_state (283-294), with the linear solve of
// solve_phase_linear_system_eq25 (437-503) handed to `solve` on the interiors (natural order):
// solve(user, alpha0, rhs, c, n, &iterations, &iterate_residual, &equation_residual) returns 0 or an
// error. Runs that do not solve CH take the reference's own path.
typedef int (*RefChSolveFn)(void *user, double alpha0, const double *rhs, double *c, int n, int *iterations,
                            double *iterate_residual, double *equation_residual);

int ref_run_phase_external_ch(void *h, RefChSolveFn solve, void *user, double *seconds, char *err, int errlen) {
  auto *run = static_cast<RefRun *>(h);
  ding::Solver &s = *run->solver;
  try {
    if (!s.should_solve_cahn_hilliard()) return ref_run_phase(h, seconds, err, errlen);
    const auto t0 = std::chrono::steady_clock::now();
    const int step = run->step;
    s.current_step_index_ = step;
    s.c_two_steps_back_ = s.c_previous_step_;
    s.c_previous_step_ = s.c_;
    s.rho_previous_step_ = s.rho_;
    s.eta_previous_step_ = s.eta_;
    s.u_previous_step_ = s.u_;
    s.v_previous_step_ = s.v_;
    s.current_coupling_iteration_ = 0;
    const ding::Config &cfg = s.cfg_;
    const int nx = cfg.nx, ny = cfg.ny;
    ding::Field2D rhs_field(nx, ny, cfg.ghost, 0.0);
    ding::Field2D explicit_n(nx, ny, cfg.ghost, 0.0);
    ding::Field2D c_state = s.c_;
    s.build_phase_explicit_operator(s.c_, s.u_, s.v_, explicit_n);
    const double alpha0 = (step == 0) ? (1.0 / cfg.dt) : (3.0 / (2.0 * cfg.dt));
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        rhs_field(i, j) = (step == 0) ? s.c_previous_step_(i, j) / cfg.dt + explicit_n(i, j)
                                      : (4.0 * s.c_previous_step_(i, j) - s.c_two_steps_back_(i, j)) / (2.0 * cfg.dt) +
                                            2.0 * explicit_n(i, j) - s.phase_explicit_operator_prev_(i, j);
    s.apply_scalar_bc(rhs_field);
    std::vector<double> rhs(static_cast<std::size_t>(nx) * ny), c(rhs.size());
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        rhs[static_cast<std::size_t>(i) * ny + j] = rhs_field(i, j);
        c[static_cast<std::size_t>(i) * ny + j] = c_state(i, j);
      }
    int iterations = 0;
    double iterate_residual = 0.0, equation_residual = 0.0;
    if (solve(user, alpha0, rhs.data(), c.data(), nx * ny, &iterations, &iterate_residual, &equation_residual) != 0)
      throw std::runtime_error("external CH solve failed");
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) c_state(i, j) = c[static_cast<std::size_t>(i) * ny + j];
    s.apply_scalar_bc(c_state);
    s.last_ch_iterations_ = iterations;
    s.c_ = c_state;
    s.apply_scalar_bc(s.c_);
    s.update_chemical_potential(s.c_, s.mu_);
    s.last_ch_equation_residual_ = equation_residual;
    s.phase_advection_rhs_prev_.fill(0.0);
    s.phase_explicit_operator_prev_ = explicit_n;
    run->ch_residual = iterate_residual;
    s.update_materials();
    s.update_midpoint_materials();
    s.update_surface_tension_force(s.c_previous_step_, s.c_);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// compute_momentum_fluxes + solve_momentum_predictor of the current step (solver.cpp:265-267).
int ref_run_momentum(void *h, double *seconds, char *err, int errlen) {
  auto *run = static_cast<RefRun *>(h);
  ding::Solver &s = *run->solver;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    ding::Field2D u_adv(s.u_.nx, s.u_.ny, s.u_.ghost, 0.0);
    ding::Field2D v_adv(s.v_.nx, s.v_.ny, s.v_.ghost, 0.0);
    s.compute_momentum_fluxes(u_adv, v_adv);
    run->momentum_residual = s.solve_momentum_predictor(u_adv, v_adv, run->step);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// The momentum predictor's result fields after an external predictor wrote u_star_ / v_star_ and the
// convection fields: records its residual and iteration count for the step's diagnostics.
void ref_run_set_momentum_report(void *h, double residual, int iterations) {
  auto *run = static_cast<RefRun *>(h);
  run->momentum_residual = residual;
  run->solver->last_momentum_iterations_ = iterations;
}

// advance_one_timestep(step) up to the pressure step (solver.cpp:224-269).
int ref_run_phase_momentum(void *h, double *seconds, char *err, int errlen) {
  auto *run = static_cast<RefRun *>(h);
  ding::Solver &s = *run->solver;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    const int step = run->step;
    s.current_step_index_ = step;
    s.c_two_steps_back_ = s.c_previous_step_;
    s.c_previous_step_ = s.c_;
    s.rho_previous_step_ = s.rho_;
    s.eta_previous_step_ = s.eta_;
    s.u_previous_step_ = s.u_;
    s.v_previous_step_ = s.v_;
    s.current_coupling_iteration_ = 0;
    const ding::Solver::PhaseStepReport phase = s.advance_phase_state(step);
    run->ch_residual = phase.ch_residual;
    ding::Field2D u_adv(s.u_.nx, s.u_.ny, s.u_.ghost, 0.0);
    ding::Field2D v_adv(s.v_.nx, s.v_.ny, s.v_.ghost, 0.0);
    s.compute_momentum_fluxes(u_adv, v_adv);
    run->momentum_residual = s.solve_momentum_predictor(u_adv, v_adv, step);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// The reference's own pressure step (solver.cpp:271-272).
int ref_run_pressure(void *h, int *iterations, double *residual, double *seconds, char *err, int errlen) {
  auto *run = static_cast<RefRun *>(h);
  ding::Solver &s = *run->solver;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    run->pressure_residual = s.solve_pressure_correction();
    s.apply_pressure_velocity_correction();
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (iterations) *iterations = s.last_pressure_iterations_;
    if (residual) *residual = run->pressure_residual;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// The rest of advance_one_timestep (solver.cpp:274-282); the step counter advances.
// out[0..5]: ch residual, momentum residual, momentum iterations, CH iterations, mass, max velocity.
int ref_run_finish(void *h, double *out, char *err, int errlen) {
  auto *run = static_cast<RefRun *>(h);
  ding::Solver &s = *run->solver;
  try {
    s.refresh_chemical_potential_for_current_phase();
    s.last_diag_ = s.compute_diagnostics();
    if (out) {
      out[0] = run->ch_residual;
      out[1] = run->momentum_residual;
      out[2] = s.last_momentum_iterations_;
      out[3] = s.last_ch_iterations_;
      out[4] = s.last_diag_.mass;
      out[5] = s.last_diag_.max_velocity;
    }
    ++run->step;
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// The Field2D buffer named `name` (nx, ny and ghost of the field in dims[0..2]); null if unknown.
double *ref_run_field(void *h, const char *name, int *dims) {
  ding::Solver &s = *static_cast<RefRun *>(h)->solver;
  const std::map<std::string, ding::Field2D *> fields = {
      {"c", &s.c_}, {"rho", &s.rho_}, {"rho_mid", &s.rho_mid_}, {"eta", &s.eta_}, {"eta_mid", &s.eta_mid_},
      {"eta_previous_step", &s.eta_previous_step_}, {"mu", &s.mu_}, {"u", &s.u_}, {"v", &s.v_},
      {"u_star", &s.u_star_}, {"v_star", &s.v_star_}, {"pressure", &s.pressure_},
      {"pressure_previous_step", &s.pressure_previous_step_}, {"pressure_correction", &s.pressure_correction_},
      {"surface_fx_u", &s.surface_fx_u_}, {"surface_fy_v", &s.surface_fy_v_},
      {"momentum_u_rhs_prev", &s.momentum_u_rhs_prev_}, {"momentum_v_rhs_prev", &s.momentum_v_rhs_prev_}};
  auto it = fields.find(name);
  if (it == fields.end()) return nullptr;
  dims[0] = it->second->nx;
  dims[1] = it->second->ny;
  dims[2] = it->second->ghost;
  return it->second->data.data();
}


// Solver::compute_momentum_fluxes + solve_momentum_predictor (momentum.cpp:55-85, 157-502) on the
// whole ghosted buffers given (G layers): u, v, rho_mid, eta, eta_prev (eta_previous_step_),
// surface_fx_u, surface_fy_v and u_prev / v_prev (momentum_u/v_rhs_prev_, replaced by this step's
// convection fields). u_star / v_star receive the predictor's fields with their ghosts.
struct RefMomentumProblem {
  int nx, ny;
  double lx, ly, dt, re;
  double body_force_x, body_force_y;
  int periodic_x, periodic_y;
  int step_index;
  int max_iterations;
  double tolerance;
};

int ref_momentum_predictor(const RefMomentumProblem *mp, const RefVelocityBc *vb, int G, const double *u,
                           const double *v, const double *rho_mid, const double *eta, const double *eta_prev,
                           const double *sfx, const double *sfy, double *u_prev, double *v_prev, double *u_star,
                           double *v_star, int *iterations, double *abs_res, int repeats, double *seconds, char *err,
                           int errlen) {
  try {
    ding::Config cfg;
    cfg.name = "oracle_ref_momentum";
    cfg.nx = mp->nx;
    cfg.ny = mp->ny;
    cfg.ghost = G;
    cfg.lx = mp->lx;
    cfg.ly = mp->ly;
    cfg.dt = mp->dt;
    cfg.re = mp->re;
    cfg.body_force_x = mp->body_force_x;
    cfg.body_force_y = mp->body_force_y;
    cfg.periodic_x = mp->periodic_x != 0;
    cfg.periodic_y = mp->periodic_y != 0;
    cfg.momentum_iterations = mp->max_iterations;
    cfg.momentum_tolerance = mp->tolerance;
    cfg.verbose = false;
    cfg.write_vtk = false;
    ding::BoundaryConditionSpec *ub[4] = {&cfg.u_bc_left, &cfg.u_bc_right, &cfg.u_bc_bottom, &cfg.u_bc_top};
    ding::BoundaryConditionSpec *vbs[4] = {&cfg.v_bc_left, &cfg.v_bc_right, &cfg.v_bc_bottom, &cfg.v_bc_top};
    for (int k = 0; k < 4; ++k) {
      ub[k]->type = static_cast<ding::BoundaryConditionType>(vb->u_type[k]);
      ub[k]->value = vb->u_value[k];
      vbs[k]->type = static_cast<ding::BoundaryConditionType>(vb->v_type[k]);
      vbs[k]->value = vb->v_value[k];
    }
    cfg.bottom_wall_velocity_x = vb->bottom_wall_velocity_x;
    cfg.top_wall_velocity_x = vb->top_wall_velocity_x;
    auto s = std::make_unique<ding::Solver>(cfg);
    s->validate_boundary_configuration();
    auto in = [](ding::Field2D &f, const double *src) { std::memcpy(f.data.data(), src, f.data.size() * sizeof(double)); };
    auto out = [](const ding::Field2D &f, double *dst) { std::memcpy(dst, f.data.data(), f.data.size() * sizeof(double)); };
    std::vector<double> times;
    for (int r = 0; r < std::max(1, repeats); ++r) {
      in(s->u_, u);
      in(s->v_, v);
      in(s->rho_mid_, rho_mid);
      in(s->eta_, eta);
      in(s->eta_previous_step_, eta_prev);
      in(s->surface_fx_u_, sfx);
      in(s->surface_fy_v_, sfy);
      in(s->momentum_u_rhs_prev_, u_prev);
      in(s->momentum_v_rhs_prev_, v_prev);
      const auto t0 = std::chrono::steady_clock::now();
      ding::Field2D u_adv(s->u_.nx, s->u_.ny, s->u_.ghost, 0.0);
      ding::Field2D v_adv(s->v_.nx, s->v_.ny, s->v_.ghost, 0.0);
      s->compute_momentum_fluxes(u_adv, v_adv);
      *abs_res = s->solve_momentum_predictor(u_adv, v_adv, mp->step_index);
      times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    if (seconds) {
      std::sort(times.begin(), times.end());
      *seconds = times[times.size() / 2];
    }
    *iterations = s->last_momentum_iterations_;
    out(s->u_star_, u_star);
    out(s->v_star_, v_star);
    out(s->momentum_u_rhs_prev_, u_prev);
    out(s->momentum_v_rhs_prev_, v_prev);
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}


// The CH linear solve through the reference Solver's own solve_phase_linear_system_eq25
// (phase_field.cpp:437-503), on a Solver that keeps its cached operator and preconditioner between
// calls (phase_field.cpp:109-135), as in a run.
struct RefChConfig {
  int nx, ny;
  double lx, ly;
  int periodic_x, periodic_y;
  double cn, pe;
  int ch_inner_iterations;
  double ch_tolerance;
};

void *ref_ch_create(const RefChConfig *cc, char *err, int errlen) {
  try {
    ding::Config cfg;
    cfg.name = "oracle_ref_ch";
    cfg.nx = cc->nx;
    cfg.ny = cc->ny;
    cfg.lx = cc->lx;
    cfg.ly = cc->ly;
    cfg.periodic_x = cc->periodic_x != 0;
    cfg.periodic_y = cc->periodic_y != 0;
    cfg.cn = cc->cn;
    cfg.pe = cc->pe;
    cfg.ch_inner_iterations = cc->ch_inner_iterations;
    cfg.ch_tolerance = cc->ch_tolerance;
    cfg.verbose = false;
    cfg.write_vtk = false;
    return new ding::Solver(cfg);
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

void ref_ch_destroy(void *h) { delete static_cast<ding::Solver *>(h); }

// rhs and c are natural-order nx*ny arrays (c: the initial guess in, the solution out). name receives
// last_ch_solver_name_. With repeats > 1 the solve runs that many times from the same c and rhs (the
// operator cached after the first) and *seconds gets the median of one solve.
int ref_ch_solve(void *h, double alpha0, const double *rhs, double *c, int *iterations, double *iterate_residual,
                 double *equation_residual, char *name, int namelen, int repeats, double *seconds, char *err,
                 int errlen) {
  auto &s = *static_cast<ding::Solver *>(h);
  try {
    const int nx = s.cfg_.nx, ny = s.cfg_.ny;
    ding::Field2D rhs_field(nx, ny, s.cfg_.ghost, 0.0);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) rhs_field(i, j) = rhs[static_cast<std::size_t>(i) * ny + j];
    std::vector<double> times;
    ding::Field2D c_state(nx, ny, s.cfg_.ghost, 0.0);
    for (int rep = 0; rep < std::max(1, repeats); ++rep) {
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) c_state(i, j) = c[static_cast<std::size_t>(i) * ny + j];
      const auto t0 = std::chrono::steady_clock::now();
      s.solve_phase_linear_system_eq25(rhs_field, alpha0, 0.0, c_state, *iterate_residual, *equation_residual);
      times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    if (seconds) {
      std::sort(times.begin(), times.end());
      *seconds = times[times.size() / 2];
    }
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) c[static_cast<std::size_t>(i) * ny + j] = c_state(i, j);
    *iterations = s.last_ch_iterations_;
    if (name && namelen > 0) std::snprintf(name, static_cast<std::size_t>(namelen), "%s", s.last_ch_solver_name_.c_str());
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// The cached operator's CSR after a solve (sizes first with null arrays) and its preconditioner kind
// (1 ILDLT, 0 diagonal); lower / D when ILDLT.
void ref_ch_operator(void *h, int *n, int *nnz, int *row_ptr, int *col, double *val, int *kind, double *lower,
                     double *D) {
  auto &s = *static_cast<ding::Solver *>(h);
  const auto &A = s.ch_linear_system_matrix_;
  const auto &P = s.ch_linear_system_preconditioner_;
  *n = A.n;
  *nnz = static_cast<int>(A.values.size());
  *kind = P.type == ding::ch_sparse_krylov::PreconditionerType::incomplete_ldlt ? 1 : 0;
  if (row_ptr) std::copy(A.row_ptr.begin(), A.row_ptr.end(), row_ptr);
  if (col) std::copy(A.col_idx.begin(), A.col_idx.end(), col);
  if (val) std::copy(A.values.begin(), A.values.end(), val);
  if (lower && *kind == 1) std::copy(P.lower.begin(), P.lower.end(), lower);
  if (D && *kind == 1) std::copy(P.diag.begin(), P.diag.end(), D);
}

}  // extern "C"

