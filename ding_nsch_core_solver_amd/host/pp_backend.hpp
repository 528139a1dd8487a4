// C++ host mirror of the reference's pressure-backend interface over the engine's C ABI.
//
// The reference selects a pressure backend by the string cfg_.pressure_scheme in
// Solver::solve_pressure_correction (src/core/coupled/pressure.cpp:1031-1089); each backend
// (pressure.cpp:302-365, 470-491) zero-fills pressure_correction_, solves, subtracts the mean when no
// pressure edge is Dirichlet, fills ghosts with apply_pressure_bc (operators.cpp:143-181), sets
// last_pressure_iterations_ and returns the relative residual; failures throw std::runtime_error
// with the backend's message. PressureBackend::solve does the same for the hip_* schemes, with
// fields given as ghosted Field2D-like views (data[(i+g)(ny+2g)+(j+g)], field.hpp:9-36).
#pragma once

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "pp_solver.h"

namespace ppamd {

// A Field2D-shaped view: nx x ny interior cells with `ghost` layers, i-major, j contiguous.
struct FieldView {
  double *data = nullptr;
  int nx = 0, ny = 0, ghost = 0;
  int ld() const { return ny + 2 * ghost; }
  double &operator()(int i, int j) const { return data[static_cast<long>(i + ghost) * ld() + (j + ghost)]; }
  double *interior() const { return &(*this)(0, 0); }
};

struct BoundarySpec {
  int type = PP_BC_UNSET;  // PP_BC_UNSET, PP_BC_DIRICHLET, PP_BC_NEUMANN
  double value = 0.0;
};

// The configuration values the pressure backends read (config.hpp pressure fields).
struct PressureSettings {
  double lx = 1.0, ly = 1.0, dt = 1.0e-3;
  bool periodic_x = false, periodic_y = false;
  BoundarySpec left, right, bottom, top;
  int poisson_iterations = 400;
  double pressure_tolerance = 1.0e-8;
  double density_ratio = 1.0;
};

// The velocity boundary values the correction step reads (config.hpp:105-106, 116-123); unset
// edges are no-slip, with the wall speeds on the u bottom / top edges (operators.cpp:62-112).
struct VelocitySettings {
  BoundarySpec u_left, u_right, u_bottom, u_top;
  BoundarySpec v_left, v_right, v_bottom, v_top;
  double bottom_wall_velocity_x = 0.0, top_wall_velocity_x = 0.0;
};

// Scheme strings served by the engine, and their engine codes; 0 when not a hip_* scheme.
int hip_scheme_code(const std::string &scheme);
// True for the split hip_* schemes: callers must also treat them as split in their velocity
// correction and analysis lists (operators.cpp:400-409, pressure.cpp:1078-1085).
bool hip_scheme_is_split(const std::string &scheme);

// Effective boundary of one edge: unset edges are homogeneous Neumann (operators.cpp:36-60).
BoundarySpec effective_bc(const BoundarySpec &b);
// pressure_has_dirichlet_boundary (operators.cpp:114-128).
bool has_dirichlet(const PressureSettings &s);

// Ghost fill of a pressure-like field (periodic copy, Dirichlet 2v - f, Neumann f + 2(g-1/2) v h).
void apply_pressure_bc(const PressureSettings &s, const FieldView &f);

// Per-iteration PCG monitor: called once per record of the solve's log, in order, after the solve
// (the scalar part of the reference's CGMonitor, ch_sparse_krylov.hpp:47-61).
using CGMonitor = std::function<void(const pp_monitor_record &)>;

// The reference's CSR system and solve report as the momentum predictor holds them
// (ch_sparse_krylov.hpp:12-18, 40-45). diag_pos is not read: the engine locates the diagonal.
struct SparseMatrixCSR {
  int n = 0;
  std::vector<int> row_ptr;
  std::vector<int> col_idx;
  std::vector<int> diag_pos;
  std::vector<double> values;
};
struct LinearSolveReport {
  int iterations = 0;
  double absolute_residual = 0.0;
  double relative_residual = 0.0;
  double iterate_residual = 0.0;
};

class PressureBackend {
 public:
  explicit PressureBackend(int device = 0);
  ~PressureBackend();
  PressureBackend(const PressureBackend &) = delete;
  PressureBackend &operator=(const PressureBackend &) = delete;

  // One pressure correction. rho_mid: nx x ny; u_star: (nx+1) x ny; v_star: nx x (ny+1) (all with
  // the same ghost count); pressure / pressure_previous are read by split schemes (may have null
  // data). Writes pressure_correction (interior + ghosts) and returns the relative residual.
  double solve(const std::string &scheme, const PressureSettings &s, const FieldView &rho_mid,
               const FieldView &u_star, const FieldView &v_star, const FieldView &pressure,
               const FieldView &pressure_previous, int step_index, const FieldView &pressure_correction);
  // The same solve with the PCG monitor log replayed through `monitor` (FP64 path).
  double solve(const std::string &scheme, const PressureSettings &s, const FieldView &rho_mid,
               const FieldView &u_star, const FieldView &v_star, const FieldView &pressure,
               const FieldView &pressure_previous, int step_index, const FieldView &pressure_correction,
               const CGMonitor &monitor);

  // The step after the solve, Solver::apply_pressure_velocity_correction (operators.cpp:480-534):
  // u, v get the rho-weighted gradient of pressure_correction (plus the split schemes' explicit
  // term) and their ghost cells, pressure_previous takes pressure, and pressure becomes the new
  // pressure with ghosts. All fields share one ghost count; u, v, pressure and pressure_previous
  // are read and written. Failures throw std::runtime_error.
  void correct_velocity(const std::string &scheme, const PressureSettings &s, const VelocitySettings &vel,
                        const FieldView &rho_mid, const FieldView &u_star, const FieldView &v_star,
                        const FieldView &pressure_correction, int step_index, const FieldView &u,
                        const FieldView &v, const FieldView &pressure, const FieldView &pressure_previous);

  // Both calls of solver.cpp:271-272 in one go (pp_pressure_step): pressure_correction is
  // written (interior and ghosts), then u, v, pressure and pressure_previous are corrected as by
  // correct_velocity; the fields cross to the device and back once. Returns the solve's relative
  // residual; a solve failure throws with the reference's message and leaves pressure_correction
  // zero and the other fields untouched.
  double pressure_step(const std::string &scheme, const PressureSettings &s, const VelocitySettings &vel,
                       const FieldView &rho_mid, const FieldView &u_star, const FieldView &v_star, int step_index,
                       const FieldView &pressure_correction, const FieldView &u, const FieldView &v,
                       const FieldView &pressure, const FieldView &pressure_previous);

  // The momentum predictor's solve (momentum.cpp:469-472): build_diagonal_preconditioner +
  // solve_preconditioned_bicgstab (ch_sparse_krylov.hpp:421-431, 839-963) on the engine, with the
  // reference's arguments, report and errors (std::runtime_error with its texts). x is the
  // initial guess, reset to n zeros when its size differs (as the reference does), then the
  // solution.
  LinearSolveReport solve_bicgstab(const SparseMatrixCSR &matrix, const std::vector<double> &rhs, int max_iterations,
                                   double tolerance, std::vector<double> &x, bool use_relative_tolerance = true);

  // Summation order of the pressure mean in correct_velocity / pressure_step: the reference's
  // sequential sum (exact = true, its bits) or the engine's tree (false, the default; within
  // rounding, microseconds instead of one dependent addition per cell). pp_set_exact_mean.
  void set_exact_mean(bool exact);

  int last_iterations() const { return last_iterations_; }
  const pp_report &last_report() const { return report_; }
  // last_boundary_speed_pre_correction_ / _post_correction_ of the last correct_velocity
  double boundary_speed_pre_correction() const { return speeds_[0]; }
  double boundary_speed_post_correction() const { return speeds_[1]; }

 private:
  pp_ctx *ctx_ = nullptr;
  int last_iterations_ = 0;
  pp_report report_{};
  double speeds_[2] = {0.0, 0.0};
};

}  // namespace ppamd
