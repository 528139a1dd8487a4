// Host side of the C ABI declared in include/pp_solver.h: argument checks, device buffers, layout
// tables, transfers and the translation of device status codes into the reference's error texts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "pp_device.hpp"
#include "pp_kernels.hpp"
#include "pp_solver.h"

struct pp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;

  struct Layout {
    int nlev = 0;
    int *base = nullptr;   // level-base table with kLevPad entries of margin on both sides
    uint32_t *cell = nullptr;
    int split[pp::kWaves + 1] = {};  // SpMV level ranges per wave
  };
  std::map<std::pair<int, int>, Layout> layouts;

  // work arrays, sized for `cap` cells
  size_t cap = 0;
  double *work = nullptr;      // 13 FP64 arrays
  float *fwork = nullptr;      // 12 FP32 arrays
  double *wrap64 = nullptr;    // LX, LY
  float *wrap32 = nullptr;     // fLX, fLY
  // staging for the host entry point
  size_t stage_cap = 0;
  double *stage = nullptr;     // rho, u, v, p_n, p_nm1, p_out
  pp::SysArgs *d_args = nullptr;
  pp::DevReport *d_rep = nullptr;
  pp::DevReport *h_rep = nullptr;  // pinned
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  double last_kernel_ms = 0.0;  // event time of the last call's kernels (pp_last_kernel_ms)

  // sweep operand ring depth per grid shape (pp_set_tuning)
  std::map<std::pair<int, int>, int> tuning;
  // sleeps a hand-off wait may spend before the solve fails (pp_set_wait_limit; 0: built-in)
  int wait_limit = 0;
  // velocity correction's pressure mean: 1 the reference's sequential order, 0 the tree (pp_set_exact_mean)
  int exact_mean = 0;

  // batched solves: per-system work slab, wrap lines, arguments, reports, order and counters
  size_t batch_cells = 0;       // capacity of batch_work in cells (sum of n + kPad)
  size_t batch_fcells = 0;      // capacity of batch_fwork in cells (mixed-precision systems only)
  int batch_cap = 0;            // capacity in systems
  double *batch_work = nullptr;
  float *batch_fwork = nullptr;
  double *batch_wrap = nullptr;
  pp::SysArgs *d_batch_args = nullptr;
  pp::DevReport *d_batch_rep = nullptr;
  pp::DevReport *h_batch_rep = nullptr;  // pinned

  // velocity correction: argument block, split-extrapolation field and speed pair
  pp::CorrArgs *d_corr_args = nullptr;
  double *corr_work = nullptr;  // 4 doubles: the two boundary speeds, the pressure mean

  // BiCGSTAB (pp_bicgstab*): 12 work vectors of bicg_cap rows, per-workgroup partials, the
  // device state block and its pinned copy, and the host entry's CSR staging
  size_t bicg_cap = 0;
  double *bicg_work = nullptr;
  double *bicg_partials = nullptr;     // 12 * chunks
  pp::BicgState *bicg_state = nullptr;
  pp::BicgState *h_bicg = nullptr;     // pinned copies of the state (3 slots)
  hipEvent_t bicg_ev[2] = {nullptr, nullptr};
  size_t bicg_stage_rows = 0, bicg_stage_nnz = 0;
  int *bicg_stage_idx = nullptr;     // row_ptr then col_idx
  double *bicg_stage_val = nullptr;  // values, rhs, x

  // momentum predictor (pp_momentum_predictor*): per-row entry slots, counts, scan block sums, CSR,
  // right-hand side and solution for mom_cap rows; host-entry field staging; its own events
  size_t mom_cap = 0;
  int *mom_int = nullptr;         // slot_col (9 cap), count (cap), block sums, row_ptr (cap + 1), col (9 cap)
  double *mom_dbl = nullptr;      // slot_val (9 cap), val (9 cap), rhs (cap), x (cap)
  size_t mom_stage_cap = 0;
  double *mom_stage = nullptr;
  hipEvent_t mom_ev[2] = {nullptr, nullptr};

  // Cahn-Hilliard solve (pp_ch_solve*): the cached operator with its schedule, factor and vectors
  struct ChOperator {
    bool valid = false;
    int nx = 0, ny = 0, px = 0, py = 0;
    double dx = 0.0, dy = 0.0, alpha0 = 0.0, beta1 = 0.0, beta2 = 0.0;
    int n = 0, nnz = 0, nlev = 0;
    int kind = 0;                  // PP_CH_ILDLT, PP_CH_DIAGONAL or PP_CH_DIAGONAL_FALLBACK
    int *d_int = nullptr;          // row_ptr, col, diag_pos, lev_ptr, lev_rows, lt_ptr, lt_row, lt_pos
    double *d_dbl = nullptr;       // val, lower, D, dinv, r, z, p, ap, y
    pp::ChState *d_state = nullptr;
    pp::ChState *h_state = nullptr;  // pinned
    pp::ChArgs args{};
    size_t stage_cap = 0;
    double *stage = nullptr;       // host entry: rhs and c
    // full-chip diagonal CG: partials, state (pinned copies: two groups and the final state), events
    double *d_part = nullptr;
    pp::ChDiagState *d_dstate = nullptr;
    pp::ChDiagState *h_dstate = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
  } ch;

  // PCG monitor log of the host entry point (pp_solve_monitored)
  pp::MonitorRecord *monitor_log = nullptr;
  size_t monitor_cap = 0;       // capacity in records
  size_t field_cap = 0;         // host entry: staging capacity in doubles
  double *field_stage = nullptr;
};

namespace {

int fail(pp_ctx *ctx, pp_report *rep, int status, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  if (rep) {
    rep->status = status;
    std::snprintf(rep->message, sizeof(rep->message), "%s", msg.c_str());
  }
  return status;
}

#define PP_HIP_CHECK(expr)                                                                     \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail(ctx, report, PP_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));    \
  } while (0)

// Wavefront layout tables: the position of cell (i, d - i) is base[d] + i, and the (i, j) of every
// position. base has kLevPad entries of margin on each side (value 0: those levels hold no cell).
void build_layout(int nx, int ny, std::vector<int> &base, std::vector<uint32_t> &cell) {
  const int nlev = nx + ny - 1;
  base.assign(static_cast<size_t>(nlev) + 2 * pp::kLevPad, 0);
  cell.clear();
  cell.reserve(static_cast<size_t>(nx) * ny);
  for (int d = 0; d < nlev; ++d) {
    const int lo = std::max(0, d - ny + 1), hi = std::min(d, nx - 1);
    base[static_cast<size_t>(d + pp::kLevPad)] = static_cast<int>(cell.size()) - lo;
    for (int i = lo; i <= hi; ++i) cell.push_back((static_cast<uint32_t>(i) << 16) | static_cast<uint32_t>(d - i));
  }
}

// SpMV level ranges: wave w takes levels [split[w], split[w+1]), the boundaries at the first levels
// where the running cost reaches w / kWaves of the total (ranges may be empty on small grids). A
// level costs what the windowed pass spends on it: its 64-row chunks that hold a row of the level
// (the others skip their arithmetic) plus a fixed per-level part. Balancing by cells instead left
// the waves with the corner levels, many and sparse, as the last to finish (192x192 SpMV 11.5 ms by
// cells, 9.7 ms by this cost).
void spmv_split(int nx, int ny, int *split) {
  const int nlev = nx + ny - 1;
  auto cost = [&](int d) {  // the level's 64-row chunks with a row on it, plus kSpmvLevelCost per level
    return 64LL * (std::min(d, nx - 1) / 64 - std::max(0, d - ny + 1) / 64 + 1) + pp::kSpmvLevelCost;
  };
  long long total = 0;
  for (int d = 0; d < nlev; ++d) total += cost(d);
  long long cum = 0;
  int w = 1;
  split[0] = 0;
  for (int d = 0; d < nlev && w < pp::kWaves; ++d) {
    cum += cost(d);
    while (w < pp::kWaves && cum * pp::kWaves >= w * total) split[w++] = d + 1;
  }
  while (w <= pp::kWaves) split[w++] = nlev;
}

const char *status_text(int status, int scheme) {
  const bool ildlt = scheme == PP_ILDLT_PCG || scheme == PP_SPLIT_ILDLT_PCG;
  switch (status) {
    case pp::kFactorFail: return ildlt ? "ILDLT preconditioner construction failed" : "ICC preconditioner construction failed";
    case pp::kPapFail: return "pcg breakdown: p^T A p is invalid";
    case pp::kRz0Fail: return "pcg preconditioner breakdown: non-positive r^T M^{-1} r";
    case pp::kRzFail: return "pcg preconditioner breakdown: non-positive new r^T M^{-1} r";
    case pp::kNoDiag: return "sparse matrix assembly failed to produce a diagonal entry";
    case pp::kSweepFail: return "wavefront hand-off did not complete";
    default: return "unknown device status";
  }
}

// Checks of the problem itself; ctx may be null (the error text then only goes to the report).
int validate_problem(pp_ctx *ctx, const pp_problem *pb, pp_report *report) {
  if (!pb) return fail(ctx, report, PP_ERR_ARG, "null problem");
  // nx: one row per lane of the 8 waves. ny: the (i << 16) | j cell code; a periodic-x grid also
  // keeps its wrap line (ny values) in LDS, sized for kMaxSide.
  if (pb->nx < 1 || pb->ny < 1 || pb->nx > pp::kMaxRows || pb->ny > pp::kMaxColumns)
    return fail(ctx, report, PP_ERR_ARG, "grid size outside nx 1..512, ny 1..65535");
  if (pb->periodic_x && pb->ny > pp::kMaxSide)
    return fail(ctx, report, PP_ERR_ARG, "a grid periodic in x takes ny 1..1024");
  if ((pb->periodic_x && pb->nx < 4) || (pb->periodic_y && pb->ny < 4))
    return fail(ctx, report, PP_ERR_ARG, "periodic direction needs at least 4 cells");
  if (pb->scheme < PP_ICPCG || pb->scheme > PP_SPLIT_ILDLT_PCG) return fail(ctx, report, PP_ERR_ARG, "unknown scheme");
  if (pb->precision != PP_FP64 && pb->precision != PP_MIXED) return fail(ctx, report, PP_ERR_ARG, "unknown precision");
  for (int s = 0; s < 4; ++s)
    if (pb->bc_type[s] < PP_BC_UNSET || pb->bc_type[s] > PP_BC_NEUMANN)
      return fail(ctx, report, PP_ERR_ARG, "unknown boundary type");
  // explicit pressure BCs on periodic axes are rejected by the reference, which checks left, right,
  // bottom, top in that order and names the first edge that is set (operators.cpp:11-34)
  static const char *const kEdgeNames[4] = {"pressure_bc_left", "pressure_bc_right", "pressure_bc_bottom",
                                            "pressure_bc_top"};
  for (int s = 0; s < 4; ++s) {
    const bool periodic = s < 2 ? pb->periodic_x != 0 : pb->periodic_y != 0;
    if (periodic && pb->bc_type[s] != PP_BC_UNSET)
      return fail(ctx, report, PP_ERR_ARG, std::string(kEdgeNames[s]) + " cannot be set when the direction is periodic");
  }
  if (pb->ring_depth != 0 && !pp::valid_ring((pb->nx + 63) / 64, pb->ring_depth))
    return fail(ctx, report, PP_ERR_ARG,
                "ring_depth must be a power of two >= 8 with ceil(nx/64) * ring_depth <= 64, or -1 for direct sweeps");
  return PP_OK;
}

int validate(pp_ctx *ctx, const pp_problem *pb, pp_report *report) {
  if (!ctx) return fail(ctx, report, PP_ERR_ARG, "null context");
  return validate_problem(ctx, pb, report);
}

bool has_dirichlet(const pp_problem *pb) {
  auto dir = [&](int s) { return pb->bc_type[s] == PP_BC_DIRICHLET; };
  if (!pb->periodic_x && (dir(0) || dir(1))) return true;
  if (!pb->periodic_y && (dir(2) || dir(3))) return true;
  return false;
}

int ensure_layout(pp_ctx *ctx, int nx, int ny, pp_ctx::Layout **out, pp_report *report) {
  auto key = std::make_pair(nx, ny);
  auto it = ctx->layouts.find(key);
  if (it == ctx->layouts.end()) {
    std::vector<int> base;
    std::vector<uint32_t> cell;
    build_layout(nx, ny, base, cell);
    pp_ctx::Layout L;
    L.nlev = nx + ny - 1;
    spmv_split(nx, ny, L.split);
    PP_HIP_CHECK(hipMalloc(&L.base, base.size() * sizeof(int)));
    PP_HIP_CHECK(hipMalloc(&L.cell, cell.size() * sizeof(uint32_t)));
    PP_HIP_CHECK(hipMemcpy(L.base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
    PP_HIP_CHECK(hipMemcpy(L.cell, cell.data(), cell.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    it = ctx->layouts.emplace(key, L).first;
  }
  *out = &it->second;
  return PP_OK;
}

constexpr int kF64Arrays = 13;
constexpr int kF32Arrays = 12;

int ensure_work(pp_ctx *ctx, size_t n, pp_report *report) {
  if (n <= ctx->cap) return PP_OK;
  if (ctx->work) (void)hipFree(ctx->work);
  if (ctx->fwork) (void)hipFree(ctx->fwork);
  ctx->work = nullptr;
  ctx->fwork = nullptr;
  ctx->cap = 0;
  const size_t stride = n + pp::kPad;
  PP_HIP_CHECK(hipMalloc(&ctx->work, stride * kF64Arrays * sizeof(double)));
  PP_HIP_CHECK(hipMalloc(&ctx->fwork, stride * kF32Arrays * sizeof(float)));
  // padding is read by lanes outside the grid; keep it finite
  PP_HIP_CHECK(hipMemset(ctx->work, 0, stride * kF64Arrays * sizeof(double)));
  PP_HIP_CHECK(hipMemset(ctx->fwork, 0, stride * kF32Arrays * sizeof(float)));
  ctx->cap = n;
  return PP_OK;
}

// FP64 work arrays of one system: kF64Arrays arrays of `stride` doubles starting at w.
void set_work(pp::SysArgs &a, double *w, size_t stride) {
  double **f64[kF64Arrays] = {&a.aP, &a.aW, &a.aS, &a.Lw, &a.Ls, &a.dg, &a.b, &a.x, &a.r, &a.pa, &a.pb, &a.z, &a.y};
  for (int k = 0; k < kF64Arrays; ++k) *f64[k] = w + static_cast<size_t>(k) * stride;
}

// FP32 work arrays of one mixed-precision system: kF32Arrays arrays of `stride` floats at w.
void set_fwork(pp::SysArgs &a, float *w, size_t stride) {
  float **f32[kF32Arrays] = {&a.fP, &a.fW, &a.fS, &a.fLw, &a.fLs, &a.fRd, &a.fx, &a.fr, &a.fpa, &a.fpb, &a.fz, &a.fy};
  for (int k = 0; k < kF32Arrays; ++k) *f32[k] = w + static_cast<size_t>(k) * stride;
}

// The problem's scalar part of the argument block (no buffers, default ring depth and wait limit).
void fill_problem_args(const pp_problem *pb, pp::SysArgs &a) {
  std::memset(&a, 0, sizeof(a));
  a.nx = pb->nx;
  a.ny = pb->ny;
  a.n = pb->nx * pb->ny;
  a.nlev = pb->nx + pb->ny - 1;
  a.dx = pb->dx;
  a.dy = pb->dy;
  a.dt = pb->dt;
  a.periodic_x = pb->periodic_x ? 1 : 0;
  a.periodic_y = pb->periodic_y ? 1 : 0;
  for (int s = 0; s < 4; ++s) {
    a.bc_type[s] = pb->bc_type[s];
    a.bc_value[s] = pb->bc_value[s];
  }
  a.split = (pb->scheme == PP_SPLIT_ICPCG || pb->scheme == PP_SPLIT_ILDLT_PCG) ? 1 : 0;
  a.ildlt = (pb->scheme == PP_ILDLT_PCG || pb->scheme == PP_SPLIT_ILDLT_PCG) ? 1 : 0;
  a.projected = has_dirichlet(pb) ? 0 : 1;
  a.max_iterations = pb->max_iterations;
  a.tolerance = pb->tolerance;
  a.rho_ref = std::min(1.0, pb->density_ratio);
  a.step_index = pb->step_index;
  a.inner_max_iterations = pb->inner_max_iterations;
  a.inner_tolerance = pb->inner_tolerance;
  a.ring = pb->ring_depth;
}

void fill_args(pp_ctx *ctx, const pp_problem *pb, const pp_ctx::Layout &L, pp::SysArgs &a) {
  fill_problem_args(pb, a);
  if (a.ring == 0) {
    auto t = ctx->tuning.find(std::make_pair(pb->nx, pb->ny));
    if (t != ctx->tuning.end()) a.ring = t->second;
  }
  a.spin_limit = ctx->wait_limit;
  a.lev_base = L.base + pp::kLevPad;
  for (int w = 0; w <= pp::kWaves; ++w) a.spmv_split[w] = L.split[w];
  a.cell = L.cell;
  set_work(a, ctx->work, ctx->cap + pp::kPad);
  set_fwork(a, ctx->fwork, ctx->cap + pp::kPad);
  a.LX = ctx->wrap64;
  a.LY = ctx->wrap64 + pp::kMaxSide;
  a.fLX = ctx->wrap32;
  a.fLY = ctx->wrap32 + pp::kMaxSide;
  a.rep = ctx->d_rep;
}

// Fills the geometry/work part of `a` for problem pb; field pointers are the caller's.
int prepare(pp_ctx *ctx, const pp_problem *pb, pp::SysArgs &a, pp_report *report) {
  pp_ctx::Layout *L = nullptr;
  if (int rc = ensure_layout(ctx, pb->nx, pb->ny, &L, report)) return rc;
  if (int rc = ensure_work(ctx, static_cast<size_t>(pb->nx) * pb->ny, report)) return rc;
  fill_args(ctx, pb, *L, a);
  return PP_OK;
}

void copy_report(const pp::DevReport &r, double kernel_ms, pp_report *report) {
  if (!report) return;
  std::memset(report, 0, sizeof(*report));
  report->iterations = r.iterations;
  report->outer_cycles = r.outer_cycles;
  report->relative_residual = r.rel;
  report->absolute_residual = r.abs;
  report->iterate_residual = r.iter_res;
  report->kernel_ms = kernel_ms;
  for (int k = 0; k < pp::kPhases; ++k) report->phase_ms[k] = static_cast<double>(r.phase_ticks[k]) * 1.0e-5;
  report->status = PP_OK;
}

// monitored: launch the kernel variant that writes the PCG monitor log and counts its calls
int run_solve(pp_ctx *ctx, const pp_problem *pb, pp::SysArgs &a, pp_report *report, hipStream_t stream,
              bool monitored = false) {
  PP_HIP_CHECK(hipMemcpyAsync(ctx->d_args, &a, sizeof(a), hipMemcpyHostToDevice, stream));
  PP_HIP_CHECK(hipMemsetAsync(ctx->d_rep, 0, sizeof(pp::DevReport), stream));
  PP_HIP_CHECK(hipEventRecord(ctx->ev_start, stream));
  if (pb->precision == PP_MIXED)
    PP_HIP_CHECK(pp::launch_mixed_solve(ctx->d_args, 1, a.ildlt != 0, stream));
  else
    PP_HIP_CHECK(pp::launch_fp64_solve(ctx->d_args, 1, a.ildlt != 0, stream, monitored));
  PP_HIP_CHECK(hipEventRecord(ctx->ev_stop, stream));
  PP_HIP_CHECK(hipMemcpyAsync(ctx->h_rep, ctx->d_rep, sizeof(pp::DevReport), hipMemcpyDeviceToHost, stream));
  PP_HIP_CHECK(hipStreamSynchronize(stream));
  float kernel_ms = 0.0f;
  PP_HIP_CHECK(hipEventElapsedTime(&kernel_ms, ctx->ev_start, ctx->ev_stop));
  ctx->last_kernel_ms = kernel_ms;
  const pp::DevReport &r = *ctx->h_rep;
  copy_report(r, kernel_ms, report);
  if (r.status != pp::kOk) return fail(ctx, report, r.status, status_text(r.status, pb->scheme));
  ctx->last_error.clear();
  return PP_OK;
}

static_assert(sizeof(pp::MonitorRecord) == sizeof(pp_monitor_record), "monitor record layout");
static_assert(offsetof(pp::MonitorRecord, prec_res) == offsetof(pp_monitor_record, preconditioned_residual),
              "monitor record layout");

int validate_monitor(pp_ctx *ctx, const pp_problem *pb, const void *records, int capacity, pp_report *report) {
  if (int rc = validate(ctx, pb, report)) return rc;
  if (pb->precision == PP_MIXED) return fail(ctx, report, PP_ERR_ARG, "the PCG monitor log covers FP64 solves");
  if (capacity < 0 || (capacity > 0 && !records)) return fail(ctx, report, PP_ERR_ARG, "bad monitor log buffer");
  return PP_OK;
}

// pp_solve_device with an optional monitor log (device records, capacity, number of calls)
int solve_device(pp_ctx *ctx, const pp_problem *pb, const double *d_rho, const double *d_u, const double *d_v,
                 const double *d_p_n, const double *d_p_nm1, double *d_p_out, pp::MonitorRecord *d_records,
                 int capacity, int *count, pp_report *report, void *stream) {
  if (int rc = validate(ctx, pb, report)) return rc;
  if (!d_rho && !(pb->scheme == PP_SPLIT_ICPCG || pb->scheme == PP_SPLIT_ILDLT_PCG))
    return fail(ctx, report, PP_ERR_ARG, "rho is required");
  if (!d_u || !d_v || !d_p_out) return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  pp::SysArgs a;
  if (int rc = prepare(ctx, pb, a, report)) return rc;
  a.rho = d_rho;
  a.u = d_u;
  a.v = d_v;
  a.p_n = d_p_n;
  a.p_nm1 = d_p_nm1;
  a.p_out = d_p_out;
  const bool monitored = d_records != nullptr || count != nullptr;
  a.monitor = d_records;
  a.monitor_cap = d_records ? capacity : 0;
  const int rc = run_solve(ctx, pb, a, report, s, monitored);
  if (count) *count = monitored ? ctx->h_rep->monitor_calls : 0;
  return rc;
}

int solve_host(pp_ctx *ctx, const pp_problem *pb, const double *rho, int ld_rho, const double *u, int ld_u,
               const double *v, int ld_v, const double *p_n, int ld_pn, const double *p_nm1, int ld_pnm1,
               double *p_out, int ld_out, pp_monitor_record *records, int capacity, int *count, pp_report *report);

void free_batch(pp_ctx *ctx) {
  if (ctx->batch_work) (void)hipFree(ctx->batch_work);
  if (ctx->batch_fwork) (void)hipFree(ctx->batch_fwork);
  if (ctx->batch_wrap) (void)hipFree(ctx->batch_wrap);
  if (ctx->d_batch_args) (void)hipFree(ctx->d_batch_args);
  if (ctx->d_batch_rep) (void)hipFree(ctx->d_batch_rep);
  if (ctx->h_batch_rep) (void)hipHostFree(ctx->h_batch_rep);
  ctx->batch_work = ctx->batch_wrap = nullptr;
  ctx->batch_fwork = nullptr;
  ctx->d_batch_args = nullptr;
  ctx->d_batch_rep = ctx->h_batch_rep = nullptr;
  ctx->batch_cells = ctx->batch_fcells = 0;
  ctx->batch_cap = 0;
}

// Batch buffers for `count` systems with `cells` FP64 work cells in total, `fcells` of them also
// needing FP32 work arrays (mixed precision).
int ensure_batch(pp_ctx *ctx, int count, size_t cells, size_t fcells, pp_report *report) {
  if (count <= ctx->batch_cap && cells <= ctx->batch_cells && fcells <= ctx->batch_fcells) return PP_OK;
  const int cap = std::max(count, ctx->batch_cap);
  const size_t cap_cells = std::max(cells, ctx->batch_cells);
  const size_t cap_fcells = std::max(fcells, ctx->batch_fcells);
  free_batch(ctx);
  PP_HIP_CHECK(hipMalloc(&ctx->batch_work, cap_cells * kF64Arrays * sizeof(double)));
  PP_HIP_CHECK(hipMemset(ctx->batch_work, 0, cap_cells * kF64Arrays * sizeof(double)));  // finite padding
  if (cap_fcells > 0) {
    PP_HIP_CHECK(hipMalloc(&ctx->batch_fwork, cap_fcells * kF32Arrays * sizeof(float)));
    PP_HIP_CHECK(hipMemset(ctx->batch_fwork, 0, cap_fcells * kF32Arrays * sizeof(float)));
  }
  PP_HIP_CHECK(hipMalloc(&ctx->batch_wrap, static_cast<size_t>(cap) * 2 * pp::kMaxSide * sizeof(double)));
  PP_HIP_CHECK(hipMalloc(&ctx->d_batch_args, static_cast<size_t>(cap) * sizeof(pp::SysArgs)));
  PP_HIP_CHECK(hipMalloc(&ctx->d_batch_rep, static_cast<size_t>(cap) * sizeof(pp::DevReport)));
  PP_HIP_CHECK(hipHostMalloc(&ctx->h_batch_rep, static_cast<size_t>(cap) * sizeof(pp::DevReport), hipHostMallocDefault));
  ctx->batch_cap = cap;
  ctx->batch_cells = cap_cells;
  ctx->batch_fcells = cap_fcells;
  return PP_OK;
}

// Relative cost of one system for largest-first scheduling: iterations grow with the grid extent;
// an iteration costs two level-sequential sweeps plus streaming passes over the n cells.
double batch_cost(const pp_problem &pb) {
  const double levels = pb.nx + pb.ny - 1, n = static_cast<double>(pb.nx) * pb.ny;
  return (pb.nx + pb.ny) * (0.63 * levels + 0.0035 * n);
}

int ensure_stage(pp_ctx *ctx, size_t need, pp_report *report) {
  if (need <= ctx->stage_cap) return PP_OK;
  if (ctx->stage) (void)hipFree(ctx->stage);
  ctx->stage = nullptr;
  ctx->stage_cap = 0;
  PP_HIP_CHECK(hipMalloc(&ctx->stage, need * sizeof(double)));
  ctx->stage_cap = need;
  return PP_OK;
}

// Coefficient-level solves and the adjoint are defined on the non-split operator.
int validate_coeffs(pp_ctx *ctx, const pp_problem *pb, pp_report *report) {
  if (int rc = validate(ctx, pb, report)) return rc;
  if (pb->scheme != PP_ICPCG && pb->scheme != PP_ILDLT_PCG)
    return fail(ctx, report, PP_ERR_ARG, "coefficient-level solves and the adjoint take PP_ICPCG or PP_ILDLT_PCG");
  return PP_OK;
}

int run_coeff_solve(pp_ctx *ctx, const pp_problem *pb, const double *d_ku, const double *d_kv, const double *d_src,
                    int rhs_bc, double *d_out, pp_report *report, hipStream_t s) {
  pp::SysArgs a;
  if (int rc = prepare(ctx, pb, a, report)) return rc;
  a.coeff_mode = 1;
  a.rhs_bc = rhs_bc;
  a.ku = d_ku;
  a.kv = d_kv;
  a.src = d_src;
  a.p_out = d_out;
  return run_solve(ctx, pb, a, report, s);
}

int run_adjoint(pp_ctx *ctx, const pp_problem *pb, const double *d_ku, const double *d_kv, const double *d_p,
                const double *d_g, double *d_lam, double *d_dku, double *d_dkv, pp_report *report, hipStream_t s) {
  // lam = A^+ P g: the same solve with g as the source and no boundary terms on the right-hand side
  if (int rc = run_coeff_solve(ctx, pb, d_ku, d_kv, d_g, 0, d_lam, report, s)) return rc;
  pp::GradArgs g;
  std::memset(&g, 0, sizeof(g));
  g.nx = pb->nx;
  g.ny = pb->ny;
  g.dx = pb->dx;
  g.dy = pb->dy;
  g.periodic_x = pb->periodic_x ? 1 : 0;
  g.periodic_y = pb->periodic_y ? 1 : 0;
  for (int k = 0; k < 4; ++k) {
    g.bc_type[k] = pb->bc_type[k];
    g.bc_value[k] = pb->bc_value[k];
  }
  g.p = d_p;
  g.lam = d_lam;
  g.dku = d_dku;
  g.dkv = d_dkv;
  PP_HIP_CHECK(pp::launch_face_gradient(g, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return PP_OK;
}

// Argument block of the correction / ghost-fill kernels: grid, effective pressure BCs
// (operators.cpp:36-60: unset -> Neumann 0) and, when vbc is given, effective velocity BCs
// (62-112: unset -> no-slip, u bottom/top at the wall speed). Field pointers are the caller's.
pp::CorrArgs corr_args(const pp_problem *pb, const pp_velocity_bc *vbc, int G) {
  pp::CorrArgs c;
  std::memset(&c, 0, sizeof(c));
  c.nx = pb->nx;
  c.ny = pb->ny;
  c.G = G;
  c.dx = pb->dx;
  c.dy = pb->dy;
  c.dt = pb->dt;
  c.periodic_x = pb->periodic_x ? 1 : 0;
  c.periodic_y = pb->periodic_y ? 1 : 0;
  for (int k = 0; k < 4; ++k) {
    const bool p_unset = pb->bc_type[k] == PP_BC_UNSET;
    c.p_type[k] = p_unset ? PP_BC_NEUMANN : pb->bc_type[k];
    c.p_value[k] = p_unset ? 0.0 : pb->bc_value[k];
    if (!vbc) continue;
    const bool u_unset = vbc->u_type[k] == PP_BC_UNSET;
    const double u_default = k == 2 ? vbc->bottom_wall_velocity_x : (k == 3 ? vbc->top_wall_velocity_x : 0.0);
    c.u_type[k] = u_unset ? PP_BC_DIRICHLET : vbc->u_type[k];
    c.u_value[k] = u_unset ? u_default : vbc->u_value[k];
    const bool v_unset = vbc->v_type[k] == PP_BC_UNSET;
    c.v_type[k] = v_unset ? PP_BC_DIRICHLET : vbc->v_type[k];
    c.v_value[k] = v_unset ? 0.0 : vbc->v_value[k];
  }
  c.split = (pb->scheme == PP_SPLIT_ICPCG || pb->scheme == PP_SPLIT_ILDLT_PCG) ? 1 : 0;
  c.step_index = pb->step_index;
  c.has_dirichlet = has_dirichlet(pb) ? 1 : 0;
  c.rho_ref = c.split ? std::min(1.0, pb->density_ratio) : 1.0;
  return c;
}

int ensure_corr(pp_ctx *ctx) {
  pp_report *report = nullptr;
  if (!ctx->corr_work) PP_HIP_CHECK(hipMalloc(&ctx->corr_work, 4 * sizeof(double)));
  if (!ctx->d_corr_args) PP_HIP_CHECK(hipMalloc(&ctx->d_corr_args, sizeof(pp::CorrArgs)));
  return PP_OK;
}

// Device staging of whole ghosted Field2D buffers for the host-pointer correction / step entries.
int ensure_fields(pp_ctx *ctx, size_t need) {
  pp_report *report = nullptr;
  if (need <= ctx->field_cap) return PP_OK;
  if (ctx->field_stage) (void)hipFree(ctx->field_stage);
  ctx->field_stage = nullptr;
  ctx->field_cap = 0;
  PP_HIP_CHECK(hipMalloc(&ctx->field_stage, need * sizeof(double)));
  ctx->field_cap = need;
  return PP_OK;
}

// Velocity correction (operators.cpp:480-534): resolved boundary conditions and the launch.
int run_velocity_correction(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost,
                            const double *d_rho_mid, const double *d_u_star, const double *d_v_star,
                            const double *d_p_corr, double *d_u, double *d_v, double *d_pressure,
                            double *d_pressure_prev, double *speeds, hipStream_t s) {
  pp_report *report = nullptr;
  const int nx = pb->nx, ny = pb->ny, G = ghost;
  if (int rc = ensure_corr(ctx)) return rc;
  pp::CorrArgs c = corr_args(pb, vbc, G);
  c.exact_mean = ctx->exact_mean;
  c.rho_mid = d_rho_mid;
  c.u_star = d_u_star;
  c.v_star = d_v_star;
  c.p_corr = d_p_corr;
  c.u = d_u;
  c.v = d_v;
  c.pressure = d_pressure;
  c.pressure_prev = d_pressure_prev;
  c.work = ctx->corr_work;
  PP_HIP_CHECK(hipMemcpyAsync(ctx->d_corr_args, &c, sizeof(c), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipEventRecord(ctx->ev_start, s));
  PP_HIP_CHECK(pp::launch_velocity_correction(ctx->d_corr_args, nx, ny, G, s));
  PP_HIP_CHECK(hipEventRecord(ctx->ev_stop, s));
  double sp[2] = {0.0, 0.0};
  PP_HIP_CHECK(hipMemcpyAsync(sp, ctx->corr_work, sizeof(sp), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  float kernel_ms = 0.0f;
  PP_HIP_CHECK(hipEventElapsedTime(&kernel_ms, ctx->ev_start, ctx->ev_stop));
  ctx->last_kernel_ms = kernel_ms;
  if (speeds) {
    speeds[0] = sp[0];
    speeds[1] = sp[1];
  }
  ctx->last_error.clear();
  return PP_OK;
}

int validate_correction(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost) {
  pp_report *report = nullptr;
  if (int rc = validate(ctx, pb, report)) return rc;
  if (!vbc) return fail(ctx, report, PP_ERR_ARG, "null velocity boundary conditions");
  if (ghost < 1 || ghost > 8) return fail(ctx, report, PP_ERR_ARG, "ghost must be 1..8");
  for (int k = 0; k < 4; ++k)
    if (vbc->u_type[k] < PP_BC_UNSET || vbc->u_type[k] > PP_BC_NEUMANN || vbc->v_type[k] < PP_BC_UNSET ||
        vbc->v_type[k] > PP_BC_NEUMANN)
      return fail(ctx, report, PP_ERR_ARG, "unknown velocity boundary type");
  return PP_OK;
}

}  // namespace

extern "C" {

int pp_velocity_correction_device(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost,
                                  const double *d_rho_mid, const double *d_u_star, const double *d_v_star,
                                  const double *d_p_corr, double *d_u, double *d_v, double *d_pressure,
                                  double *d_pressure_prev, double *speeds, void *stream) {
  if (int rc = validate_correction(ctx, pb, vbc, ghost)) return rc;
  pp_report *report = nullptr;
  if (!d_rho_mid || !d_u_star || !d_v_star || !d_p_corr || !d_u || !d_v || !d_pressure || !d_pressure_prev)
    return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return run_velocity_correction(ctx, pb, vbc, ghost, d_rho_mid, d_u_star, d_v_star, d_p_corr, d_u, d_v, d_pressure,
                                 d_pressure_prev, speeds, s);
}

int pp_velocity_correction(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost,
                           const double *rho_mid, const double *u_star, const double *v_star, const double *p_corr,
                           double *u, double *v, double *pressure, double *pressure_prev, double *speeds) {
  if (int rc = validate_correction(ctx, pb, vbc, ghost)) return rc;
  pp_report *report = nullptr;
  if (!rho_mid || !u_star || !v_star || !p_corr || !u || !v || !pressure || !pressure_prev)
    return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  const int nx = pb->nx, ny = pb->ny, G = ghost;
  const size_t nc = static_cast<size_t>(nx + 2 * G) * (ny + 2 * G);
  const size_t nu = static_cast<size_t>(nx + 1 + 2 * G) * (ny + 2 * G);
  const size_t nv = static_cast<size_t>(nx + 2 * G) * (ny + 1 + 2 * G);
  if (int rc = ensure_fields(ctx, 5 * nc + 2 * nu + 2 * nv)) return rc;
  double *d_rho = ctx->field_stage, *d_pc = d_rho + nc, *d_p = d_pc + nc, *d_pp = d_p + nc;
  double *d_us = d_pp + nc, *d_u = d_us + nu, *d_vs = d_u + nu, *d_v = d_vs + nv;
  hipStream_t s = ctx->stream;
  auto up = [&](double *dst, const double *src, size_t count) {
    return hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, s);
  };
  PP_HIP_CHECK(up(d_rho, rho_mid, nc));
  PP_HIP_CHECK(up(d_pc, p_corr, nc));
  PP_HIP_CHECK(up(d_p, pressure, nc));
  PP_HIP_CHECK(up(d_pp, pressure_prev, nc));
  PP_HIP_CHECK(up(d_us, u_star, nu));
  PP_HIP_CHECK(up(d_u, u, nu));
  PP_HIP_CHECK(up(d_vs, v_star, nv));
  PP_HIP_CHECK(up(d_v, v, nv));
  if (int rc = run_velocity_correction(ctx, pb, vbc, ghost, d_rho, d_us, d_vs, d_pc, d_u, d_v, d_p, d_pp, speeds, s))
    return rc;
  auto down = [&](double *dst, const double *src, size_t count) {
    return hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, s);
  };
  PP_HIP_CHECK(down(u, d_u, nu));
  PP_HIP_CHECK(down(v, d_v, nv));
  PP_HIP_CHECK(down(pressure, d_p, nc));
  PP_HIP_CHECK(down(pressure_prev, d_pp, nc));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_pressure_step(pp_ctx *ctx, const pp_problem *pb, const pp_velocity_bc *vbc, int ghost, const double *rho_mid,
                     const double *u_star, const double *v_star, double *p_corr, double *u, double *v,
                     double *pressure, double *pressure_prev, double *speeds, pp_report *report) {
  if (int rc = validate_correction(ctx, pb, vbc, ghost)) return rc;
  if (!rho_mid || !u_star || !v_star || !p_corr || !u || !v || !pressure || !pressure_prev)
    return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  const int nx = pb->nx, ny = pb->ny, G = ghost;
  const size_t nc = static_cast<size_t>(nx + 2 * G) * (ny + 2 * G);
  const size_t nu = static_cast<size_t>(nx + 1 + 2 * G) * (ny + 2 * G);
  const size_t nv = static_cast<size_t>(nx + 2 * G) * (ny + 1 + 2 * G);
  if (int rc = ensure_fields(ctx, 5 * nc + 2 * nu + 2 * nv)) return rc;
  double *d_rho = ctx->field_stage, *d_pc = d_rho + nc, *d_p = d_pc + nc, *d_pp = d_p + nc;
  double *d_us = d_pp + nc, *d_u = d_us + nu, *d_vs = d_u + nu, *d_v = d_vs + nv;
  hipStream_t s = ctx->stream;
  auto up = [&](double *dst, const double *src, size_t count) {
    return hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, s);
  };
  auto down = [&](double *dst, const double *src, size_t count) {
    return hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, s);
  };
  PP_HIP_CHECK(up(d_rho, rho_mid, nc));
  PP_HIP_CHECK(up(d_p, pressure, nc));
  PP_HIP_CHECK(up(d_pp, pressure_prev, nc));
  PP_HIP_CHECK(up(d_us, u_star, nu));
  PP_HIP_CHECK(up(d_u, u, nu));
  PP_HIP_CHECK(up(d_vs, v_star, nv));
  PP_HIP_CHECK(up(d_v, v, nv));
  const int rc = pp_solve_device_field(ctx, pb, ghost, d_rho, d_us, d_vs, d_p, d_pp, d_pc, report, s);
  if (rc == PP_OK) {
    if (int rc2 = run_velocity_correction(ctx, pb, vbc, ghost, d_rho, d_us, d_vs, d_pc, d_u, d_v, d_p, d_pp, speeds,
                                          s))
      return rc2;
    PP_HIP_CHECK(down(u, d_u, nu));
    PP_HIP_CHECK(down(v, d_v, nv));
    PP_HIP_CHECK(down(pressure, d_p, nc));
    PP_HIP_CHECK(down(pressure_prev, d_pp, nc));
  }
  // pressure_correction_ as the reference leaves it: solved with ghosts, or all zeros after a failure
  PP_HIP_CHECK(down(p_corr, d_pc, nc));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return rc;
}

int pp_create(pp_ctx **out, int device) {
  if (!out) return PP_ERR_ARG;
  *out = nullptr;
  auto *ctx = new pp_ctx();
  ctx->device = device;
  pp_report *report = nullptr;
  PP_HIP_CHECK(hipSetDevice(device));
  PP_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  PP_HIP_CHECK(hipMalloc(&ctx->wrap64, 2 * pp::kMaxSide * sizeof(double)));
  PP_HIP_CHECK(hipMalloc(&ctx->wrap32, 2 * pp::kMaxSide * sizeof(float)));
  PP_HIP_CHECK(hipMalloc(&ctx->d_args, sizeof(pp::SysArgs)));
  PP_HIP_CHECK(hipMalloc(&ctx->d_rep, sizeof(pp::DevReport)));
  PP_HIP_CHECK(hipHostMalloc(&ctx->h_rep, sizeof(pp::DevReport), hipHostMallocDefault));
  PP_HIP_CHECK(hipEventCreate(&ctx->ev_start));
  PP_HIP_CHECK(hipEventCreate(&ctx->ev_stop));
  *out = ctx;
  return PP_OK;
}

void pp_destroy(pp_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->layouts) {
    (void)hipFree(kv.second.base);
    (void)hipFree(kv.second.cell);
  }
  if (ctx->work) (void)hipFree(ctx->work);
  if (ctx->fwork) (void)hipFree(ctx->fwork);
  if (ctx->wrap64) (void)hipFree(ctx->wrap64);
  if (ctx->wrap32) (void)hipFree(ctx->wrap32);
  if (ctx->stage) (void)hipFree(ctx->stage);
  if (ctx->d_args) (void)hipFree(ctx->d_args);
  if (ctx->d_rep) (void)hipFree(ctx->d_rep);
  if (ctx->h_rep) (void)hipHostFree(ctx->h_rep);
  free_batch(ctx);
  if (ctx->d_corr_args) (void)hipFree(ctx->d_corr_args);
  if (ctx->corr_work) (void)hipFree(ctx->corr_work);
  if (ctx->monitor_log) (void)hipFree(ctx->monitor_log);
  if (ctx->field_stage) (void)hipFree(ctx->field_stage);
  if (ctx->bicg_work) (void)hipFree(ctx->bicg_work);
  if (ctx->bicg_partials) (void)hipFree(ctx->bicg_partials);
  if (ctx->bicg_state) (void)hipFree(ctx->bicg_state);
  if (ctx->h_bicg) (void)hipHostFree(ctx->h_bicg);
  if (ctx->bicg_ev[0]) (void)hipEventDestroy(ctx->bicg_ev[0]);
  if (ctx->bicg_ev[1]) (void)hipEventDestroy(ctx->bicg_ev[1]);
  if (ctx->ch.d_int) (void)hipFree(ctx->ch.d_int);
  if (ctx->ch.d_dbl) (void)hipFree(ctx->ch.d_dbl);
  if (ctx->ch.d_state) (void)hipFree(ctx->ch.d_state);
  if (ctx->ch.h_state) (void)hipHostFree(ctx->ch.h_state);
  if (ctx->ch.stage) (void)hipFree(ctx->ch.stage);
  if (ctx->ch.d_part) (void)hipFree(ctx->ch.d_part);
  if (ctx->ch.d_dstate) (void)hipFree(ctx->ch.d_dstate);
  if (ctx->ch.h_dstate) (void)hipHostFree(ctx->ch.h_dstate);
  if (ctx->ch.ev[0]) (void)hipEventDestroy(ctx->ch.ev[0]);
  if (ctx->ch.ev[1]) (void)hipEventDestroy(ctx->ch.ev[1]);
  if (ctx->mom_int) (void)hipFree(ctx->mom_int);
  if (ctx->mom_dbl) (void)hipFree(ctx->mom_dbl);
  if (ctx->mom_stage) (void)hipFree(ctx->mom_stage);
  if (ctx->mom_ev[0]) (void)hipEventDestroy(ctx->mom_ev[0]);
  if (ctx->mom_ev[1]) (void)hipEventDestroy(ctx->mom_ev[1]);
  if (ctx->bicg_stage_idx) (void)hipFree(ctx->bicg_stage_idx);
  if (ctx->bicg_stage_val) (void)hipFree(ctx->bicg_stage_val);
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *pp_last_error(const pp_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

double pp_last_kernel_ms(const pp_ctx *ctx) { return ctx ? ctx->last_kernel_ms : 0.0; }

int pp_solve_device(pp_ctx *ctx, const pp_problem *pb, const double *d_rho, const double *d_u, const double *d_v,
                    const double *d_p_n, const double *d_p_nm1, double *d_p_out, pp_report *report, void *stream) {
  return solve_device(ctx, pb, d_rho, d_u, d_v, d_p_n, d_p_nm1, d_p_out, nullptr, 0, nullptr, report, stream);
}

int pp_solve_device_monitored(pp_ctx *ctx, const pp_problem *pb, const double *d_rho, const double *d_u,
                              const double *d_v, const double *d_p_n, const double *d_p_nm1, double *d_p_out,
                              pp_monitor_record *d_records, int capacity, int *count, pp_report *report,
                              void *stream) {
  if (int rc = validate_monitor(ctx, pb, d_records, capacity, report)) return rc;
  return solve_device(ctx, pb, d_rho, d_u, d_v, d_p_n, d_p_nm1, d_p_out, reinterpret_cast<pp::MonitorRecord *>(d_records),
                      capacity, count, report, stream);
}

int pp_solve_device_field(pp_ctx *ctx, const pp_problem *pb, int ghost, const double *d_rho_mid,
                          const double *d_u_star, const double *d_v_star, const double *d_pressure,
                          const double *d_pressure_prev, double *d_p_corr, pp_report *report, void *stream) {
  if (int rc = validate(ctx, pb, report)) return rc;
  if (ghost < 1 || ghost > 8) return fail(ctx, report, PP_ERR_ARG, "ghost must be 1..8");
  const bool split = pb->scheme == PP_SPLIT_ICPCG || pb->scheme == PP_SPLIT_ILDLT_PCG;
  if (!d_rho_mid && !split) return fail(ctx, report, PP_ERR_ARG, "rho is required");
  if (!d_u_star || !d_v_star || !d_p_corr) return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  const int G = ghost, nx = pb->nx, ny = pb->ny;
  const int lc = ny + 2 * G, lv = ny + 1 + 2 * G;
  // the solve kernel reads contiguous fields: pack the interiors of the Field2D buffers into the
  // context's staging area with strided device-to-device copies, and unpack the result after
  const size_t n = static_cast<size_t>(nx) * ny;
  const size_t need = 4 * n + static_cast<size_t>(nx + 1) * ny + static_cast<size_t>(nx) * (ny + 1);
  if (int rc = ensure_stage(ctx, need, report)) return rc;
  double *d_rho = ctx->stage;
  double *d_u = d_rho + n;
  double *d_v = d_u + static_cast<size_t>(nx + 1) * ny;
  double *d_pn = d_v + static_cast<size_t>(nx) * (ny + 1);
  double *d_pnm1 = d_pn + n;
  double *d_out = d_pnm1 + n;
  auto at00 = [&](const double *f, int ld) { return f + static_cast<size_t>(G) * ld + G; };
  auto pack = [&](double *dst, const double *field, int rows, int cols, int ld) {
    return hipMemcpy2DAsync(dst, static_cast<size_t>(cols) * sizeof(double), at00(field, ld),
                            static_cast<size_t>(ld) * sizeof(double), static_cast<size_t>(cols) * sizeof(double),
                            static_cast<size_t>(rows), hipMemcpyDeviceToDevice, s);
  };
  const bool has_pn = split && d_pressure, has_pnm1 = split && d_pressure_prev;
  if (d_rho_mid) PP_HIP_CHECK(pack(d_rho, d_rho_mid, nx, ny, lc));
  PP_HIP_CHECK(pack(d_u, d_u_star, nx + 1, ny, lc));
  PP_HIP_CHECK(pack(d_v, d_v_star, nx, ny + 1, lv));
  if (has_pn) PP_HIP_CHECK(pack(d_pn, d_pressure, nx, ny, lc));
  if (has_pnm1) PP_HIP_CHECK(pack(d_pnm1, d_pressure_prev, nx, ny, lc));
  pp::SysArgs a;
  if (int rc = prepare(ctx, pb, a, report)) return rc;
  a.rho = d_rho_mid ? d_rho : nullptr;
  a.u = d_u;
  a.v = d_v;
  a.p_n = has_pn ? d_pn : nullptr;
  a.p_nm1 = has_pnm1 ? d_pnm1 : nullptr;
  a.p_out = d_out;
  // the reference zero-fills pressure_correction_ first: the ghost passes read deeper ghost cells
  // of that zero field before rewriting them, and a failed solve leaves zeros
  PP_HIP_CHECK(hipMemsetAsync(d_p_corr, 0, static_cast<size_t>(nx + 2 * G) * lc * sizeof(double), s));
  if (int rc = run_solve(ctx, pb, a, report, s)) return rc;
  double *out00 = d_p_corr + static_cast<size_t>(G) * lc + G;
  PP_HIP_CHECK(hipMemcpy2DAsync(out00, static_cast<size_t>(lc) * sizeof(double), d_out,
                                static_cast<size_t>(ny) * sizeof(double), static_cast<size_t>(ny) * sizeof(double),
                                static_cast<size_t>(nx), hipMemcpyDeviceToDevice, s));
  if (int rc = ensure_corr(ctx)) return rc;
  const pp::CorrArgs c = corr_args(pb, nullptr, G);
  PP_HIP_CHECK(hipMemcpyAsync(ctx->d_corr_args, &c, sizeof(c), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(pp::launch_pressure_ghosts(ctx->d_corr_args, d_p_corr, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_solve(pp_ctx *ctx, const pp_problem *pb, const double *rho, int ld_rho, const double *u, int ld_u,
             const double *v, int ld_v, const double *p_n, int ld_pn, const double *p_nm1, int ld_pnm1,
             double *p_out, int ld_out, pp_report *report) {
  return solve_host(ctx, pb, rho, ld_rho, u, ld_u, v, ld_v, p_n, ld_pn, p_nm1, ld_pnm1, p_out, ld_out, nullptr, 0,
                    nullptr, report);
}

int pp_solve_monitored(pp_ctx *ctx, const pp_problem *pb, const double *rho, int ld_rho, const double *u, int ld_u,
                       const double *v, int ld_v, const double *p_n, int ld_pn, const double *p_nm1, int ld_pnm1,
                       double *p_out, int ld_out, pp_monitor_record *records, int capacity, int *count,
                       pp_report *report) {
  if (int rc = validate_monitor(ctx, pb, records, capacity, report)) return rc;
  return solve_host(ctx, pb, rho, ld_rho, u, ld_u, v, ld_v, p_n, ld_pn, p_nm1, ld_pnm1, p_out, ld_out, records,
                    capacity, count, report);
}

}  // extern "C"

namespace {

int solve_host(pp_ctx *ctx, const pp_problem *pb, const double *rho, int ld_rho, const double *u, int ld_u,
               const double *v, int ld_v, const double *p_n, int ld_pn, const double *p_nm1, int ld_pnm1,
               double *p_out, int ld_out, pp_monitor_record *records, int capacity, int *count, pp_report *report) {
  if (int rc = validate(ctx, pb, report)) return rc;
  const int nx = pb->nx, ny = pb->ny;
  if (!rho || !u || !v || !p_out) return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  if (ld_rho < ny || ld_u < ny || ld_v < ny + 1 || ld_out < ny || (p_n && ld_pn < ny) || (p_nm1 && ld_pnm1 < ny))
    return fail(ctx, report, PP_ERR_ARG, "leading dimension smaller than the row length");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  const size_t n = static_cast<size_t>(nx) * ny;
  const size_t need = 4 * n + static_cast<size_t>(nx + 1) * ny + static_cast<size_t>(nx) * (ny + 1);
  if (int rc = ensure_stage(ctx, need, report)) return rc;
  double *d_rho = ctx->stage;
  double *d_u = d_rho + n;
  double *d_v = d_u + static_cast<size_t>(nx + 1) * ny;
  double *d_pn = d_v + static_cast<size_t>(nx) * (ny + 1);
  double *d_pnm1 = d_pn + n;
  double *d_out = d_pnm1 + n;
  hipStream_t s = ctx->stream;
  auto up = [&](double *dst, const double *src, int ld, int rows, int cols) {
    return hipMemcpy2DAsync(dst, static_cast<size_t>(cols) * sizeof(double), src, static_cast<size_t>(ld) * sizeof(double),
                            static_cast<size_t>(cols) * sizeof(double), static_cast<size_t>(rows), hipMemcpyHostToDevice, s);
  };
  PP_HIP_CHECK(up(d_rho, rho, ld_rho, nx, ny));
  PP_HIP_CHECK(up(d_u, u, ld_u, nx + 1, ny));
  PP_HIP_CHECK(up(d_v, v, ld_v, nx, ny + 1));
  if (p_n) PP_HIP_CHECK(up(d_pn, p_n, ld_pn, nx, ny));
  if (p_nm1) PP_HIP_CHECK(up(d_pnm1, p_nm1, ld_pnm1, nx, ny));
  pp::SysArgs a;
  if (int rc = prepare(ctx, pb, a, report)) return rc;
  a.rho = d_rho;
  a.u = d_u;
  a.v = d_v;
  a.p_n = p_n ? d_pn : nullptr;
  a.p_nm1 = p_nm1 ? d_pnm1 : nullptr;
  a.p_out = d_out;
  if (records && capacity > 0) {
    if (static_cast<size_t>(capacity) > ctx->monitor_cap) {
      if (ctx->monitor_log) (void)hipFree(ctx->monitor_log);
      ctx->monitor_log = nullptr;
      ctx->monitor_cap = 0;
      PP_HIP_CHECK(hipMalloc(&ctx->monitor_log, static_cast<size_t>(capacity) * sizeof(pp::MonitorRecord)));
      ctx->monitor_cap = static_cast<size_t>(capacity);
    }
    a.monitor = ctx->monitor_log;
    a.monitor_cap = capacity;
  }
  const bool monitored = records != nullptr || count != nullptr;
  const int rc = run_solve(ctx, pb, a, report, s, monitored);
  const int calls = monitored ? ctx->h_rep->monitor_calls : 0;
  if (count) *count = calls;
  if (a.monitor && calls > 0)
    PP_HIP_CHECK(hipMemcpyAsync(records, ctx->monitor_log,
                                static_cast<size_t>(std::min(calls, capacity)) * sizeof(pp::MonitorRecord),
                                hipMemcpyDeviceToHost, s));
  if (rc == PP_OK)
    PP_HIP_CHECK(hipMemcpy2DAsync(p_out, static_cast<size_t>(ld_out) * sizeof(double), d_out,
                                  static_cast<size_t>(ny) * sizeof(double), static_cast<size_t>(ny) * sizeof(double),
                                  static_cast<size_t>(nx), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return rc;
}

}  // namespace

extern "C" {

int pp_solve_coeffs_device(pp_ctx *ctx, const pp_problem *pb, const double *d_ku, const double *d_kv,
                           const double *d_src, double *d_p_out, pp_report *report, void *stream) {
  if (int rc = validate_coeffs(ctx, pb, report)) return rc;
  if (!d_ku || !d_kv || !d_src || !d_p_out) return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return run_coeff_solve(ctx, pb, d_ku, d_kv, d_src, 1, d_p_out, report, s);
}

int pp_adjoint_device(pp_ctx *ctx, const pp_problem *pb, const double *d_ku, const double *d_kv, const double *d_p,
                      const double *d_dLdp, double *d_dLdsrc, double *d_dLdku, double *d_dLdkv, pp_report *report,
                      void *stream) {
  if (int rc = validate_coeffs(ctx, pb, report)) return rc;
  if (!d_ku || !d_kv || !d_p || !d_dLdp || !d_dLdsrc || !d_dLdku || !d_dLdkv)
    return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return run_adjoint(ctx, pb, d_ku, d_kv, d_p, d_dLdp, d_dLdsrc, d_dLdku, d_dLdkv, report, s);
}

int pp_solve_coeffs(pp_ctx *ctx, const pp_problem *pb, const double *ku, const double *kv, const double *src,
                    double *p_out, pp_report *report) {
  if (int rc = validate_coeffs(ctx, pb, report)) return rc;
  if (!ku || !kv || !src || !p_out) return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  const size_t n = static_cast<size_t>(pb->nx) * pb->ny;
  const size_t nu = static_cast<size_t>(pb->nx + 1) * pb->ny, nv = static_cast<size_t>(pb->nx) * (pb->ny + 1);
  if (int rc = ensure_stage(ctx, nu + nv + 2 * n, report)) return rc;
  double *d_ku = ctx->stage, *d_kv = d_ku + nu, *d_src = d_kv + nv, *d_out = d_src + n;
  hipStream_t s = ctx->stream;
  PP_HIP_CHECK(hipMemcpyAsync(d_ku, ku, nu * sizeof(double), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_kv, kv, nv * sizeof(double), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_src, src, n * sizeof(double), hipMemcpyHostToDevice, s));
  if (int rc = run_coeff_solve(ctx, pb, d_ku, d_kv, d_src, 1, d_out, report, s)) return rc;
  PP_HIP_CHECK(hipMemcpyAsync(p_out, d_out, n * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_adjoint(pp_ctx *ctx, const pp_problem *pb, const double *ku, const double *kv, const double *p,
               const double *dLdp, double *dLdsrc, double *dLdku, double *dLdkv, pp_report *report) {
  if (int rc = validate_coeffs(ctx, pb, report)) return rc;
  if (!ku || !kv || !p || !dLdp || !dLdsrc || !dLdku || !dLdkv) return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  const size_t n = static_cast<size_t>(pb->nx) * pb->ny;
  const size_t nu = static_cast<size_t>(pb->nx + 1) * pb->ny, nv = static_cast<size_t>(pb->nx) * (pb->ny + 1);
  if (int rc = ensure_stage(ctx, 2 * (nu + nv) + 3 * n, report)) return rc;
  double *d_ku = ctx->stage, *d_kv = d_ku + nu, *d_p = d_kv + nv, *d_g = d_p + n, *d_lam = d_g + n;
  double *d_dku = d_lam + n, *d_dkv = d_dku + nu;
  hipStream_t s = ctx->stream;
  PP_HIP_CHECK(hipMemcpyAsync(d_ku, ku, nu * sizeof(double), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_kv, kv, nv * sizeof(double), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_p, p, n * sizeof(double), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_g, dLdp, n * sizeof(double), hipMemcpyHostToDevice, s));
  if (int rc = run_adjoint(ctx, pb, d_ku, d_kv, d_p, d_g, d_lam, d_dku, d_dkv, report, s)) return rc;
  PP_HIP_CHECK(hipMemcpyAsync(dLdsrc, d_lam, n * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipMemcpyAsync(dLdku, d_dku, nu * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipMemcpyAsync(dLdkv, d_dkv, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return PP_OK;
}

int pp_set_tuning(pp_ctx *ctx, int count, const int *nx, const int *ny, const int *ring_depth) {
  if (!ctx) return PP_ERR_ARG;
  pp_report *report = nullptr;
  if (count < 0 || (count > 0 && (!nx || !ny || !ring_depth))) return fail(ctx, report, PP_ERR_ARG, "bad tuning table");
  std::map<std::pair<int, int>, int> table;
  for (int k = 0; k < count; ++k) {
    if (nx[k] < 1 || ny[k] < 1 || nx[k] > pp::kMaxRows || ny[k] > pp::kMaxColumns ||
        !pp::valid_ring((nx[k] + 63) / 64, ring_depth[k]))
      return fail(ctx, report, PP_ERR_ARG, "invalid tuning entry");
    table[std::make_pair(nx[k], ny[k])] = ring_depth[k];
  }
  ctx->tuning.swap(table);
  return PP_OK;
}

int pp_set_exact_mean(pp_ctx *ctx, int exact) {
  if (!ctx) return PP_ERR_ARG;
  ctx->exact_mean = exact ? 1 : 0;
  return PP_OK;
}

int pp_set_wait_limit(pp_ctx *ctx, int sleeps) {
  if (!ctx) return PP_ERR_ARG;
  pp_report *report = nullptr;
  if (sleeps < 0) return fail(ctx, report, PP_ERR_ARG, "wait limit must be >= 0");
  ctx->wait_limit = sleeps;
  return PP_OK;
}

int pp_solve_batched_device(pp_ctx *ctx, int count, const pp_problem *problems, const double *const *d_rho,
                            const double *const *d_u, const double *const *d_v, const double *const *d_p_n,
                            const double *const *d_p_nm1, double *const *d_p_out, pp_report *reports,
                            void *stream) {
  pp_report *report = nullptr;
  if (!ctx) return PP_ERR_ARG;
  if (count < 0) return fail(ctx, report, PP_ERR_ARG, "negative batch size");
  if (count == 0) return PP_OK;
  if (!problems || !d_rho || !d_u || !d_v || !d_p_out) return fail(ctx, report, PP_ERR_ARG, "null batch array");
  for (int k = 0; k < count; ++k) {
    pp_report *rk = reports ? reports + k : nullptr;
    const pp_problem *pb = problems + k;
    if (int rc = validate(ctx, pb, rk)) return rc;
    const bool split = pb->scheme == PP_SPLIT_ICPCG || pb->scheme == PP_SPLIT_ILDLT_PCG;
    if ((!d_rho[k] && !split) || !d_u[k] || !d_v[k] || !d_p_out[k]) return fail(ctx, rk, PP_ERR_ARG, "null field pointer");
  }
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  size_t cells = 0, fcells = 0;
  std::vector<pp_ctx::Layout *> layouts(static_cast<size_t>(count));
  for (int k = 0; k < count; ++k) {
    if (int rc = ensure_layout(ctx, problems[k].nx, problems[k].ny, &layouts[static_cast<size_t>(k)], report)) return rc;
    const size_t stride = static_cast<size_t>(problems[k].nx) * problems[k].ny + pp::kPad;
    cells += stride;
    if (problems[k].precision == PP_MIXED) fcells += stride;
  }
  if (int rc = ensure_batch(ctx, count, cells, fcells, report)) return rc;

  std::vector<pp::SysArgs> args(static_cast<size_t>(count));
  size_t at = 0, fat = 0;
  for (int k = 0; k < count; ++k) {
    const pp_problem *pb = problems + k;
    pp::SysArgs &a = args[static_cast<size_t>(k)];
    fill_args(ctx, pb, *layouts[static_cast<size_t>(k)], a);
    const size_t stride = static_cast<size_t>(pb->nx) * pb->ny + pp::kPad;
    set_work(a, ctx->batch_work + at * kF64Arrays, stride);
    at += stride;
    a.fP = a.fW = a.fS = a.fLw = a.fLs = a.fRd = nullptr;
    a.fx = a.fr = a.fpa = a.fpb = a.fz = a.fy = nullptr;
    if (pb->precision == PP_MIXED) {
      set_fwork(a, ctx->batch_fwork + fat * kF32Arrays, stride);
      fat += stride;
    }
    a.fLX = a.fLY = nullptr;
    a.LX = ctx->batch_wrap + static_cast<size_t>(k) * 2 * pp::kMaxSide;
    a.LY = a.LX + pp::kMaxSide;
    a.rep = ctx->d_batch_rep + k;
    a.rho = d_rho[k];
    a.u = d_u[k];
    a.v = d_v[k];
    a.p_n = d_p_n ? d_p_n[k] : nullptr;
    a.p_nm1 = d_p_nm1 ? d_p_nm1[k] : nullptr;
    a.p_out = d_p_out[k];
  }
  // Systems grouped into kernel families (FP64 ICC, FP64 ILDLT, mixed ICC, mixed ILDLT), each
  // family largest first; the device copy of the arguments is in that order, reports stay indexed
  // by system. One launch per non-empty family, one workgroup per system.
  auto family = [&](int k) {
    return (problems[k].precision == PP_MIXED ? 2 : 0) + (args[static_cast<size_t>(k)].ildlt ? 1 : 0);
  };
  std::vector<int> order(static_cast<size_t>(count));
  for (int k = 0; k < count; ++k) order[static_cast<size_t>(k)] = k;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const int fx = family(x), fy = family(y);
    if (fx != fy) return fx < fy;
    return batch_cost(problems[x]) > batch_cost(problems[y]);
  });
  std::vector<pp::SysArgs> sorted(static_cast<size_t>(count));
  int family_count[4] = {0, 0, 0, 0};
  for (int k = 0; k < count; ++k) {
    sorted[static_cast<size_t>(k)] = args[static_cast<size_t>(order[static_cast<size_t>(k)])];
    ++family_count[family(order[static_cast<size_t>(k)])];
  }

  PP_HIP_CHECK(hipMemcpyAsync(ctx->d_batch_args, sorted.data(), sorted.size() * sizeof(pp::SysArgs),
                              hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemsetAsync(ctx->d_batch_rep, 0, static_cast<size_t>(count) * sizeof(pp::DevReport), s));
  PP_HIP_CHECK(hipEventRecord(ctx->ev_start, s));
  int first_sys = 0;
  for (int f = 0; f < 4; ++f) {
    if (family_count[f] == 0) continue;
    const pp::SysArgs *d_first = ctx->d_batch_args + first_sys;
    if (f >= 2)
      PP_HIP_CHECK(pp::launch_mixed_solve(d_first, family_count[f], (f & 1) != 0, s));
    else
      PP_HIP_CHECK(pp::launch_fp64_solve(d_first, family_count[f], (f & 1) != 0, s));
    first_sys += family_count[f];
  }
  PP_HIP_CHECK(hipEventRecord(ctx->ev_stop, s));
  PP_HIP_CHECK(hipMemcpyAsync(ctx->h_batch_rep, ctx->d_batch_rep, static_cast<size_t>(count) * sizeof(pp::DevReport),
                              hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  float kernel_ms = 0.0f;
  PP_HIP_CHECK(hipEventElapsedTime(&kernel_ms, ctx->ev_start, ctx->ev_stop));
  ctx->last_kernel_ms = kernel_ms;
  int first = PP_OK;
  for (int k = 0; k < count; ++k) {
    const pp::DevReport &r = ctx->h_batch_rep[k];
    pp_report *rk = reports ? reports + k : nullptr;
    copy_report(r, kernel_ms, rk);
    if (r.status != pp::kOk) {
      const int rc = fail(ctx, rk, r.status, status_text(r.status, problems[k].scheme));
      if (first == PP_OK) first = rc;
    }
  }
  if (first == PP_OK) ctx->last_error.clear();
  return first;
}

}  // extern "C"

// ---- BiCGSTAB for the momentum predictor's system ----------------------------------------------

namespace {

int ensure_bicg(pp_ctx *ctx, size_t n, pp_report *report) {
  if (!ctx->bicg_state) {
    PP_HIP_CHECK(hipMalloc(&ctx->bicg_state, sizeof(pp::BicgState)));
    PP_HIP_CHECK(hipHostMalloc(&ctx->h_bicg, 3 * sizeof(pp::BicgState), hipHostMallocDefault));
    PP_HIP_CHECK(hipEventCreateWithFlags(&ctx->bicg_ev[0], hipEventDisableTiming));
    PP_HIP_CHECK(hipEventCreateWithFlags(&ctx->bicg_ev[1], hipEventDisableTiming));
  }
  if (n <= ctx->bicg_cap) return PP_OK;
  if (ctx->bicg_work) (void)hipFree(ctx->bicg_work);
  if (ctx->bicg_partials) (void)hipFree(ctx->bicg_partials);
  ctx->bicg_work = nullptr;
  ctx->bicg_partials = nullptr;
  ctx->bicg_cap = 0;
  const size_t chunks = (n + pp::kBicgRows - 1) / pp::kBicgRows;
  PP_HIP_CHECK(hipMalloc(&ctx->bicg_work, 12 * n * sizeof(double)));
  PP_HIP_CHECK(hipMalloc(&ctx->bicg_partials, pp::kBicgSlots * chunks * sizeof(double)));
  ctx->bicg_cap = n;
  return PP_OK;
}

// Iterations per queued group. The host reads the state of group g while group g + 1 is queued,
// so a solve that stops at iteration k leaves fewer than two groups of passes that return at once.
constexpr int kBicgQueue = 4;

// solve_preconditioned_bicgstab (ch_sparse_krylov.hpp:839-963). The device passes run the
// reference's recurrence and tests; the host queues iterations and reads the state between
// groups of kBicgQueue.
int run_bicgstab(pp_ctx *ctx, int n, const int *rp, const int *ci, const double *val, const double *rhs, double *x,
                 int max_iterations, double tolerance, bool use_rel, pp_report *report, hipStream_t s) {
  if (int rc = ensure_bicg(ctx, static_cast<size_t>(n), report)) return rc;
  pp::BicgArgs a{};
  a.n = n;
  a.chunks = (n + pp::kBicgRows - 1) / pp::kBicgRows;
  a.use_rel = use_rel ? 1 : 0;
  a.tolerance = tolerance;
  a.row_ptr = rp;
  a.col_idx = ci;
  a.values = val;
  a.rhs = rhs;
  a.x = x;
  double **vec[12] = {&a.r, &a.r_hat, &a.p[0], &a.p[1], &a.v[0], &a.v[1], &a.p_hat, &a.s, &a.s_hat, &a.t,
                      &a.update, &a.dinv};
  for (int k = 0; k < 12; ++k) *vec[k] = ctx->bicg_work + static_cast<size_t>(k) * n;
  double *part = ctx->bicg_partials;
  const size_t c = static_cast<size_t>(a.chunks);
  a.partials = part;
  a.p_setup = part + pp::kSlotSetup * c;
  a.p_v = part + pp::kSlotV * c;
  a.p_b = part + pp::kSlotB * c;
  a.p_x = part + pp::kSlotX * c;
  a.p_exit = part + pp::kSlotExit * c;
  a.presummed = a.chunks > pp::kBicgPresumChunks ? 1 : 0;
  a.state = ctx->bicg_state;
  const int limit = std::max(1, max_iterations);
  PP_HIP_CHECK(hipMemsetAsync(a.state, 0, sizeof(pp::BicgState), s));
  PP_HIP_CHECK(hipEventRecord(ctx->ev_start, s));
  PP_HIP_CHECK(pp::launch_bicg_setup(a, s));
  // Group g (kBicgQueue iterations) is followed by a copy of the state into pinned slot g & 1 and
  // an event. The host waits for group g's copy only after queueing group g + 1, so the device
  // always has the next group queued while the host reads the state.
  pp::BicgState *hs = ctx->h_bicg;  // slots 0, 1: groups; 2: the final state
  const int groups = (limit + kBicgQueue - 1) / kBicgQueue;
  auto queue_group = [&](int g) -> int {
    for (int iter = g * kBicgQueue; iter < std::min(limit, (g + 1) * kBicgQueue); ++iter)
      PP_HIP_CHECK(pp::launch_bicg_iteration(a, iter, s));
    PP_HIP_CHECK(hipMemcpyAsync(&hs[g & 1], a.state, sizeof(pp::BicgState), hipMemcpyDeviceToHost, s));
    PP_HIP_CHECK(hipEventRecord(ctx->bicg_ev[g & 1], s));
    return PP_OK;
  };
  if (int rc = queue_group(0)) return rc;
  bool stopped = false;
  for (int g = 0; g < groups && !stopped; ++g) {
    if (g + 1 < groups)
      if (int rc = queue_group(g + 1)) return rc;
    PP_HIP_CHECK(hipEventSynchronize(ctx->bicg_ev[g & 1]));
    stopped = hs[g & 1].status != pp::kBicgRunning;
  }
  if (!stopped) PP_HIP_CHECK(pp::launch_bicg_report(a, limit, s));
  PP_HIP_CHECK(hipEventRecord(ctx->ev_stop, s));
  pp::BicgState &st = hs[2];
  PP_HIP_CHECK(hipMemcpyAsync(&st, a.state, sizeof(pp::BicgState), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  PP_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  ctx->last_kernel_ms = ms;
  pp_report rep{};
  rep.iterations = st.iterations;
  rep.outer_cycles = 1;
  rep.absolute_residual = st.abs_res;
  rep.relative_residual = st.rel_res;
  rep.iterate_residual = st.iter_res;
  rep.kernel_ms = ms;
  if (report) *report = rep;
  switch (st.status) {
    case pp::kBicgDone: ctx->last_error.clear(); return PP_OK;
    case pp::kBicgNoDiag: return fail(ctx, report, PP_ERR_NODIAG, "sparse matrix assembly failed to produce a diagonal entry");
    case pp::kBicgRho: return fail(ctx, report, PP_ERR_BICG_RHO, "bicgstab breakdown: invalid rho");
    case pp::kBicgRhatV: return fail(ctx, report, PP_ERR_BICG_RHAT_V, "bicgstab breakdown: invalid r_hat^T A M^{-1} p");
    case pp::kBicgTT: return fail(ctx, report, PP_ERR_BICG_TT, "bicgstab breakdown: invalid t^T t");
    case pp::kBicgOmega: return fail(ctx, report, PP_ERR_BICG_OMEGA, "bicgstab breakdown: invalid omega");
    default: return fail(ctx, report, PP_ERR_HIP, "bicgstab: the device state did not reach a stop");
  }
}

int validate_bicg(pp_ctx *ctx, int n, const void *rp, const void *ci, const void *val, const void *rhs, const void *x,
                  pp_report *report) {
  if (!ctx) return fail(ctx, report, PP_ERR_ARG, "null context");
  if (n <= 0) return fail(ctx, report, PP_ERR_ARG, "cannot solve an empty linear system");
  if (!rp || !ci || !val || !rhs || !x) return fail(ctx, report, PP_ERR_ARG, "null CSR, right-hand side or x pointer");
  return PP_OK;
}

}  // namespace

extern "C" {

int pp_bicgstab_device(pp_ctx *ctx, int n, const int *d_row_ptr, const int *d_col_idx, const double *d_values,
                       const double *d_rhs, double *d_x, int max_iterations, double tolerance,
                       int use_relative_tolerance, pp_report *report, void *stream) {
  if (int rc = validate_bicg(ctx, n, d_row_ptr, d_col_idx, d_values, d_rhs, d_x, report)) return rc;
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return run_bicgstab(ctx, n, d_row_ptr, d_col_idx, d_values, d_rhs, d_x, max_iterations, tolerance,
                      use_relative_tolerance != 0, report, s);
}

int pp_bicgstab(pp_ctx *ctx, int n, const int *row_ptr, const int *col_idx, const double *values, const double *rhs,
                double *x, int max_iterations, double tolerance, int use_relative_tolerance, pp_report *report) {
  if (int rc = validate_bicg(ctx, n, row_ptr, col_idx, values, rhs, x, report)) return rc;
  if (row_ptr[0] != 0) return fail(ctx, report, PP_ERR_ARG, "row_ptr must start at 0");
  for (int r = 0; r < n; ++r)
    if (row_ptr[r + 1] < row_ptr[r]) return fail(ctx, report, PP_ERR_ARG, "row_ptr must not decrease");
  for (int pos = 0; pos < row_ptr[n]; ++pos)
    if (col_idx[pos] < 0 || col_idx[pos] >= n) return fail(ctx, report, PP_ERR_ARG, "column index out of range");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  const size_t N = static_cast<size_t>(n), nnz = static_cast<size_t>(row_ptr[n]);
  if (N > ctx->bicg_stage_rows || nnz > ctx->bicg_stage_nnz) {
    if (ctx->bicg_stage_idx) (void)hipFree(ctx->bicg_stage_idx);
    if (ctx->bicg_stage_val) (void)hipFree(ctx->bicg_stage_val);
    ctx->bicg_stage_idx = nullptr;
    ctx->bicg_stage_val = nullptr;
    ctx->bicg_stage_rows = ctx->bicg_stage_nnz = 0;
    PP_HIP_CHECK(hipMalloc(&ctx->bicg_stage_idx, (N + 1 + nnz) * sizeof(int)));
    PP_HIP_CHECK(hipMalloc(&ctx->bicg_stage_val, (nnz + 2 * N) * sizeof(double)));
    ctx->bicg_stage_rows = N;
    ctx->bicg_stage_nnz = nnz;
  }
  hipStream_t s = ctx->stream;
  int *d_rp = ctx->bicg_stage_idx, *d_ci = d_rp + N + 1;
  double *d_val = ctx->bicg_stage_val, *d_rhs = d_val + nnz, *d_x = d_rhs + N;
  PP_HIP_CHECK(hipMemcpyAsync(d_rp, row_ptr, (N + 1) * sizeof(int), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_ci, col_idx, nnz * sizeof(int), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_val, values, nnz * sizeof(double), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_rhs, rhs, N * sizeof(double), hipMemcpyHostToDevice, s));
  PP_HIP_CHECK(hipMemcpyAsync(d_x, x, N * sizeof(double), hipMemcpyHostToDevice, s));
  const int rc = run_bicgstab(ctx, n, d_rp, d_ci, d_val, d_rhs, d_x, max_iterations, tolerance,
                              use_relative_tolerance != 0, report, s);
  const std::string err = ctx->last_error;
  PP_HIP_CHECK(hipMemcpyAsync(x, d_x, N * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  ctx->last_error = err;
  return rc;
}

}  // extern "C"

// ---- pp_solve_device_noalloc: the device solve without any allocation -----------------------------
//
// Every buffer the solve kernels need beyond the caller's inputs and output lives in one file-scope
// device array, `scratch`, of kNoAllocBytes bytes: the argument block, the report, the wavefront
// layout tables (filled by a kernel on each call), the periodic wrap lines and the FP64 and FP32
// work arrays, laid out for at most kNoAllocMaxCells cells. Each GPU the library runs on holds its own
// copy until the process ends. Host code never uses the array's name as an address: the address on
// each device comes from hipGetSymbolAddress, taken once per device on the first call there.

namespace {

constexpr size_t kNoAllocMaxCells = PP_NOALLOC_MAX_CELLS;
constexpr size_t kNoAllocStride = kNoAllocMaxCells + pp::kPad;  // work-array stride, as a context's cap + kPad
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kOffArgs = 0;
constexpr size_t kOffRep = align_up(kOffArgs + sizeof(pp::SysArgs));
constexpr size_t kOffBase = align_up(kOffRep + sizeof(pp::DevReport));
constexpr size_t kBaseEntries = pp::kMaxRows + pp::kMaxSide - 1 + 2 * pp::kLevPad;
constexpr size_t kOffCell = align_up(kOffBase + kBaseEntries * sizeof(int));
constexpr size_t kOffWrap64 = align_up(kOffCell + kNoAllocMaxCells * sizeof(uint32_t));
constexpr size_t kOffWrap32 = align_up(kOffWrap64 + 2 * pp::kMaxSide * sizeof(double));
constexpr size_t kOffWork = align_up(kOffWrap32 + 2 * pp::kMaxSide * sizeof(float));
constexpr size_t kOffFwork = align_up(kOffWork + kF64Arrays * kNoAllocStride * sizeof(double));
constexpr size_t kNoAllocBytes = align_up(kOffFwork + kF32Arrays * kNoAllocStride * sizeof(float));

static_assert(kNoAllocBytes == PP_NOALLOC_SCRATCH_BYTES, "pp_solver.h states the scratch size");

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

__global__ void noalloc_zero_kernel(uint4 *p, size_t words) {
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < words) p[t] = make_uint4(0u, 0u, 0u, 0u);
}

// The device address of `scratch` on the current device, looked up (and the array zeroed) once per
// device, with that device current.
int scratch_address(int device, unsigned char **out, pp_report *report) {
  pp_ctx *ctx = nullptr;
  if (device < 0 || device >= kMaxDevices) return fail(ctx, report, PP_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    PP_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(ctx, report, PP_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    // the work arrays' padding is read by lanes outside the grid: start the whole array at zero once
    const size_t words = kNoAllocBytes / sizeof(uint4);
    hipLaunchKernelGGL(noalloc_zero_kernel, dim3(static_cast<unsigned>((words + 255) / 256)), dim3(256), 0, nullptr,
                       static_cast<uint4 *>(ptr), words);
    PP_HIP_CHECK(hipGetLastError());
    PP_HIP_CHECK(hipDeviceSynchronize());
    g_scratch_ptr[device] = static_cast<unsigned char *>(ptr);
  }
  *out = g_scratch_ptr[device];
  return PP_OK;
}

// Layout tables of an nx x ny grid into the scratch (the values build_layout computes on the host:
// base[d + kLevPad] = cells on levels < d minus the level's first row, 0 on the margins; cell[pos(i, j)]
// = (i << 16) | j), and a zeroed report.
__global__ void noalloc_layout_kernel(int nx, int ny, int *base, uint32_t *cell, pp::DevReport *rep) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nlev = nx + ny - 1, n = nx * ny;
  if (t < nlev + 2 * pp::kLevPad) {
    const int d = t - pp::kLevPad;
    base[t] = (d >= 0 && d < nlev) ? pp::level_offset(d, nx, ny) - pp::level_lo(d, ny) : 0;
  }
  if (t < n) {
    const int i = t / ny, j = t - i * ny;
    cell[pp::grid_pos(i, j, nx, ny)] = (static_cast<uint32_t>(i) << 16) | static_cast<uint32_t>(j);
  }
  if (t == 0) {
    int *r = reinterpret_cast<int *>(rep);
    for (size_t k = 0; k < sizeof(pp::DevReport) / sizeof(int); ++k) r[k] = 0;
  }
}

}  // namespace

extern "C" {

int pp_solve_device_noalloc(const pp_problem *pb, const double *d_rho, const double *d_u, const double *d_v,
                            const double *d_p_n, const double *d_p_nm1, double *d_p_out, pp_report *report,
                            void *stream) {
  pp_ctx *ctx = nullptr;
  if (int rc = validate_problem(ctx, pb, report)) return rc;
  if (static_cast<size_t>(pb->nx) * static_cast<size_t>(pb->ny) > kNoAllocMaxCells)
    return fail(ctx, report, PP_ERR_ARG,
                "grid has more cells than the allocation-free solve supports (nx * ny <= " +
                    std::to_string(kNoAllocMaxCells) + ")");
  if (pb->ny > pp::kMaxSide)  // its level-base table is sized for nx + ny - 1 <= 512 + 1024 - 1 levels
    return fail(ctx, report, PP_ERR_ARG, "the allocation-free solve takes ny 1..1024");
  if (!d_rho && !(pb->scheme == PP_SPLIT_ICPCG || pb->scheme == PP_SPLIT_ILDLT_PCG))
    return fail(ctx, report, PP_ERR_ARG, "rho is required");
  if (!d_u || !d_v || !d_p_out) return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  int device = 0;
  PP_HIP_CHECK(hipGetDevice(&device));
  unsigned char *base = nullptr;
  if (int rc = scratch_address(device, &base, report)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);

  pp::SysArgs a;
  fill_problem_args(pb, a);
  spmv_split(pb->nx, pb->ny, a.spmv_split);
  int *lev = reinterpret_cast<int *>(base + kOffBase);
  uint32_t *cell = reinterpret_cast<uint32_t *>(base + kOffCell);
  auto *d_rep = reinterpret_cast<pp::DevReport *>(base + kOffRep);
  a.lev_base = lev + pp::kLevPad;
  a.cell = cell;
  set_work(a, reinterpret_cast<double *>(base + kOffWork), kNoAllocStride);
  set_fwork(a, reinterpret_cast<float *>(base + kOffFwork), kNoAllocStride);
  double *w64 = reinterpret_cast<double *>(base + kOffWrap64);
  float *w32 = reinterpret_cast<float *>(base + kOffWrap32);
  a.LX = w64;
  a.LY = w64 + pp::kMaxSide;
  a.fLX = w32;
  a.fLY = w32 + pp::kMaxSide;
  a.rep = d_rep;
  a.rho = d_rho;
  a.u = d_u;
  a.v = d_v;
  a.p_n = d_p_n;
  a.p_nm1 = d_p_nm1;
  a.p_out = d_p_out;

  const int cells = pb->nx * pb->ny, entries = pb->nx + pb->ny - 1 + 2 * pp::kLevPad;
  const int threads = 256, blocks = (std::max(cells, entries) + threads - 1) / threads;
  hipLaunchKernelGGL(noalloc_layout_kernel, dim3(blocks), dim3(threads), 0, s, pb->nx, pb->ny, lev, cell, d_rep);
  PP_HIP_CHECK(hipGetLastError());
  auto *d_args = reinterpret_cast<pp::SysArgs *>(base + kOffArgs);
  PP_HIP_CHECK(hipMemcpyAsync(d_args, &a, sizeof(a), hipMemcpyHostToDevice, s));
  if (pb->precision == PP_MIXED)
    PP_HIP_CHECK(pp::launch_mixed_solve(d_args, 1, a.ildlt != 0, s));
  else
    PP_HIP_CHECK(pp::launch_fp64_solve(d_args, 1, a.ildlt != 0, s, false));
  pp::DevReport r;
  PP_HIP_CHECK(hipMemcpyAsync(&r, d_rep, sizeof(r), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  copy_report(r, 0.0, report);
  if (r.status != pp::kOk) return fail(ctx, report, r.status, status_text(r.status, pb->scheme));
  return PP_OK;
}

}  // extern "C"

// ---- momentum predictor (momentum.cpp:55-85, 157-502) ----------------------------------------------

namespace {

int validate_momentum(pp_ctx *ctx, const pp_momentum_problem *mp, const pp_velocity_bc *vbc, int ghost,
                      pp_report *report) {
  if (!ctx) return fail(ctx, report, PP_ERR_ARG, "null context");
  if (!mp || !vbc) return fail(ctx, report, PP_ERR_ARG, "null momentum problem or velocity boundary conditions");
  if (mp->nx < 3 || mp->ny < 3) return fail(ctx, report, PP_ERR_ARG, "momentum predictor needs nx, ny >= 3");
  if (ghost < 1 || ghost > 8) return fail(ctx, report, PP_ERR_ARG, "ghost must be 1..8");
  for (int k = 0; k < 4; ++k)
    if (vbc->u_type[k] < PP_BC_UNSET || vbc->u_type[k] > PP_BC_NEUMANN || vbc->v_type[k] < PP_BC_UNSET ||
        vbc->v_type[k] > PP_BC_NEUMANN)
      return fail(ctx, report, PP_ERR_ARG, "unknown velocity boundary type");
  // explicit velocity BCs on periodic axes are rejected by the reference (operators.cpp:11-34)
  static const char *const kU[4] = {"u_bc_left", "u_bc_right", "u_bc_bottom", "u_bc_top"};
  static const char *const kV[4] = {"v_bc_left", "v_bc_right", "v_bc_bottom", "v_bc_top"};
  for (int axis = 0; axis < 2; ++axis) {
    if (!(axis == 0 ? mp->periodic_x : mp->periodic_y)) continue;
    const int lo = 2 * axis;
    const char *names[4] = {kU[lo], kU[lo + 1], kV[lo], kV[lo + 1]};
    const int types[4] = {vbc->u_type[lo], vbc->u_type[lo + 1], vbc->v_type[lo], vbc->v_type[lo + 1]};
    for (int k = 0; k < 4; ++k)
      if (types[k] != PP_BC_UNSET)
        return fail(ctx, report, PP_ERR_ARG, std::string(names[k]) + " cannot be set when the direction is periodic");
  }
  return PP_OK;
}

int ensure_momentum(pp_ctx *ctx, size_t n, pp_report *report) {
  if (!ctx->mom_ev[0]) PP_HIP_CHECK(hipEventCreate(&ctx->mom_ev[0]));
  if (!ctx->mom_ev[1]) PP_HIP_CHECK(hipEventCreate(&ctx->mom_ev[1]));
  if (n <= ctx->mom_cap) return PP_OK;
  if (ctx->mom_int) PP_HIP_CHECK(hipFree(ctx->mom_int));
  if (ctx->mom_dbl) PP_HIP_CHECK(hipFree(ctx->mom_dbl));
  ctx->mom_int = nullptr;
  ctx->mom_dbl = nullptr;
  ctx->mom_cap = 0;
  const size_t blocks = (n + pp::kMomScanBlock - 1) / pp::kMomScanBlock;
  PP_HIP_CHECK(hipMalloc(&ctx->mom_int, (2 * pp::kMomRow * n + 2 * n + blocks + 1) * sizeof(int)));
  PP_HIP_CHECK(hipMalloc(&ctx->mom_dbl, (2 * pp::kMomRow * n + 2 * n) * sizeof(double)));
  ctx->mom_cap = n;
  return PP_OK;
}

// The whole predictor on device Field2D buffers: system, BiCGSTAB from x0 = u / v, scatter.
int run_momentum(pp_ctx *ctx, const pp_momentum_problem *mp, const pp_velocity_bc *vbc, int G, const double *d_u,
                 const double *d_v, const double *d_rho, const double *d_eta, const double *d_eta_prev,
                 const double *d_sfx, const double *d_sfy, double *d_u_prev, double *d_v_prev, double *d_u_star,
                 double *d_v_star, pp_report *report, hipStream_t s) {
  pp::MomArgs a{};
  a.nx = mp->nx;
  a.ny = mp->ny;
  a.G = G;
  a.dx = mp->dx;
  a.dy = mp->dy;
  a.dt = mp->dt;
  a.alpha = 0.5 * mp->dt / mp->re;
  a.body_force_x = mp->body_force_x;
  a.body_force_y = mp->body_force_y;
  a.periodic_x = mp->periodic_x ? 1 : 0;
  a.periodic_y = mp->periodic_y ? 1 : 0;
  a.step_index = mp->step_index;
  for (int k = 0; k < 4; ++k) {  // effective_u_bc / effective_v_bc (operators.cpp:62-112)
    const bool u_unset = vbc->u_type[k] == PP_BC_UNSET;
    const double u_default = k == 2 ? vbc->bottom_wall_velocity_x : (k == 3 ? vbc->top_wall_velocity_x : 0.0);
    a.u_type[k] = u_unset ? PP_BC_DIRICHLET : vbc->u_type[k];
    a.u_value[k] = u_unset ? u_default : vbc->u_value[k];
    const bool v_unset = vbc->v_type[k] == PP_BC_UNSET;
    a.v_type[k] = v_unset ? PP_BC_DIRICHLET : vbc->v_type[k];
    a.v_value[k] = v_unset ? 0.0 : vbc->v_value[k];
  }
  a.iu0 = a.periodic_x ? 0 : 1;
  a.jv0 = a.periodic_y ? 0 : 1;
  a.niu = a.nx - a.iu0;
  a.njv = a.ny - a.jv0;
  a.nu = a.niu * a.ny;
  a.n = a.nu + a.nx * a.njv;
  const size_t n = static_cast<size_t>(a.n);
  if (int rc = ensure_momentum(ctx, n, report)) return rc;
  const size_t blocks = (n + pp::kMomScanBlock - 1) / pp::kMomScanBlock;
  int *ib = ctx->mom_int;
  a.slot_col = ib;
  a.count = ib + pp::kMomRow * n;
  a.block_sum = a.count + n;
  a.row_ptr = a.block_sum + blocks;
  a.col = a.row_ptr + n + 1;
  double *db = ctx->mom_dbl;
  a.slot_val = db;
  a.val = db + pp::kMomRow * n;
  a.rhs = a.val + pp::kMomRow * n;
  a.x = a.rhs + n;
  a.u = d_u;
  a.v = d_v;
  a.rho = d_rho;
  a.eta = d_eta;
  a.eta_prev = d_eta_prev;
  a.sfx = d_sfx;
  a.sfy = d_sfy;
  a.u_prev = d_u_prev;
  a.v_prev = d_v_prev;
  a.u_star = d_u_star;
  a.v_star = d_v_star;
  PP_HIP_CHECK(hipEventRecord(ctx->mom_ev[0], s));
  PP_HIP_CHECK(pp::launch_momentum_system(a, s));
  const int rc = run_bicgstab(ctx, a.n, a.row_ptr, a.col, a.val, a.rhs, a.x, mp->max_iterations, mp->tolerance, false,
                              report, s);
  if (rc != PP_OK) return rc;  // the reference throws before writing u_star_ / v_star_
  PP_HIP_CHECK(pp::launch_momentum_scatter(a, s));
  PP_HIP_CHECK(hipEventRecord(ctx->mom_ev[1], s));
  PP_HIP_CHECK(hipEventSynchronize(ctx->mom_ev[1]));
  float ms = 0.0f;
  PP_HIP_CHECK(hipEventElapsedTime(&ms, ctx->mom_ev[0], ctx->mom_ev[1]));
  ctx->last_kernel_ms = ms;
  if (report) report->kernel_ms = ms;
  ctx->last_error.clear();
  return PP_OK;
}

}  // namespace

extern "C" {

int pp_momentum_predictor_device(pp_ctx *ctx, const pp_momentum_problem *mp, const pp_velocity_bc *vbc, int ghost,
                                 const double *d_u, const double *d_v, const double *d_rho_mid, const double *d_eta,
                                 const double *d_eta_prev, const double *d_surface_fx_u, const double *d_surface_fy_v,
                                 double *d_u_rhs_prev, double *d_v_rhs_prev, double *d_u_star, double *d_v_star,
                                 pp_report *report, void *stream) {
  if (int rc = validate_momentum(ctx, mp, vbc, ghost, report)) return rc;
  if (!d_u || !d_v || !d_rho_mid || !d_eta || !d_eta_prev || !d_surface_fx_u || !d_surface_fy_v || !d_u_rhs_prev ||
      !d_v_rhs_prev || !d_u_star || !d_v_star)
    return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return run_momentum(ctx, mp, vbc, ghost, d_u, d_v, d_rho_mid, d_eta, d_eta_prev, d_surface_fx_u, d_surface_fy_v,
                      d_u_rhs_prev, d_v_rhs_prev, d_u_star, d_v_star, report, s);
}

int pp_momentum_predictor(pp_ctx *ctx, const pp_momentum_problem *mp, const pp_velocity_bc *vbc, int ghost,
                          const double *u, const double *v, const double *rho_mid, const double *eta,
                          const double *eta_prev, const double *surface_fx_u, const double *surface_fy_v,
                          double *u_rhs_prev, double *v_rhs_prev, double *u_star, double *v_star, pp_report *report) {
  if (int rc = validate_momentum(ctx, mp, vbc, ghost, report)) return rc;
  if (!u || !v || !rho_mid || !eta || !eta_prev || !surface_fx_u || !surface_fy_v || !u_rhs_prev || !v_rhs_prev ||
      !u_star || !v_star)
    return fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_HIP_CHECK(hipSetDevice(ctx->device));
  const int G = ghost;
  const size_t nc = static_cast<size_t>(mp->nx + 2 * G) * (mp->ny + 2 * G);
  const size_t nu = static_cast<size_t>(mp->nx + 1 + 2 * G) * (mp->ny + 2 * G);
  const size_t nv = static_cast<size_t>(mp->nx + 2 * G) * (mp->ny + 1 + 2 * G);
  const size_t need = 3 * nc + 4 * nu + 4 * nv;
  if (need > ctx->mom_stage_cap) {
    if (ctx->mom_stage) PP_HIP_CHECK(hipFree(ctx->mom_stage));
    ctx->mom_stage = nullptr;
    ctx->mom_stage_cap = 0;
    PP_HIP_CHECK(hipMalloc(&ctx->mom_stage, need * sizeof(double)));
    ctx->mom_stage_cap = need;
  }
  hipStream_t s = ctx->stream;
  double *p = ctx->mom_stage;
  double *d_rho = p, *d_eta = p + nc, *d_eta_prev = p + 2 * nc;
  double *d_u = p + 3 * nc, *d_sfx = d_u + nu, *d_u_prev = d_sfx + nu, *d_u_star = d_u_prev + nu;
  double *d_v = d_u_star + nu, *d_sfy = d_v + nv, *d_v_prev = d_sfy + nv, *d_v_star = d_v_prev + nv;
  const std::pair<double *, const double *> up[] = {{d_rho, rho_mid}, {d_eta, eta}, {d_eta_prev, eta_prev}};
  for (const auto &f : up) PP_HIP_CHECK(hipMemcpyAsync(f.first, f.second, nc * sizeof(double), hipMemcpyHostToDevice, s));
  const std::pair<double *, const double *> uu[] = {{d_u, u}, {d_sfx, surface_fx_u}, {d_u_prev, u_rhs_prev}};
  for (const auto &f : uu) PP_HIP_CHECK(hipMemcpyAsync(f.first, f.second, nu * sizeof(double), hipMemcpyHostToDevice, s));
  const std::pair<double *, const double *> vv[] = {{d_v, v}, {d_sfy, surface_fy_v}, {d_v_prev, v_rhs_prev}};
  for (const auto &f : vv) PP_HIP_CHECK(hipMemcpyAsync(f.first, f.second, nv * sizeof(double), hipMemcpyHostToDevice, s));
  const int rc = run_momentum(ctx, mp, vbc, G, d_u, d_v, d_rho, d_eta, d_eta_prev, d_sfx, d_sfy, d_u_prev, d_v_prev,
                              d_u_star, d_v_star, report, s);
  if (rc != PP_OK) {
    (void)hipStreamSynchronize(s);
    return rc;
  }
  PP_HIP_CHECK(hipMemcpyAsync(u_rhs_prev, d_u_prev, nu * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipMemcpyAsync(v_rhs_prev, d_v_prev, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipMemcpyAsync(u_star, d_u_star, nu * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipMemcpyAsync(v_star, d_v_star, nv * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_HIP_CHECK(hipStreamSynchronize(s));
  return PP_OK;
}

}  // extern "C"

// ---- Cahn-Hilliard linear solve (phase_field.cpp:437-503, ch_sparse_krylov.hpp:135-233) ------------

namespace {

// One CSR row built as std::map rows are: each column's value is the sum of its contributions in
// insertion order starting from 0.0; emitted in ascending columns.
struct MapRow {
  std::vector<std::pair<int, double>> e;
  void add(int col, double value) {
    for (auto &x : e)
      if (x.first == col) {
        x.second += value;
        return;
      }
    e.push_back({col, 0.0 + value});
  }
  void sort() { std::sort(e.begin(), e.end()); }
};

struct HostCsr {
  std::vector<int> rp{0}, ci, diag;
  std::vector<double> v;
  bool push(MapRow &row, int r) {  // finalize_row_maps for one row (drop |v| < 1e-18, find the diagonal)
    row.sort();
    int d = -1;
    for (const auto &x : row.e) {
      if (std::abs(x.second) < 1.0e-18) continue;
      if (x.first == r) d = static_cast<int>(v.size());
      ci.push_back(x.first);
      v.push_back(x.second);
    }
    diag.push_back(d);
    rp.push_back(static_cast<int>(v.size()));
    return d >= 0;
  }
};

// build_laplacian_matrix, multiply and build_cahn_hilliard_operator (ch_sparse_krylov.hpp:135-233):
// A = alpha0 I - beta1 L + beta2 L L with L the 5-point Laplacian of the grid.
bool ch_assemble(int nx, int ny, double dx, double dy, bool px, bool py, double alpha0, double beta1, double beta2,
                 HostCsr &A) {
  const double ix = 1.0 / (dx * dx), iy = 1.0 / (dy * dy);
  HostCsr L;
  bool ok = true;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      MapRow row;
      double diag = 0.0;
      if (px || i > 0) {
        row.add((px && i == 0 ? nx - 1 : i - 1) * ny + j, ix);
        diag -= ix;
      }
      if (px || i < nx - 1) {
        row.add((px && i == nx - 1 ? 0 : i + 1) * ny + j, ix);
        diag -= ix;
      }
      if (py || j > 0) {
        row.add(i * ny + (py && j == 0 ? ny - 1 : j - 1), iy);
        diag -= iy;
      }
      if (py || j < ny - 1) {
        row.add(i * ny + (py && j == ny - 1 ? 0 : j + 1), iy);
        diag -= iy;
      }
      row.add(i * ny + j, diag);
      ok = L.push(row, i * ny + j) && ok;
    }
  const int n = nx * ny;
  for (int r = 0; r < n; ++r) {
    MapRow l2;
    for (int pa = L.rp[r]; pa < L.rp[r + 1]; ++pa) {
      const int mid = L.ci[pa];
      for (int pb = L.rp[mid]; pb < L.rp[mid + 1]; ++pb) l2.add(L.ci[pb], L.v[pa] * L.v[pb]);
    }
    l2.sort();
    MapRow row;
    row.add(r, alpha0);
    for (int p = L.rp[r]; p < L.rp[r + 1]; ++p) row.add(L.ci[p], -beta1 * L.v[p]);
    for (const auto &x : l2.e)
      if (!(std::abs(x.second) < 1.0e-18)) row.add(x.first, beta2 * x.second);
    ok = A.push(row, r) && ok;
  }
  return ok;
}

int validate_ch(pp_ctx *ctx, const pp_ch_problem *cp, pp_ch_report *report) {
  pp_report *rep = nullptr;
  (void)rep;
  auto bad = [&](const char *msg) {
    if (ctx) ctx->last_error = msg;
    if (report) {
      report->status = PP_ERR_ARG;
      std::snprintf(report->message, sizeof(report->message), "%s", msg);
    }
    return PP_ERR_ARG;
  };
  if (!ctx) return bad("null context");
  if (!cp) return bad("null problem");
  if (cp->nx < 1 || cp->ny < 1) return bad("cannot solve an empty linear system");
  return PP_OK;
}

int ch_fail(pp_ctx *ctx, pp_ch_report *report, int status, const std::string &msg) {
  ctx->last_error = msg;
  if (report) {
    report->status = status;
    std::snprintf(report->message, sizeof(report->message), "%s", msg.c_str());
  }
  return status;
}

#define PP_CH_CHECK(expr)                                                                            \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess) return ch_fail(ctx, report, PP_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// The operator, its level schedule, lower-transpose lists and factor, rebuilt when the grid or the
// coefficients change; alpha0 within 1e-14 max(1, |alpha0|) of the cached one reuses the cache, as
// ensure_ch_linear_system does (phase_field.cpp:121-135).
int ensure_ch(pp_ctx *ctx, const pp_ch_problem *cp, pp_ch_report *report, hipStream_t s) {
  auto &C = ctx->ch;
  const bool same = C.valid && C.nx == cp->nx && C.ny == cp->ny && C.px == (cp->periodic_x != 0) &&
                    C.py == (cp->periodic_y != 0) && C.dx == cp->dx && C.dy == cp->dy && C.beta1 == cp->beta1 &&
                    C.beta2 == cp->beta2 &&
                    std::abs(cp->alpha0 - C.alpha0) <= 1.0e-14 * std::max(1.0, std::abs(cp->alpha0));
  if (same) return PP_OK;
  C.valid = false;
  HostCsr A;
  if (!ch_assemble(cp->nx, cp->ny, cp->dx, cp->dy, cp->periodic_x != 0, cp->periodic_y != 0, cp->alpha0, cp->beta1,
                   cp->beta2, A))
    return ch_fail(ctx, report, PP_ERR_NODIAG, "sparse matrix assembly failed to produce a diagonal entry");
  const int n = cp->nx * cp->ny, nnz = static_cast<int>(A.v.size());
  // levels: 1 + the largest level of the row's lower columns
  std::vector<int> lvl(static_cast<size_t>(n), 0);
  int nlev = 0;
  for (int r = 0; r < n; ++r) {
    int l = 0;
    for (int p = A.rp[r]; p < A.rp[r + 1] && A.ci[p] < r; ++p) l = std::max(l, lvl[static_cast<size_t>(A.ci[p])] + 1);
    lvl[static_cast<size_t>(r)] = l;
    nlev = std::max(nlev, l + 1);
  }
  std::vector<int> lev_ptr(static_cast<size_t>(nlev) + 1, 0), lev_rows(static_cast<size_t>(n));
  for (int r = 0; r < n; ++r) ++lev_ptr[static_cast<size_t>(lvl[static_cast<size_t>(r)]) + 1];
  for (int k = 0; k < nlev; ++k) lev_ptr[static_cast<size_t>(k) + 1] += lev_ptr[static_cast<size_t>(k)];
  {
    std::vector<int> at(lev_ptr.begin(), lev_ptr.end() - 1);
    for (int r = 0; r < n; ++r) lev_rows[static_cast<size_t>(at[static_cast<size_t>(lvl[static_cast<size_t>(r)])]++)] = r;
  }
  // children of each row in ascending order (lower_transpose, ch_sparse_krylov.hpp:322-332)
  std::vector<int> lt_ptr(static_cast<size_t>(n) + 1, 0), lt_row, lt_pos;
  for (int r = 0; r < n; ++r)
    for (int p = A.rp[r]; p < A.rp[r + 1] && A.ci[p] < r; ++p) ++lt_ptr[static_cast<size_t>(A.ci[p]) + 1];
  for (int r = 0; r < n; ++r) lt_ptr[static_cast<size_t>(r) + 1] += lt_ptr[static_cast<size_t>(r)];
  lt_row.resize(static_cast<size_t>(lt_ptr[static_cast<size_t>(n)]));
  lt_pos.resize(lt_row.size());
  {
    std::vector<int> at(lt_ptr.begin(), lt_ptr.end() - 1);
    for (int r = 0; r < n; ++r)
      for (int p = A.rp[r]; p < A.rp[r + 1] && A.ci[p] < r; ++p) {
        const int k = at[static_cast<size_t>(A.ci[p])]++;
        lt_row[static_cast<size_t>(k)] = r;
        lt_pos[static_cast<size_t>(k)] = p;
      }
  }
  // Batched sequential sweeps when the level schedule is close to a chain (periodic y makes row
  // (i, 0) depend on (i-1, ny-1)) and the vectors fit in LDS: the largest batch B <= 64 for which every
  // row's in-batch lower entries are L(r, r-2), L(r, r-1), its in-batch children are r+1, r+2, and
  // at most kChSeqWidth entries of each list lie outside the batch. The operands are packed per row.
  constexpr size_t W = pp::kChSeqWidth;
  int seq_batch = 0;
  std::vector<int> fo_cnt, fo_col, fo_pos, fin_pos, bf_cnt, bf_row, bf_pos, bin_pos;
  if (n <= pp::kChSeqMaxRows && nlev > n / 16) {
    for (int B = 64; B >= 8; --B) {
      const size_t nn = static_cast<size_t>(n);
      fo_cnt.assign(nn, 0);
      fo_col.assign(nn * W, 0);
      fo_pos.assign(nn * W, -1);
      fin_pos.assign(2 * nn, -1);
      bf_cnt.assign(nn, 0);
      bf_row.assign(nn * W, 0);
      bf_pos.assign(nn * W, -1);
      bin_pos.assign(2 * nn, -1);
      bool ok = true;
      for (int r = 0; r < n && ok; ++r) {
        const size_t ru = static_cast<size_t>(r);
        const int b0 = (r / B) * B, hi = std::min(b0 + B, n) - 1;
        int older = 0;
        for (int p = A.rp[r]; p < A.rp[r + 1] && A.ci[p] < r; ++p) {
          const int c = A.ci[p];
          if (c < b0) {  // before the batch: packed in CSR order (they precede the in-batch ones)
            if (older >= static_cast<int>(W)) {
              ok = false;
              break;
            }
            fo_col[ru * W + static_cast<size_t>(older)] = c;
            fo_pos[ru * W + static_cast<size_t>(older)] = p;
            ++older;
          } else if (c == r - 2 && fin_pos[2 * ru] < 0 && fin_pos[2 * ru + 1] < 0) {
            fin_pos[2 * ru] = p;
          } else if (c == r - 1 && fin_pos[2 * ru + 1] < 0) {
            fin_pos[2 * ru + 1] = p;
          } else {
            ok = false;
          }
        }
        fo_cnt[ru] = older;
        int t = lt_ptr[ru];
        const int t1 = lt_ptr[ru + 1];
        if (t < t1 && lt_row[static_cast<size_t>(t)] == r + 1 && r + 1 <= hi) bin_pos[2 * ru] = lt_pos[static_cast<size_t>(t++)];
        if (t < t1 && lt_row[static_cast<size_t>(t)] == r + 2 && r + 2 <= hi) bin_pos[2 * ru + 1] = lt_pos[static_cast<size_t>(t++)];
        if (t < t1 && lt_row[static_cast<size_t>(t)] <= hi) ok = false;  // another in-batch child
        if (t1 - t > static_cast<int>(W)) ok = false;
        for (int e = 0; ok && t + e < t1; ++e) {
          bf_row[ru * W + static_cast<size_t>(e)] = lt_row[static_cast<size_t>(t + e)];
          bf_pos[ru * W + static_cast<size_t>(e)] = lt_pos[static_cast<size_t>(t + e)];
        }
        bf_cnt[ru] = t1 - t;
      }
      if (ok) {
        seq_batch = B;
        break;
      }
    }
  }
  if (!seq_batch)
    for (auto *v : {&fo_cnt, &fo_col, &fo_pos, &fin_pos, &bf_cnt, &bf_row, &bf_pos, &bin_pos}) v->clear();
  // The batched factorization's cross-term positions: L(c, k) for each pair of a row's entries before
  // its batch (find_lower_entry of the factor loop), and where L(r-2, k), L(r-1, k) sit among the
  // entries of rows r-2, r-1 before the same batch.
  std::vector<int> fx_pos, fab;
  if (seq_batch) {
    const size_t nn = static_cast<size_t>(n);
    fx_pos.assign(nn * W * W, -1);
    fab.assign(nn * W, 0);
    auto lower_pos = [&](int row, int target) {
      for (int p = A.rp[row]; p < A.rp[row + 1] && A.ci[p] < row && A.ci[p] <= target; ++p)
        if (A.ci[p] == target) return p;
      return -1;
    };
    auto far_index = [&](int row, int k) {
      for (int e = 0; e < fo_cnt[static_cast<size_t>(row)]; ++e)
        if (fo_col[static_cast<size_t>(row) * W + static_cast<size_t>(e)] == k) return e;
      return -1;
    };
    for (int r = 0; r < n; ++r) {
      const size_t base = static_cast<size_t>(r) * W;
      const int fc = fo_cnt[static_cast<size_t>(r)];
      for (int e = 0; e < fc; ++e)
        for (int e2 = 0; e2 < e; ++e2)
          fx_pos[(base + static_cast<size_t>(e)) * W + static_cast<size_t>(e2)] =
              lower_pos(fo_col[base + static_cast<size_t>(e)], fo_col[base + static_cast<size_t>(e2)]);
      for (int e2 = 0; e2 < fc; ++e2) {
        const int k = fo_col[base + static_cast<size_t>(e2)];
        const int i2 = fin_pos[2 * static_cast<size_t>(r)] >= 0 ? far_index(r - 2, k) : -1;
        const int i1 = fin_pos[2 * static_cast<size_t>(r) + 1] >= 0 ? far_index(r - 1, k) : -1;
        fab[base + static_cast<size_t>(e2)] = (i2 + 1) | ((i1 + 1) << 8);
      }
    }
  }
  // device buffers
  if (C.d_int) (void)hipFree(C.d_int);
  if (C.d_dbl) (void)hipFree(C.d_dbl);
  if (C.d_part) (void)hipFree(C.d_part);
  C.d_int = nullptr;
  C.d_dbl = nullptr;
  C.d_part = nullptr;
  const size_t nl = lt_row.size();
  const size_t seq_ints = fo_cnt.size() + fo_col.size() + fo_pos.size() + fin_pos.size() + bf_cnt.size() +
                          bf_row.size() + bf_pos.size() + bin_pos.size() + fx_pos.size() + fab.size();
  const size_t ints = (n + 1) + nnz + n + (nlev + 1) + n + (n + 1) + 2 * nl + seq_ints;
  const size_t seq_dbls = seq_batch ? (2 * W + 4) * static_cast<size_t>(n) : 0;  // fo_val, bf_val, fin_val, bin_val
  const size_t dbls = 2 * static_cast<size_t>(nnz) + 7 * static_cast<size_t>(n) + seq_dbls;
  PP_CH_CHECK(hipMalloc(&C.d_int, ints * sizeof(int)));
  PP_CH_CHECK(hipMalloc(&C.d_dbl, dbls * sizeof(double)));
  if (!C.d_state) PP_CH_CHECK(hipMalloc(&C.d_state, sizeof(pp::ChState)));
  if (!C.h_state) PP_CH_CHECK(hipHostMalloc(&C.h_state, sizeof(pp::ChState), hipHostMallocDefault));
  const size_t chunks = (static_cast<size_t>(n) + pp::kChRows - 1) / pp::kChRows;
  PP_CH_CHECK(hipMalloc(&C.d_part, pp::kChSlots * chunks * sizeof(double)));
  if (!C.d_dstate) PP_CH_CHECK(hipMalloc(&C.d_dstate, sizeof(pp::ChDiagState)));
  if (!C.h_dstate) PP_CH_CHECK(hipHostMalloc(&C.h_dstate, 3 * sizeof(pp::ChDiagState), hipHostMallocDefault));
  for (hipEvent_t &e : C.ev)
    if (!e) PP_CH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int *ip = C.d_int;
  double *dp = C.d_dbl;
  pp::ChArgs &a = C.args;
  a = pp::ChArgs{};
  a.n = n;
  a.nlev = nlev;
  auto put_int = [&](const std::vector<int> &h, const int *&dst) -> int {
    PP_CH_CHECK(hipMemcpyAsync(ip, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, s));
    dst = ip;
    ip += h.size();
    return PP_OK;
  };
  if (int rc = put_int(A.rp, a.row_ptr)) return rc;
  if (int rc = put_int(A.ci, a.col)) return rc;
  if (int rc = put_int(A.diag, a.diag_pos)) return rc;
  if (int rc = put_int(lev_ptr, a.lev_ptr)) return rc;
  if (int rc = put_int(lev_rows, a.lev_rows)) return rc;
  if (int rc = put_int(lt_ptr, a.lt_ptr)) return rc;
  if (int rc = put_int(lt_row, a.lt_row)) return rc;
  if (int rc = put_int(lt_pos, a.lt_pos)) return rc;
  a.seq_batch = seq_batch;
  if (seq_batch) {
    if (int rc = put_int(fo_cnt, a.fo_cnt)) return rc;
    if (int rc = put_int(fo_col, a.fo_col)) return rc;
    if (int rc = put_int(fo_pos, a.fo_pos)) return rc;
    if (int rc = put_int(fin_pos, a.fin_pos)) return rc;
    if (int rc = put_int(bf_cnt, a.bf_cnt)) return rc;
    if (int rc = put_int(bf_row, a.bf_row)) return rc;
    if (int rc = put_int(bf_pos, a.bf_pos)) return rc;
    if (int rc = put_int(bin_pos, a.bin_pos)) return rc;
    if (int rc = put_int(fx_pos, a.fx_pos)) return rc;
    if (int rc = put_int(fab, a.fab)) return rc;
  }
  PP_CH_CHECK(hipMemcpyAsync(dp, A.v.data(), A.v.size() * sizeof(double), hipMemcpyHostToDevice, s));
  a.val = dp;
  a.lower = dp + nnz;
  a.D = a.lower + nnz;
  a.dinv = a.D + n;
  a.r = a.dinv + n;
  a.z = a.r + n;
  a.p = a.z + n;
  a.ap = a.p + n;
  a.y = a.ap + n;
  if (seq_batch) {
    a.fo_val = a.y + n;
    a.bf_val = a.fo_val + W * static_cast<size_t>(n);
    a.fin_val = a.bf_val + W * static_cast<size_t>(n);
    a.bin_val = a.fin_val + 2 * static_cast<size_t>(n);
  }
  a.state = C.d_state;
  PP_CH_CHECK(pp::launch_ch_factor(a, s));
  PP_CH_CHECK(hipMemcpyAsync(C.h_state, C.d_state, sizeof(pp::ChState), hipMemcpyDeviceToHost, s));
  PP_CH_CHECK(hipStreamSynchronize(s));
  C.kind = C.h_state->ildlt_ok ? PP_CH_ILDLT : PP_CH_DIAGONAL;
  if (seq_batch && C.kind == PP_CH_ILDLT) {
    PP_CH_CHECK(pp::launch_ch_pack(a, s));
    PP_CH_CHECK(hipStreamSynchronize(s));
  }
  C.nx = cp->nx;
  C.ny = cp->ny;
  C.px = cp->periodic_x != 0;
  C.py = cp->periodic_y != 0;
  C.dx = cp->dx;
  C.dy = cp->dy;
  C.alpha0 = cp->alpha0;
  C.beta1 = cp->beta1;
  C.beta2 = cp->beta2;
  C.n = n;
  C.nnz = nnz;
  C.nlev = nlev;
  C.valid = true;
  return PP_OK;
}

const char *ch_status_text(int status) {
  switch (status) {
    case pp::kChRz0Fail: return "pcg preconditioner breakdown: non-positive r^T M^{-1} r";
    case pp::kChRzFail: return "pcg preconditioner breakdown: non-positive new r^T M^{-1} r";
    case pp::kChPapFail: return "pcg breakdown: p^T A p is invalid";
    default: return "unknown device status";
  }
}

int ch_status_code(int status) {
  switch (status) {
    case pp::kChRz0Fail: return PP_ERR_RZ0;
    case pp::kChRzFail: return PP_ERR_RZ;
    case pp::kChPapFail: return PP_ERR_PAP;
    default: return PP_ERR_HIP;
  }
}

constexpr int kChQueue = 8;  // CG iterations queued between the host's state checks

// The diagonal-preconditioned CG as full-chip passes from x = d_c (solve_preconditioned_cg with the
// diagonal preconditioner): the passes run the reference's recurrence and tests on the device; the
// host queues iterations in groups and reads the state between them, with the next group queued.
// Leaves the outcome in C.h_dstate[2].
int run_ch_diag(pp_ctx *ctx, const pp_ch_problem *cp, const double *d_rhs, double *d_c, pp_ch_report *report,
                hipStream_t s) {
  auto &C = ctx->ch;
  pp::ChDiagArgs a{};
  a.n = C.n;
  a.chunks = (C.n + pp::kChRows - 1) / pp::kChRows;
  a.presummed = a.chunks > pp::kChPresumChunks ? 1 : 0;
  a.max_iterations = cp->max_iterations;
  a.tolerance = cp->tolerance;
  a.row_ptr = C.args.row_ptr;
  a.col = C.args.col;
  a.val = C.args.val;
  a.dinv = C.args.dinv;
  a.rhs = d_rhs;
  a.x = d_c;
  a.r = C.args.r;
  a.z = C.args.z;
  a.p[0] = C.args.p;
  a.p[1] = C.args.y;  // y is the ILDLT sweeps' vector; the diagonal CG uses it as p of odd iterations
  a.ap = C.args.ap;
  a.partials = C.d_part;
  a.state = C.d_dstate;
  const int limit = std::max(1, cp->max_iterations);
  PP_CH_CHECK(hipMemsetAsync(a.state, 0, sizeof(pp::ChDiagState), s));
  PP_CH_CHECK(pp::launch_ch_diag_setup(a, s));
  pp::ChDiagState *hs = C.h_dstate;
  const int groups = (limit + kChQueue - 1) / kChQueue;
  auto queue_group = [&](int g) -> int {
    for (int k = g * kChQueue; k < std::min(limit, (g + 1) * kChQueue); ++k)
      PP_CH_CHECK(pp::launch_ch_diag_iteration(a, k, s));
    PP_CH_CHECK(hipMemcpyAsync(&hs[g & 1], a.state, sizeof(pp::ChDiagState), hipMemcpyDeviceToHost, s));
    PP_CH_CHECK(hipEventRecord(C.ev[g & 1], s));
    return PP_OK;
  };
  if (int rc = queue_group(0)) return rc;
  bool stopped = false;
  for (int g = 0; g < groups && !stopped; ++g) {
    if (g + 1 < groups)
      if (int rc = queue_group(g + 1)) return rc;
    PP_CH_CHECK(hipEventSynchronize(C.ev[g & 1]));
    stopped = hs[g & 1].stopped != 0;
  }
  if (!stopped) PP_CH_CHECK(pp::launch_ch_diag_report(a, limit, s));
  PP_CH_CHECK(pp::launch_ch_diag_finish(a, s));
  PP_CH_CHECK(hipMemcpyAsync(&hs[2], a.state, sizeof(pp::ChDiagState), hipMemcpyDeviceToHost, s));
  PP_CH_CHECK(hipStreamSynchronize(s));
  return PP_OK;
}

int run_ch(pp_ctx *ctx, const pp_ch_problem *cp, const double *d_rhs, double *d_c, pp_ch_report *report,
           hipStream_t s) {
  if (int rc = ensure_ch(ctx, cp, report, s)) return rc;
  auto &C = ctx->ch;
  pp::ChArgs a = C.args;
  a.rhs = d_rhs;
  a.x = d_c;
  a.max_iterations = cp->max_iterations;
  a.tolerance = cp->tolerance;
  PP_CH_CHECK(hipEventRecord(ctx->ev_start, s));
  // With ILDLT(0) the one-workgroup PCG; a breakdown there switches the operator to the diagonal
  // preconditioner for good, and the diagonal CG runs again from the x the failed attempt left
  // (phase_field.cpp:465-474). The diagonal CG runs as full-chip passes.
  bool fell_back = false;
  int status = pp::kChOk, iterations = 0;
  double abs_res = 0.0, rel_res = 0.0, iter_res = 0.0, eq_res = 0.0;
  if (C.kind == PP_CH_ILDLT) {
    PP_CH_CHECK(pp::launch_ch_pcg(a, true, s));
    PP_CH_CHECK(hipMemcpyAsync(C.h_state, C.d_state, sizeof(pp::ChState), hipMemcpyDeviceToHost, s));
    PP_CH_CHECK(hipStreamSynchronize(s));
    const pp::ChState &st = *C.h_state;
    status = st.status;
    iterations = st.iterations;
    abs_res = st.abs_res;
    rel_res = st.rel_res;
    iter_res = st.iter_res;
    eq_res = st.eq_res;
    if (status != pp::kChOk) {
      C.kind = PP_CH_DIAGONAL;  // later solves report SparsePCG[Diagonal], as the reference's cached one
      fell_back = true;
    }
  }
  if (C.kind != PP_CH_ILDLT) {
    if (int rc = run_ch_diag(ctx, cp, d_rhs, d_c, report, s)) return rc;
    const pp::ChDiagState &st = C.h_dstate[2];
    status = st.status;
    iterations = st.iterations;
    abs_res = st.abs_res;
    rel_res = st.rel_res;
    iter_res = st.iter_res;
    eq_res = st.eq_res;
  }
  PP_CH_CHECK(hipEventRecord(ctx->ev_stop, s));
  PP_CH_CHECK(hipEventSynchronize(ctx->ev_stop));
  float ms = 0.0f;
  PP_CH_CHECK(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  ctx->last_kernel_ms = ms;
  if (report) {
    std::memset(report, 0, sizeof(*report));
    report->iterations = iterations;
    report->preconditioner = fell_back ? PP_CH_DIAGONAL_FALLBACK : C.kind;
    report->iterate_residual = iter_res;
    report->absolute_residual = abs_res;
    report->relative_residual = rel_res;
    report->equation_residual = eq_res;
    report->kernel_ms = ms;
    report->levels = C.nlev;
  }
  if (status != pp::kChOk) return ch_fail(ctx, report, ch_status_code(status), ch_status_text(status));
  ctx->last_error.clear();
  return PP_OK;
}

}  // namespace

extern "C" {

int pp_ch_solve_device(pp_ctx *ctx, const pp_ch_problem *cp, const double *d_rhs, double *d_c, pp_ch_report *report,
                       void *stream) {
  if (int rc = validate_ch(ctx, cp, report)) return rc;
  if (!d_rhs || !d_c) return ch_fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_CH_CHECK(hipSetDevice(ctx->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  return run_ch(ctx, cp, d_rhs, d_c, report, s);
}

int pp_ch_solve(pp_ctx *ctx, const pp_ch_problem *cp, const double *rhs, double *c, pp_ch_report *report) {
  if (int rc = validate_ch(ctx, cp, report)) return rc;
  if (!rhs || !c) return ch_fail(ctx, report, PP_ERR_ARG, "null field pointer");
  PP_CH_CHECK(hipSetDevice(ctx->device));
  auto &C = ctx->ch;
  const size_t n = static_cast<size_t>(cp->nx) * cp->ny;
  if (2 * n > C.stage_cap) {
    if (C.stage) (void)hipFree(C.stage);
    C.stage = nullptr;
    C.stage_cap = 0;
    PP_CH_CHECK(hipMalloc(&C.stage, 2 * n * sizeof(double)));
    C.stage_cap = 2 * n;
  }
  hipStream_t s = ctx->stream;
  PP_CH_CHECK(hipMemcpyAsync(C.stage, rhs, n * sizeof(double), hipMemcpyHostToDevice, s));
  PP_CH_CHECK(hipMemcpyAsync(C.stage + n, c, n * sizeof(double), hipMemcpyHostToDevice, s));
  const int rc = run_ch(ctx, cp, C.stage, C.stage + n, report, s);
  const std::string err = ctx->last_error;
  PP_CH_CHECK(hipMemcpyAsync(c, C.stage + n, n * sizeof(double), hipMemcpyDeviceToHost, s));
  PP_CH_CHECK(hipStreamSynchronize(s));
  ctx->last_error = err;
  return rc;
}

}  // extern "C"
