// Arguments and launch entry points shared between the host layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pp {

struct DevReport;

// One PCG monitor record: the scalar fields of one call of the reference's CGMonitor (layout of
// pp_monitor_record in pp_solver.h).
struct MonitorRecord {
  int iteration, pad;
  double alpha, beta, abs_res, rel_res, iter_res, prec_res;
};

// One pressure system as the kernels see it. Field inputs/outputs are in the natural i-major
// layout; every work array is in the wavefront layout described in pp_device.hpp.
struct SysArgs {
  int nx, ny, n, nlev;
  double dx, dy, dt;
  int periodic_x, periodic_y;
  int bc_type[4];
  double bc_value[4];
  int split, ildlt, projected;
  int max_iterations;
  double tolerance;
  double rho_ref;
  int step_index;
  int ring;                // sweep operand ring depth in levels (0: default for the wave count)
  int spin_limit;          // sleeps a hand-off wait may spend before the solve fails (0: kSpinLimit)
  // mixed precision controls
  int inner_max_iterations;
  double inner_tolerance;
  // layout tables
  const int *lev_base;     // level_base(d) for d in [-kLevPad, nlev + kLevPad) (0 outside the grid)
  int spmv_split[9];       // SpMV level ranges of the 8 waves: wave w takes [spmv_split[w], spmv_split[w+1])
  const uint32_t *cell;    // n, (i << 16) | j of each wavefront position
  // fields (natural layout)
  const double *rho, *u, *v, *p_n, *p_nm1;
  double *p_out;
  // coefficient mode (coeff_mode = 1): face inverse densities k_u ((nx+1) x ny), k_v (nx x (ny+1))
  // and a source term src (nx x ny) replace rho, u and v; c_f = k_f / h^2, b = src (+ the boundary
  // terms when rhs_bc = 1)
  int coeff_mode, rhs_bc;
  const double *ku, *kv, *src;
  // FP64 work arrays (wavefront layout, n each)
  double *aP, *aW, *aS;    // CSR diagonal, west entry, south entry (0 when absent or dropped)
  double *Lw, *Ls, *dg;    // factor: west/south lower entries, pivot (ICC d or ILDLT D)
  double *b, *x, *r, *pa, *pb, *z, *y;
  double *LX, *LY;         // periodic wrap lower entries: LX[j] row (nx-1,j) col (0,j); LY[i] row (i,ny-1) col (i,0)
  // FP32 work arrays for mixed precision (wavefront layout)
  float *fP, *fW, *fS, *fLw, *fLs, *fRd;   // operator, factor, reciprocal pivot
  float *fx, *fr, *fpa, *fpb, *fz, *fy;
  float *fLX, *fLY;
  DevReport *rep;
  // PCG monitor log (FP64 solve): record k written for monitor call k while k < monitor_cap
  MonitorRecord *monitor;
  int monitor_cap;
};

// One workgroup per system: workgroup k solves d_systems[k]. With more systems than CUs the
// dispatcher starts workgroup k as soon as a CU is free, in index order, so an array sorted by
// decreasing cost is scheduled largest first.
// monitor: the variant that writes the PCG monitor log (SysArgs::monitor) of every system.
hipError_t launch_fp64_solve(const SysArgs *d_systems, int count, bool ildlt, hipStream_t stream,
                             bool monitor = false);

// Mixed precision (FP32 inner PCG inside FP64 refinement), one workgroup per system.
hipError_t launch_mixed_solve(const SysArgs *d_systems, int count, bool ildlt, hipStream_t stream);

// Default and valid operand ring depths for S sweep waves: a power of two, at least one 8-level
// sweep chunk (the sweep releases ring space once per chunk, so a shallower ring would stall its
// own loaders), and S * ring * 256 doubles within the LDS stage.
inline int default_ring(int S) { return S <= 4 ? 16 : 8; }
// Sweep wave roles: a ring depth picks the staged sweeps (8 - S loader waves fill an LDS ring of
// that many levels per sweep wave); kDirectSweeps picks the direct ones (every sweep wave loads its
// own operands, no loaders), which wide grids use anyway. Both give the same bits.
constexpr int kDirectSweeps = -1;

inline bool valid_ring(int S, int ring) {
  return ring == kDirectSweeps || (ring >= 8 && ring <= 64 && (ring & (ring - 1)) == 0 && S * ring <= 64);
}

// Face gradients of the coefficient-mode adjoint: given the solution p and the adjoint solution lam
// (both nx x ny, natural layout), writes dL/dk_u ((nx+1) x ny) and dL/dk_v (nx x (ny+1)).
struct GradArgs {
  int nx, ny;
  double dx, dy;
  int periodic_x, periodic_y;
  int bc_type[4];
  double bc_value[4];
  const double *p, *lam;
  double *dku, *dkv;
};

hipError_t launch_face_gradient(const GradArgs &g, hipStream_t stream);

// Velocity correction and pressure-history shift on ghosted Field2D buffers (G ghost layers).
// Boundary conditions arrive resolved: unset pressure edges as Neumann 0, unset velocity edges as
// their no-slip / wall-speed Dirichlet defaults.
struct CorrArgs {
  int nx, ny, G;
  double dx, dy, dt;
  int periodic_x, periodic_y;
  int p_type[4];
  double p_value[4];
  int u_type[4];
  double u_value[4];
  int v_type[4];
  double v_value[4];
  int split, step_index, has_dirichlet;
  int exact_mean;  // 1: the pressure mean in the reference's sequential order; 0: the 256-thread tree
  double rho_ref;
  const double *rho_mid, *u_star, *v_star, *p_corr;
  double *u, *v, *pressure, *pressure_prev;
  double *work;    // [4]: boundary speed before / after, the pressure mean, unused
};

// Faces, history and finish kernels of the velocity correction on `stream` (nx, ny, ghost as in
// the argument block, for the grid sizes).
hipError_t launch_velocity_correction(const CorrArgs *d_args, int nx, int ny, int ghost, hipStream_t stream);

// apply_pressure_bc on d_field (nx x ny with G ghost layers, Field2D layout) per the pressure BCs
// and periodicity of the argument block; one block on `stream`.
hipError_t launch_pressure_ghosts(const CorrArgs *d_args, double *d_field, hipStream_t stream);

// BiCGSTAB with the diagonal preconditioner on a CSR system (the momentum predictor's solve).
// Workgroup c of `chunks` owns rows [kBicgRows c, kBicgRows (c + 1)); each pass leaves its dot
// products as per-workgroup partials (partials[k * chunks + c]) that the next pass sums, and the
// scalar recurrence and stopping tests run on the device against the state block.
// Per-level part of the windowed SpMV's cost in the host's level split, in lane units (a level costs
// 64 per 64-row chunk holding a row of it, plus this). Weights of 64 to 256 split the 192x192 SpMV
// equally fast (9.7 ms); 160 keeps the canonical p^T A p order at the reference's iteration count on
// every timestep state and parity case, as the cell-count split did. The oracle uses the same value.
constexpr int kSpmvLevelCost = 160;

constexpr int kBicgRows = 512;
enum BicgStatus : int {
  kBicgRunning = 0, kBicgDone = 1, kBicgNoDiag = 2, kBicgRho = 3, kBicgRhatV = 4, kBicgTT = 5, kBicgOmega = 6
};
struct BicgState {
  int status;      // BicgStatus
  int half_exit;   // converged at the half step; the next iteration's pass A finishes the report
  int iterations;  // report (LinearSolveReport)
  int pad;
  double rho[2];   // rho of the iterations of each parity
  double alpha, omega;
  double abs_res, rel_res, iter_res;
  double rhs_norm;  // |b|, from the first iteration's pass on
  double sums[12];  // large systems: the dot products of the partial slots, from the sum pass
};
// Partial slots (each `chunks` long, in this order): setup b.b, r.r; pass A r_hat.v; pass B s.s,
// t.t, t.s; pass C update.update, x.x, r.r, r_hat.r; half exit update.update, x.x.
constexpr int kSlotSetup = 0, kSlotV = 2, kSlotB = 3, kSlotX = 6, kSlotExit = 10, kBicgSlots = 12;
// Above this many workgroups the partials are summed once by a one-workgroup pass.
constexpr int kBicgPresumChunks = 256;
struct BicgArgs {
  int n, chunks, use_rel;
  int presummed;  // chunks > kBicgPresumChunks: passes read state->sums
  double tolerance;
  const int *row_ptr, *col_idx;
  const double *values, *rhs;
  double *x;
  double *r, *r_hat, *p[2], *v[2], *p_hat, *s, *s_hat, *t, *update, *dinv;  // n each; p, v by parity
  double *partials;  // kBicgSlots x chunks
  double *p_setup, *p_v, *p_b, *p_x, *p_exit;  // the slots' starts
  BicgState *state;                                                    // zeroed before the setup pass
};
// The setup pass; the three passes of iteration iter; the report after `iterations` iterations
// when the limit ended the solve.
hipError_t launch_bicg_setup(const BicgArgs &a, hipStream_t stream);
hipError_t launch_bicg_iteration(const BicgArgs &a, int iter, hipStream_t stream);
hipError_t launch_bicg_report(const BicgArgs &a, int iterations, hipStream_t stream);

// Cahn-Hilliard linear solve (solve_phase_linear_system_eq25): the operator's CSR with its level
// schedule and lower-transpose lists (set on the host), the ILDLT(0) factor and the PCG vectors, all
// natural order. One workgroup of kThreads runs the factorization, and one the whole PCG.
enum ChStatus : int { kChOk = 0, kChRz0Fail = 1, kChRzFail = 2, kChPapFail = 3 };
struct ChState {
  int status, iterations, ildlt_ok, pad;
  double floor;                        // ILDLT pivot floor max(1, max |a_ii|) 1e-12
  double abs_res, rel_res, iter_res, eq_res;
};
struct ChArgs {
  int n, nlev;
  const int *row_ptr, *col, *diag_pos;
  const double *val;
  const int *lev_ptr, *lev_rows;       // rows of level k: lev_rows[lev_ptr[k] .. lev_ptr[k+1])
  const int *lt_ptr, *lt_row, *lt_pos; // children of row r in ascending order, with the CSR position of L(child, r)
  double *lower, *D, *dinv;            // factor (aligned with the CSR) and the diagonal preconditioner
  const double *rhs;
  double *x, *r, *z, *p, *ap, *y;
  int max_iterations;
  double tolerance;
  ChState *state;
  // Batched sequential sweeps (seq_batch = B > 0, used when the level schedule degenerates into a
  // chain, as periodic y makes it): rows in batches of B, the vectors in LDS, every row's operands
  // packed (kChSeqWidth per row, in the reference's order) so a batch's loads are independent and
  // can be issued while the previous batch runs. Forward: fo_cnt[r] lower entries before the batch
  // (columns fo_col, CSR positions fo_pos), then the in-batch L(r, r-2), L(r, r-1) (CSR positions
  // fin_pos[2r], fin_pos[2r+1], -1 when absent). Backward: the in-batch children r+1, r+2 (positions
  // bin_pos[2r], bin_pos[2r+1] of L(r+1, r), L(r+2, r)), which lead the ascending child list, then
  // bf_cnt[r] children outside the batch (rows bf_row, positions bf_pos). The values are gathered
  // from the factor once per operator (launch_ch_pack) into fo_val, fin_val, bf_val, bin_val.
  // The batched factorization also takes, for entry e of row r before the batch and each earlier
  // entry e2 of the row, the CSR position of L(fo_col[e], fo_col[e2]) (fx_pos[(r W + e) W + e2],
  // -1 when absent), and for entry e2 the index + 1 of its column among the entries of rows r-2 and
  // r-1 before the batch (fab[r W + e2], bits 0-7 and 8-15, 0 when absent).
  int seq_batch;
  const int *fo_cnt, *fo_col, *fo_pos, *fin_pos, *bf_cnt, *bf_row, *bf_pos, *bin_pos, *fx_pos, *fab;
  double *fo_val, *fin_val, *bf_val, *bin_val;
};
constexpr int kChSeqWidth = 12;       // entries of a row outside its batch (the 13-point stencil has at most 12)
constexpr int kChSeqMaxRows = 6912;   // the sequential sweeps keep n doubles (and 30 KB of row records and staging) in LDS
hipError_t launch_ch_pack(const ChArgs &a, hipStream_t stream);
hipError_t launch_ch_factor(const ChArgs &a, hipStream_t stream);
hipError_t launch_ch_pcg(const ChArgs &a, bool ildlt, hipStream_t stream);

// The CH solve's CG with the diagonal preconditioner as full-chip passes (the chunked reductions of
// pp_chunked.hpp): a setup pass, then per iteration pass A (the stop tests of the previous step,
// beta, p = z + beta p and A p) and pass B (alpha, the x / r update, z = M^-1 r), and after the
// stop the residual of the returned x. Scalars live in the state block; a stop recorded there
// makes every later pass return at once.
// Partial slots (each `chunks` long): setup b.b, r.r, r.z; pass A p.Ap; pass B delta.delta, x.x,
// r.r, r.z; the closing |rhs - A x|^2.
constexpr int kChRows = 512;  // rows per workgroup of the full-chip passes
constexpr int kChSlotSetup = 0, kChSlotPap = 3, kChSlotUpd = 4, kChSlotEq = 8, kChSlots = 9;
constexpr int kChPresumChunks = 256;  // above this many workgroups a one-workgroup pass sums the partials
struct ChDiagState {
  int stopped, status, iterations, pad;  // status: ChStatus of the stop
  double rz[2];                          // r.z of the iterations of each parity
  double bnorm, bsq;                     // |b| and b.b (the setup's sums)
  double abs_res, rel_res, iter_res, eq_res;
  double sums[kChSlots];                 // presummed slots
};
struct ChDiagArgs {
  int n, chunks, presummed, max_iterations;
  double tolerance;
  const int *row_ptr, *col;
  const double *val, *dinv, *rhs;
  double *x, *r, *z, *p[2], *ap;  // p by iteration parity
  double *partials;               // kChSlots x chunks
  ChDiagState *state;             // zeroed before the setup pass
};
hipError_t launch_ch_diag_setup(const ChDiagArgs &a, hipStream_t stream);
hipError_t launch_ch_diag_iteration(const ChDiagArgs &a, int k, hipStream_t stream);
// The stop test after `iterations` iterations when the limit ended the loop.
hipError_t launch_ch_diag_report(const ChDiagArgs &a, int iterations, hipStream_t stream);
// |rhs - A x| of the returned x over |b|.
hipError_t launch_ch_diag_finish(const ChDiagArgs &a, hipStream_t stream);

// Momentum predictor (Solver::compute_momentum_fluxes + solve_momentum_predictor) on ghosted Field2D
// buffers (G layers). Unknowns: the active u faces (i-major; i = iu0 .. nx-1, iu0 = 0 periodic else 1)
// then the active v faces (j = jv0 .. ny-1); each row keeps up to kMomRow entries.
constexpr int kMomRow = 9;
constexpr int kMomScanBlock = 1024;  // rows per block of the row-pointer scan
struct MomArgs {
  int nx, ny, G;
  double dx, dy, dt, alpha;           // alpha = dt / (2 re)
  double body_force_x, body_force_y;
  int periodic_x, periodic_y, step_index;
  int u_type[4];                      // resolved velocity BCs (left, right, bottom, top)
  double u_value[4];
  int v_type[4];
  double v_value[4];
  int iu0, jv0, niu, njv, nu, n;      // row layout: u rows (i - iu0) * ny + j, v rows nu + i * njv + (j - jv0)
  const double *u, *v, *rho, *eta, *eta_prev, *sfx, *sfy;
  double *u_prev, *v_prev;            // in: the previous step's convection; out: this step's
  double *u_star, *v_star;
  // work (n rows): merged row entries, counts, CSR, right-hand side, solution (initial guess x0)
  int *slot_col;
  double *slot_val;
  int *count;
  int *block_sum;                     // per scan block, then its exclusive prefix
  int *row_ptr, *col;
  double *val, *rhs, *x;
};

// Assembly, row-pointer scan and CSR compaction, and this step's convection fields (with their ghost
// fills) into u_prev / v_prev; all on `stream`. A row whose diagonal was dropped is emitted without
// it, and the BiCGSTAB setup pass reports it (the reference's finalize_row_maps text).
hipError_t launch_momentum_system(const MomArgs &a, hipStream_t stream);
// u_star / v_star from u / v and the solution x, then their ghost fills.
hipError_t launch_momentum_scatter(const MomArgs &a, hipStream_t stream);

}  // namespace pp
