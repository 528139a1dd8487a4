// Test driver for the C++ host mirror (ding_nsch_core_solver_amd/host/pp_backend.*), built by
// tests/test_host_mirror.py with g++ against libppsolve.so.
//
//   driver schemes                         scheme string -> code / split, one per line
//   driver ghost  <cfg> <field.bin>        ghost-fill a (nx+2g)(ny+2g) field in place
//   driver solve  <cfg> <scheme> <in.bin> <out.bin>
//        in.bin: rho (nx*ny), u ((nx+1)*ny), v (nx*(ny+1)), p_n, p_nm1 (nx*ny each), doubles
//        out.bin: pressure_correction with 3 ghost layers; prints "iterations residual"
//   driver monitor <cfg> <scheme> <in.bin> <out.bin>
//        as solve, with the PCG monitor: one line per record "k alpha beta abs rel iterate precond"
//        after the "iterations residual" line
//   driver step <cfg> <vel> <scheme> <in.bin> <out.bin>
//        as correct, through PressureBackend::pressure_step (p_corr in in.bin is ignored); out.bin:
//        p_corr, u, v, pressure, pressure_prev; prints "iterations residual speed_pre speed_post"
//   driver correct <cfg> <vel> <scheme> <in.bin> <out.bin>
//        in.bin: whole ghosted buffers (3 layers) rho_mid, u_star, v_star, p_corr, u, v, pressure,
//        pressure_prev; out.bin: u, v, pressure, pressure_prev; prints "speed_pre speed_post"
// <cfg> = "nx ny lx ly dt px py t0 v0 t1 v1 t2 v2 t3 v3 maxit tol dr step" as one argument;
// <vel> = "ut0 uv0 .. ut3 uv3 vt0 vv0 .. vt3 vv3 bottom_wall top_wall" as one argument.
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "pp_backend.hpp"

using ppamd::FieldView;
using ppamd::PressureSettings;

namespace {

struct Cfg {
  int nx = 0, ny = 0, max_it = 0, step = 0;
  PressureSettings s;
};

Cfg parse(const char *text) {
  std::istringstream in(text);
  Cfg c;
  int px = 0, py = 0;
  in >> c.nx >> c.ny >> c.s.lx >> c.s.ly >> c.s.dt >> px >> py;
  ppamd::BoundarySpec *e[4] = {&c.s.left, &c.s.right, &c.s.bottom, &c.s.top};
  for (auto *b : e) in >> b->type >> b->value;
  in >> c.max_it >> c.s.pressure_tolerance >> c.s.density_ratio >> c.step;
  c.s.periodic_x = px != 0;
  c.s.periodic_y = py != 0;
  c.s.poisson_iterations = c.max_it;
  if (!in) {
    std::fprintf(stderr, "bad cfg\n");
    std::exit(2);
  }
  return c;
}

ppamd::VelocitySettings parse_vel(const char *text) {
  std::istringstream in(text);
  ppamd::VelocitySettings v;
  ppamd::BoundarySpec *e[8] = {&v.u_left, &v.u_right, &v.u_bottom, &v.u_top,
                               &v.v_left, &v.v_right, &v.v_bottom, &v.v_top};
  for (auto *b : e) in >> b->type >> b->value;
  in >> v.bottom_wall_velocity_x >> v.top_wall_velocity_x;
  if (!in) {
    std::fprintf(stderr, "bad velocity cfg\n");
    std::exit(2);
  }
  return v;
}

std::vector<double> read_all(const char *path) {
  std::vector<double> v;
  FILE *f = std::fopen(path, "rb");
  if (!f) std::exit(3);
  double x;
  while (std::fread(&x, sizeof(double), 1, f) == 1) v.push_back(x);
  std::fclose(f);
  return v;
}

void write_all(const char *path, const std::vector<double> &v) {
  FILE *f = std::fopen(path, "wb");
  if (!f) std::exit(3);
  std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
}

// A ghosted field of rows x cols interior cells filled from a contiguous interior array.
struct Owned {
  std::vector<double> buf;
  FieldView view;
  Owned(int rows, int cols, int g, const double *interior) : buf(static_cast<size_t>(rows + 2 * g) * (cols + 2 * g), 0.0) {
    view = FieldView{buf.data(), rows, cols, g};
    if (interior)
      for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) view(i, j) = interior[static_cast<size_t>(i) * cols + j];
  }
};

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  if (mode == "schemes") {
    const char *names[] = {"hip_icpcg", "hip_ildlt_pcg", "hip_split_icpcg", "hip_liu_split_icpcg",
                           "hip_paper_split_icpcg", "hip_split_ildlt_pcg", "hip_liu_split_ildlt_pcg",
                           "hip_paper_split_ildlt_pcg", "icpcg", "petsc_pcg"};
    for (const char *n : names) std::printf("%s %d %d\n", n, ppamd::hip_scheme_code(n), ppamd::hip_scheme_is_split(n) ? 1 : 0);
    return 0;
  }
  if (mode == "ghost" && argc == 4) {
    const Cfg c = parse(argv[2]);
    std::vector<double> f = read_all(argv[3]);
    ppamd::apply_pressure_bc(c.s, FieldView{f.data(), c.nx, c.ny, 3});
    write_all(argv[3], f);
    return 0;
  }
  if ((mode == "solve" || mode == "monitor") && argc == 6) {
    const Cfg c = parse(argv[2]);
    const int nx = c.nx, ny = c.ny, g = 3;
    const std::vector<double> in = read_all(argv[4]);
    const size_t n = static_cast<size_t>(nx) * ny;
    const double *rho = in.data(), *u = rho + n, *v = u + static_cast<size_t>(nx + 1) * ny;
    const double *pn = v + static_cast<size_t>(nx) * (ny + 1), *pm = pn + n;
    Owned R(nx, ny, g, rho), U(nx + 1, ny, g, u), V(nx, ny + 1, g, v), P(nx, ny, g, pn), PM(nx, ny, g, pm);
    Owned out(nx, ny, g, nullptr);
    try {
      ppamd::PressureBackend backend(0);
      std::vector<pp_monitor_record> log;
      const double res =
          mode == "monitor"
              ? backend.solve(argv[3], c.s, R.view, U.view, V.view, P.view, PM.view, c.step, out.view,
                              [&](const pp_monitor_record &m) { log.push_back(m); })
              : backend.solve(argv[3], c.s, R.view, U.view, V.view, P.view, PM.view, c.step, out.view);
      write_all(argv[5], out.buf);
      std::printf("%d %.17g\n", backend.last_iterations(), res);
      for (const pp_monitor_record &m : log)
        std::printf("%d %.17g %.17g %.17g %.17g %.17g %.17g\n", m.iteration, m.alpha, m.beta, m.absolute_residual,
                    m.relative_residual, m.iterate_residual, m.preconditioned_residual);
    } catch (const std::exception &e) {
      std::printf("error %s\n", e.what());
      return 1;
    }
    return 0;
  }
  if ((mode == "correct" || mode == "step") && argc == 7) {
    const Cfg c = parse(argv[2]);
    const ppamd::VelocitySettings vel = parse_vel(argv[3]);
    const int nx = c.nx, ny = c.ny, g = 3;
    std::vector<double> in = read_all(argv[5]);
    const size_t cell = static_cast<size_t>(nx + 2 * g) * (ny + 2 * g);
    const size_t uf = static_cast<size_t>(nx + 1 + 2 * g) * (ny + 2 * g);
    const size_t vf = static_cast<size_t>(nx + 2 * g) * (ny + 1 + 2 * g);
    if (in.size() != 4 * cell + 2 * uf + 2 * vf) return 3;
    double *at = in.data();
    auto take = [&](size_t n) {
      double *p = at;
      at += n;
      return p;
    };
    const FieldView rho{take(cell), nx, ny, g}, us{take(uf), nx + 1, ny, g}, vs{take(vf), nx, ny + 1, g};
    const FieldView pc{take(cell), nx, ny, g}, u{take(uf), nx + 1, ny, g}, v{take(vf), nx, ny + 1, g};
    const FieldView p{take(cell), nx, ny, g}, pm{take(cell), nx, ny, g};
    try {
      ppamd::PressureBackend backend(0);
      backend.set_exact_mean(true);  // compared with the reference bit for bit
      std::vector<double> out;
      if (mode == "step") {
        const double res = backend.pressure_step(argv[4], c.s, vel, rho, us, vs, c.step, pc, u, v, p, pm);
        std::printf("%d %.17g ", backend.last_iterations(), res);
        out.insert(out.end(), pc.data, pc.data + cell);
      } else {
        backend.correct_velocity(argv[4], c.s, vel, rho, us, vs, pc, c.step, u, v, p, pm);
      }
      out.insert(out.end(), u.data, u.data + uf);
      out.insert(out.end(), v.data, v.data + vf);
      out.insert(out.end(), p.data, p.data + cell);
      out.insert(out.end(), pm.data, pm.data + cell);
      write_all(argv[6], out);
      std::printf("%.17g %.17g\n", backend.boundary_speed_pre_correction(), backend.boundary_speed_post_correction());
    } catch (const std::exception &e) {
      std::printf("error %s\n", e.what());
      return 1;
    }
    return 0;
  }
  if (mode == "bicgstab" && argc == 4) {
    // input: n, nnz, max_iterations, tolerance, use_relative, row_ptr[n+1], col_idx[nnz], values[nnz],
    // rhs[n], x0[n] (all as doubles); output: x
    const std::vector<double> in = read_all(argv[2]);
    const int n = static_cast<int>(in[0]), nnz = static_cast<int>(in[1]);
    ppamd::SparseMatrixCSR m;
    m.n = n;
    size_t o = 5;
    for (int k = 0; k <= n; ++k) m.row_ptr.push_back(static_cast<int>(in[o++]));
    for (int k = 0; k < nnz; ++k) m.col_idx.push_back(static_cast<int>(in[o++]));
    m.values.assign(in.begin() + o, in.begin() + o + nnz);
    o += nnz;
    const std::vector<double> rhs(in.begin() + o, in.begin() + o + n);
    std::vector<double> x(in.begin() + o + n, in.begin() + o + 2 * n);
    try {
      ppamd::PressureBackend backend(0);
      const ppamd::LinearSolveReport r = backend.solve_bicgstab(m, rhs, static_cast<int>(in[2]), in[3], x, in[4] != 0.0);
      write_all(argv[3], x);
      std::printf("%d %.17g %.17g %.17g\n", r.iterations, r.absolute_residual, r.relative_residual, r.iterate_residual);
    } catch (const std::exception &e) {
      std::printf("error %s\n", e.what());
      return 1;
    }
    return 0;
  }
  return 2;
}
