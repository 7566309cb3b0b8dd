// drop_in.cc -- the reference's C++ library API (lib/*.h) implemented over
// the C ABI of libdachshund_gpu.so.  Hot loops (matvec, residual, PCG,
// S_mp projections) run in the gfx950 kernels; only O(n) record handling,
// file input and scalar element evaluation stay on the host.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../../include/dachshund_gpu.h"
#include "dachshund.h"

// ---- GPU context ---------------------------------------------------------

namespace {

ds_ctx* g_ctx = nullptr;
std::once_flag g_once;

[[noreturn]] void die(const char* what, int rc) {
  fprintf(stderr, "[dachshund-gpu] %s failed (status %d): %s\n", what, rc,
          g_ctx ? ds_ctx_last_error(g_ctx) : "no context");
  exit(1);
}

void check(int rc, const char* what) {
  if (rc != DS_OK) die(what, rc);
}

struct CovarView {
  ds_covar c;
  explicit CovarView(const SignalCovarParams* p) {
    c.var_s = p->var_s;
    c.c_perp = p->gaussian_const_perp;
    c.c_para = p->gaussian_const_para;
    c.dx_inv = p->func_table_dx_inv;
    c.table_n = p->func_table_n;
    c.table = p->func_table;
  }
};

ds_problem* make_problem(int form, int n, const double* recs,
                         const SignalCovarParams* sp) {
  CovarView cv(sp);
  ds_problem* p = nullptr;
  check(ds_problem_create(dachshund_gpu_context(), form, n, recs, &cv.c, &p),
        "ds_problem_create");
  return p;
}

PCGResult run_solve(ds_problem* p, const PCGParams& prm, double* x,
                    bool mixed, double drop) {
  ds_pcg_params c{prm.n, prm.max_iter, prm.step_r, prm.tol, prm.verbose};
  ds_pcg_result r{};
  std::vector<double> norms((size_t)std::max(prm.max_iter, 1) + 1);
  const int rc = mixed ? ds_solve_mixed(p, &c, drop, x, norms.data(),
                                        (int)norms.size(), &r)
                       : ds_solve_pcg(p, &c, x, norms.data(),
                                      (int)norms.size(), &r);
  check(rc, mixed ? "ds_solve_mixed" : "ds_solve_pcg");
  PCGResult res;
  res.success = r.success != 0;
  res.num_iters = r.num_iters;
  res.residual_norms.assign(norms.begin(), norms.begin() + r.n_norms);
  // lib/pcg.h:84,112-139,154 messages
  if (r.norm_b <= 0.0) {
    puts("[WARNING] Requested PCG solve where |b| <= 0.0. Returning early.\n");
    return res;
  }
  if (prm.verbose) {
    // lib/pcg.h:112-139's trace, printed once the device solve returns
    std::vector<double> deltas(norms.size());
    int nd = 0;
    check(ds_problem_get_deltas(p, deltas.data(), (int)deltas.size(), &nd),
          "ds_problem_get_deltas");
    const double tol_norm_b = prm.tol * r.norm_b;
    printf("[PCG] Solving %i x %i problem.\n", prm.n, prm.n);
    printf("    Goal |r| < (tol) |b| = %e,  delta = %e\n", tol_norm_b,
           deltas[0]);
    const int last = (int)res.residual_norms.size() - 1;
    for (int i = 1; i <= last && i < nd; ++i) {
      if (i == last && res.success)
        printf("    iter %i reached |r| = %e < tol |b| = %e\n", i,
               res.residual_norms[i], tol_norm_b);
      else
        printf("    iter %i, |r| %e, delta %e\n", i, res.residual_norms[i],
               deltas[i]);
    }
  }
  if (!res.success)
    puts("[WARNING] PCG reached max iteration before stop condition.");
  return res;
}

}  // namespace

ds_ctx* dachshund_gpu_context() {
  std::call_once(g_once, []() {
    const char* dev = getenv("DACHSHUND_DEVICE");
    const int d = dev ? atoi(dev) : 0;
    const int rc = ds_ctx_create(d, &g_ctx);
    if (rc != DS_OK) {
      fprintf(stderr, "[dachshund-gpu] no usable GPU %d (status %d)\n", d,
              rc);
      exit(1);
    }
  });
  return g_ctx;
}

// ---- timer.h -------------------------------------------------------------

Timer total_timer;

Timer::Timer() { reset(); }
tv Timer::get_time() {
  tv t;
  gettimeofday(&t, nullptr);
  return t;
}
double Timer::interval(const tv a, const tv b) {
  return (double)(b.tv_sec - a.tv_sec) + 1e-6 * (double)(b.tv_usec - a.tv_usec);
}
void Timer::reset() { t0 = get_time(); }
double Timer::elapsed() const { return interval(t0, get_time()); }

// ---- linalg.h ------------------------------------------------------------

double ds_vector_dot(const int n, const double* const a,
                     const double* const b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
double ds_vector_norm(const int n, const double* const x) {
  return std::sqrt(ds_vector_dot(n, x, x));
}
void ds_vector_add_cv(const int n, const double a, const double* const b,
                      double* const c) {
  for (int i = 0; i < n; ++i) c[i] += a * b[i];
}
void ds_vector_sub_cv(const int n, const double a, const double* const b,
                      double* const c) {
  for (int i = 0; i < n; ++i) c[i] -= a * b[i];
}

// ---- signal_covar.h ------------------------------------------------------

SignalCovarParams::SignalCovarParams(const double var_s_, const double l_perp,
                                     const double l_para)
    : var_s(var_s_) {
  gaussian_const_perp = 1.0 / (2.0 * l_perp * l_perp);
  gaussian_const_para = 1.0 / (2.0 * l_para * l_para);
  // table over [0, x_max^2 / 2] with func_table_n points
  const double fx_max =
      default_gaussian_table_x_max * default_gaussian_table_x_max / 2.0;
  func_table_n = default_gaussian_table_n;
  const double ft_dx = fx_max / (func_table_n - 1);
  func_table_dx_inv = 1.0 / ft_dx;
  func_table = new double[func_table_n];
  for (int i = 0; i < func_table_n; ++i) func_table[i] = std::exp(-(ft_dx * i));
}

SignalCovarParams::SignalCovarParams(const SignalCovarParams& o)
    : var_s(o.var_s),
      gaussian_const_perp(o.gaussian_const_perp),
      gaussian_const_para(o.gaussian_const_para),
      func_table_n(o.func_table_n),
      func_table_dx_inv(o.func_table_dx_inv),
      func_table(new double[o.func_table_n]) {
  std::memcpy(func_table, o.func_table, sizeof(double) * func_table_n);
}

SignalCovarParams& SignalCovarParams::operator=(const SignalCovarParams& o) {
  if (this != &o) {
    double* t = new double[o.func_table_n];
    std::memcpy(t, o.func_table, sizeof(double) * o.func_table_n);
    delete[] func_table;
    func_table = t;
    var_s = o.var_s;
    gaussian_const_perp = o.gaussian_const_perp;
    gaussian_const_para = o.gaussian_const_para;
    func_table_n = o.func_table_n;
    func_table_dx_inv = o.func_table_dx_inv;
  }
  return *this;
}

SignalCovarParams::~SignalCovarParams() { delete[] func_table; }

// ---- pixel.h -------------------------------------------------------------

Point* points_from_pixels_alloc(const int n, Pixel* const pixels) {
  Point* pts = new Point[n];
  for (int i = 0; i < n; ++i) pts[i] = Point{pixels[i].x, pixels[i].y,
                                             pixels[i].z};
  return pts;
}

void pixel_to_npixel(const int n, Pixel* const p, NPixel* const np) {
  for (int i = 0; i < n; ++i) {
    const Pixel a = p[i];
    np[i] = NPixel{a.x, a.y, a.z, a.n * a.n, a.d};
  }
}
void npixel_to_pixel(const int n, NPixel* const np, Pixel* const p) {
  for (int i = 0; i < n; ++i) {
    const NPixel a = np[i];
    p[i] = Pixel{a.x, a.y, a.z, std::sqrt(a.N), a.d};
  }
}
void pixel_to_wpixel(const int n, Pixel* const p, WPixel* const wp) {
  for (int i = 0; i < n; ++i) {
    const Pixel a = p[i];
    wp[i] = WPixel{a.x, a.y, a.z, 1.0 / a.n, a.d / a.n};
  }
}
void wpixel_to_pixel(const int n, WPixel* const wp, Pixel* const p) {
  for (int i = 0; i < n; ++i) {
    const WPixel a = wp[i];
    p[i] = Pixel{a.x, a.y, a.z, 1.0 / a.w, a.wd / a.w};
  }
}

// ---- util.h --------------------------------------------------------------

void read_pixel_data(const std::string path, const int num_pixels,
                     Pixel* const pixels) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) {
    printf("Could not open file %s\n", path.c_str());
    exit(1);
  }
  const size_t got = fread(pixels, sizeof(Pixel), (size_t)num_pixels, f);
  fclose(f);
  if (got != (size_t)num_pixels)
    fprintf(stderr, "[WARNING] %s holds %zu of %d pixels\n", path.c_str(),
            got, num_pixels);
}

void init_uniform_map_coords(const double lx, const double ly, const double lz,
                             const int nx, const int ny, const int nz,
                             Point* const coords) {
  const double dx = lx / nx, dy = ly / ny, dz = lz / nz;
  for (int ix = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy)
      for (int iz = 0; iz < nz; ++iz) {
        const size_t i = ((size_t)ix * ny + iy) * nz + iz;
        coords[i] = Point{dx * (ix + 0.5), dy * (iy + 0.5), dz * (iz + 0.5)};
      }
}

void check_finite(const int n, const double* const v) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) printf("non-finite value v[%i] = %e\n", i, v[i]);
}

// ---- wf_vector.h ---------------------------------------------------------

namespace {

template <class Pix>
double element(const Pix* px, int i, int j, const SignalCovarParams* sp,
               bool exact) {
  const double dx = px[i].x - px[j].x;
  const double dy = px[i].y - px[j].y;
  const double dz = px[i].z - px[j].z;
  const double xp2 = (dx * dx + dy * dy);
  const double xl2 = dz * dz;
  return exact ? signal_covar_exact(xp2, xl2, sp) : signal_covar(xp2, xl2, sp);
}

}  // namespace

#define DS_BIND_OPS(T)                                                        \
  this_pc_solve = [this](const double* b, double* x) {                       \
    preconditioner_solve(b, x);                                               \
  };                                                                          \
  this_A_product = [this](const double* a, double* b) { A_product(a, b); }; \
  this_A_residual = [this](const double* x, double* r) { A_residual(x, r); }

WFVectorSN::WFVectorSN(const int num_pixels, const NPixel* const pixels_,
                       const SignalCovarParams* const s_params_)
    : WFVector(num_pixels, s_params_), pixels(pixels_) {
  DS_BIND_OPS(WFVectorSN);
}
WFVectorSN::~WFVectorSN() {
  if (gpu_) ds_problem_destroy(gpu_);
}
ds_problem* WFVectorSN::gpu() {
  if (!gpu_)
    gpu_ = make_problem(DS_FORM_SN, n, (const double*)pixels, s_params);
  return gpu_;
}
double WFVectorSN::A(const int i, const int j, const bool exact) {
  return element(pixels, i, j, s_params, exact) + pixels[i].N * (i == j);
}
double WFVectorSN::norm_b() {
  double v = 0.0;
  check(ds_norm_b(gpu(), &v), "ds_norm_b");
  return v;
}
void WFVectorSN::preconditioner_solve(const double* const b, double* const x) {
  check(ds_precond(gpu(), b, x), "ds_precond");
}
void WFVectorSN::A_product(const double* const a, double* const b) {
  check(ds_A_product(gpu(), DS_PREC_F64, a, b), "ds_A_product");
}
void WFVectorSN::A_residual(const double* const x, double* const r) {
  check(ds_A_residual(gpu(), x, r), "ds_A_residual");
}
PCGResult WFVectorSN::solve_pcg(const PCGParams& prm, double* const x) {
  return run_solve(gpu(), prm, x, false, 0.0);
}

WFVectorW::WFVectorW(const int num_pixels, const WPixel* const pixels_,
                     const SignalCovarParams* const s_params_)
    : WFVector(num_pixels, s_params_), pixels(pixels_) {
  DS_BIND_OPS(WFVectorW);
}
WFVectorW::~WFVectorW() {
  if (gpu_) ds_problem_destroy(gpu_);
}
ds_problem* WFVectorW::gpu() {
  if (!gpu_)
    gpu_ = make_problem(DS_FORM_W, n, (const double*)pixels, s_params);
  return gpu_;
}
double WFVectorW::A(const int i, const int j, const bool exact) {
  return pixels[i].w * element(pixels, i, j, s_params, exact) * pixels[j].w +
         (i == j);
}
double WFVectorW::norm_b() {
  double v = 0.0;
  check(ds_norm_b(gpu(), &v), "ds_norm_b");
  return v;
}
void WFVectorW::preconditioner_solve(const double* const b, double* const x) {
  check(ds_precond(gpu(), b, x), "ds_precond");
}
void WFVectorW::A_product(const double* const a, double* const b) {
  check(ds_A_product(gpu(), DS_PREC_F64, a, b), "ds_A_product");
}
void WFVectorW::A_residual(const double* const x, double* const r) {
  check(ds_A_residual(gpu(), x, r), "ds_A_residual");
}
PCGResult WFVectorW::solve_pcg(const PCGParams& prm, double* const x) {
  return run_solve(gpu(), prm, x, false, 0.0);
}
PCGResult WFVectorW::solve_pcg_mixed(const PCGParams& prm, double* const x,
                                     const double drop) {
  return run_solve(gpu(), prm, x, true, drop);
}

// ---- map.h / map_covar.h -------------------------------------------------

void smp_product(const int num_pixel_points, const Point* const pixel_coords,
                 const int num_map_points, const Point* const map_coords,
                 const SignalCovarParams* const s_params,
                 const double* const x, double* const m) {
  CovarView cv(s_params);
  check(ds_smp_product(dachshund_gpu_context(), num_pixel_points,
                       (const double*)pixel_coords, num_map_points,
                       (const double*)map_coords, &cv.c, x, m),
        "ds_smp_product");
}

void sigma_m_noise_dom(const int n_pix, const NPixel* pixels, const int n_map,
                       const Point* map_coords,
                       const SignalCovarParams* s_params, double* sigma_m) {
  CovarView cv(s_params);
  check(ds_sigma_m_noise_dom(dachshund_gpu_context(), n_pix,
                             (const double*)pixels, n_map,
                             (const double*)map_coords, &cv.c, sigma_m),
        "ds_sigma_m_noise_dom");
}

// One map cell's column solve (lib/map_covar.cc:40-71).  The column of
// S(map_i, .) is O(n) and is built here on the host; the solve runs on the
// GPU.  MapCovarDiag::solve_pcg below does every cell on the device.
PCGResult MapCovarDiag::solve_column_pcg(const int i,
                                         const PCGParams& pcg_params,
                                         double* const x) {
  std::vector<NPixel> col(n_pix);
  const Point mp = map_coords[i];
  for (int k = 0; k < n_pix; ++k) {
    const double dx = mp.x - pixels[k].x;
    const double dy = mp.y - pixels[k].y;
    const double dz = mp.z - pixels[k].z;
    col[k] = NPixel{pixels[k].x, pixels[k].y, pixels[k].z, pixels[k].N,
                    signal_covar(dx * dx + dy * dy, dz * dz, s_params)};
  }
  WFVectorSN wf(n_pix, col.data(), s_params);
  return wf.solve_pcg(pcg_params, x);
}

std::vector<PCGResult> MapCovarDiag::solve_pcg(const PCGParams& pcg_params,
                                               double* const sigma_m) {
  CovarView cv(s_params);
  ds_pcg_params c{pcg_params.n, pcg_params.max_iter, pcg_params.step_r,
                  pcg_params.tol, pcg_params.verbose};
  std::vector<int> iters(n_map), ok(n_map);
  check(ds_map_covar_diag(dachshund_gpu_context(), n_pix,
                          (const double*)pixels, n_map,
                          (const double*)map_coords, &cv.c, &c, sigma_m,
                          iters.data(), ok.data()),
        "ds_map_covar_diag");
  std::vector<PCGResult> out(n_map);
  for (int i = 0; i < n_map; ++i) {
    out[i].success = ok[i] != 0;
    out[i].num_iters = iters[i];
  }
  return out;
}

// ---- closed-form solves (lib/wf_vector.cc:123-155, :276-319) -------------

void WFVectorSN::solve_cf(double* const x, const bool verbose,
                          const bool exact) {
  if (verbose) puts("Dense solve on the GPU (Cholesky).");
  check(ds_solve_cf(gpu(), exact ? 1 : 0, x), "ds_solve_cf");
}

void WFVectorW::solve_cf(double* const x, const bool verbose,
                         const bool exact) {
  if (verbose) puts("Dense solve on the GPU (Cholesky).");
  check(ds_solve_cf(gpu(), exact ? 1 : 0, x), "ds_solve_cf");
}
