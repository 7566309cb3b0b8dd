/*
 * ds_oracle.c -- CPU FP64 restatement of dachshund's Wiener-filter hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the CPU-baseline leg of the benchmark may load this; the product never does.
 *
 * Build with -ffp-contract=off: the reference is built for x86-64 without
 * FMA, so every a*b+c below is a rounded multiply followed by a rounded add.
 * Each function cites the reference routine whose operation order it keeps.
 */
#include "ds_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* lib/signal_covar.h:19-41: c = 1/(2 l^2); table over [0, 4^2/2] with 200
 * points, f_i = exp(-dx*i) where dx = 8/199. */
void ora_covar_init(ora_covar* p, double var_s, double l_perp, double l_para) {
  p->var_s = var_s;
  p->c_perp = 1.0 / (2.0 * l_perp * l_perp);
  p->c_para = 1.0 / (2.0 * l_para * l_para);
  const double fx_max = 4.0 * 4.0 / 2.0;
  p->table_n = ORA_TABLE_N;
  const double ft_dx = fx_max / (ORA_TABLE_N - 1);
  p->dx_inv = 1.0 / ft_dx;
  for (int i = 0; i < ORA_TABLE_N; ++i) {
    double xi = ft_dx * i;
    p->table[i] = exp(-xi);
  }
}

/* lib/signal_covar.h:53-65.  The reference truncates xx to int and returns 0
 * when that index exceeds n-2; for xx >= 0 that is exactly xx >= n-1, which
 * is tested first here so huge arguments never overflow the int cast. */
double ora_linterp(double x, int n, double dx_inv, const double* f) {
  double xx = x * dx_inv;
  if (!(xx < (double)(n - 1))) return 0.0;
  int ix = (int)xx;
  return f[ix] * ((double)(ix + 1) - xx) + f[ix + 1] * (xx - (double)ix);
}

/* lib/signal_covar.h:67-77 */
double ora_signal_covar(double x_perp_2, double x_para_2, const ora_covar* p) {
  const double ga_perp = p->c_perp * x_perp_2;
  const double ga_para = p->c_para * x_para_2;
  const double gp = ora_linterp(ga_perp + ga_para, p->table_n, p->dx_inv,
                                p->table);
  return p->var_s * gp;
}

/* lib/signal_covar.h:79-86 */
double ora_signal_covar_exact(double x_perp_2, double x_para_2,
                              const ora_covar* p) {
  const double ga_perp = p->c_perp * x_perp_2;
  const double ga_para = p->c_para * x_para_2;
  return p->var_s * exp(-(ga_perp + ga_para));
}

/* Pair geometry exactly as lib/wf_vector.cc:62-66: dx,dy,dz then
 * (dx*dx + dy*dy) and dz*dz. */
static inline double pair_S(const double* pi, const double* pj,
                            const ora_covar* p, int exact) {
  const double dx = pi[0] - pj[0];
  const double dy = pi[1] - pj[1];
  const double dz = pi[2] - pj[2];
  const double x_perp_2 = (dx * dx + dy * dy);
  const double x_para_2 = dz * dz;
  return exact ? ora_signal_covar_exact(x_perp_2, x_para_2, p)
               : ora_signal_covar(x_perp_2, x_para_2, p);
}

/* lib/wf_vector.cc:16-40 (SN: S_ij + N_ii d_ij), :157-183 (W: w_i S w_j + d_ij) */
double ora_A_elem(int form, int n, const double* pix, const ora_covar* p,
                  int i, int j, int exact) {
  (void)n;
  const int delta_ij = (i == j);
  const double* pi = pix + 5 * (size_t)i;
  const double* pj = pix + 5 * (size_t)j;
  const double S_ij = pair_S(pi, pj, p, exact);
  if (form == ORA_FORM_SN) return S_ij + pi[3] * delta_ij;
  return pi[3] * S_ij * pj[3] + delta_ij;
}

/* lib/wf_vector.cc:51-74 (SN) and :196-220 (W) */
void ora_A_product(int form, int n, const double* pix, const ora_covar* p,
                   const double* a, double* b) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n; ++i) {
    const double* pi = pix + 5 * (size_t)i;
    double bi = 0.0;
    for (int j = 0; j < n; ++j) {
      const int delta_ij = (i == j);
      const double* pj = pix + 5 * (size_t)j;
      const double S_ij = pair_S(pi, pj, p, 0);
      double A_ij;
      if (form == ORA_FORM_SN) A_ij = S_ij + pi[3] * delta_ij;
      else A_ij = pi[3] * S_ij * pj[3] + delta_ij;
      bi += A_ij * a[j];
    }
    b[i] = bi;
  }
}

/* lib/wf_vector.cc:76-103 (SN, b = d) and :222-251 (W, b = wd) */
void ora_A_residual(int form, int n, const double* pix, const ora_covar* p,
                    const double* x, double* r) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n; ++i) {
    const double* pi = pix + 5 * (size_t)i;
    double sum_Ax_i = 0.0;
    for (int j = 0; j < n; ++j) {
      const int delta_ij = (i == j);
      const double* pj = pix + 5 * (size_t)j;
      const double S_ij = pair_S(pi, pj, p, 0);
      double A_ij;
      if (form == ORA_FORM_SN) A_ij = S_ij + pi[3] * delta_ij;
      else A_ij = pi[3] * S_ij * pj[3] + delta_ij;
      sum_Ax_i += A_ij * x[j];
    }
    r[i] = pi[4] - sum_Ax_i;
  }
}

/* lib/wf_vector.cc:42-49 (SN: 1/(var_s+N)), :185-194 (W: 1/(w^2 var_s + 1)) */
void ora_precond(int form, int n, const double* pix, const ora_covar* p,
                 const double* b, double* x) {
  const double var_s = p->var_s;
  for (int i = 0; i < n; ++i) {
    const double* pi = pix + 5 * (size_t)i;
    double Apc_ii;
    if (form == ORA_FORM_SN) Apc_ii = var_s + pi[3];
    else Apc_ii = pi[3] * pi[3] * var_s + 1.0;
    x[i] = b[i] / Apc_ii;
  }
}

/* lib/wf_vector.cc:105-112, :253-260 */
double ora_norm_b(int form, int n, const double* pix) {
  (void)form;
  double sum_b2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double bi = pix[5 * (size_t)i + 4];
    sum_b2 += bi * bi;
  }
  return sqrt(sum_b2);
}

/* lib/linalg.cc:9-17 */
double ora_dot(int n, const double* a, const double* b) {
  double dot = 0.0;
  for (int i = 0; i < n; ++i) dot += a[i] * b[i];
  return dot;
}
/* lib/linalg.cc:21-29 */
double ora_norm(int n, const double* x) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += x[i] * x[i];
  return sqrt(s);
}
/* lib/linalg.cc:32-39 (computes c += a*b despite its comment) */
void ora_add_cv(int n, double a, const double* b, double* c) {
  for (int i = 0; i < n; ++i) c[i] += a * b[i];
}
/* lib/linalg.cc:42-49 */
void ora_sub_cv(int n, double a, const double* b, double* c) {
  for (int i = 0; i < n; ++i) c[i] -= a * b[i];
}

/* lib/pcg.h:27-65 */
static void pcg_step(int form, int n, const double* pix, const ora_covar* p,
                     double* d, double* q, double* r, double* s, double* x,
                     int full_residual, double* delta) {
  ora_A_product(form, n, pix, p, d, q);
  double alpha = *delta / ora_dot(n, d, q);
  ora_add_cv(n, alpha, d, x);
  if (full_residual) ora_A_residual(form, n, pix, p, x, r);
  else ora_sub_cv(n, alpha, q, r);
  ora_precond(form, n, pix, p, r, s);
  double delta_old = *delta;
  *delta = ora_dot(n, r, s);
  double beta = *delta / delta_old;
  for (int i = 0; i < n; ++i) d[i] = s[i] + beta * d[i];
}

/* lib/pcg.h:67-160, plus the W-form x = w .* y of lib/wf_vector.cc:268-271 */
int ora_solve_pcg(int form, int n, const double* pix, const ora_covar* p,
                  int max_iter, int step_r, double tol, double* x,
                  int* num_iters, double* residual_norms, int* n_norms) {
  int nn = 0;
  const double norm_b = ora_norm_b(form, n, pix);
  int success;
  int iter = 0;
  if (norm_b <= 0.0) {
    success = 0;
  } else {
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* s = (double*)malloc(sizeof(double) * (size_t)n);
    double* q = (double*)malloc(sizeof(double) * (size_t)n);
    const double tol_norm_b = tol * norm_b;
    ora_A_residual(form, n, pix, p, x, r);
    ora_precond(form, n, pix, p, r, d);
    double delta = ora_dot(n, r, d);
    double sum_r2 = 0.0;
    for (int i = 0; i < n; ++i) sum_r2 += r[i] * r[i];
    double norm_r = sqrt(sum_r2);
    if (residual_norms) residual_norms[nn] = norm_r;
    nn++;
    for (iter = 1; iter < max_iter; ++iter) {
      int full_residual = iter % step_r == 0;
      pcg_step(form, n, pix, p, d, q, r, s, x, full_residual, &delta);
      sum_r2 = 0.0;
      for (int i = 0; i < n; ++i) sum_r2 += r[i] * r[i];
      norm_r = sqrt(sum_r2);
      if (residual_norms) residual_norms[nn] = norm_r;
      nn++;
      if (norm_r <= tol_norm_b) break;
    }
    free(d); free(r); free(s); free(q);
    success = 1;
    if (iter == max_iter) success = 0;
  }
  if (form == ORA_FORM_W) {
    for (int i = 0; i < n; ++i) x[i] *= pix[5 * (size_t)i + 3];
  }
  if (num_iters) *num_iters = iter;
  if (n_norms) *n_norms = nn;
  return success;
}

void ora_A_dense(int form, int n, const double* pix, const ora_covar* p,
                 int exact, double* A) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      A[(size_t)i * n + j] = ora_A_elem(form, n, pix, p, i, j, exact);
}

/* In-place Cholesky A = L L^T (lower triangle), then two triangular solves. */
int ora_cholesky_solve(int n, double* A, const double* b, double* x) {
  for (int k = 0; k < n; ++k) {
    double* Ak = A + (size_t)k * n;
    double dkk = Ak[k];
    for (int t = 0; t < k; ++t) dkk -= Ak[t] * Ak[t];
    if (!(dkk > 0.0)) return -1;
    const double lkk = sqrt(dkk);
    Ak[k] = lkk;
    int i;
#pragma omp parallel for schedule(static)
    for (i = k + 1; i < n; ++i) {
      double* Ai = A + (size_t)i * n;
      double v = Ai[k];
      for (int t = 0; t < k; ++t) v -= Ai[t] * Ak[t];
      Ai[k] = v / lkk;
    }
  }
  for (int i = 0; i < n; ++i) {
    const double* Ai = A + (size_t)i * n;
    double v = b[i];
    for (int t = 0; t < i; ++t) v -= Ai[t] * x[t];
    x[i] = v / Ai[i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double v = x[i];
    for (int t = i + 1; t < n; ++t) v -= A[(size_t)t * n + i] * x[t];
    x[i] = v / A[(size_t)i * n + i];
  }
  return 0;
}

/* lib/wf_vector.cc:123-155 (SN) and :276-319 (W, x = w .* y) */
int ora_solve_cf(int form, int n, const double* pix, const ora_covar* p,
                 int exact, double* x) {
  double* A = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* b = (double*)malloc(sizeof(double) * (size_t)n);
  ora_A_dense(form, n, pix, p, exact, A);
  for (int i = 0; i < n; ++i) b[i] = pix[5 * (size_t)i + 4];
  int rc = ora_cholesky_solve(n, A, b, x);
  if (form == ORA_FORM_W)
    for (int i = 0; i < n; ++i) x[i] = pix[5 * (size_t)i + 3] * x[i];
  free(A);
  free(b);
  return rc;
}

/* lib/map.cc:20-43 */
void ora_smp_product(int n_pix, const double* pix_pts, int pix_stride,
                     int n_map, const double* map_pts, int map_stride,
                     const ora_covar* p, const double* x, double* m) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n_map; ++i) {
    const double* mi = map_pts + (size_t)map_stride * i;
    double mv = 0.0;
    for (int j = 0; j < n_pix; ++j) {
      const double S_ij = pair_S(mi, pix_pts + (size_t)pix_stride * j, p, 0);
      mv += S_ij * x[j];
    }
    m[i] = mv;
  }
}

/* lib/map_covar.cc:11-38: sigma_i = sum_j S_ij * (1/N_j) * S_ij */
void ora_sigma_m_noise_dom(int n_pix, const double* npix, int n_map,
                           const double* map_pts, int map_stride,
                           const ora_covar* p, double* sigma_m) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n_map; ++i) {
    const double* mi = map_pts + (size_t)map_stride * i;
    double sv = 0.0;
    for (int j = 0; j < n_pix; ++j) {
      const double* pj = npix + 5 * (size_t)j;
      const double W_jj = 1.0 / pj[3];
      const double S_ij = pair_S(mi, pj, p, 0);
      sv += S_ij * W_jj * S_ij;
    }
    sigma_m[i] = sv;
  }
}

/* lib/map_covar.cc:47-61 */
void ora_spm_column(int n_pix, const double* npix, const double* map_pt,
                    const ora_covar* p, double* col) {
  for (int k = 0; k < n_pix; ++k)
    col[k] = pair_S(map_pt, npix + 5 * (size_t)k, p, 0);
}

/* lib/map_covar.cc:73-113 */
void ora_map_covar_diag(int n_pix, const double* npix, int n_map,
                        const double* map_pts, int map_stride,
                        const ora_covar* p, int max_iter, int step_r,
                        double tol, double* sigma_m, int* num_iters_out,
                        int* success_out) {
  double* x = (double*)malloc(sizeof(double) * (size_t)n_pix);
  double* col = (double*)malloc(sizeof(double) * 5 * (size_t)n_pix);
  for (int i = 0; i < n_map; ++i) {
    const double* mi = map_pts + (size_t)map_stride * i;
    sigma_m[i] = 0.0;
    for (int j = 0; j < n_pix; ++j) x[j] = 0.0;
    for (int k = 0; k < n_pix; ++k) {
      const double* pk = npix + 5 * (size_t)k;
      col[5 * (size_t)k + 0] = pk[0];
      col[5 * (size_t)k + 1] = pk[1];
      col[5 * (size_t)k + 2] = pk[2];
      col[5 * (size_t)k + 3] = pk[3];
      col[5 * (size_t)k + 4] = pair_S(mi, pk, p, 0);
    }
    int it = 0;
    int ok = ora_solve_pcg(ORA_FORM_SN, n_pix, col, p, max_iter, step_r, tol,
                           x, &it, NULL, NULL);
    if (num_iters_out) num_iters_out[i] = it;
    if (success_out) success_out[i] = ok;
    if (ok) {
      double sv = 0.0;
      for (int j = 0; j < n_pix; ++j)
        sv += pair_S(mi, npix + 5 * (size_t)j, p, 0) * x[j];
      sigma_m[i] += sv;
    }
  }
  free(x);
  free(col);
}

/* lib/util.cc:25-46: cell centres, index (ix*ny + iy)*nz + iz */
void ora_init_uniform_map_coords(double lx, double ly, double lz, int nx,
                                 int ny, int nz, double* coords) {
  const double dx = lx / nx, dy = ly / ny, dz = lz / nz;
  for (int ix = 0; ix < nx; ++ix) {
    const double x = dx * (ix + 0.5);
    for (int iy = 0; iy < ny; ++iy) {
      const double y = dy * (iy + 0.5);
      for (int iz = 0; iz < nz; ++iz) {
        const double z = dz * (iz + 0.5);
        size_t i = ((size_t)ix * ny + iy) * nz + iz;
        coords[3 * i + 0] = x;
        coords[3 * i + 1] = y;
        coords[3 * i + 2] = z;
      }
    }
  }
}

/* lib/pixel.cc:23-71; each record is read fully before it is written, so the
 * conversions may run in place like the reference's. */
void ora_pixel_to_npixel(int n, const double* p, double* np) {
  for (int i = 0; i < n; ++i) {
    const double* a = p + 5 * (size_t)i;
    double* o = np + 5 * (size_t)i;
    const double x = a[0], y = a[1], z = a[2], nn = a[3], d = a[4];
    o[0] = x; o[1] = y; o[2] = z; o[3] = nn * nn; o[4] = d;
  }
}
void ora_npixel_to_pixel(int n, const double* np, double* p) {
  for (int i = 0; i < n; ++i) {
    const double* a = np + 5 * (size_t)i;
    double* o = p + 5 * (size_t)i;
    const double x = a[0], y = a[1], z = a[2], N = a[3], d = a[4];
    o[0] = x; o[1] = y; o[2] = z; o[3] = sqrt(N); o[4] = d;
  }
}
void ora_pixel_to_wpixel(int n, const double* p, double* wp) {
  for (int i = 0; i < n; ++i) {
    const double* a = p + 5 * (size_t)i;
    double* o = wp + 5 * (size_t)i;
    const double x = a[0], y = a[1], z = a[2], ni = a[3], d = a[4];
    o[0] = x; o[1] = y; o[2] = z; o[3] = 1.0 / ni; o[4] = d / ni;
  }
}
void ora_wpixel_to_pixel(int n, const double* wp, double* p) {
  for (int i = 0; i < n; ++i) {
    const double* a = wp + 5 * (size_t)i;
    double* o = p + 5 * (size_t)i;
    const double x = a[0], y = a[1], z = a[2], wi = a[3], wd = a[4];
    o[0] = x; o[1] = y; o[2] = z; o[3] = 1.0 / wi; o[4] = wd / wi;
  }
}

int ora_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
