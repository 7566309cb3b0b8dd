/*
 * ds_oracle.h -- CPU FP64 restatement of dachshund's Wiener-filter hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libdachshund_gpu.so and the C++ drop-in above it) never links,
 * loads or falls back to it.
 *
 * Every function restates one reference routine, cited file:line, in the same
 * order of floating-point operations (compile with -ffp-contract=off), so that
 * on the same inputs it reproduces the reference bit for bit.  Row loops are
 * OpenMP-parallel like the reference's; each row is summed in j order.
 *
 * Pixel arrays are the reference's AoS records of 5 doubles:
 *   form DS_FORM_SN : NPixel {x, y, z, N = n^2, d}      (lib/pixel.h:27-33)
 *   form DS_FORM_W  : WPixel {x, y, z, w = 1/n, wd=d/n} (lib/pixel.h:39-45)
 *   raw             : Pixel  {x, y, z, n, d}            (lib/pixel.h:16-22)
 * Point arrays are {x, y, z} (lib/pixel.h:8-12); functions taking points also
 * take a stride in doubles so Pixel arrays can be read as points.
 */
#ifndef DS_ORACLE_H
#define DS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_TABLE_N 200

enum { ORA_FORM_SN = 0, ORA_FORM_W = 1 };

typedef struct {
  double var_s;
  double c_perp;   /* 1/(2 l_perp^2) */
  double c_para;   /* 1/(2 l_para^2) */
  int table_n;
  double dx_inv;
  double table[ORA_TABLE_N];
} ora_covar;

/* lib/signal_covar.h:19-41 */
void ora_covar_init(ora_covar* p, double var_s, double l_perp, double l_para);
/* lib/signal_covar.h:53-65 */
double ora_linterp(double x, int n, double dx_inv, const double* f);
/* lib/signal_covar.h:67-77 */
double ora_signal_covar(double x_perp_2, double x_para_2, const ora_covar* p);
/* lib/signal_covar.h:79-86 */
double ora_signal_covar_exact(double x_perp_2, double x_para_2,
                              const ora_covar* p);

/* lib/wf_vector.cc:16-40 (SN) and :157-183 (W): one element of A. */
double ora_A_elem(int form, int n, const double* pix, const ora_covar* p,
                  int i, int j, int exact);
/* lib/wf_vector.cc:51-74 (SN), :196-220 (W) */
void ora_A_product(int form, int n, const double* pix, const ora_covar* p,
                   const double* a, double* b);
/* lib/wf_vector.cc:76-103 (SN), :222-251 (W) */
void ora_A_residual(int form, int n, const double* pix, const ora_covar* p,
                    const double* x, double* r);
/* lib/wf_vector.cc:42-49 (SN), :185-194 (W) */
void ora_precond(int form, int n, const double* pix, const ora_covar* p,
                 const double* b, double* x);
/* lib/wf_vector.cc:105-112 (SN), :253-260 (W) */
double ora_norm_b(int form, int n, const double* pix);

/* lib/linalg.cc:9-49 */
double ora_dot(int n, const double* a, const double* b);
double ora_norm(int n, const double* x);
void ora_add_cv(int n, double a, const double* b, double* c);
void ora_sub_cv(int n, double a, const double* b, double* c);

/*
 * lib/pcg.h:67-160 driven by the operators of `form`, and for the W form the
 * final x *= w of lib/wf_vector.cc:263-274 (SN: :114-121).
 * residual_norms must hold max_iter+1 doubles (or be NULL); *n_norms gets the
 * number written.  Returns success (0/1).  num_iters is 0 on the |b|<=0 early
 * return (the reference leaves it uninitialised there).
 */
int ora_solve_pcg(int form, int n, const double* pix, const ora_covar* p,
                  int max_iter, int step_r, double tol, double* x,
                  int* num_iters, double* residual_norms, int* n_norms);

/* Dense A (row-major n*n) via ora_A_elem, and a dense SPD solve by Cholesky
 * (the closed-form path of lib/wf_vector.cc:123-155,276-319; the reference
 * uses Eigen LDLT, any exact factorisation of an SPD matrix agrees to
 * rounding).  Returns 0 on success, -1 if A is not positive definite. */
void ora_A_dense(int form, int n, const double* pix, const ora_covar* p,
                 int exact, double* A);
int ora_cholesky_solve(int n, double* A, const double* b, double* x);
/* solve_cf: W form returns x = w .* y like lib/wf_vector.cc:316-318. */
int ora_solve_cf(int form, int n, const double* pix, const ora_covar* p,
                 int exact, double* x);

/* lib/map.cc:20-43.  pix_pts / map_pts are {x,y,z,...} records with the given
 * stride in doubles (3 for Point, 5 for Pixel). */
void ora_smp_product(int n_pix, const double* pix_pts, int pix_stride,
                     int n_map, const double* map_pts, int map_stride,
                     const ora_covar* p, const double* x, double* m);
/* lib/map_covar.cc:11-38 (NPixel input) */
void ora_sigma_m_noise_dom(int n_pix, const double* npix, int n_map,
                           const double* map_pts, int map_stride,
                           const ora_covar* p, double* sigma_m);
/* lib/map_covar.cc:40-71 column RHS build: d_k = S(map_i, pix_k) */
void ora_spm_column(int n_pix, const double* npix, const double* map_pt,
                    const ora_covar* p, double* col);
/* lib/map_covar.cc:73-113: one SN-form PCG per map cell, then S_mp row . x.
 * num_iters_out / success_out may be NULL, else hold n_map entries. */
void ora_map_covar_diag(int n_pix, const double* npix, int n_map,
                        const double* map_pts, int map_stride,
                        const ora_covar* p, int max_iter, int step_r,
                        double tol, double* sigma_m, int* num_iters_out,
                        int* success_out);

/* lib/util.cc:25-46 */
void ora_init_uniform_map_coords(double lx, double ly, double lz, int nx,
                                 int ny, int nz, double* coords);
/* lib/pixel.cc:23-71, in place allowed */
void ora_pixel_to_npixel(int n, const double* p, double* np);
void ora_npixel_to_pixel(int n, const double* np, double* p);
void ora_pixel_to_wpixel(int n, const double* p, double* wp);
void ora_wpixel_to_pixel(int n, const double* wp, double* p);

/* Number of OpenMP threads the row loops use (1 if built without OpenMP). */
int ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
