/*
 * dachshund_gpu.h -- C ABI of the gfx950 Wiener-filter library
 * (libdachshund_gpu.so).
 *
 * Plain pointers and sizes only.  Every function returns a ds_status (0 = OK);
 * on error ds_ctx_last_error() describes it and nothing is exit()ed.  A context
 * owns one HIP stream; contexts and problems are not internally locked (use
 * one per host thread, like the reference's single-threaded callers).
 *
 * Pixel records are the reference's 5-double AoS structs:
 *   DS_FORM_W : WPixel {x, y, z, w = 1/n, wd = d/n}  (lib/pixel.h:39-45)
 *   DS_FORM_SN: NPixel {x, y, z, N = n^2, d}          (lib/pixel.h:27-33)
 * Points are {x, y, z} (lib/pixel.h:8-12).  Host vectors are in the caller's
 * pixel order; the library sorts pixels internally and undoes the permutation
 * on the way out.
 */
#ifndef DACHSHUND_GPU_H
#define DACHSHUND_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ds_ctx ds_ctx;
typedef struct ds_problem ds_problem;

typedef enum {
  DS_OK = 0,
  DS_E_ARG = 1,       /* invalid argument */
  DS_E_HIP = 2,       /* HIP runtime error */
  DS_E_NOMEM = 3,     /* device allocation failed */
  DS_E_STATE = 4      /* call not valid in this state */
} ds_status;

enum { DS_FORM_SN = 0, DS_FORM_W = 1 };
/* Arithmetic of the pair evaluation and matvec.
 * DS_PREC_F64: the reference's FP64 table interpolation, per-pair values
 *              bit-identical to lib/signal_covar.h:53-77.
 * DS_PREC_F32: FP32 evaluation of the same interpolant on pre-scaled
 *              block-relative coordinates, FP32 accumulation. */
enum { DS_PREC_F64 = 0, DS_PREC_F32 = 1 };

/* Covariance parameters exactly as SignalCovarParams holds them
 * (lib/signal_covar.h:10-46): var_s, gaussian_const_perp/para (= 1/(2 l^2)),
 * func_table_dx_inv, func_table_n, func_table. */
typedef struct {
  double var_s;
  double c_perp;
  double c_para;
  double dx_inv;
  int table_n;           /* 2 .. 1024 (the reference uses 200) */
  const double* table;   /* table_n values */
} ds_covar;

/* Mirrors PCGParams (lib/pcg.h:13-19). */
typedef struct {
  int n;
  int max_iter;
  int step_r;
  double tol;
  int verbose;
} ds_pcg_params;

/* Mirrors PCGResult (lib/pcg.h:21-25) plus solver statistics. */
typedef struct {
  int success;
  int num_iters;
  int n_norms;           /* residual norms written (<= norms_cap) */
  double norm_b;
  double final_norm;     /* last |r| recorded */
  double true_norm;      /* |b - A x| in FP64 at exit (mixed solver) */
  int n_matvec64;        /* FP64 pair-evaluation passes */
  int n_matvec32;        /* FP32 pair-evaluation passes */
  double seconds;        /* device time of the solve (HIP events) */
} ds_pcg_result;

/* Statistics of a problem's blocking and culling. */
typedef struct {
  int n;
  int n_blocks;
  long long block_pairs;     /* (target, source) block pairs kept */
  long long pairs_per_matvec;/* pair evaluations per matvec = block_pairs*64^2 */
  double setup_seconds;      /* host time of sort + culling + upload */
} ds_problem_stats;

/* ---- context ---------------------------------------------------------- */
int ds_ctx_create(int device, ds_ctx** out);
int ds_ctx_destroy(ds_ctx* ctx);
const char* ds_ctx_last_error(const ds_ctx* ctx);
int ds_ctx_synchronize(ds_ctx* ctx);
/* Number of compute units of the context's device. */
int ds_ctx_num_cu(const ds_ctx* ctx);

/* ---- pixel problem (replaces WFVectorSN / WFVectorW construction,
 *      lib/wf_vector.h:33-96; the pixels and covariance are copied to HBM,
 *      sorted into blocks and the cutoff-culled block-pair list is built) --- */
int ds_problem_create(ds_ctx* ctx, int form, int n, const double* pixels,
                      const ds_covar* covar, ds_problem** out);
int ds_problem_destroy(ds_problem* p);
int ds_problem_get_stats(const ds_problem* p, ds_problem_stats* out);

/* WFVector*::A_product (lib/wf_vector.cc:51-74 SN, :196-220 W): b = A a. */
int ds_A_product(ds_problem* p, int prec, const double* a, double* b);
/* WFVector*::A_residual (lib/wf_vector.cc:76-103 SN, :222-251 W):
 * r = b - A x with b = d (SN) or wd (W). */
int ds_A_residual(ds_problem* p, int prec, const double* x, double* r);
/* WFVector*::preconditioner_solve (lib/wf_vector.cc:42-49, :185-194). */
int ds_precond(ds_problem* p, const double* b, double* x);
/* WFVector*::norm_b (lib/wf_vector.cc:105-112, :253-260). */
int ds_norm_b(ds_problem* p, double* out);

/* WFVector*::solve_pcg (lib/wf_vector.cc:114-121, :263-274) over pcg_solve
 * (lib/pcg.h:67-160), run on the device in FP64 with the reference's
 * iteration numbering, step_r true-residual refresh and stop rule.
 * x: in = initial guess (W form: y0 in W space), out = solution (W form:
 * x = w .* y).  residual_norms (may be NULL) receives up to norms_cap norms. */
int ds_solve_pcg(ds_problem* p, const ds_pcg_params* params, double* x,
                 double* residual_norms, int norms_cap, ds_pcg_result* res);

/* Mixed-precision solve of the same system: FP32 pair evaluation inside
 * PCG, FP64 vectors, dot products and true residuals; whenever the
 * recurrence residual has fallen by `refine_drop` since the last true
 * residual, or meets the tolerance, the FP64 true residual replaces it (an
 * outer refinement step).  Stops only when the FP64 true residual meets
 * |r| <= tol |b|.  No reference counterpart; same x in/out convention. */
int ds_solve_mixed(ds_problem* p, const ds_pcg_params* params,
                   double refine_drop, double* x, double* residual_norms,
                   int norms_cap, ds_pcg_result* res);

/* Device-resident variants: the problem keeps its own solution vector in HBM
 * (in sorted order).  set/get move it from/to host order (x in W space for
 * the W form, like the solve_pcg input); the *_resident solves start from it
 * and leave the solution (x = w .* y for W) in it. */
int ds_problem_set_x(ds_problem* p, const double* x);
int ds_problem_get_x(ds_problem* p, double* x);
int ds_solve_pcg_resident(ds_problem* p, const ds_pcg_params* params,
                          ds_pcg_result* res);
int ds_solve_mixed_resident(ds_problem* p, const ds_pcg_params* params,
                            double refine_drop, ds_pcg_result* res);
/* Copy the residual history of the last resident solve. */
int ds_problem_get_norms(ds_problem* p, double* norms, int cap, int* count);

/* smp_product (lib/map.cc:20-43) onto the uniform grid of
 * init_uniform_map_coords (lib/util.cc:25-46), using the problem's pixels
 * as sources and its resident solution as x:  m_i = sum_j S(map_i, pix_j) x_j.
 * m is a host array of nx*ny*nz doubles, index (ix*ny + iy)*nz + iz. */
int ds_problem_smp_grid(ds_problem* p, int prec, double lx, double ly,
                        double lz, int nx, int ny, int nz, double* m);

/* smp_product (lib/map.cc:20-43) with arbitrary pixel and map points
 * (Point arrays), host in/out. */
int ds_smp_product(ds_ctx* ctx, int n_pix, const double* pix_pts, int n_map,
                   const double* map_pts, const ds_covar* covar,
                   const double* x, double* m);
/* smp_product onto the uniform map grid (lib/util.cc:25-46). */
int ds_smp_product_grid(ds_ctx* ctx, int prec, int n_pix,
                        const double* pix_pts, double lx, double ly, double lz,
                        int nx, int ny, int nz, const ds_covar* covar,
                        const double* x, double* m);
/* sigma_m_noise_dom (lib/map_covar.cc:11-38): sigma_i = sum_j S_ij^2 / N_j,
 * NPixel input, arbitrary map points. */
int ds_sigma_m_noise_dom(ds_ctx* ctx, int n_pix, const double* npix,
                         int n_map, const double* map_pts,
                         const ds_covar* covar, double* sigma);

#ifdef __cplusplus
}
#endif
#endif
