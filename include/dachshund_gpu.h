/*
 * dachshund_gpu.h -- C ABI of the gfx950 Wiener-filter library
 * (libdachshund_gpu.so).
 *
 * Plain pointers and sizes only.  Every int-returning function returns a
 * ds_status (0 = DS_OK); on error ds_ctx_last_error() describes it and
 * nothing is exit()ed.  A context owns one HIP stream; contexts and problems
 * are not internally locked (one per host thread, like the reference's
 * single-threaded callers).  All host arrays are caller-owned.
 *
 * Pixel records are the reference's 5-double AoS structs:
 *   DS_FORM_W : WPixel {x, y, z, w = 1/n, wd = d/n}  (lib/pixel.h:39-45)
 *   DS_FORM_SN: NPixel {x, y, z, N = n^2, d}          (lib/pixel.h:27-33)
 * Points are {x, y, z} records (lib/pixel.h:8-12).  Host vectors are in the
 * caller's pixel order; the library orders pixels spatially on the device
 * and undoes the permutation on the way out.
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
/* FP32 pair kernel: table argument expanded on the VALU, or on the matrix
 * cores (v_mfma_f32_16x16x4f32, with 16-row sub-block culling; the
 * default; DS_F32_MFMA_CLAMPED is accepted as the same kernel).  A
 * subvolume with a point far from its tile frame always uses direct
 * differences on the VALU for its own blocks (decided per subvolume). */
enum { DS_F32_VALU = 0, DS_F32_MFMA = 1, DS_F32_MFMA_CLAMPED = 2 };
/* Preconditioner of a problem's device solves: the diagonal of A (the
 * reference's Jacobi; the default) or the exact inverse of each 64-row
 * target block of A (block-Jacobi). */
enum { DS_PRECOND_JACOBI = 0, DS_PRECOND_BLOCK = 1 };

/* Arithmetic of the pair evaluation.
 * DS_PREC_F64: the reference's FP64 table interpolation; the table argument
 *              and cutoff decision of every pair are bit-identical to
 *              lib/signal_covar.h:53-77 on the same coordinates.
 * DS_PREC_F32: the same interpolant evaluated in FP32 on pre-scaled,
 *              block-relative coordinates, FP32 row sums, FP64 epilogue. */
enum { DS_PREC_F64 = 0, DS_PREC_F32 = 1 };

/* Covariance parameters as SignalCovarParams holds them
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
  int n_norms;           /* residual norms recorded */
  int n_true;            /* mixed solver: FP64 true residuals taken */
  int n_matvec64;        /* FP64 pair-evaluation passes */
  int n_matvec32;        /* FP32 pair-evaluation passes */
  double norm_b;
  double final_norm;     /* last |r| recorded */
  double true_norm;      /* mixed solver: last FP64 true residual |b - A x| */
  double seconds;        /* device time of the solve (HIP events) */
} ds_pcg_result;

/* Statistics of a problem's blocking and culling. */
typedef struct {
  int n;
  int n_blocks;              /* target blocks of 64 rows */
  int n_tiles;               /* source tiles of 16 slots */
  long long tile_pairs;      /* (block, tile) pairs kept after the cutoff */
  long long pairs_per_matvec;/* evaluated pairs per matvec = tile_pairs*1024 */
  long long sub_pairs_per_matvec; /* the same with 16-row sub-block culling
                                     (pair products the MFMA FP32 kernel
                                     delivers) */
  double setup_seconds;      /* host time of ordering, culling and upload */
  long long sub_pairs_evaluated;  /* of those, evaluated by the MFMA FP32
                                     kernel (each tile pair once, used for
                                     both products) */
  long long f64_pairs_per_matvec; /* pairs the FP64 kernel evaluates per
                                     matvec at its current tile shape */
} ds_problem_stats;

/* ---- host-only planning (no device needed) ---------------------------- */
/* The spatial ordering and cutoff culling a problem of these points would
 * use: perm_out[npad] (sorted slot -> point index, padding repeats the last
 * point), tile_ptr[n_blocks+1] and up to idx_cap source-tile indices per
 * the CSR list.  *n_blocks, *n_idx receive the sizes (call once with
 * idx_cap = 0 to size the buffers; perm_out may then be NULL).  Points are
 * records of `stride` doubles starting {x, y, z}.  No reference
 * counterpart: it exposes the exact culling that replaces the reference's
 * all-pairs loops (lib/wf_vector.cc:54-73) for inspection. */
int ds_plan_tiles(int n, const double* pts, int stride, const ds_covar* covar,
                  int* perm_out, int* tile_ptr, int* tile_idx, int idx_cap,
                  int* n_blocks, int* n_idx);

/* ---- context ---------------------------------------------------------- */
int ds_ctx_create(int device, ds_ctx** out);
/* Releases the caller's hold on ctx; the context (its stream) is freed once
 * every problem and warm start created on it has been destroyed too, in
 * whatever order the caller (or a garbage collector) destroys them.  The
 * caller must not use ctx after this call. */
int ds_ctx_destroy(ds_ctx* ctx);
const char* ds_ctx_last_error(const ds_ctx* ctx);
/* Waits for all work the context's calls have enqueued (its main stream
 * and the second stream that block-Jacobi builds run on). */
int ds_ctx_synchronize(ds_ctx* ctx);
int ds_ctx_num_cu(const ds_ctx* ctx);
/* Where the problems created on ctx get their ordering, culled tile lists,
 * per-slot arrays and column-sum slots: on the device (0, the default) or
 * from the host threads (1), the reference implementation of the device
 * builder, which gives identical arrays (ds_problem_get_geometry).  No
 * reference counterpart: both replace the all-pairs loops of
 * lib/wf_vector.cc:54-73. */
int ds_ctx_set_host_geometry(ds_ctx* ctx, int on);

/* ---- pixel problem: replaces the construction of WFVectorSN / WFVectorW
 *      (lib/wf_vector.h:33-96).  Pixels and covariance are copied to HBM,
 *      ordered into blocks, and the cutoff-culled tile lists are built. --- */
int ds_problem_create(ds_ctx* ctx, int form, int n, const double* pixels,
                      const ds_covar* covar, ds_problem** out);
int ds_problem_destroy(ds_problem* p);
int ds_problem_get_stats(const ds_problem* p, ds_problem_stats* out);
/* A problem's geometry arrays, as the pair and CG kernels read them (no
 * reference counterpart; for inspection and for comparing the device and
 * host builders).  Copies array `which` (DS_GEOM_*) to out when cap_bytes
 * holds it; *bytes receives its size (out may be NULL to ask for it). */
enum {
  DS_GEOM_PERM = 0,      /* int per slot: caller index, -1 for padding */
  DS_GEOM_POS = 1,       /* double4 per slot: x, y, z, w (W) or N (SN) */
  DS_GEOM_RHS = 2,       /* double per slot */
  DS_GEOM_DIAG = 3,      /* double per slot: diagonal of A */
  DS_GEOM_WVEC = 4,      /* double per slot */
  DS_GEOM_LOC32 = 5,     /* float4 per slot: source, tile frame */
  DS_GEOM_TGT32 = 6,     /* float4 per slot: target, block frame */
  DS_GEOM_TCOUNT = 7,    /* int per tile */
  DS_GEOM_TBOX_LO = 8,   /* 3 doubles per tile */
  DS_GEOM_TBOX_HI = 9,
  DS_GEOM_TPTR = 10,     /* int per block + 1: list offsets */
  DS_GEOM_TIDX = 11,     /* int per entry: source tile */
  DS_GEOM_TMASK = 12,    /* byte per entry: live 16-row sub-blocks */
  DS_GEOM_TOFF32 = 13,   /* float4 per entry: frame offset, mask | tile<<4 */
  DS_GEOM_HPTR = 14,     /* int per block + 1: half-list offsets */
  DS_GEOM_HREC = 15,     /* float4 per half-list record */
  DS_GEOM_GPTR = 16,     /* int per tile + 1: column-sum slot offsets */
  DS_GEOM_GSLOT = 17,    /* int per half-list record (-1: none) */
  DS_GEOM_BLK_ORDER = 18,/* int per block: FP64 launch order */
  DS_GEOM_ORD32 = 19,    /* int per 2 blocks: FP32 launch lists */
  DS_GEOM_KNN = 20,      /* int per slot x 32: the warm start's nearest
                            neighbours (slots, nearest first, -1 = none);
                            built by the first prediction */
  DS_GEOM_MINV = 21      /* float x 33 x 64 per block: the block-Jacobi
                            inverse by diagonals, row k holding
                            M_{i,(i+k) mod 64} (DS_PRECOND_BLOCK) */
};
int ds_problem_get_geometry(const ds_problem* p, int which, void* out,
                            long long cap_bytes, long long* bytes);
/* Pair-kernel launch shape: waves (of 64 lanes) per target block, 1, 2 or 4.
 * Changes only the order of floating-point sums. */
int ds_problem_set_waves(ds_problem* p, int waves_per_block);

/* ---- batches of independent subvolumes (no reference counterpart) ------
 * n_sub subvolumes with sub_n[k] pixels each (records consecutive in
 * `pixels`) and covariance covars[k] (all with the same table), held in one
 * problem: each is ordered and culled on its own, every pair kernel and CG
 * kernel processes all of them in one launch, and each has its own CG
 * scalars and stopping test.  The operator entry points (A_product, ...,
 * set_x/get_x) act on the concatenated vector; a one-subvolume batch is an
 * ordinary problem. */
int ds_batch_create(ds_ctx* ctx, int form, int n_sub, const int* sub_n,
                    const double* pixels, const ds_covar* covars,
                    ds_problem** out);
int ds_batch_num_sub(const ds_problem* p);
/* Evaluated pairs per matvec of each subvolume: FP64 kernel (at its
 * current tile shape, ds_problem_set_tile) and MFMA FP32 kernel (16-row
 * sub-block culling). */
int ds_batch_sub_pairs(const ds_problem* p, long long* pairs64,
                       long long* pairs32);
/* Solve every subvolume from the resident x (W space for the W form); res
 * gets one entry per subvolume.  mixed: the FP32/FP64 solver of
 * ds_solve_mixed with refine_drop.  use_graph: capture the iterations into
 * a hipGraph and replay it (mixed solver). */
int ds_batch_solve(ds_problem* p, const ds_pcg_params* params, int mixed,
                   double refine_drop, int use_graph, ds_pcg_result* res);
/* Pair-kernel launch shape: waves per block (1, 2, 4; changes the order of
 * a row's floating-point sums) and whether blocks are launched longest tile
 * list first (the default; changes no result). */
int ds_problem_set_launch(ds_problem* p, int waves_per_block,
                          int order_by_cost);
/* Preconditioner of the problem's solves and of ds_precond (no reference
 * counterpart for DS_PRECOND_BLOCK: the reference's preconditioner is
 * Jacobi, lib/wf_vector.cc:42-49, :185-194).  The block inverses are built
 * on the device on first selection (8.4 KB of FP32 per 64 rows: the
 * symmetric inverse stored by diagonals), on the context's second stream:
 * the call returns at once and the build overlaps the work enqueued next
 * until the first solve or ds_precond, which waits for it. */
int ds_problem_set_precond(ds_problem* p, int kind);
/* FP32 pair kernel variant (DS_F32_VALU or DS_F32_MFMA). */
int ds_problem_set_f32_kernel(ds_problem* p, int kind);
/* FP64 pair kernel tile shape (no reference counterpart; the loop it
 * tiles is lib/wf_vector.cc:199-219): rows per group x 16 source slots.
 * 64: one group per 64-row target block, every listed (block, tile) pair
 * evaluated; 32 or 16: the block's 16-row sub-blocks that fail the cutoff
 * box test against a tile are skipped (lanes hold 16 or 32 rows x 4 or 2
 * source columns).  Changes only the order of a row's sums. */
int ds_problem_set_tile(ds_problem* p, int f64_rows);
/* The current launch settings (any pointer may be NULL). */
int ds_problem_get_launch(const ds_problem* p, int* waves, int* order,
                          int* f64_rows, int* f32_kind);
/* Time every launch shape on this problem's pair kernel of precision prec
 * and keep the fastest: waves per block (1, 2, 4) x launch order x, for
 * FP32, the kernel form (VALU 64 x 16 tiles or MFMA 16 x 16 sub-block
 * tiles) or, for FP64, the tile rows (64, 32, 16).  On an error the
 * previous settings are restored.  ds_problem_get_launch reports the
 * choice. */
int ds_problem_autotune(ds_problem* p, int prec, int reps, int* waves_out,
                        int* order_out, double* ms_out);

/* Explicit BF16 covariance block for a batch whose subvolumes (at most 32)
 * share one geometry, e.g. several noise realizations of the same pixels
 * (no reference counterpart; the dense matrix the reference builds only in
 * solve_cf, lib/wf_vector.cc:276-290).  S is built once on the device in
 * BF16 (2 * n^2 bytes) and the mixed solver's inner operator becomes one
 * MFMA GEMM S [w .* d_1 ... w .* d_R] with FP32 accumulation per iteration;
 * FP64 matrix-free true residuals keep the stopping rule.  set_dense
 * switches between it and the matrix-free FP32 operator.  The block's
 * relative error is about 2^-9 (BF16 S and directions), so its solves want
 * a refine_drop of about 0.1 (1e-3 suits the FP32 operator). */
int ds_problem_build_dense_bf16(ds_problem* p);
int ds_problem_set_dense(ds_problem* p, int on);
/* The same multi-RHS operator without materializing S (no reference
 * counterpart): per (64-row block, 16-source tile) of the shared geometry
 * the interpolant tile is evaluated once in FP32 on the VALU, rounded to
 * BF16 and multiplied into all right-hand sides by MFMA with FP32
 * accumulation.  Needs no explicit block; switching it on switches the
 * explicit block off. */
int ds_problem_set_fused(ds_problem* p, int on);
/* b = A a through the selected multi-RHS operator (all subvolumes). */
int ds_problem_dense_product(ds_problem* p, const double* a, double* b);
/* Average duration of the selected multi-RHS operator's kernel (the block
 * GEMM, or the fused kernel with its finalize; HIP events). */
int ds_problem_time_dense(ds_problem* p, int reps, double* ms_per_launch);

/* WFVector*::A_product (lib/wf_vector.cc:51-74 SN, :196-220 W): b = A a. */
int ds_A_product(ds_problem* p, int prec, const double* a, double* b);
/* WFVector*::A_residual (lib/wf_vector.cc:76-103 SN, :222-251 W):
 * r = b - A x with b = d (SN) or wd (W).  FP64 only. */
int ds_A_residual(ds_problem* p, const double* x, double* r);
/* WFVector*::preconditioner_solve (lib/wf_vector.cc:42-49, :185-194). */
int ds_precond(ds_problem* p, const double* b, double* x);
/* x = M^-1 b with the block-Jacobi preconditioner as the mixed solver
 * applies it (b rounded to FP32, products and sums in FP64, its own
 * cross-lane order).  No reference counterpart: exposed so that tests can
 * pin the mixed solver's application.  DS_E_STATE unless DS_PRECOND_BLOCK
 * is selected. */
int ds_precond_r32(ds_problem* p, const double* b, double* x);
/* WFVector*::norm_b (lib/wf_vector.cc:105-112, :253-260), summed in the
 * reference's order on the host. */
int ds_norm_b(ds_problem* p, double* out);

/* WFVector*::solve_pcg (lib/wf_vector.cc:114-121, :263-274) over pcg_solve
 * (lib/pcg.h:67-160), run on the device in FP64 with the reference's
 * iteration numbering, step_r true-residual refresh and stop rule.
 * x: in = initial guess (W form: y0 in W space), out = solution (W form:
 * x = w .* y).  residual_norms (may be NULL) receives up to norms_cap norms. */
int ds_solve_pcg(ds_problem* p, const ds_pcg_params* params, double* x,
                 double* residual_norms, int norms_cap, ds_pcg_result* res);

/* Mixed-precision solve of the same system (no reference counterpart):
 * FP32 pair evaluation inside PCG with FP64 vectors and dot products; the
 * initial residual b - A x0 is formed with the FP32 operator (matrix-free
 * problems; the shared-geometry block paths form it in FP64); when the
 * recurrence residual has fallen by `refine_drop` since the last
 * (initial or FP64 true) residual, or meets the tolerance, or step_r
 * iterations have passed since that one, an FP64 true residual replaces
 * it.  Stops only when the FP64 true residual meets |r| <= tol |b|.  Same x
 * in/out convention as ds_solve_pcg. */
int ds_solve_mixed(ds_problem* p, const ds_pcg_params* params,
                   double refine_drop, double* x, double* residual_norms,
                   int norms_cap, ds_pcg_result* res);

/* Device-resident solution vector (sorted order in HBM).  set/get move it
 * from/to host order; the *_resident solves start from it (W space for the
 * W form) and leave the solution (x = w .* y for W) in it. */
int ds_problem_set_x(ds_problem* p, const double* x);
int ds_problem_get_x(ds_problem* p, double* x);
int ds_solve_pcg_resident(ds_problem* p, const ds_pcg_params* params,
                          ds_pcg_result* res);
int ds_solve_mixed_resident(ds_problem* p, const ds_pcg_params* params,
                            double refine_drop, ds_pcg_result* res);
/* Set every entry of the resident x to `value` on the device (no host
 * transfer).  No reference counterpart (the reference's callers zero their
 * host x, app/dachshund.cc:237-238). */
int ds_problem_fill_x(ds_problem* p, double value);
/* Time `reps` back-to-back launches of the pair matvec kernel (prec
 * DS_PREC_F64 or DS_PREC_F32) with HIP events on the problem's stream; the
 * resident x is the input.  *ms_per_launch = average launch duration.  No
 * reference counterpart: measurement hook. */
int ds_problem_time_matvec(ds_problem* p, int prec, int reps,
                           double* ms_per_launch);
/* Residual history of the last solve on this problem. */
int ds_problem_get_norms(ds_problem* p, double* norms, int cap, int* count);
/* The delta = r . M^-1 r after each iteration of the last solve (entry 0:
 * the initial one), the value lib/pcg.h:112-139's verbose trace prints
 * beside |r| (pcg_solve keeps it in a local).  Same count as the norms. */
int ds_problem_get_deltas(ds_problem* p, double* deltas, int cap, int* count);

/* smp_product (lib/map.cc:20-43) onto the uniform grid of
 * init_uniform_map_coords (lib/util.cc:25-46), with the problem's pixels as
 * sources and its resident solution as x:  m_i = sum_j S(map_i, pix_j) x_j.
 * m: host array of nx*ny*nz doubles, index (ix*ny + iy)*nz + iz. */
int ds_problem_smp_grid(ds_problem* p, double lx, double ly, double lz,
                        int nx, int ny, int nz, double* m);

/* ds_problem_smp_grid without device allocations.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * per-call device buffer -- the source tiles' boxes, the map tiles' list
 * counts, offsets and entries, and the map before its copy to m -- lives in
 * one static device array of DS_NOALLOC_SCRATCH_BYTES that each GPU the
 * library runs on holds until the process ends; the call makes no device
 * or host allocation.  Returns DS_E_ARG, having written nothing to m, when
 * the grid and its lists do not fit (8 bytes per map cell plus 4 per list
 * entry plus the boxes and offsets; the 2.53M-cell 60 x 48 x 880 grid of
 * a 76,800-pixel problem fits).  Calls on one GPU share the array: calls
 * that overlap on one GPU race and must be serialized by the caller.  The
 * map equals ds_problem_smp_grid's bit for bit. */
#define DS_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int ds_problem_smp_grid_noalloc(ds_problem* p, double lx, double ly,
                                double lz, int nx, int ny, int nz, double* m);

/* smp_product (lib/map.cc:20-43) with arbitrary pixel and map points. */
int ds_smp_product(ds_ctx* ctx, int n_pix, const double* pix_pts, int n_map,
                   const double* map_pts, const ds_covar* covar,
                   const double* x, double* m);
/* smp_product onto the uniform map grid (lib/util.cc:25-46). */
int ds_smp_product_grid(ds_ctx* ctx, int n_pix, const double* pix_pts,
                        double lx, double ly, double lz, int nx, int ny,
                        int nz, const ds_covar* covar, const double* x,
                        double* m);
/* sigma_m_noise_dom (lib/map_covar.cc:11-38): sigma_i = sum_j S_ij^2 / N_j,
 * NPixel input, arbitrary map points. */
int ds_sigma_m_noise_dom(ds_ctx* ctx, int n_pix, const double* npix,
                         int n_map, const double* map_pts,
                         const ds_covar* covar, double* sigma);

/* solve_cf (lib/wf_vector.cc:123-155 SN, :276-319 W): dense A built on the
 * device (table interpolant, or the exact exponential when exact != 0, as
 * WFVector*::A(i, j, exact)), Cholesky-factored on the device (the
 * reference uses Eigen's LDLT), x = A^-1 b (W form: x = w .* y).  Needs
 * 8 n^2 bytes of device memory; DS_E_STATE if A is not positive definite. */
int ds_solve_cf(ds_problem* p, int exact, double* x);

/* MapCovarDiag::solve_pcg (lib/map_covar.cc:73-113): for every map point,
 * an SN-form PCG solve with the column S(map_i, .) as right-hand side from
 * x = 0, then sigma_i = S(map_i, .) . x (0 when the solve did not succeed).
 * NPixel input; num_iters / success (may be NULL) get n_map entries. */
int ds_map_covar_diag(ds_ctx* ctx, int n_pix, const double* npix, int n_map,
                      const double* map_pts, const ds_covar* covar,
                      const ds_pcg_params* params, double* sigma,
                      int* num_iters, int* success);

/* ---- learned warm start (no reference counterpart) --------------------
 * A point network predicts each pixel's solution value x_i from its 32
 * nearest neighbours (offsets / l, noise / sqrt(var_s), data / sqrt(var_s))
 * and its own noise and datum; the prediction becomes the initial guess of
 * the solve (W form: y0 = x0 / w), the x in/out argument the reference's
 * solve already takes (lib/pcg.h:98, app/dachshund.cc:237-244).  The
 * stopping rule is unchanged.  Network: 162 -> 64 -> 64 -> 64 -> 1, ReLU,
 * FP32 weights, BF16 activations; trained on the device by its own forward,
 * backward and Adam kernels on solutions the solver produced. */
typedef struct ds_warmstart ds_warmstart;
int ds_ws_num_params(void);
int ds_ws_feature_dim(void);
int ds_ws_create(ds_ctx* ctx, unsigned long long seed, ds_warmstart** out);
int ds_ws_destroy(ds_warmstart* ws);
int ds_ws_set_params(ds_warmstart* ws, const float* params, int n);
int ds_ws_get_params(ds_warmstart* ws, float* params, int n);
/* BF16 rounding of hidden activations (default on). */
int ds_ws_set_quantized(ds_warmstart* ws, int on);
/* Features of every pixel of p, host order, n x ds_ws_feature_dim(). */
int ds_ws_features(ds_warmstart* ws, ds_problem* p, float* feat_out);
/* Predict the initial guess into p's resident x (W space for the W form).
 * Returns once the work is enqueued: it is ordered before every later call
 * on p's context (also when ws lives on another context); errors of the
 * kernels surface at the next call that waits for the device. */
int ds_ws_predict(ds_warmstart* ws, ds_problem* p);
/* Add p as a training sample: its features and, as target, its resident x,
 * which must hold a solution (x = w .* y after a W-form solve).  Samples
 * stay in device memory, rows shuffled with `seed`. */
int ds_ws_add_sample(ds_warmstart* ws, ds_problem* p, unsigned long long seed);
int ds_ws_num_samples(ds_warmstart* ws, long long* count);
int ds_ws_clear_samples(ds_warmstart* ws);
/* Adam on the mean squared error over all samples; lr * lr_decay^epoch;
 * loss_per_epoch (may be NULL) gets the mean loss of each epoch. */
int ds_ws_train(ds_warmstart* ws, int epochs, int batch, float lr,
                float lr_decay, unsigned long long seed,
                double* loss_per_epoch);
/* Forward and backward on n given feature rows (host, rounded to BF16 on
 * upload) and targets: predictions, the gradient of the mean squared error
 * (ds_ws_num_params() floats) and the loss.  For gradient checks. */
int ds_ws_eval_batch(ds_warmstart* ws, int n, const float* feat,
                     const float* target, float* pred, float* grad,
                     double* loss);

#ifdef __cplusplus
}
#endif
#endif
