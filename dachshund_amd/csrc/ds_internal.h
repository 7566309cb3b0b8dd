// ds_internal.h -- shared types of the gfx950 Wiener-filter library.
//
// Layout in HBM (per problem, all in *sorted* pixel order, npad entries):
//   pos64   double4  {x, y, z, w}   w = 1/n (W form) or N = n^2 (SN form)
//   rhs64   double   wd (W form) or d (SN form); 0 on pad slots
//   loc32   float4   source coordinates {X, Y, Z, X^2+Y^2+Z^2} relative to
//                    the frame (per-axis median) of the pixel's tile,
//                    pre-scaled so that squared distances are table
//                    arguments (X = x sqrt(c_perp dx_inv), Z = z sqrt(...))
//   tgt32   float4   target coordinates relative to the pixel's block frame
//   toff32  float4   per list entry: block frame - tile frame
//   tptr/tidx        CSR list, per target block of DS_BS = 64 rows, of the
//                    source tiles (DS_TS = 16 consecutive slots) that can
//                    hold a nonzero covariance under the exact table cutoff
// Pixels are ordered by a k-d tree into blocks of 64, and each block again
// into 4 tiles of 16.  A problem may hold several independent subvolumes
// (a batch): each occupies a contiguous range of blocks, ordered and culled
// on its own, with its own covariance constants and CG scalars; a single
// volume is the one-subvolume case.  The last block of each subvolume is
// padded (copies of its last pixel); every kernel writes 0 to padded rows
// and every vector keeps 0 there, so padded columns contribute nothing.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/dachshund_gpu.h"

#define DS_BS 64          // target rows per block (one wavefront)
#define DS_TS 16          // source slots per tile
#define DS_TPB (DS_BS / DS_TS)
// block-Jacobi inverse of one block, stored by diagonals (symmetric):
// DS_BJ_ROWS rows of DS_BS floats
#define DS_BJ_ROWS (DS_BS / 2 + 1)
#define DS_BJ_SIZE (DS_BJ_ROWS * DS_BS)
#define DS_MAX_TABLE 1024
#define DS_MAP_T 4        // map tiles are 4x4x4 cells
#define DS_DENSE_R 32     // right-hand sides of the explicit BF16 block path
#define DS_DENSE_SPLIT_MAX 8   // largest split-K count of its GEMM

namespace ds {

// Covariance parameters as the kernels see them.
struct CovarDev {
  double var_s;
  double c_perp;
  double c_para;
  double dx_inv;
  int table_n;
  double s_perp;   // sqrt(c_perp * dx_inv)
  double s_para;   // sqrt(c_para * dx_inv)
};

// Quantization of a point set for the k-d ordering: per axis the set's
// minimum, 2^32 - 1 over its extent (0 for a flat axis) and the weight that
// turns a quantized extent back into a metric one (scale x extent / 2^32-1).
struct KdQuant {
  double lo[3], inv[3], w[3];
};
KdQuant kd_quant(const double* pts, int stride, int n, const double scale[3]);
void kd_quant_axis(double lo, double hi, double scale, KdQuant& Q, int a);

// Quantized coordinate (0 .. 2^32 - 1), monotone in x; host and device
// evaluate the same two IEEE operations.
__host__ __device__ inline unsigned kd_q(double x, double lo, double inv) {
  const double t = (x - lo) * inv;
  return t >= 4294967295.0 ? 0xFFFFFFFFu : (unsigned)t;
}
// Split key: quantized coordinate, then point index (unique per point).
__host__ __device__ inline unsigned long long kd_key(unsigned q, int idx) {
  return ((unsigned long long)q << 32) | (unsigned)idx;
}
// Split axis: the largest weighted quantized extent, the first on ties.
__host__ __device__ inline int kd_axis(const unsigned lo[3],
                                       const unsigned hi[3],
                                       const double w[3]) {
  int axis = 0;
  double best = -1.0;
  for (int a = 0; a < 3; ++a) {
    const double ext = (double)(hi[a] - lo[a]) * w[a];
    if (ext > best) {
      best = ext;
      axis = a;
    }
  }
  return axis;
}

// Host-side result of spatial blocking.
struct Blocking {
  int n = 0;
  int nblk = 0;                 // blocks of DS_BS
  int ntile = 0;                // tiles of DS_TS (= nblk * DS_TPB)
  int npad = 0;
  std::vector<int> perm;        // sorted slot -> original index (npad)
  std::vector<double> blo, bhi; // nblk * 3, physical AABB of real pixels
  std::vector<double> tlo, thi; // ntile * 3 (empty tiles: lo > hi)
  std::vector<int> tcount;      // real pixels per tile
};

// Sort points (stride doubles apart, x,y,z first) into k-d blocks of DS_BS,
// each split again into DS_TPB tiles.  scale[3] weights the axes when
// choosing splits (the covariance metric).
void build_blocking(const double* pts, int stride, int n, const double scale[3],
                    Blocking& out);

// Exact culling: for each target box keep the source tiles that may hold a
// pair with table index < table_n - 1 under the reference's arithmetic.
// With sublo/subhi (DS_TPB sub-boxes per target, 3 x DS_TPB x ntgt) and
// mask, each kept entry also gets the bit mask of the sub-boxes that pass
// the same test, and entries no sub-box passes are dropped.
void build_tile_lists(int ntgt, const double* tlo, const double* thi,
                      const Blocking& src, const CovarDev& cv,
                      std::vector<int>& ptr, std::vector<int>& idx,
                      const double* sublo = nullptr,
                      const double* subhi = nullptr,
                      std::vector<unsigned char>* mask = nullptr);

// Lower bound of the table argument xx for any pair drawn from two boxes,
// with the operation order of lib/signal_covar.h:67-77 and
// lib/wf_vector.cc:62-66.
double box_gap_xx(const double* alo, const double* ahi, const double* blo,
                  const double* bhi, const CovarDev& cv);
// Culling threshold on box_gap_xx: an entry is kept when xx < cull_limit.
double cull_limit(const CovarDev& cv);

CovarDev make_covar(const ds_covar* c);

}  // namespace ds

// Device-resident scalars of one CG solve (FP64).
struct ds_cg_scalars {
  double delta;        // r . M^-1 r
  double alpha;
  double beta;
  double dq;           // d . A d
  double norm_r;       // |r|_2 of the current residual
  double tol_norm_b;
  double ref_norm;     // mixed solver: |r| at the last true residual
  double true_norm;    // mixed solver: last FP64 true residual norm
  double saved_delta_old;  // mixed solver: delta before the step whose
                           // residual a true residual replaces
  int iter;            // iterations completed (reference numbering)
  int done;            // 1 once converged or out of iterations
  int success;
  int num_iters;
  int n_norms;
  int max_iter;
  int n_true;          // mixed solver: FP64 true residuals taken
  int want_true;       // mixed solver: next step recomputes r in FP64
  int last_true;       // mixed solver: iteration of the last true residual
  int refresh_every;   // mixed solver: true residual at least this often
  int pad;
};

struct ds_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;   // second stream (default priority): work
                               // that may overlap the main stream's
  char err[512] = {0};         // last error (fixed storage: setting it
                               // allocates nothing)
  int host_geometry = 0;       // build problems' geometry on the host
  int num_cu = 0;
  double* scratch = nullptr;   // reduction partials
  size_t scratch_n = 0;
  // the device address of the no-alloc entry points' static scratch array
  // on this context's device (looked up once per device at creation)
  unsigned char* noalloc = nullptr;
  // owner + every live problem and warm start created on this context: the
  // context is freed when the last of them lets go (so destroying it before
  // its problems, as a garbage collector may, frees nothing they still use)
  std::atomic<int> refs{1};
  // Device buffers of destroyed problems (records, lists, vectors, block
  // inverses, neighbour lists), kept for the next problems created on this
  // context; at most kBufCacheBytes in all.  Reuse saves the hipMalloc
  // calls, and a problem of the same size gets the same memory back: the
  // pair kernels run up to ~7 % slower on memory newly placed by hipMalloc
  // later in a process than on the first problem's.
  std::mutex buf_mu;
  std::vector<std::pair<size_t, void*>> buf_cache;
  size_t buf_cache_bytes = 0;
  // the same for pinned host buffers (hipHostMalloc pins pages: ~1 ms for
  // the staging buffer of a 300k-pixel problem); at most kHostCacheBytes
  std::vector<std::pair<size_t, void*>> host_cache;
  size_t host_cache_bytes = 0;
  // DS_POOL_POISON=1 at creation: buffers returned to the cache are filled
  // with 0xFF bytes, so that code reading a buffer it did not write first
  // sees NaNs and wild indices instead of a previous problem's values
  int pool_poison = 0;
};

struct ds_problem;

namespace ds {
void ctx_retain(ds_ctx* c);
void ctx_release(ds_ctx* c);
constexpr size_t kBufCacheBytes = size_t(1) << 30;
// A device buffer of at least `bytes`: the smallest cached one of at most
// bytes + bytes / 8 + 64 KiB, else a new hipMalloc; *cap is its size.
hipError_t ctx_buf_get(ds_ctx* c, void** p, size_t bytes, size_t* cap);
// Return a buffer of size `cap` (the stream work that used it has
// completed) to the cache, or free it when it does not fit.
void ctx_buf_put(ds_ctx* c, void* p, size_t cap);
constexpr size_t kHostCacheBytes = size_t(256) << 20;
// A problem's own device buffers (host: pinned host buffers): taken from
// its context's cache and registered with the problem, which returns them
// when it is freed.
hipError_t pool_alloc(ds_problem* p, void** ptr, size_t bytes,
                      bool host = false);
void pool_free(ds_problem* p, void* ptr);
template <class T>
hipError_t palloc(ds_problem* p, T** ptr, size_t count) {
  *ptr = nullptr;
  return pool_alloc(p, (void**)ptr, sizeof(T) * (count ? count : 1));
}
template <class T>
void pfree(ds_problem* p, T*& ptr) {
  if (ptr) pool_free(p, (void*)ptr);
  ptr = nullptr;
}
// A temporary device buffer from the context's pool, returned when it goes
// out of scope once the context's stream has finished with it.
template <class T>
struct PoolTmp {
  ds_ctx* c = nullptr;
  T* p = nullptr;
  size_t cap = 0;
  hipError_t alloc(ds_ctx* ctx, size_t n) {
    c = ctx;
    return ctx_buf_get(ctx, (void**)&p, sizeof(T) * (n ? n : 1), &cap);
  }
  ~PoolTmp() {
    if (!p) return;
    (void)hipStreamSynchronize(c->stream);
    ctx_buf_put(c, (void*)p, cap);
  }
};
// A copy of device memory into pageable host memory (dst, bytes) in parts
// enqueued on st as their data is produced (enqueue: bytes [off, off + len)
// of src), complete when finish() returns.  Large copies go through a
// pinned buffer of the context's host pool: each part's DMA runs on the
// context's second stream once st has produced it, and host threads copy
// it on as soon as it arrives, while st computes the next part.
struct HostCopy {
  ds_ctx* c = nullptr;
  void* dst = nullptr;
  void* stage = nullptr;
  size_t cap = 0, bytes = 0;
  hipStream_t st = nullptr;
  bool direct = true;
  std::vector<std::pair<size_t, size_t>> parts;
  std::vector<hipEvent_t> ev;
  hipError_t begin(ds_ctx* ctx, void* d, size_t n, hipStream_t s);
  hipError_t enqueue(const void* src, size_t off, size_t len);
  hipError_t finish();
  ~HostCopy();
};
}  // namespace ds

struct ds_problem {
  ds_ctx* ctx = nullptr;
  // buffers taken with ds::pool_alloc: pointer, size, pinned host memory
  struct Pooled {
    void* ptr;
    size_t cap;
    bool host;
  };
  std::vector<Pooled> pooled;
  int form = DS_FORM_W;
  int n = 0, nblk = 0, ntile = 0, npad = 0;
  ds::CovarDev cv{};                   // subvolume 0 (table shared by all)
  std::vector<double> table;           // host copy, table_n entries
  std::vector<int> perm;               // slot -> caller index, -1 for padding
  int* perm_dev = nullptr;             // the same on the device
  double* tbox_lo = nullptr;           // per tile: box of its real pixels
  double* tbox_hi = nullptr;           // (3 each; empty tiles lo > hi)
  long long nslot = 0;                 // column-sum slots (gptr[ntile])
  long long ntile_pairs = 0;           // (block, tile) pairs after culling
  long long nsub_pairs = 0;            // (16-row sub-block, tile) pairs
  long long nhalf_pairs = 0;           // (32-row half block, tile) pairs
  std::vector<long long> sub_pairs64, sub_pairs32;   // per subvolume
  std::vector<long long> sub_half64;   // per subvolume, 32-row FP64 tiles
  double norm_b = 0.0;                 // |b| of subvolume 0
  // subvolumes
  int n_sub = 1;
  std::vector<int> sub_n;              // pixels per subvolume
  std::vector<int> sub_off;            // first caller index per subvolume
  std::vector<int> sub_blk0_h;         // first block per subvolume (n_sub+1)
  std::vector<ds::CovarDev> sub_cv_h;
  std::vector<double> sub_norm_b;      // |b| per subvolume, reference order
  std::vector<int> blk_len_h;          // list length per block
  int* blk_sub = nullptr;              // subvolume of each block
  int* blk_nreal = nullptr;            // real rows of each block
  unsigned char* tmask = nullptr;      // per list entry: live sub-blocks
  float* minv = nullptr;               // block-Jacobi inverses, 64 x 64 each
  int minv_mod = 1;                    // block I uses inverse I % minv_mod
                                       // (shared geometry: one set)
  int precond_kind = DS_PRECOND_JACOBI;
  bool use_minv() const { return precond_kind == DS_PRECOND_BLOCK && minv; }
  // the inverses are built on the context's second stream; the first
  // kernel on the main stream that reads them waits for this event
  hipEvent_t minv_ev = nullptr;
  mutable bool minv_pending = false;
  int* knn_nbr = nullptr;              // learned warm start: K nearest
                                       // rows per row, built on first use
  size_t minv_bytes = 0;               // the inverses' bytes
  float knn_r0 = 1.0f;                 // its first ring radius
  int* sub_blk0 = nullptr;             // first block per subvolume
  double4* sub_cv = nullptr;           // {var_s, c_perp, c_para, 0}
  int* tcount = nullptr;               // real slots per tile
  int* done_count = nullptr;           // subvolumes finished
  int* blk_order = nullptr;            // launch order of target blocks
  bool shared_geometry = false;        // all subvolumes: same pixels/weights
  // explicit BF16 block path (shared geometry, <= DS_DENSE_R subvolumes)
  bool use_dense = false;
  bool use_fused = false;              // fused S tiles x all RHS (MFMA)
  unsigned short* dense_s = nullptr;   // S, dense_mpad x dense_ld, BF16
  float* dense_y = nullptr;            // split-K partials of S U
  unsigned short* du16 = nullptr;      // BF16 scaled directions, R x ld
  int dense_ld = 0, dense_mpad = 0, dense_split = 4;
  double setup_seconds = 0.0;
  int waves_per_block = 4;             // pair-kernel launch shape
  int f64_rows = 16;                   // FP64 kernel tile: rows per group
                                       // (64, or 16 / 32 with sub-block
                                       // culling) x 16 source slots
  int order_by_cost = 1;               // launch blocks longest list first
  // device buffers
  double4* pos64 = nullptr;
  double* rhs64 = nullptr;
  double* diag = nullptr;              // Jacobi diagonal of A
  float4* loc32 = nullptr;             // sources: tile frame, .w = |X|^2
  float4* tgt32 = nullptr;             // targets: block frame
  // FP32 kernel launch lists: blocks of subvolumes whose points all sit
  // near their tile frames first (MFMA or expanded VALU form), then those
  // of the rest (direct differences); longest list first in [0, nblk),
  // block order in [nblk, 2 nblk)
  int* ord32 = nullptr;
  int n32_fast = 0;                    // blocks in the first part
  // Symmetric MFMA FP32 kernel: each unordered pair of 16-slot tiles is
  // evaluated once.  Block B keeps the entries of tiles T >= 4 B, live
  // bits restricted to its sub-blocks s with 4 B + s <= T (records as
  // toff32); sub-blocks with 4 B + s < T also give the transposed product,
  // whose 16 column sums per entry go to slot gslot[e] of pbuf.  The slots
  // of a tile are consecutive (gptr[T] .. gptr[T + 1]), in block order, so
  // the gather reads one contiguous run per tile.
  int* hptr = nullptr;
  float4* hrec = nullptr;
  int* gslot = nullptr;                // per half-list entry (-1: none)
  float* pbuf = nullptr;               // 16 floats per slot
  double* ydir = nullptr;              // per row: the block's own sums
  int* gptr = nullptr;
  int* sub_sym = nullptr;              // per subvolume: blocks run the
                                       // symmetric kernel (no far points)
  long long nsub_eval = 0;             // live (entry, sub-block) evaluated
  int f32_kind = DS_F32_MFMA;          // FP32 pair kernel: MFMA or VALU form
  int* tptr = nullptr;
  int* tidx = nullptr;
  float4* toff32 = nullptr;            // per list entry: centre offsets
  double2* table64 = nullptr;          // {T[k], T[k+1]-T[k]}, last = {0, 0}
  float2* table32 = nullptr;
  // FP32 MFMA kernel: {g_k, df_k}, the interval's interpolant as the line
  // g_k + df_k xx
  float2* table32g = nullptr;
  double* table_raw = nullptr;         // T[k] as given (dense/column paths)
  // CG work vectors (npad each)
  double *x = nullptr, *r = nullptr, *d = nullptr, *s = nullptr, *q = nullptr,
         *tmp = nullptr;
  double* wvec = nullptr;              // w (W form), 0 on pad slots
  double *du = nullptr, *xu = nullptr; // w .* d, w .* x (W form)
  float* du32 = nullptr;               // FP32 copy of du
  double* partials = nullptr;          // 3 * nblk: two partial vectors,
                                       // then a single volume's d . q ones
  ds_cg_scalars* sc = nullptr;         // n_sub entries
  double* norms = nullptr;             // residual histories, n_sub x cap
  double* deltas = nullptr;            // delta histories (after norms)
  int norms_cap = 0;
  ds_cg_scalars* sc_host = nullptr;    // pinned, n_sub entries (final)
  int* done_host = nullptr;            // pinned poll slots (2)
  double* norm_host = nullptr;         // pinned: |r| at chunk ends (2)
  double* host_tmp = nullptr;          // pinned npad doubles
};

namespace ds {
// A problem's ordering, culled lists, per-slot arrays and column-sum slots
// built on the device from the pixel records (geometry_dev.hip), which
// dpix holds on the device (n x 5 doubles, copied there on p's stream);
// p's sizes, subvolume arrays and covariance constants are set.  Fills
// p->perm, p->perm_dev, p->nslot and the device arrays, bstat (per block:
// list length, half-list length, live sub-block entries, live 32-row
// halves, evaluated sub-block entries) and max_src (per subvolume: largest
// |source coordinate| in its tile frame).  *error: 1 = more than INT_MAX
// entries, 2 = inconsistent symmetric lists.
hipError_t build_geometry_device(ds_problem* p, const double* dpix,
                                 const std::vector<KdQuant>& quant,
                                 std::vector<float>& max_src,
                                 std::vector<int>& bstat, int* error);
// Order `st` after the block-Jacobi build (once per build).
inline hipError_t minv_join(const ds_problem* p, hipStream_t st) {
  if (!p->minv_pending) return hipSuccess;
  p->minv_pending = false;
  return hipStreamWaitEvent(st, p->minv_ev, 0);
}
hipError_t launch_scale(const double* v, const double* w, double* u,
                        float* u32, int npad, hipStream_t st);
// Block-Jacobi: build the inverses (minv allocated); apply the problem's
// preconditioner (Jacobi or block) to a device vector.
hipError_t bj_build(ds_problem* p, hipStream_t st);
// The block-Jacobi product as the mixed solver's update applies it (b
// rounded to FP32, bj_apply_regs_r32).
hipError_t launch_precond_r32(const ds_problem* p, const double* b, double* x,
                              hipStream_t st);
hipError_t launch_precond_p(const ds_problem* p, const double* b, double* x,
                            hipStream_t st);
// Shared-geometry batches: the true residual of every subvolume that wants
// one (one pass over the tile lists for all right-hand sides), then its
// preconditioned residual and dot-product partials.  sync (the mixed
// solver's refinement): once any subvolume wants one, every live subvolume
// is marked and gets its true residual in the same pass.
hipError_t launch_res64_mrhs(ds_problem* p, const double* in,
                             const double* u, double* out, int all,
                             hipStream_t st, int sync = 0);
// s = M^-1 r and the r.s, r.r partials of the subvolumes that asked for a
// true residual (all: every live one); with q, r = rhs - q is formed first
// and stored to r.
hipError_t launch_res_precond(ds_problem* p, double* out2, int all,
                              hipStream_t st, const double* q = nullptr);
hipError_t launch_precond(const double* b, const double* diag, double* x,
                          int npad, hipStream_t st);
hipError_t run_solver(ds_problem* p, const ds_pcg_params* prm, int mixed,
                      double drop, ds_pcg_result* res, int use_graph);
hipError_t launch_fill(double* v, double value, int n, hipStream_t st);
hipError_t launch_to_bf16(const float* src, unsigned short* dst, int n,
                          hipStream_t st);
hipError_t dense_build(ds_problem* p, hipStream_t st);
hipError_t dense_gemm(ds_problem* p, hipStream_t st);
hipError_t dense_epilogue(ds_problem* p, hipStream_t st);
// Fused multi-RHS operator of a shared-geometry batch: q and the d . q
// partials of every subvolume from one pass over subvolume 0's tile lists.
hipError_t fused_mrhs(ds_problem* p, hipStream_t st);
// k_column over subvolume k of a batch (its slots, its pixel count).
hipError_t launch_column_sub(const ds_problem* p, int k,
                             const double* table_dev, const double* map_pt,
                             double* col, hipStream_t st);
hipError_t launch_column(const ds_problem* p, const double* table_dev,
                         const double* map_pt, double* col, hipStream_t st);
hipError_t launch_dot1(const double* a, const double* b, int n, double* out,
                       hipStream_t st);
hipError_t dense_solve(ds_problem* p, int exact, const double* table_dev,
                       int* info_out);
}  // namespace ds

namespace ds {

// Pair-kernel launchers (pairs64.hip / pairs32.hip).  `in` is the matvec
// input a in sorted order and `u` its scaled copy (w .* a for the W form, a
// itself for SN); out = A a, or rhs - A a with `residual`.  The CG epilogues
// fuse the reductions that follow the matvec in lib/pcg.h:35-37 (d . q and
// alpha) or :52-61 plus :123-135 (precondition, r . M^-1 r, |r|, stop rule).
enum EpiKind { EK_NONE = 0, EK_CG_DQ = 1, EK_CG_RES = 2 };
struct CgEpilogue {
  int kind = EK_NONE;
  double* out2 = nullptr;   // EK_CG_RES: M^-1 r
  int init = 0, mixed = 0, is_true = 0, need_want = 0;
  double drop = 0.0;
  // the launcher skips its finalize: the direction kernel that follows
  // folds it in (launch_dir with fold)
  int fold = 0;
};
hipError_t launch_matvec64_ex(ds_problem* p, int residual, const double* in,
                              const double* u, double* out,
                              const CgEpilogue& e, hipStream_t st);
// FP32 A in.  With cg, in = the CG direction and the d . q partials are
// stored for launch_finalize; the symmetric MFMA kernel then leaves `out`
// of its subvolumes (sub_sym) to the update kernel, which forms each row
// from ydir and the column-sum slots (launch_update with sym).  With
// fold_alpha (a single volume) the partials go to the third nblk of
// `partials` and no finalize is launched: the update kernel sums them.
hipError_t launch_matvec32_ex(ds_problem* p, const float* u32,
                              const double* in, double* out, int cg,
                              hipStream_t st, int fold_alpha = 0);
// Sum the per-block partials of the kernel just launched and finalize the
// CG scalars of every subvolume: EK_CG_DQ -> alpha; EK_CG_RES -> delta,
// beta, |r|, stop rule (init / mixed / is_true / need_want as in e).
hipError_t launch_finalize(ds_problem* p, const CgEpilogue& e,
                           hipStream_t st);

// Map-side launchers.  Sources are a blocked pixel set (pos64 / loc32 with
// w or 1/N in .w), targets the uniform grid of lib/util.cc:25-46 cut into
// 4x4x4 tiles, lists per map tile.
struct MapGrid {
  double lx, ly, lz;
  int nx, ny, nz;
  int tx, ty, tz;      // tiles per axis
};
enum MapMode { MM_SMP = 0, MM_SIGMA = 1 };
// (the map tiles with x tile index in [x0, x1); x1 < 0: all of them)
hipError_t launch_map64(int mode, const MapGrid& g, const double4* pos,
                        const double* u, const int* tptr, const int* tidx,
                        const double2* table64, const CovarDev& cv,
                        double* m, hipStream_t st, int x0 = 0, int x1 = -1);
// Source-tile lists of the grid's tiles on the device: count[T] when idx is
// null, else idx at the offsets ptr; equal to build_tile_lists' lists.
hipError_t launch_grid_lists(const MapGrid& g, const double* slo,
                             const double* shi, int ntile,
                             const CovarDev& cv, int* count, const int* ptr,
                             int* idx, hipStream_t st);
// Arbitrary target points (Point records), targets sorted by a blocking.
hipError_t launch_points64(int mode, int n_tgt, const double4* tgt,
                           const int* tgt_perm, const double4* pos,
                           const double* u, const int* tptr, const int* tidx,
                           int n_tblk, const double2* table64,
                           const CovarDev& cv, double* m, hipStream_t st);

}  // namespace ds
