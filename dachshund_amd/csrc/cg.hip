// cg.hip -- Jacobi-preconditioned CG on the device (lib/pcg.h:27-160).
//
// COMPILE WITH -ffp-contract=off: the elementwise updates below are the
// reference's own rounded operations (x += alpha d, r -= alpha q,
// s = r / diag, d = s + beta d; lib/pcg.h:39,48,52,62-64, lib/linalg.cc:32-49).
//
// All CG scalars live in device memory (ds_cg_scalars); every kernel first
// checks `done` and returns at once after convergence, so the host can
// enqueue iterations in chunks and poll the flag without a round trip per
// iteration.  Kernels per iteration:
//   plain FP64:  matvec64+dq/alpha | update(+precond, dots, stop) | dir
//                (every step_r-th: x-update | residual64+precond+dots+stop)
//   mixed:       matvec32+dq/alpha | update(+dots, refine test) |
//                residual64 if requested (true residual) | dir
//                (a single volume: alpha formed inside the update kernel)
#include <algorithm>
#include <chrono>
#include <cstring>

#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

constexpr int kVecThreads = 256;
constexpr int kVecWaves = 4;   // target blocks per CG vector workgroup

// The CG vector kernels run one wave per target block (kVecWaves waves per
// workgroup), so every wave belongs to one subvolume and stores the partial
// sums of its target block.
struct UpdArgs {
  const double* d;
  const double* q;
  const double* diag;
  const float* minv;      // block-Jacobi inverses, or nullptr (Jacobi)
  int minv_mod;
  const double* wvec;     // w (W form) or nullptr
  const int* blk_sub;
  double* x;
  double* xu;             // w .* x (W form); may alias x
  double* r;
  double* s;
  double* partials;
  const ds_cg_scalars* sc;
  int nblk;
  int only_x;
  // q of the subvolumes whose rows the symmetric FP32 operator left to
  // this kernel (sub_sym non-null): formed here from ydir and the slots
  SymRows sym;
  const int* sub_sym;
  const double4* sub_cv;
  // FA (single volume): the pair kernel's d . q partials, one per block,
  // summed here by every workgroup instead of by a finalize kernel
  const double* dq_part;
  ds_cg_scalars* sc_w;
};

// q_i of row `lane` of block I, subvolume sv (see UpdArgs::sym).
__device__ __forceinline__ double update_q(const UpdArgs& a, int I, int sv,
                                           int lane, int i) {
  if (a.sub_sym && a.sub_sym[sv])
    return sym_row(a.sym, I, lane, a.sub_cv[sv].x, a.d[i]);
  return a.q[i];
}

// lib/pcg.h:37-58: x += alpha d; r -= alpha q; s = M^-1 r; r.s and r.r.
// BJ: block-Jacobi (a.minv set, a.only_x clear; 1: r broadcast in FP64,
// 2: in FP32, for the mixed solver), 0: Jacobi or x only;
// W: a.wvec set; SYM: a.sub_sym set (q of the flagged subvolumes formed
// here, as sym_row does, from the block's own sums and the tile's
// column-sum slots, in slot order).  The kernel is bound by the latency of
// its dependent loads (one wave per block, a few waves per SIMD), so every
// load that depends only on the row is issued first, the block's inverse
// in one batch, and the register budget keeps every block resident.
// FA (a single volume): alpha = delta / (d . q) is formed here (lib/pcg.h:
// 35-37).  Every workgroup sums the nblk partials itself, in k_finalize<1>'s
// order at 256 threads (the order a batch's finalize uses, so a volume
// solved alone still equals its solve in a batch), and the first
// workgroup records dq and alpha as k_finalize<1> does.  No other
// workgroup reads those scalars in this kernel.
template <int BJ, bool W, bool SYM, bool FA>
__global__ __launch_bounds__(64 * kVecWaves)
__attribute__((amdgpu_waves_per_eu(BJ ? 5 : 1))) void k_cg_update(
    UpdArgs a) {
  // wave-uniform: the per-block scalars come by scalar loads
  const int I = __builtin_amdgcn_readfirstlane(blockIdx.x * kVecWaves +
                                               (threadIdx.x >> 6));
  double alpha_fa = 0.0;
  if (FA) {
    __shared__ double red[kVecWaves];
    // gather_partials<1>'s sums (each thread adds its partials in index
    // order), with fewer loads in flight to stay inside this kernel's
    // register budget
    double v[1] = {0.0};
    {
      const double* pk = a.dq_part;
      const int n = a.nblk;
      int b = threadIdx.x;
      for (; b + 3 * kVecThreads < n; b += 4 * kVecThreads) {
        double t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) t[u] = pk[b + u * kVecThreads];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[0] += t[u];
      }
      for (; b < n; b += kVecThreads) v[0] += pk[b];
    }
    block_sum<1>(v, red);
    const double delta = a.sc->delta;
    // uniform: held in scalar registers across the row work
    const double al = delta / v[0];
    const long long bits = __double_as_longlong(al);
    alpha_fa = __longlong_as_double(
        ((long long)__builtin_amdgcn_readfirstlane((int)(bits >> 32)) << 32) |
        (unsigned)__builtin_amdgcn_readfirstlane((int)bits));
    if (blockIdx.x == 0 && threadIdx.x == 0 && !a.sc->done)
      finalize_alpha(a.sc_w, v[0]);
  }
  if (I >= a.nblk) return;
  const int lane = threadIdx.x & 63;
  const int i = I * DS_BS + lane;
  // The row's loads go out before the subvolume's state is known (its
  // scalars are a chain of dependent loads).  No branch on that state: a
  // finished subvolume's rows store x, xu and r back unchanged (its s and
  // partials are read by nobody).
  int gb = 0, ge = 0;
  double ys = 0.0;
  if (SYM) {
    const int T = DS_TPB * I + (lane >> 4);
    gb = a.sym.gptr[T];
    ge = a.sym.gptr[T + 1];
    ys = a.sym.ydir[i];
  }
  const double di = a.d[i], x0 = a.x[i];
  const double wi = W ? a.wvec[i] : 0.0;
  const double qd = a.q[i];
  const double r0 = a.r[i];
  // keep these loads ahead of the inverse's (loads complete in order)
  __builtin_amdgcn_sched_barrier(0);
  float m[DS_BJ_ROWS];
  if (BJ)
    load_bj_diagonals(a.minv + (size_t)(I % a.minv_mod) * DS_BJ_SIZE, m);
  const int sv = a.blk_sub[I];
  const ds_cg_scalars* sc = a.sc + sv;
  const int symq = SYM ? a.sub_sym[sv] : 0;
  const double var_s = SYM ? a.sub_cv[sv].x : 0.0;
  const int nreal = SYM ? a.sym.blk_nreal[I] : 0;
  const bool live = !sc->done;
  const double alpha = FA ? alpha_fa : sc->alpha;
  const double xi = x0 + alpha * di;
  a.x[i] = live ? xi : x0;
  if (W) a.xu[i] = wi * (live ? xi : x0);
  if (!BJ && a.only_x) return;
  double qi = qd;
  if (SYM) {
    double y = ys;
    const unsigned c = lane & 15;
    const char* pb = reinterpret_cast<const char*>(a.sym.pbuf);
    for (int k = gb; k < ge; k += 16) {
      float pv[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const unsigned off = 4u * (16u * (unsigned)min(k + q, ge - 1) + c);
        pv[q] = *reinterpret_cast<const float*>(pb + off);
      }
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (k + q < ge) y += (double)pv[q];
    }
    const double wn = W ? wi : a.sym.pos[i].w;
    double q = W ? wn * (var_s * y) + di : var_s * y + wn * di;
    if (lane >= nreal) q = 0.0;
    if (symq) qi = q;
  }
  const double ri = r0 - alpha * qi;
  a.r[i] = live ? ri : r0;
  const double si = BJ == 2   ? bj_apply_regs_r32(m, ri)
                    : BJ == 1 ? bj_apply_regs(m, ri)
                              : ri / a.diag[i];
  a.s[i] = si;
  double v[2] = {wave_sum(ri * si), wave_sum(ri * ri)};
  if ((threadIdx.x & 63) == 0) {
    a.partials[I] = v[0];
    a.partials[(size_t)a.nblk + I] = v[1];
  }
}

// The preconditioned true residual and its dot-product partials for the
// subvolumes that asked for a true residual (after launch_res64_mrhs):
// the per-subvolume residual epilogue's operations, one wave per block.
struct ResPArgs {
  const double* r;
  const double* diag;
  const float* minv;
  int minv_mod;
  const int* blk_sub;
  const int* blk_nreal;
  const ds_cg_scalars* sc;
  double* s;
  double* partials;
  int nblk;
  int all;   // every live subvolume (initial residual)
  // set: the residual is formed here, r = rhs - q, and stored to r_out
  const double* q;
  const double* rhs;
  double* r_out;
};

__global__ __launch_bounds__(64 * kVecWaves) void k_res_precond(ResPArgs a) {
  const int I = blockIdx.x * kVecWaves + (threadIdx.x >> 6);
  if (I >= a.nblk) return;
  const ds_cg_scalars* sc = a.sc + a.blk_sub[I];
  if (sc->done || !(a.all || sc->want_true)) return;
  const int lane = threadIdx.x & 63;
  const int i = I * DS_BS + lane;
  const bool real = lane < a.blk_nreal[I];
  const double bi = a.q ? a.rhs[i] - a.q[i] : a.r[i];
  if (a.q) a.r_out[i] = bi;
  const double si =
      a.minv ? bj_apply(a.minv + (size_t)(I % a.minv_mod) * DS_BJ_SIZE, bi)
             : (real ? bi / a.diag[i] : 0.0);
  a.s[i] = si;
  const double v0 = wave_sum(bi * si), v1 = wave_sum(bi * bi);
  if (lane == 0) {
    a.partials[I] = v0;
    a.partials[(size_t)a.nblk + I] = v1;
  }
}

struct FinArgs {
  const double* partials;
  const int* sub_blk0;
  ds_cg_scalars* sc;
  double* norms;
  double* deltas;
  int* done_count;
  double drop;
  int nblk, norms_cap;
  int init, mixed, is_true, need_want;
};

// One block per subvolume: sum its blocks' partials in block order and
// finalize its scalars (alpha for NV = 1, the step for NV = 2).  Skips the
// subvolumes whose producing kernel skipped them.
template <int NV>
__global__ __launch_bounds__(1024) void k_finalize(FinArgs a) {
  __shared__ double red[NV * 16];
  const int sv = blockIdx.x;
  ds_cg_scalars* sc = a.sc + sv;
  if (sc->done) return;
  if (a.need_want && !sc->want_true) return;
  const int b0 = a.sub_blk0[sv], nb = a.sub_blk0[sv + 1] - b0;
  double v[NV];
  gather_partials<NV>(v, a.partials + b0, a.nblk, nb, red);
  if (threadIdx.x != 0) return;
  if (NV == 1)
    finalize_alpha(sc, v[0]);
  else
    finalize_step(sc, a.norms + (size_t)sv * a.norms_cap,
                  a.deltas + (size_t)sv * a.norms_cap, v[0], v[NV - 1],
                  a.init, a.mixed, a.drop, a.is_true, a.done_count);
}

// Shared-geometry batches with block-Jacobi: the subvolumes share the
// block inverses, so one workgroup per geometry block stages its 8.4 KB
// inverse (33 stored diagonals) in LDS once and applies it to that block of every subvolume
// (wave w takes subvolumes w, w + kVecWaves, ...).  Same operations as
// k_cg_update, so the results are the same bit for bit.
// R32: the mixed solver's block-Jacobi (r broadcast in FP32), as
// k_cg_update<2>.
template <bool R32>
__global__ __launch_bounds__(64 * kVecWaves) void k_cg_update_shared(
    UpdArgs a, int nblk1, int n_sub) {
  __shared__ float mt[DS_BJ_SIZE];
  const int J = blockIdx.x;
  const float* src = a.minv + (size_t)J * DS_BJ_SIZE;
  for (int k = threadIdx.x; k < DS_BJ_SIZE; k += blockDim.x)
    mt[k] = src[k];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // Two subvolumes per pass: both rows' loads go out before either
  // subvolume's scalars are known (a chain of dependent loads), so their
  // latencies overlap.  As in k_cg_update, a finished subvolume's rows store
  // x, xu and r back unchanged (its s and partials are read by nobody).
  constexpr int kP = 2;
  for (int sv0 = wave; sv0 < n_sub; sv0 += kP * kVecWaves) {
    double xv[kP], dv[kP], wv[kP], qv[kP], rv[kP];
#pragma unroll
    for (int h = 0; h < kP; ++h) {
      const int sv = min(sv0 + h * kVecWaves, n_sub - 1);
      const int I = sv * nblk1 + J;
      const int i = I * DS_BS + lane;
      xv[h] = a.x[i];
      dv[h] = a.d[i];
      wv[h] = a.wvec ? a.wvec[i] : 0.0;
      qv[h] = a.only_x ? 0.0 : update_q(a, I, sv, lane, i);
      rv[h] = a.only_x ? 0.0 : a.r[i];
    }
#pragma unroll
    for (int h = 0; h < kP; ++h) {
      const int sv = sv0 + h * kVecWaves;
      if (sv >= n_sub) break;   // uniform
      const ds_cg_scalars* sc = a.sc + sv;
      const bool live = !sc->done;
      const double alpha = sc->alpha;
      const int I = sv * nblk1 + J;
      const int i = I * DS_BS + lane;
      const double xi = live ? xv[h] + alpha * dv[h] : xv[h];
      a.x[i] = xi;
      if (a.wvec) a.xu[i] = wv[h] * xi;
      if (a.only_x) continue;
      const double ri = live ? rv[h] - alpha * qv[h] : rv[h];
      a.r[i] = ri;
      const double si = (R32 ? bj_apply_r32(mt, ri) : bj_apply(mt, ri));
      a.s[i] = si;
      const double v0 = wave_sum(ri * si), v1 = wave_sum(ri * ri);
      if (lane == 0) {
        a.partials[I] = v0;
        a.partials[(size_t)a.nblk + I] = v1;
      }
    }
  }
}

struct DirArgs {
  const double* s;
  const double* wvec;
  const int* blk_sub;
  double* d;
  double* du;
  float* du32;
  unsigned short* du16;   // BF16 copy for the explicit block path
  ds_cg_scalars* sc;
  int nblk;
  // FOLD: the true residual's finalize (mixed solver) done here
  const double* partials;
  const int* sub_blk0;
  double* norms;
  double* deltas;
  int* done_count;
  double drop;
  int norms_cap;
};

__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return (unsigned short)(u >> 16);
}

// lib/pcg.h:61-64: d = s + beta d (and its scaled copies).  As in
// k_cg_update, the row's loads go out before the subvolume's scalars and
// a finished subvolume's rows store their values back unchanged.
// FOLD (mixed solver): when the subvolume took an FP64 true residual in
// this step (want_true), each wave sums its r.s and r.r partials itself,
// in one fixed order, and takes beta from them; the wave of the
// subvolume's first block also records the step (finalize_step with
// is_true), which leaves want_true set for the other waves.
template <bool W, bool FOLD>
__global__ __launch_bounds__(64 * kVecWaves) void k_cg_dir(DirArgs a) {
  const int I = __builtin_amdgcn_readfirstlane(blockIdx.x * kVecWaves +
                                               (threadIdx.x >> 6));
  if (I >= a.nblk) return;
  const int i = I * DS_BS + (threadIdx.x & 63);
  const double si = a.s[i], d0 = a.d[i];
  const double wi = W ? a.wvec[i] : 1.0;
  const int sv = a.blk_sub[I];
  ds_cg_scalars* sc = a.sc + sv;
  const bool live = !sc->done;
  double beta = sc->beta;
  if (FOLD && live && sc->want_true) {
    const int b0 = a.sub_blk0[sv], nb = a.sub_blk0[sv + 1] - b0;
    const int lane = threadIdx.x & 63;
    double rs = 0.0, rr = 0.0;
    for (int b = lane; b < nb; b += 64) {
      rs += a.partials[b0 + b];
      rr += a.partials[(size_t)a.nblk + b0 + b];
    }
    rs = wave_sum(rs);
    rr = wave_sum(rr);
    beta = rs / sc->saved_delta_old;
    if (I == b0 && lane == 0)
      finalize_step(sc, a.norms + (size_t)sv * a.norms_cap,
                    a.deltas + (size_t)sv * a.norms_cap, rs, rr, 0, 1,
                    a.drop, 1, a.done_count);
  }
  const double di = live ? si + beta * d0 : d0;
  a.d[i] = di;
  const double ui = W ? wi * di : di;
  if (W) a.du[i] = ui;
  if (a.du32) a.du32[i] = (float)ui;
  if (a.du16) a.du16[i] = to_bf16((float)ui);
}

__global__ __launch_bounds__(kVecThreads) void k_to_bf16(const float* src,
                                                         unsigned short* dst,
                                                         int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = to_bf16(src[i]);
}

// u = w .* v (w may be null: u = v), optional FP32 copy.
__global__ __launch_bounds__(kVecThreads) void k_scale(const double* v,
                                                       const double* w,
                                                       double* u, float* u32,
                                                       int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const double ui = w ? w[i] * v[i] : v[i];
  if (u) u[i] = ui;
  if (u32) u32[i] = (float)ui;
}

// lib/wf_vector.cc:42-49 / :185-194: x = b / diag.
__global__ __launch_bounds__(kVecThreads) void k_precond(const double* b,
                                                         const double* diag,
                                                         double* x, int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  x[i] = b[i] / diag[i];
}

inline dim3 vec_grid(int npad) {
  return dim3((unsigned)((npad + kVecThreads - 1) / kVecThreads));
}

}  // namespace

hipError_t launch_scale(const double* v, const double* w, double* u,
                        float* u32, int npad, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, vec_grid(npad), dim3(kVecThreads), 0, st, v, w,
                     u, u32, npad);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* src, unsigned short* dst, int n,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, vec_grid(n), dim3(kVecThreads), 0, st, src,
                     dst, n);
  return hipGetLastError();
}

hipError_t launch_precond(const double* b, const double* diag, double* x,
                          int npad, hipStream_t st) {
  hipLaunchKernelGGL(k_precond, vec_grid(npad), dim3(kVecThreads), 0, st, b,
                     diag, x, npad);
  return hipGetLastError();
}

hipError_t launch_res_precond(ds_problem* p, double* out2, int all,
                              hipStream_t st, const double* q) {
  ResPArgs a{};
  if (q) {
    a.q = q;
    a.rhs = p->rhs64;
    a.r_out = p->r;
  }
  a.r = p->r;
  a.diag = p->diag;
  a.minv = p->use_minv() ? p->minv : nullptr;
  a.minv_mod = p->minv_mod;
  a.blk_sub = p->blk_sub;
  a.blk_nreal = p->blk_nreal;
  a.sc = p->sc;
  a.s = out2;
  a.partials = p->partials;
  a.nblk = p->nblk;
  a.all = all;
  const dim3 g((p->nblk + kVecWaves - 1) / kVecWaves), b(64 * kVecWaves);
  hipLaunchKernelGGL(k_res_precond, g, b, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_finalize(ds_problem* p, const CgEpilogue& e,
                           hipStream_t st) {
  FinArgs a{};
  a.partials = p->partials;
  a.sub_blk0 = p->sub_blk0;
  a.sc = p->sc;
  a.norms = p->norms;
  a.deltas = p->deltas;
  a.done_count = p->done_count;
  a.drop = e.drop;
  a.nblk = p->nblk;
  a.norms_cap = p->norms_cap;
  a.init = e.init;
  a.mixed = e.mixed;
  a.is_true = e.is_true;
  a.need_want = e.need_want;
  // one block per subvolume, sized to its number of partials
  const int threads = p->n_sub == 1 ? 1024 : 256;
  if (e.kind == EK_CG_DQ)
    hipLaunchKernelGGL(k_finalize<1>, dim3(p->n_sub), dim3(threads), 0, st,
                       a);
  else
    hipLaunchKernelGGL(k_finalize<2>, dim3(p->n_sub), dim3(threads), 0, st,
                       a);
  return hipGetLastError();
}

namespace {

#define DS_TRY(expr)                     \
  do {                                   \
    hipError_t e_ = (expr);              \
    if (e_ != hipSuccess) return e_;     \
  } while (0)

// sym: the FP32 MFMA operator produced q's sums only (its symmetric form
// on the subvolumes flagged in sub_sym); the update kernel forms q there.
hipError_t launch_update(ds_problem* p, int only_x, int mixed, double drop,
                         hipStream_t st, int sym = 0, int fold_alpha = 0) {
  UpdArgs a{};
  a.d = p->d;
  a.q = p->q;
  a.diag = p->diag;
  a.minv = p->use_minv() ? p->minv : nullptr;
  a.minv_mod = p->minv_mod;
  a.wvec = p->form == DS_FORM_W ? p->wvec : nullptr;
  a.blk_sub = p->blk_sub;
  a.x = p->x;
  a.xu = p->xu;
  a.r = p->r;
  a.s = p->s;
  a.partials = p->partials;
  a.sc = p->sc;
  a.nblk = p->nblk;
  a.only_x = only_x;
  if (sym) {
    a.sym = SymRows{p->ydir, p->pbuf, p->gptr, p->pos64, p->wvec,
                    p->blk_nreal, p->form};
    a.sub_sym = p->sub_sym;
    a.sub_cv = p->sub_cv;
  }
  if (fold_alpha) {
    a.dq_part = p->partials + 2 * (size_t)p->nblk;
    a.sc_w = p->sc;
  }
  const int nblk1 = p->nblk / p->n_sub;
  if (a.minv && p->shared_geometry && p->n_sub > 1 && p->minv_mod == nblk1) {
    if (mixed)
      hipLaunchKernelGGL(k_cg_update_shared<true>, dim3(nblk1),
                         dim3(64 * kVecWaves), 0, st, a, nblk1, p->n_sub);
    else
      hipLaunchKernelGGL(k_cg_update_shared<false>, dim3(nblk1),
                         dim3(64 * kVecWaves), 0, st, a, nblk1, p->n_sub);
  } else {
    const dim3 g((p->nblk + kVecWaves - 1) / kVecWaves), b(64 * kVecWaves);
    // block-Jacobi mode: 0 Jacobi (or x only), 1 FP64 broadcast of r,
    // 2 FP32 broadcast of r (mixed solver)
    const int bj = (a.minv && !only_x) ? (mixed ? 2 : 1) : 0;
    const bool w = a.wvec, sy = a.sub_sym;
#define DS_UPD(B, W_, S)                                             \
  do {                                                               \
    if (fold_alpha)                                                  \
      hipLaunchKernelGGL((k_cg_update<B, W_, S, true>), g, b, 0, st, a); \
    else                                                             \
      hipLaunchKernelGGL((k_cg_update<B, W_, S, false>), g, b, 0, st, a); \
  } while (0)
#define DS_UPD_B(B)                                        \
  do {                                                     \
    if (w) { if (sy) DS_UPD(B, true, true); else DS_UPD(B, true, false); } \
    else { if (sy) DS_UPD(B, false, true); else DS_UPD(B, false, false); } \
  } while (0)
    if (bj == 2) DS_UPD_B(2);
    else if (bj == 1) DS_UPD_B(1);
    else DS_UPD_B(0);
#undef DS_UPD_B
#undef DS_UPD
  }
  DS_TRY(hipGetLastError());
  if (only_x) return hipSuccess;
  CgEpilogue e;
  e.kind = EK_CG_RES;
  e.mixed = mixed;
  e.drop = drop;
  return launch_finalize(p, e, st);
}

hipError_t launch_dir(ds_problem* p, int mixed, hipStream_t st,
                      int fold = 0, double drop = 0.0) {
  DirArgs a{};
  a.partials = p->partials;
  a.sub_blk0 = p->sub_blk0;
  a.norms = p->norms;
  a.deltas = p->deltas;
  a.done_count = p->done_count;
  a.drop = drop;
  a.norms_cap = p->norms_cap;
  a.s = p->s;
  a.wvec = p->form == DS_FORM_W ? p->wvec : nullptr;
  a.blk_sub = p->blk_sub;
  a.d = p->d;
  a.du = p->du;
  a.du32 = mixed ? p->du32 : nullptr;
  a.du16 = mixed && (p->use_dense || p->use_fused) ? p->du16 : nullptr;
  a.sc = p->sc;
  a.nblk = p->nblk;
  const dim3 g((p->nblk + kVecWaves - 1) / kVecWaves), b(64 * kVecWaves);
  if (fold) {
    if (a.wvec) hipLaunchKernelGGL((k_cg_dir<true, true>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_cg_dir<false, true>), g, b, 0, st, a);
  } else {
    if (a.wvec) hipLaunchKernelGGL((k_cg_dir<true, false>), g, b, 0, st, a);
    else hipLaunchKernelGGL((k_cg_dir<false, false>), g, b, 0, st, a);
  }
  return hipGetLastError();
}

// One iteration's kernels, in stream order.
hipError_t enqueue_iteration(ds_problem* p, int iter, int step_r, int mixed,
                             double drop, hipStream_t st) {
  CgEpilogue dq;
  dq.kind = EK_CG_DQ;
  if (mixed) {
    if (p->use_fused) {
      // shared-geometry batch: S tiles on the fly x all RHS on MFMA
      DS_TRY(fused_mrhs(p, st));
    } else if (p->use_dense) {
      // explicit BF16 block: q = w .* (S (w .* d)) + d for every RHS at once
      DS_TRY(dense_gemm(p, st));
      DS_TRY(dense_epilogue(p, st));
    } else {
      DS_TRY(launch_matvec32_ex(p, p->du32, p->d, p->q, 1, st,
                                p->n_sub == 1));
    }
    const int sym = !p->use_fused && !p->use_dense &&
                    p->f32_kind != DS_F32_VALU && p->n32_fast > 0;
    DS_TRY(launch_update(p, 0, 1, drop, st, sym,
                         !p->use_fused && !p->use_dense && p->n_sub == 1));
    CgEpilogue tr;
    tr.kind = EK_CG_RES;
    tr.out2 = p->s;
    tr.mixed = 1;
    tr.is_true = 1;
    tr.need_want = 1;
    tr.drop = drop;
    if (p->use_fused || p->use_dense) {
      // shared geometry: one pass for every subvolume's true residual
      DS_TRY(launch_res64_mrhs(p, p->x, p->form == DS_FORM_W ? p->xu : p->x,
                               p->r, 0, st, 1));
      DS_TRY(launch_res_precond(p, p->s, 0, st));
      DS_TRY(launch_finalize(p, tr, st));
    } else {
      // the true residual's finalize is folded into the direction kernel
      tr.fold = 1;
      DS_TRY(launch_matvec64_ex(p, 1, p->x, p->xu, p->r, tr, st));
      return launch_dir(p, mixed, st, 1, drop);
    }
  } else {
    DS_TRY(launch_matvec64_ex(p, 0, p->d, p->du, p->q, dq, st));
    if (iter % step_r == 0) {
      // lib/pcg.h:42-45: full residual r = b - A x after the x update.
      DS_TRY(launch_update(p, 1, 0, 0.0, st));
      CgEpilogue res;
      res.kind = EK_CG_RES;
      res.out2 = p->s;
      DS_TRY(launch_matvec64_ex(p, 1, p->x, p->xu, p->r, res, st));
    } else {
      DS_TRY(launch_update(p, 0, 0, 0.0, st));
    }
  }
  return launch_dir(p, mixed, st);
}

}  // namespace

// Solve every subvolume with the problem's resident x (sorted order, W
// space on entry) as the initial guess; on return x holds the solutions
// (x = w .* y for the W form, lib/wf_vector.cc:268-271).  res receives one
// entry per subvolume.  With use_graph the iterations of each chunk are
// captured once into a hipGraph and replayed (mixed solver only: its
// iterations all launch the same kernels with the same arguments).
hipError_t run_solver(ds_problem* p, const ds_pcg_params* prm, int mixed,
                      double drop, ds_pcg_result* res, int use_graph) {
  hipStream_t st = p->ctx->stream;
  DS_TRY(minv_join(p, st));   // before any capture below
  const int max_iter = prm->max_iter;
  const int ns = p->n_sub;
  const int cap = std::max(max_iter, 1) + 1;
  if (cap > p->norms_cap) {
    // the previous solve has completed (run_solver drains its stream)
    ds::pfree(p, p->norms);
    // residual norms, then deltas, cap per subvolume each
    DS_TRY(ds::palloc(p, &p->norms, 2 * (size_t)cap * ns));
    p->norms_cap = cap;
    p->deltas = p->norms + (size_t)cap * ns;
  }
  hipEvent_t e0, e1;
  DS_TRY(hipEventCreate(&e0));
  DS_TRY(hipEventCreate(&e1));
  DS_TRY(hipEventRecord(e0, st));
  ds_cg_scalars* h = p->sc_host;
  std::memset(h, 0, sizeof(ds_cg_scalars) * ns);
  int live = 0;
  for (int k = 0; k < ns; ++k) {
    h[k].tol_norm_b = prm->tol * p->sub_norm_b[k];
    h[k].max_iter = max_iter;
    h[k].refresh_every = prm->step_r;
    // lib/pcg.h:82-86: |b| <= 0 returns at once, x untouched (then x *= w)
    if (p->sub_norm_b[k] <= 0.0) h[k].done = 1;
    else ++live;
  }
  const double* w = p->form == DS_FORM_W ? p->wvec : nullptr;
  // the matrix-free mixed solver starts from an FP32 residual
  const bool init32 = mixed && !p->use_fused && !p->use_dense;
  p->done_host[0] = ns - live;
  DS_TRY(hipMemcpyAsync(p->done_count, &p->done_host[0], sizeof(int),
                        hipMemcpyHostToDevice, st));
  DS_TRY(hipMemcpyAsync(p->sc, h, sizeof(ds_cg_scalars) * ns,
                        hipMemcpyHostToDevice, st));
  if (live > 0) {
    // lib/pcg.h:97-110: r = b - A x; d = M^-1 r; delta = r . d; |r|.
    DS_TRY(launch_scale(p->x, w, w ? p->xu : nullptr,
                        init32 ? p->du32 : nullptr, p->npad, st));
    CgEpilogue init;
    init.kind = EK_CG_RES;
    init.out2 = p->d;
    init.init = 1;
    init.mixed = mixed;
    if (mixed && (p->use_fused || p->use_dense)) {
      // shared geometry: every subvolume's initial residual in one pass
      DS_TRY(launch_res64_mrhs(p, p->x, w ? p->xu : p->x, p->r, 1, st));
      DS_TRY(launch_res_precond(p, p->d, 1, st));
      DS_TRY(launch_finalize(p, init, st));
    } else if (init32) {
      // The mixed solver's initial residual with the FP32 operator: its
      // recurrence starts from b - A32 x0 (error ~1e-6 |A x0|, far above
      // the point where the first FP64 true residual replaces it), and no
      // subvolume stops before an FP64 true residual meets the tolerance.
      DS_TRY(launch_matvec32_ex(p, p->du32, p->x, p->q, 0, st));
      DS_TRY(launch_res_precond(p, p->d, 1, st, p->q));
      DS_TRY(launch_finalize(p, init, st));
    } else {
      DS_TRY(launch_matvec64_ex(p, 1, p->x, w ? p->xu : p->x, p->r, init,
                                st));
    }
    DS_TRY(launch_scale(p->d, w, w ? p->du : nullptr,
                        mixed ? p->du32 : nullptr, p->npad, st));
    if (mixed && (p->use_dense || p->use_fused))
      DS_TRY(launch_to_bf16(p->du32, p->du16, p->npad, st));
    // Enqueue iterations in chunks; after each chunk the number of finished
    // subvolumes is copied to pinned memory (two slots, alternating) and the
    // host inspects the previous chunk's count while this one runs.
    const int chunk = 8;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const bool graphs = use_graph && mixed;
    hipEvent_t ev[2];
    DS_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    DS_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int iter = 1, k = 0;
    double norm_prev = 0.0;
    hipError_t err = hipSuccess;
    while (iter < max_iter && err == hipSuccess) {
      const int end = std::min(max_iter, iter + chunk);
      if (graphs && end - iter == chunk) {
        if (!gexec) {
          err = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
          for (int it = 0; it < chunk && err == hipSuccess; ++it)
            err = enqueue_iteration(p, 1, prm->step_r, mixed, drop, st);
          hipError_t e2 = hipStreamEndCapture(st, &graph);
          if (err == hipSuccess) err = e2;
          if (err == hipSuccess)
            err = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        }
        if (err == hipSuccess) err = hipGraphLaunch(gexec, st);
        iter = end;
      } else {
        for (; iter < end && err == hipSuccess; ++iter)
          err = enqueue_iteration(p, iter, prm->step_r, mixed, drop, st);
      }
      if (err != hipSuccess) break;
      err = hipMemcpyAsync(&p->done_host[k & 1], p->done_count, sizeof(int),
                           hipMemcpyDeviceToHost, st);
      if (err == hipSuccess && ns == 1)
        err = hipMemcpyAsync(&p->norm_host[k & 1], &p->sc[0].norm_r,
                             sizeof(double), hipMemcpyDeviceToHost, st);
      if (err == hipSuccess) err = hipEventRecord(ev[k & 1], st);
      if (err == hipSuccess && k > 0) {
        err = hipEventSynchronize(ev[(k - 1) & 1]);
        if (p->done_host[(k - 1) & 1] >= ns) break;
        // Single volume near the end: when the decay over the last chunk
        // predicts the tolerance within the chunk just enqueued, wait for
        // it before enqueuing another, so that no whole chunk of launches
        // runs after convergence.  Scheduling only; results are unchanged.
        const double n2 = ns == 1 ? p->norm_host[(k - 1) & 1] : 0.0;
        if (err == hipSuccess && ns == 1 && k > 1 && norm_prev > 0.0 &&
            n2 > 0.0 && n2 < norm_prev &&
            n2 * (n2 / norm_prev) <= 2.0 * h[0].tol_norm_b) {
          err = hipEventSynchronize(ev[k & 1]);
          if (err == hipSuccess && p->done_host[k & 1] >= ns) break;
        }
        norm_prev = n2;
      }
      ++k;
    }
    // The loop leaves with the last chunk still in flight (the host has
    // seen only the previous chunk's count): its graph launch, pinned
    // copies and event records finish before the graph and the events are
    // destroyed.  The copy and wait below would wait for the same work.
    const hipError_t drained = hipStreamSynchronize(st);
    if (err == hipSuccess) err = drained;
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    DS_TRY(err);
    DS_TRY(hipMemcpyAsync(h, p->sc, sizeof(ds_cg_scalars) * ns,
                          hipMemcpyDeviceToHost, st));
    DS_TRY(hipStreamSynchronize(st));
    if (max_iter <= 1) {
      // lib/pcg.h:118,152-158: no step taken, iter == 1.
      for (int s = 0; s < ns; ++s)
        if (p->sub_norm_b[s] > 0.0) {
          h[s].num_iters = 1;
          h[s].success = (max_iter != 1);
          h[s].n_norms = 1;
          h[s].done = 1;
        }
    }
  }
  // lib/wf_vector.cc:268-271: x = w .* y.
  if (w) DS_TRY(launch_scale(p->x, w, p->x, nullptr, p->npad, st));
  DS_TRY(hipEventRecord(e1, st));
  DS_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  DS_TRY(hipEventElapsedTime(&ms, e0, e1));
  DS_TRY(hipEventDestroy(e0));
  DS_TRY(hipEventDestroy(e1));
  for (int s = 0; s < ns; ++s) {
    ds_pcg_result& r = res[s];
    std::memset(&r, 0, sizeof(r));
    r.norm_b = p->sub_norm_b[s];
    r.success = h[s].success;
    r.num_iters = h[s].num_iters;
    r.n_norms = h[s].n_norms;
    r.final_norm = h[s].norm_r;
    r.true_norm = mixed ? h[s].true_norm : h[s].norm_r;
    r.n_true = h[s].n_true;
    r.seconds = ms * 1e-3;
    // pair-evaluation passes: initial residual (FP32 in the matrix-free
    // mixed solver) + one per step (+ refreshes)
    if (r.norm_b > 0.0) {
      const int steps = std::max(0, std::min(h[s].num_iters, max_iter - 1));
      if (mixed) {
        r.n_matvec32 = steps + (init32 ? 1 : 0);
        r.n_matvec64 = (init32 ? 0 : 1) + h[s].n_true;
      } else {
        r.n_matvec64 = 1 + steps + steps / std::max(prm->step_r, 1);
      }
    }
  }
  return hipSuccess;
}

}  // namespace ds
