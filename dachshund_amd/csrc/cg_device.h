// cg_device.h -- device helpers shared by the pair kernels and the CG vector
// kernels: deterministic reductions, the last-block-done ticket, and the
// finalisation of the FP64 CG scalars of lib/pcg.h:27-65.
//
// Every reduction sums in a fixed order that depends only on the grid shape,
// so repeated solves are bit-for-bit reproducible.  A kernel that ends a
// reduction writes its per-block partials, takes a ticket from
// ds_cg_scalars::counter, and the block holding the last ticket sums the
// partials and updates the scalars; it then resets the ticket, so no
// memset is ever needed between launches.
#pragma once
#include <hip/hip_runtime.h>

#include "ds_internal.h"

namespace ds {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double load_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide sum of NV values per thread; result valid in every thread.
// `red` holds NV * (blockDim.x / 64) doubles of LDS.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k * nw + w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[k * nw + i];
    v[k] = s;
  }
  __syncthreads();
}

// Publish this block's NV partial sums (thread 0 holds them) into slot
// `slot` of a reduction group of `count` blocks (component k at
// partials[k * stride + slot]) and return true in every thread of the block
// that took the group's last ticket.
template <int NV>
__device__ __forceinline__ bool publish_partials(const double (&v)[NV],
                                                 double* partials,
                                                 size_t stride, int slot,
                                                 int count, unsigned* counter,
                                                 bool* is_last_lds) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      __hip_atomic_store(&partials[(size_t)k * stride + slot], v[k],
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    *is_last_lds = (t == (unsigned)count - 1);
  }
  __syncthreads();
  const bool last = *is_last_lds;
  if (last) __threadfence();
  return last;
}

// In the last block: sum the group's NV partial vectors in a fixed order.
// The acquire fence of publish_partials has invalidated this CU's caches,
// so plain loads see every block's partial; they are issued eight at a time
// so their memory latencies overlap.
template <int NV>
__device__ __forceinline__ void gather_partials(double (&v)[NV],
                                                const double* partials,
                                                size_t stride, int count,
                                                double* red) {
  const int step = blockDim.x;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double* pk = partials + (size_t)k * stride;
    double s = 0.0;
    int b = threadIdx.x;
    for (; b + 7 * step < count; b += 8 * step) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = pk[b + u * step];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += t[u];
    }
    for (; b < count; b += step) s += pk[b];
    v[k] = s;
  }
  block_sum<NV>(v, red);
}

__device__ __forceinline__ void release_ticket(unsigned* counter) {
  if (threadIdx.x == 0)
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Where a kernel's per-block partial sums go.  Blocks of a subvolume are
// grouped by 64; level 1 reduces each group into a group partial (<= 64
// tickets per counter), level 2 reduces a subvolume's group partials.
// Slots, groups and summation order depend only on the layout, so results
// are reproducible.
struct RedCtx {
  double* partials;       // NV x nblk
  double* gpart;          // NV x ngrp
  unsigned* gcount;       // ngrp tickets
  const int* sub_blk0;    // first block per subvolume
  const int* sub_grp0;    // first group per subvolume
  int nblk, ngrp;
};

// Two-level deterministic reduction of the block sums v (thread 0 holds
// them) over subvolume sv's blocks; true (in every thread of the block)
// for the one block that ends up holding the subvolume's total in v.
// `red` holds NV * blockDim.x / 64 doubles of LDS.
template <int NV>
__device__ __forceinline__ bool reduce_subvolume(double (&v)[NV],
                                                 const RedCtx& R, int sv,
                                                 int I, unsigned* sub_counter,
                                                 double* red, bool* flag) {
  const int b0 = R.sub_blk0[sv], nb = R.sub_blk0[sv + 1] - b0;
  const int slot = I - b0, g = slot >> 6;
  const int gsize = min(64, nb - (g << 6));
  const int ng = (nb + 63) >> 6;
  const int G0 = R.sub_grp0[sv];
  double* pg = R.partials + b0 + (g << 6);
  if (!publish_partials<NV>(v, pg, R.nblk, slot & 63, gsize, &R.gcount[G0 + g],
                            flag))
    return false;
  gather_partials<NV>(v, pg, R.nblk, gsize, red);
  release_ticket(&R.gcount[G0 + g]);
  if (!publish_partials<NV>(v, R.gpart + G0, R.ngrp, g, ng, sub_counter,
                            flag))
    return false;
  gather_partials<NV>(v, R.gpart + G0, R.ngrp, ng, red);
  release_ticket(sub_counter);
  return true;
}

inline RedCtx red_ctx(const ds_problem* p) {
  RedCtx r;
  r.partials = p->partials;
  r.gpart = p->gpart;
  r.gcount = p->gcount;
  r.sub_blk0 = p->sub_blk0;
  r.sub_grp0 = p->sub_grp0;
  r.nblk = p->nblk;
  r.ngrp = p->ngrp;
  return r;
}

// lib/pcg.h:35-37: alpha = delta / (d . q).
__device__ __forceinline__ void finalize_alpha(ds_cg_scalars* sc, double dq) {
  sc->dq = dq;
  sc->alpha = sc->delta / dq;
}

// lib/pcg.h:54-61 and :123-135 (or, with init, :100-110): new delta, beta,
// |r|, residual history and the stop rule.  rs = r . M^-1 r, rr = r . r.
// Mixed solver (mixed != 0): the recurrence residual never stops the solve
// itself; when it meets the tolerance or has dropped by `drop` since the
// last FP64 true residual, the next true-residual kernel runs (want_true).
__device__ __forceinline__ void mark_done(ds_cg_scalars* sc, int success,
                                          int num_iters, int* done_count) {
  sc->done = 1;
  sc->success = success;
  sc->num_iters = num_iters;
  atomicAdd(done_count, 1);
}

__device__ __forceinline__ void finalize_step(ds_cg_scalars* sc, double* norms,
                                              double rs, double rr, int init,
                                              int mixed, double drop,
                                              int is_true, int* done_count) {
  const double norm_r = sqrt(rr);
  if (init) {
    sc->delta = rs;
    sc->norm_r = norm_r;
    sc->iter = 0;
    sc->n_norms = 1;
    norms[0] = norm_r;
    if (mixed) {
      sc->ref_norm = norm_r;
      sc->true_norm = norm_r;
      sc->want_true = 0;
    }
    return;
  }
  // A true residual of the mixed solver replaces the recurrence residual of
  // the step that just ended: same iteration, new delta/beta/|r|.
  const double delta_old = is_true ? sc->saved_delta_old : sc->delta;
  sc->delta = rs;
  sc->beta = rs / delta_old;
  sc->norm_r = norm_r;
  const int iter = is_true ? sc->iter : sc->iter + 1;
  sc->iter = iter;
  norms[iter] = norm_r;
  sc->n_norms = iter + 1;
  if (!mixed) {
    if (norm_r <= sc->tol_norm_b) {
      mark_done(sc, 1, iter, done_count);
    } else if (iter + 1 >= sc->max_iter) {
      // lib/pcg.h:118,151-158: the loop ends with iter == max_iter.
      mark_done(sc, 0, sc->max_iter, done_count);
    }
    return;
  }
  if (is_true) {
    sc->n_true += 1;
    sc->true_norm = norm_r;
    sc->ref_norm = norm_r;
    sc->want_true = 0;
    if (norm_r <= sc->tol_norm_b) {
      mark_done(sc, 1, iter, done_count);
      return;
    }
  } else {
    if (norm_r <= sc->tol_norm_b || norm_r <= drop * sc->ref_norm) {
      sc->want_true = 1;
      sc->saved_delta_old = delta_old;
      return;
    }
  }
  if (iter + 1 >= sc->max_iter) mark_done(sc, 0, sc->max_iter, done_count);
}

}  // namespace ds
