// cg_device.h -- device helpers shared by the pair kernels and the CG vector
// kernels: deterministic reductions and the finalisation of the FP64 CG
// scalars of lib/pcg.h:27-65.
//
// Every reduction sums in a fixed order that depends only on the block
// layout, so repeated solves are bit-for-bit reproducible.  A kernel that
// ends a reduction stores one partial sum per block; the finalize kernel
// launched after it on the same stream (one block per subvolume) sums the
// subvolume's partials and updates its scalars.  The kernel boundary makes
// the partials visible across the XCDs' separate L2 caches, so no block
// pays for a device-scope fence.
#pragma once
#include <hip/hip_runtime.h>

#include "ds_internal.h"

namespace ds {

// Rows of the symmetric FP32 operator (MFMA kernel): the block's own sums
// plus the column sums in its tile's consecutive slots (slot order), then
// the epilogue q = w (var_s y) + a (W form) or var_s y + N a (SN form), 0
// on padding rows.  Row `lane` of target block I, `a` its input value.
struct SymRows {
  const double* ydir;
  const float* pbuf;
  const int* gptr;
  const double4* pos;    // SN form: N = pos.w
  const double* wvec;    // W form: w
  const int* blk_nreal;
  int form;
};

__device__ __forceinline__ double sym_row(const SymRows& s, int I, int lane,
                                          double var_s, double a) {
  const int row = I * DS_BS + lane;
  const int T = DS_TPB * I + (lane >> 4), c = lane & 15;
  double y = s.ydir[row];
  const int gb = s.gptr[T], ge = s.gptr[T + 1];
  // sixteen slots' loads in flight at a time (indices clamped to the last
  // slot, so every load is in bounds), summed in slot order
  const float* pb = s.pbuf + c;
  for (int k = gb; k < ge; k += 16) {
    float pv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      pv[q] = pb[16 * (size_t)min(k + q, ge - 1)];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (k + q < ge) y += (double)pv[q];
  }
  const double wn = s.form == DS_FORM_W ? s.wvec[row] : s.pos[row].w;
  double q = s.form == DS_FORM_W ? wn * (var_s * y) + a : var_s * y + wn * a;
  if (lane >= s.blk_nreal[I]) q = 0.0;
  return q;
}

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

// Thread 0 stores this block's NV sums (block I of nblk) for the finalize
// kernel that runs next on the stream.
template <int NV>
__device__ __forceinline__ void store_partials(const double (&v)[NV],
                                               double* partials, int nblk,
                                               int I) {
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)k * nblk + I] = v[k];
}

// Sum `count` partial vectors (component k at partials[k * stride + b]) in
// a fixed order; result valid in every thread.  Loads are issued eight at a
// time so their memory latencies overlap.
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
    // the rest (fewer than 8 per thread) as one batch of loads, indices
    // clamped in bounds, summed in the same order
    if (b < count) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = pk[min(b + u * step, count - 1)];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (b + u * step < count) s += t[u];
    }
    v[k] = s;
  }
  block_sum<NV>(v, red);
}

// Block-Jacobi application for one 64-row target block, one wave: lane i
// returns sum_k M_{i,j} r_j, j = (i + k) mod 64, k = 0..63 in order.  The
// symmetric M is stored by diagonals, DS_BJ_ROWS rows of 64 floats:
// mt[k * 64 + i] = M_{i,(i+k) mod 64} for k = 0..32; for k > 32 the element
// is M_{j,i} = mt[(64 - k) * 64 + j].  Every load is one 256-byte row (the
// rows k > 32 permuted within it, and already read for 64 - k); r_j comes
// from lane j by a shuffle.  All 64 lanes must be active.
__device__ __forceinline__ double bj_apply(const float* __restrict__ mt,
                                           double r) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
#pragma unroll 8
  for (int k = 0; k < DS_BS; ++k) {
    const int j = (lane + k) & (DS_BS - 1);
    const float m = k <= DS_BS / 2 ? mt[k * DS_BS + lane]
                                   : mt[(DS_BS - k) * DS_BS + j];
    s = fma((double)m, __shfl(r, j, 64), s);
  }
  return s;
}

// bj_apply with the lane's 33 stored diagonal values in registers (loaded
// by load_bj_diagonals; the elements of rows k > 32 come from lane j by a
// shuffle): the same values in the same order, so the same result.
__device__ __forceinline__ void load_bj_diagonals(
    const float* __restrict__ mt, float (&m)[DS_BJ_ROWS]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < DS_BJ_ROWS; ++k) m[k] = mt[k * DS_BS + lane];
}

__device__ __forceinline__ double bj_apply_regs(const float (&m)[DS_BJ_ROWS],
                                                double r) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < DS_BS; ++k) {
    const int j = (lane + k) & (DS_BS - 1);
    const float mk = k <= DS_BS / 2 ? m[k] : __shfl(m[DS_BS - k], j, 64);
    s = fma((double)mk, __shfl(r, j, 64), s);
  }
  return s;
}

// Whole-wave rotations by one lane (DPP wave_rol:1 / wave_ror:1): lane i
// receives lane i + 1, resp. i - 1 (mod 64).  All 64 lanes must be active.
__device__ __forceinline__ float wave_rol1(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF,
                                  false));
}
__device__ __forceinline__ double wave_ror1(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x13C, 0xF,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x13C, 0xF,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}

// bj_apply with r rounded to FP32 (the mixed solver, whose operator is FP32
// already), the same terms in the same order (k = 0..63), products and sum
// in FP64, the cross-lane moves on the VALU instead of LDS shuffles:
// k <= 32: the stored row k times r_{i+k}, from a copy of r rotated one
// lane per step; k > 32: the term is P_{k'}(j) = M_{j,(j+k') mod 64} r_j of
// lane j = i - k' (k' = 64 - k), accumulated Horner-fashion over lanes:
// A_31 = (the k <= 32 sum, rotated left by 31) + P_31, then
// A_{k'} = ror1(A_{k'+1}) + P_{k'}, and lane i's result is ror1(A_1).
// (Every product of two floats is exact in FP64, so each step rounds once,
// as the fma of the sequential form does.)  m(k): the lane's stored value of
// diagonal row k.
template <class Row>
__device__ __forceinline__ double bj_apply_r32_rot(Row m, double r) {
  const int lane = threadIdx.x & 63;
  const float rf = (float)r;
  float rr = rf;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k <= DS_BS / 2; ++k) {
    s = fma((double)m(k), (double)rr, s);
    if (k < DS_BS / 2) rr = wave_rol1(rr);
  }
  double acc = __shfl(s, (lane + DS_BS / 2 - 1) & (DS_BS - 1), 64) +
               (double)m(DS_BS / 2 - 1) * (double)rf;
#pragma unroll
  for (int kp = DS_BS / 2 - 2; kp >= 1; --kp)
    acc = wave_ror1(acc) + (double)m(kp) * (double)rf;
  return wave_ror1(acc);
}

// ... with the stored rows read from mt (LDS or global), or held in
// registers (loaded by load_bj_diagonals).
__device__ __forceinline__ double bj_apply_r32(const float* __restrict__ mt,
                                               double r) {
  const int lane = threadIdx.x & 63;
  return bj_apply_r32_rot([&](int k) { return mt[k * DS_BS + lane]; }, r);
}

__device__ __forceinline__ double bj_apply_regs_r32(
    const float (&m)[DS_BJ_ROWS], double r) {
  return bj_apply_r32_rot([&](int k) { return m[k]; }, r);
}

// lib/pcg.h:35-37: alpha = delta / (d . q).
// (Also ends a mixed solver's true-residual request: every reader of
// want_true in the step that asked for it has run by now.)
__device__ __forceinline__ void finalize_alpha(ds_cg_scalars* sc, double dq) {
  sc->dq = dq;
  sc->alpha = sc->delta / dq;
  sc->want_true = 0;
}

__device__ __forceinline__ void mark_done(ds_cg_scalars* sc, int success,
                                          int num_iters, int* done_count) {
  sc->done = 1;
  sc->success = success;
  sc->num_iters = num_iters;
  atomicAdd(done_count, 1);
}

// lib/pcg.h:54-61 and :123-135 (or, with init, :100-110): new delta, beta,
// |r|, residual and delta histories (what the verbose trace of
// lib/pcg.h:112-139 prints) and the stop rule.  rs = r . M^-1 r, rr = r . r.
// Mixed solver (mixed != 0): the recurrence residual never stops the solve
// itself; the next true-residual kernel runs (want_true) when it meets the
// tolerance, has dropped by `drop` since the last FP64 true residual, or
// refresh_every iterations have passed since that one (a recurrence that
// stagnates at the inner operator's precision is refreshed all the same).
__device__ __forceinline__ void finalize_step(ds_cg_scalars* sc, double* norms,
                                              double* deltas,
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
    deltas[0] = rs;
    if (mixed) {
      sc->ref_norm = norm_r;
      sc->true_norm = norm_r;
      sc->want_true = 0;
      sc->last_true = 0;
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
  deltas[iter] = rs;
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
    // want_true stays set until the next step's finalize_alpha: the
    // direction kernel's waves of this subvolume still read it
    sc->last_true = iter;
    if (norm_r <= sc->tol_norm_b) {
      mark_done(sc, 1, iter, done_count);
      return;
    }
  } else {
    // kept for a true residual that replaces this step's residual, asked
    // for here or by another subvolume of a shared-geometry batch
    sc->saved_delta_old = delta_old;
    if (norm_r <= sc->tol_norm_b || norm_r <= drop * sc->ref_norm ||
        iter - sc->last_true >= sc->refresh_every) {
      sc->want_true = 1;
      return;
    }
  }
  if (iter + 1 >= sc->max_iter) mark_done(sc, 0, sc->max_iter, done_count);
}

}  // namespace ds
