// precond.hip -- block-Jacobi preconditioner for the device solvers.
//
// The reference preconditions with the diagonal of A (Jacobi,
// lib/wf_vector.cc:42-49, :185-194), and the parity solver keeps it.  A
// problem may instead select the exact inverse of each 64-row target block
// of A (the k-d blocks of the culling, spatially compact): the local
// couplings the diagonal misses, for one extra 8.4 KB read per block and
// application (the symmetric inverse stored by diagonals).  At the 300k-pixel skewer density it cuts the iterations to
// 1e-6 by about a third.  The preconditioner is fixed and symmetric
// positive definite (the inverse of SPD blocks, symmetrized and stored in
// FP32), so PCG's theory is unchanged; solves stop on the same FP64 true
// residual.
//
// k_bj_build: one wave per target block; lane (bi, bj) = (lane / 8,
// lane mod 8) holds in registers the 64 entries of rows bi + 8 u and
// columns bj + 8 v (u, v = 0..7), built as A_ij = w_i S_ij w_j + delta_ij
// (W form) or S_ij + N_i delta_ij (SN form) with the FP64 table interpolant
// of lib/signal_covar.h:53-77 (padding rows: identity).  The block is
// inverted in place by Gauss-Jordan elimination without pivoting (the
// block is SPD): step k divides row k by its pivot (its pivot entry
// becoming 1 / pivot) and subtracts a_ik times it from every other row i
// (column k taken as 0 there first), so that after the last step the
// registers hold the inverse.  Each step the 8 lanes of row k and the 8
// lanes of column k publish them in LDS (two buffers, so one barrier per
// step), and every lane reads the 8 row values of its columns and the 8
// column values of its rows: one fma per entry and step.  The inverse is
// rounded to FP32, each entry averaged with its transpose (fetched from
// lane (bj, bi) by ds_bpermute), staged in LDS and stored by diagonals in
// coalesced rows (bj_apply's layout): minv[I][k][i] = M_{i,(i+k) mod 64},
// k = 0..32.  (Interleaved rows and columns make both the staging writes
// and the diagonal reads free of LDS bank conflicts.)
// COMPILE WITH -ffp-contract=off.
#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

struct BjArgs {
  const double4* pos;
  const int* blk_sub;
  const int* blk_nreal;
  const double4* sub_cv;
  const double* table;
  float* minv;
  double dx_inv;
  int table_n, form;
};

__global__ __launch_bounds__(64) void k_bj_build(BjArgs a) {
  constexpr int TB = 8;                  // entries per lane and axis
  constexpr int RS = TB + 2;             // padded LDS row (16-byte aligned)
  constexpr int MS = 72;                 // staging row stride (floats)
  struct Elim {
    double rowb[2][TB][RS];              // row k, by owning column lane
    double colb[2][TB][RS];              // column k, by owning row lane
    double spare[DS_BS][2 * TB + 2];     // unowned values (see below;
                                         // padded: lanes start on
                                         // different LDS banks)
    double4 bpos[DS_BS];                 // the block's pixels
  };
  __shared__ __align__(16) union {
    Elim e;
    float mt[DS_BS * MS];                // symmetrized FP32 inverse
  } sm;
  const int I = blockIdx.x, t = threadIdx.x;
  const int bi = t / TB, bj = t % TB;
  const double4 cv = a.sub_cv[a.blk_sub[I]];
  const int nreal = a.blk_nreal[I];
  sm.e.bpos[t] = a.pos[I * DS_BS + t];
  const double tlast = (double)(a.table_n - 1);
  __syncthreads();
  double m[TB][TB];   // m[u][v] = A_{bi + 8u, bj + 8v}
#pragma unroll
  for (int u = 0; u < TB; ++u) {
    const int i = bi + TB * u;
    const double4 pi = sm.e.bpos[i];
#pragma unroll
    for (int v = 0; v < TB; ++v) {
      const int j = bj + TB * v;
      // branch-free: every entry is evaluated (the index clamped into the
      // table), then the cutoff and the padding rows select
      const double4 pj = sm.e.bpos[j];
      const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const double x_perp_2 = (dx * dx + dy * dy);
      const double x_para_2 = dz * dz;
      const double xx = (cv.y * x_perp_2 + cv.z * x_para_2) * a.dx_inv;
      const int ix = min((int)fmin(xx, tlast), a.table_n - 2);
      double S = cv.x * (a.table[ix] * ((double)(ix + 1) - xx) +
                         a.table[ix + 1] * (xx - (double)ix));
      S = xx < tlast ? S : 0.0;
      const double id = (i == j) ? 1.0 : 0.0;
      const double val = a.form == DS_FORM_W ? pi.w * S * pj.w + id
                                             : S + (i == j ? pi.w : 0.0);
      m[u][v] = (i < nreal && j < nreal) ? val : id;
    }
    // one row's table loads in flight at a time (register budget)
    __builtin_amdgcn_sched_barrier(0);
  }
  // Steps k = 8 u + o: row k is held by the lanes bi = o (entries m[u][.]),
  // column k by the lanes bj = o (entries m[.][u]).  Branch-free (a branch
  // here costs the register allocator dearly): every lane stores its
  // candidate row and column values, the lanes that do not own them into a
  // spare slot of their own, and the special row and column are handled by
  // exact factors of 0 and 1.
#pragma unroll
  for (int u = 0; u < TB; ++u) {
    for (int o = 0; o < TB; ++o) {
      const int buf = o & 1;
      const bool own_row = bi == o, own_col = bj == o;
      const double zr = own_row ? 0.0 : 1.0, zc = own_col ? 0.0 : 1.0;
      double* rdst = own_row ? &sm.e.rowb[buf][bj][0] : &sm.e.spare[t][0];
      double* cdst = own_col ? &sm.e.colb[buf][bi][0] : &sm.e.spare[t][TB];
#pragma unroll
      for (int v = 0; v < TB; ++v) rdst[v] = m[u][v];
#pragma unroll
      for (int w = 0; w < TB; ++w) cdst[w] = m[w][u];
      __syncthreads();
      const double p = 1.0 / sm.e.rowb[buf][o][u];   // 1 / pivot
      double rv[TB], cl[TB];
#pragma unroll
      for (int v = 0; v < TB; ++v) rv[v] = sm.e.rowb[buf][bj][v] * p;
      rv[u] = own_col ? p : rv[u];
#pragma unroll
      for (int w = 0; w < TB; ++w) cl[w] = sm.e.colb[buf][bi][w];
      // column k taken as 0 (it becomes -a_ik / pivot); row k as 0 with
      // factor -1 (it becomes the scaled row)
      cl[u] = own_row ? -1.0 : cl[u];
#pragma unroll
      for (int w = 0; w < TB; ++w) m[w][u] *= zc;
#pragma unroll
      for (int v = 0; v < TB; ++v) m[u][v] *= zr;
#pragma unroll
      for (int w = 0; w < TB; ++w)
#pragma unroll
        for (int v = 0; v < TB; ++v) m[w][v] = fma(-cl[w], rv[v], m[w][v]);
    }
  }
  // FP32, averaged with the transpose: lane (bj, bi) holds M_{j,i} of this
  // lane's (i, j) = (bi + 8 u, bj + 8 v) as its m[v][u]
  const int tp = bj * TB + bi;
  float f[TB][TB];
#pragma unroll
  for (int u = 0; u < TB; ++u)
#pragma unroll
    for (int v = 0; v < TB; ++v) f[u][v] = (float)m[u][v];
  __syncthreads();   // the elimination's LDS is reused below
#pragma unroll
  for (int u = 0; u < TB; ++u) {
#pragma unroll
    for (int v = 0; v < TB; ++v) {
      const float tr = __int_as_float(
          __builtin_amdgcn_ds_bpermute(tp * 4, __float_as_int(f[v][u])));
      sm.mt[(bi + TB * u) * MS + bj + TB * v] =
          (float)(0.5 * ((double)f[u][v] + (double)tr));
    }
  }
  __syncthreads();
  float* out = a.minv + (size_t)I * DS_BJ_SIZE;
  for (int k = 0; k < DS_BJ_ROWS; ++k)
    out[k * DS_BS + t] = sm.mt[t * MS + ((t + k) & (DS_BS - 1))];
}

// x = M^-1 b for every target block (one wave per block).
__global__ __launch_bounds__(256) void k_bj_apply(const double* b,
                                                  const float* minv, double* x,
                                                  int nblk, int mod) {
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (I >= nblk) return;
  const int i = I * DS_BS + (threadIdx.x & 63);
  x[i] = bj_apply(minv + (size_t)(I % mod) * DS_BJ_SIZE, b[i]);
}

// The same with b rounded to FP32 and the inverse's diagonals held in
// registers: the mixed solver's update kernel's application
// (bj_apply_regs_r32), one wave per block.
__global__ __launch_bounds__(256) void k_bj_apply_r32(const double* b,
                                                      const float* minv,
                                                      double* x, int nblk,
                                                      int mod) {
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (I >= nblk) return;
  const int i = I * DS_BS + (threadIdx.x & 63);
  float m[DS_BJ_ROWS];
  load_bj_diagonals(minv + (size_t)(I % mod) * DS_BJ_SIZE, m);
  x[i] = bj_apply_regs_r32(m, b[i]);
}

}  // namespace

hipError_t launch_precond_r32(const ds_problem* p, const double* b, double* x,
                              hipStream_t st) {
  if (hipError_t e = minv_join(p, st)) return e;
  hipLaunchKernelGGL(k_bj_apply_r32, dim3((p->nblk + 3) / 4), dim3(256), 0,
                     st, b, p->minv, x, p->nblk, p->minv_mod);
  return hipGetLastError();
}

hipError_t bj_build(ds_problem* p, hipStream_t st) {
  BjArgs a{};
  a.pos = p->pos64;
  a.blk_sub = p->blk_sub;
  a.blk_nreal = p->blk_nreal;
  a.sub_cv = p->sub_cv;
  a.table = p->table_raw;
  a.minv = p->minv;
  a.dx_inv = p->cv.dx_inv;
  a.table_n = p->cv.table_n;
  a.form = p->form;
  hipLaunchKernelGGL(k_bj_build, dim3(p->minv_mod), dim3(64), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_precond_p(const ds_problem* p, const double* b, double* x,
                            hipStream_t st) {
  if (!p->use_minv()) return launch_precond(b, p->diag, x, p->npad, st);
  if (hipError_t e = minv_join(p, st)) return e;
  hipLaunchKernelGGL(k_bj_apply, dim3((p->nblk + 3) / 4), dim3(256), 0, st, b,
                     p->minv, x, p->nblk, p->minv_mod);
  return hipGetLastError();
}

}  // namespace ds
