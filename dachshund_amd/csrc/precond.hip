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
// k_bj_build: one 128-thread workgroup per target block; row i of the
// block is held in registers by threads i (columns 0..31) and 64 + i
// (columns 32..63), built as A_ij = w_i S_ij w_j + delta_ij (W form) or
// S_ij + N_i delta_ij (SN form) with the FP64 table interpolant of
// lib/signal_covar.h:53-77 (padding rows: identity).  The block is inverted
// by the in-place Gauss-Jordan sweep (no pivoting: the block is SPD), with
// each swept row kept unscaled: row i holds its true values divided by its
// pivot's reciprocal p_i (times 1 while unswept) and is scaled once at the
// end.  Step k: the two threads of row k publish it in LDS (its true values:
// it is unswept) and every row applies m_c -= g raw_c with one factor
// g = a_ik d_i p_k (d_i = 1 / p_i, or 1), a_ik taken from the published row
// by the sweep's structure (a_ik = a_ki while both are unswept, -a_ki once
// i < k is swept) so that no thread indexes its registers by k; column k
// becomes -g, row k keeps its values (g = 0) with 1 in column k.  One FMA
// per entry and step.  The inverse goes through LDS once as FP32 to be
// symmetrised, rounded to FP32 and stored by diagonals (bj_apply's layout):
// minv[I][k][i] = M_{i,(i+k) mod 64}, k = 0..32.
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

__global__ __launch_bounds__(128) void k_bj_build(BjArgs a) {
  constexpr int HC = DS_BS / 2;        // columns per thread
  __shared__ double raw[2][DS_BS];     // step k (buffer k & 1): row k
  __shared__ float mt[DS_BS][DS_BS + 1];
  const int I = blockIdx.x, t = threadIdx.x;
  const int i = t & (DS_BS - 1);
  const int h = __builtin_amdgcn_readfirstlane(t >> 6), c0 = h * HC;
  const double4 cv = a.sub_cv[a.blk_sub[I]];
  const int nreal = a.blk_nreal[I];
  const double4 pi = a.pos[I * DS_BS + i];
  double m[HC];
#pragma unroll
  for (int c = 0; c < HC; ++c) {
    const int j = c0 + c;
    double v = (i == j) ? 1.0 : 0.0;
    if (i < nreal && j < nreal) {
      const double4 pj = a.pos[I * DS_BS + j];
      const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const double x_perp_2 = (dx * dx + dy * dy);
      const double x_para_2 = dz * dz;
      const double xx = (cv.y * x_perp_2 + cv.z * x_para_2) * a.dx_inv;
      double S = 0.0;
      if (xx < (double)(a.table_n - 1)) {
        const int ix = (int)xx;
        S = cv.x * (a.table[ix] * ((double)(ix + 1) - xx) +
                    a.table[ix + 1] * (xx - (double)ix));
      }
      if (a.form == DS_FORM_W) v = pi.w * S * pj.w + v;
      else v = S + (i == j ? pi.w : 0.0);
    }
    m[c] = v;
  }
  double d_own = 1.0, p_own = 1.0;   // this row's pivot and its reciprocal
  for (int k = 0; k < DS_BS; ++k) {
    double* rw = raw[k & 1];
    if (i == k) {
#pragma unroll
      for (int c = 0; c < HC; ++c) rw[c0 + c] = m[c];
    }
    __syncthreads();
    const double piv = rw[k];
    const double p = 1.0 / piv;
    const double fr = rw[i];
    const double g = i == k ? 0.0 : ((i < k ? -fr : fr) * d_own) * p;
    if (i == k) {
      d_own = piv;
      p_own = p;
    }
#pragma unroll
    for (int c = 0; c < HC; ++c) m[c] = fma(-g, rw[c0 + c], m[c]);
    const int kc = k - c0;   // wave-uniform: column k in this half?
    if (kc >= 0 && kc < HC) {
      const double mk = i == k ? 1.0 : -g;
#pragma unroll
      for (int c = 0; c < HC; ++c)
        if (c == kc) m[c] = mk;
    }
  }
#pragma unroll
  for (int c = 0; c < HC; ++c) mt[i][c0 + c] = (float)(m[c] * p_own);
  __syncthreads();
  float* out = a.minv + (size_t)I * DS_BJ_SIZE;
  for (int k = h; k < DS_BJ_ROWS; k += 2) {
    const int j = (i + k) & (DS_BS - 1);
    out[k * DS_BS + i] =
        (float)(0.5 * ((double)mt[i][j] + (double)mt[j][i]));
  }
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

}  // namespace

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
  hipLaunchKernelGGL(k_bj_build, dim3(p->minv_mod), dim3(2 * DS_BS), 0, st,
                     a);
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
