// pairs32.hip -- FP32 pixel-pair matvec: the inner operator of the mixed
// solver.
//
// Same interpolant as the FP64 parity kernel (lib/signal_covar.h:53-77):
// S = var_s * (f[i] + (f[i+1]-f[i]) * (xx - i)), 0 beyond the cutoff, but
// evaluated in FP32 on pre-scaled coordinates, so that the table argument is
// just xx = dX^2 + dY^2 + dZ^2 with X = x * sqrt(c_perp * dx_inv),
// Z = z * sqrt(c_para * dx_inv).  Sources are stored relative to the frame
// (per-axis median) of their 16-slot tile, targets relative to that of their
// block; each (target block, source tile) list entry carries the FP32
// offset between the two frames (differenced in FP64 on the host), so large
// box coordinates never lose precision in FP32.
//
// Per pair the argument is expanded as |Xi|^2 + |Xj|^2 - 2 Xi.Xj with |Xj|^2
// stored beside the source coordinates and -2 Xi formed once per tile: one
// add and three fma; one med3 clamps it to [0, table_n - 1] (the expansion
// can leave tiny negative values for coincident points), then index
// conversion, fraction, the LDS gather of {f, df} and two fma: 10 VALU
// operations and one 8-byte LDS read per pair.  Problems with a point far
// from its tile's frame (isolated outliers, where the expansion would
// cancel badly) use direct differences instead (12 operations).  The sum of
// each lane is FP32; the epilogue (identity or noise term, dot products) is
// FP64.
#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

struct Matvec32Args {
  const float4* loc;      // sources: tile-frame coordinates, .w = |X|^2
  const float4* tgt;      // targets: block-frame coordinates
  const float* u32;       // scaled input (w .* a, or a) in FP32
  const double4* pos;     // w (W form) or N (SN form) in .w
  const double* in;       // FP64 input, own rows
  const int* tptr;
  const int* tidx;
  const float4* toff;     // per list entry: block frame - tile frame
  const float2* table;    // {f[k], f[k+1]-f[k]}, last = {0, 0}
  const int* blk_sub;
  const int* blk_nreal;
  const int* sub_blk0;
  const double4* sub_cv;
  const int* blk_order;
  double* out;
  double* partials;
  const ds_cg_scalars* sc;
  float tlast;
  int nblk, table_n;
};

// The per-tile data (list entry, frame offset, source records and vector)
// is wave-uniform; passing it as restrict-qualified kernel arguments lets
// the compiler fetch it with scalar loads.
template <int FORM, bool CG, bool DIRECT>
__global__ __launch_bounds__(256) void k_matvec32(
    const float4* __restrict__ loc, const float* __restrict__ u32,
    const int* __restrict__ tidx, const float4* __restrict__ toff,
    Matvec32Args a) {
  extern __shared__ float2 tab32[];
  __shared__ double red[4 * 64];
  const int I = a.blk_order ? a.blk_order[blockIdx.x] : blockIdx.x;
  const int sv = a.blk_sub[I];
  const ds_cg_scalars* sc = a.sc + sv;
  if (CG && sc->done) return;
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x)
    tab32[k] = a.table[k];
  __syncthreads();
  const double var_s = a.sub_cv[sv].x;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int row = I * DS_BS + lane;
  const bool real = lane < a.blk_nreal[I];
  const float4 li = a.tgt[row];
  const float tlast = a.tlast;
  float acc = 0.0f;
  const int kb = __builtin_amdgcn_readfirstlane(a.tptr[I]) + wave;
  const int ke = __builtin_amdgcn_readfirstlane(a.tptr[I + 1]);
  for (int k = kb; k < ke; k += nw) {
    const int j0 = __builtin_amdgcn_readfirstlane(tidx[k]) * DS_TS;
    const float4 o = toff[k];
    const float xi = li.x + o.x, yi = li.y + o.y, zi = li.z + o.z;
    const float ri = fmaf(zi, zi, fmaf(yi, yi, xi * xi));
    const float mx = -2.0f * xi, my = -2.0f * yi, mz = -2.0f * zi;
#pragma unroll
    for (int jj = 0; jj < DS_TS; ++jj) {
      const float4 pj = loc[j0 + jj];   // {X, Y, Z, X^2+Y^2+Z^2}
      const float uj = u32[j0 + jj];
      float xx;
      if (DIRECT) {
        const float dX = xi - pj.x, dY = yi - pj.y, dZ = zi - pj.z;
        xx = fmaf(dZ, dZ, fmaf(dY, dY, dX * dX));
      } else {
        xx = fmaf(mz, pj.z, fmaf(my, pj.y, fmaf(mx, pj.x, ri + pj.w)));
      }
      xx = __builtin_amdgcn_fmed3f(xx, 0.0f, tlast);
      const int ix = (int)xx;
      const float t = __builtin_amdgcn_fractf(xx);
      const float2 f = tab32[ix];
      acc = fmaf(fmaf(f.y, t, f.x), uj, acc);
    }
  }
  red[wave * 64 + lane] = (double)acc;
  __syncthreads();
  double v[1] = {0.0};
  if (wave == 0) {
    double y = 0.0;
    for (int w = 0; w < nw; ++w) y += red[w * 64 + lane];
    const double ai = a.in[row];
    const double wi = a.pos[row].w;
    double bi = (FORM == DS_FORM_W) ? wi * (var_s * y) + ai
                                    : var_s * y + wi * ai;
    if (!real) bi = 0.0;
    a.out[row] = bi;
    v[0] = wave_sum(ai * bi);
  }
  if (CG) store_partials<1>(v, a.partials, a.nblk, I);
}

}  // namespace

hipError_t launch_matvec32_ex(ds_problem* p, const float* u32,
                              const double* in, double* out, int cg,
                              hipStream_t st) {
  Matvec32Args a{};
  a.loc = p->loc32;
  a.tgt = p->tgt32;
  a.u32 = u32;
  a.pos = p->pos64;
  a.in = in;
  a.tptr = p->tptr;
  a.tidx = p->tidx;
  a.toff = p->toff32;
  a.table = p->table32;
  a.out = out;
  a.blk_sub = p->blk_sub;
  a.blk_nreal = p->blk_nreal;
  a.sub_blk0 = p->sub_blk0;
  a.sub_cv = p->sub_cv;
  a.blk_order = p->order_by_cost ? p->blk_order : nullptr;
  a.partials = p->partials;
  a.sc = p->sc;
  a.tlast = (float)(p->cv.table_n - 1);
  a.nblk = p->nblk;
  a.table_n = p->cv.table_n;
  const dim3 grid(p->nblk), block(64 * p->waves_per_block);
  const size_t shm = sizeof(float2) * p->cv.table_n;
  const bool w = p->form == DS_FORM_W;
#define DS_L32(F, C, D) \
  hipLaunchKernelGGL((k_matvec32<F, C, D>), grid, block, shm, st, a.loc, \
                     a.u32, a.tidx, a.toff, a)
  const bool d = p->f32_direct;
  if (cg) {
    if (w) { if (d) DS_L32(DS_FORM_W, true, true); else DS_L32(DS_FORM_W, true, false); }
    else { if (d) DS_L32(DS_FORM_SN, true, true); else DS_L32(DS_FORM_SN, true, false); }
  } else {
    if (w) { if (d) DS_L32(DS_FORM_W, false, true); else DS_L32(DS_FORM_W, false, false); }
    else { if (d) DS_L32(DS_FORM_SN, false, true); else DS_L32(DS_FORM_SN, false, false); }
  }
#undef DS_L32
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess || !cg) return err;
  CgEpilogue dq;
  dq.kind = EK_CG_DQ;
  return launch_finalize(p, dq, st);
}

}  // namespace ds
