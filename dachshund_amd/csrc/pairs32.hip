// pairs32.hip -- FP32 pixel-pair matvec: the inner operator of the mixed
// solver.
//
// Same interpolant as the FP64 parity kernel (lib/signal_covar.h:53-77):
// S = var_s * (f[i] + (f[i+1]-f[i]) * (xx - i)), 0 beyond the cutoff, but
// evaluated in FP32 on pre-scaled coordinates, so that the table argument is
// just xx = dX^2 + dY^2 + dZ^2 with X = x * sqrt(c_perp * dx_inv),
// Z = z * sqrt(c_para * dx_inv).  Coordinates are stored relative to their
// block's centre; each (target block, source tile) list entry carries the
// FP32 offset between the two block centres (differenced in FP64 on the
// host), so large box coordinates never lose precision in FP32.
//
// Per pair: 3 subtractions, 1 mul + 2 fma (xx), clamp, index conversion,
// fraction, LDS gather of {f, df}, 2 fma: 12 VALU operations and one 8-byte
// LDS read.  The sum of each lane is FP32; the epilogue (identity or noise
// term, dot products) is FP64.
#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

struct Matvec32Args {
  const float4* loc;      // sorted, block-relative pre-scaled coordinates
  const float* u32;       // scaled input (w .* a, or a) in FP32
  const double4* pos;     // w (W form) or N (SN form) in .w
  const double* in;       // FP64 input, own rows
  const int* tptr;
  const int* tidx;
  const float4* toff;     // per list entry: target centre - source centre
  const float2* table;    // {f[k], f[k+1]-f[k]}, last = {0, 0}
  double* out;
  double* partials;
  ds_cg_scalars* sc;
  double var_s;
  float tlast;
  int n, table_n;
};

template <int FORM, bool CG>
__global__ __launch_bounds__(256) void k_matvec32(Matvec32Args a) {
  extern __shared__ float2 tab32[];
  __shared__ double red[4 * 64];
  __shared__ double red2[4];
  __shared__ bool is_last;
  if (CG && a.sc->done) return;
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x)
    tab32[k] = a.table[k];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int I = blockIdx.x;
  const int row = I * DS_BS + lane;
  const float4 li = a.loc[row];
  const float tlast = a.tlast;
  float acc = 0.0f;
  const int ke = a.tptr[I + 1];
  for (int k = a.tptr[I] + wave; k < ke; k += nw) {
    const int j0 = __builtin_amdgcn_readfirstlane(a.tidx[k]) * DS_TS;
    const float4 o = a.toff[k];
    const float xi = li.x + o.x, yi = li.y + o.y, zi = li.z + o.z;
#pragma unroll
    for (int jj = 0; jj < DS_TS; ++jj) {
      const float4 pj = a.loc[j0 + jj];
      const float uj = a.u32[j0 + jj];
      const float dX = xi - pj.x;
      const float dY = yi - pj.y;
      const float dZ = zi - pj.z;
      float xx = fmaf(dZ, dZ, fmaf(dY, dY, dX * dX));
      xx = fminf(xx, tlast);
      const int ix = (int)xx;
      const float t = xx - (float)ix;
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
    double bi = (FORM == DS_FORM_W) ? wi * (a.var_s * y) + ai
                                    : a.var_s * y + wi * ai;
    if (row >= a.n) bi = 0.0;
    a.out[row] = bi;
    v[0] = wave_sum(ai * bi);
  }
  if (!CG) return;
  if (!publish_partials<1>(v, a.partials, &a.sc->counter, &is_last)) return;
  gather_partials<1>(v, a.partials, red2);
  if (threadIdx.x == 0) finalize_alpha(a.sc, v[0]);
  release_ticket(&a.sc->counter);
}

}  // namespace

hipError_t launch_matvec32_ex(ds_problem* p, const float* u32,
                              const double* in, double* out, int cg,
                              hipStream_t st) {
  Matvec32Args a{};
  a.loc = p->loc32;
  a.u32 = u32;
  a.pos = p->pos64;
  a.in = in;
  a.tptr = p->tptr;
  a.tidx = p->tidx;
  a.toff = p->toff32;
  a.table = p->table32;
  a.out = out;
  a.partials = p->partials;
  a.sc = p->sc;
  a.var_s = p->cv.var_s;
  a.tlast = (float)(p->cv.table_n - 1);
  a.n = p->n;
  a.table_n = p->cv.table_n;
  const dim3 grid(p->nblk), block(64 * p->waves_per_block);
  const size_t shm = sizeof(float2) * p->cv.table_n;
  const bool w = p->form == DS_FORM_W;
  if (cg) {
    if (w) hipLaunchKernelGGL((k_matvec32<DS_FORM_W, true>), grid, block, shm, st, a);
    else hipLaunchKernelGGL((k_matvec32<DS_FORM_SN, true>), grid, block, shm, st, a);
  } else {
    if (w) hipLaunchKernelGGL((k_matvec32<DS_FORM_W, false>), grid, block, shm, st, a);
    else hipLaunchKernelGGL((k_matvec32<DS_FORM_SN, false>), grid, block, shm, st, a);
  }
  return hipGetLastError();
}

}  // namespace ds
