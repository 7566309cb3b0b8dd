// dense_bf16.hip -- explicit BF16 covariance block times many right-hand
// sides on the matrix cores (the multi-RHS inner operator).
//
// A batch whose subvolumes share one geometry (several data vectors over the
// same pixels, e.g. noise realizations) has one S for all of them.  S is
// built once, explicitly, in BF16 (k_build_s: the FP64 parity interpolant
// of lib/signal_covar.h:53-77 rounded to nearest even), and the inner matvec
// of the mixed solver becomes Y = S U for the n x R block U of scaled
// directions (w .* d per right-hand side, BF16), on MFMA with FP32
// accumulation (k_gemm_bf16), followed by the per-right-hand-side epilogue
// q = w .* Y + d (W form) or Y + N .* d (SN form) with the d . q reduction
// and alpha (k_dense_epi).  FP64 true residuals (the matrix-free parity
// kernel) keep the stopping rule exact.
//
// k_gemm_bf16: a wave owns rows [row0, row0+32) x all R = 32 columns of a K
// range (split K: blockIdx.y).  Per K-chunk of 64 a lane of half h = lane/32
// reads the contiguous run k0 + 32h .. k0 + 32h + 31 of its row of S and of
// its column of U (the MFMA's k-slots of a half are fed from that run; A
// and B use the same assignment, so every product S[row][k] U[k][col] of
// the chunk is formed exactly once).  Each row's 128 bytes per chunk are
// read by two lanes, and fed to v_mfma_f32_32x32x16_bf16 (gfx950's BF16
// MFMA, 16 k-slots: 8 per lane) in 4 steps.  Output tile layout: register
// j of lane l holds row 8*(j/4) + 4*(l/32) + j%4, column l%32.
//
// k_fused_mrhs: the same multi-RHS product without materializing S.  A
// batch of subvolumes sharing one geometry has one set of tile lists; per
// (64-row target block, 16-source tile) entry the VALU evaluates the 64 x 16
// interpolant values g_ij in FP32 (the expanded table argument of the FP32
// pair kernel: lane = target row), rounds them to BF16, and one K = 16
// step of v_mfma_f32_32x32x16_bf16 per 32-row half multiplies them into all
// R right-hand sides at once (A = g tile, B = the BF16 directions of the 16
// sources for the 32 RHS).  The MFMA A layout wants, in lane l, the 8
// sources 8 (l/32) .. 8 (l/32) + 7 of row 32 rb + l%32; half of those
// values live in lane l ^ 32, which sends them with cross-half shuffles
// (4 per tile).  The epilogue
// scales by var_s in FP64, adds the identity or noise term per RHS and
// stores the d . q partial of each subvolume's block.
// COMPILE WITH -ffp-contract=off (the S build is FP64 parity arithmetic).
#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kR = DS_DENSE_R;     // right-hand sides per GEMM

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// S[i][j] for i < mpad, j < ld (row-major BF16); zero outside [0, n)^2.
__global__ __launch_bounds__(256) void k_build_s(const double4* pos, int n,
                                                 int ld, double var_s,
                                                 double c_perp, double c_para,
                                                 double dx_inv,
                                                 const double* table,
                                                 int table_n,
                                                 unsigned short* S) {
  const int i = blockIdx.x;   // rows on x: grid y is limited to 65536
  const int j0 = (blockIdx.y * blockDim.x + threadIdx.x) * 8;
  if (j0 >= ld) return;
  unsigned short v[8];
  const double4 pi = i < n ? pos[i] : make_double4(0, 0, 0, 0);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int j = j0 + t;
    double s = 0.0;
    if (i < n && j < n) {
      const double4 pj = pos[j];
      const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const double x_perp_2 = (dx * dx + dy * dy);
      const double x_para_2 = dz * dz;
      const double xx = (c_perp * x_perp_2 + c_para * x_para_2) * dx_inv;
      if (xx < (double)(table_n - 1)) {
        const int ix = (int)xx;
        s = var_s * (table[ix] * ((double)(ix + 1) - xx) +
                     table[ix + 1] * (xx - (double)ix));
      }
    }
    v[t] = bf16_rne((float)s);
  }
  uint4 w;
  w.x = v[0] | ((unsigned)v[1] << 16);
  w.y = v[2] | ((unsigned)v[3] << 16);
  w.z = v[4] | ((unsigned)v[5] << 16);
  w.w = v[6] | ((unsigned)v[7] << 16);
  *reinterpret_cast<uint4*>(S + (size_t)i * ld + j0) = w;
}

// Yp[split][col][row] = sum over the split's K range of S[row][k] U[col][k]
//
// A workgroup of kW waves covers kW x 32 rows and one K range (split K:
// blockIdx.y).  The 64-wide K chunks of U (32 columns x 64 k, 4 KB) are
// staged once per workgroup in LDS (double-buffered, rows padded by 16 B
// so the 32 column reads of a wave spread over the banks), so U is read
// from L2 once per kW x 32 rows instead of once per wave; each wave streams
// its 32 rows of S from HBM with the next chunk's loads in flight while the
// current chunk's 8 MFMAs run.
constexpr int kW = 8;                 // waves per workgroup
constexpr int kUPitch = 64 * 2 + 16;  // LDS bytes per U column chunk

// The lane's 32 BF16 values of the chunk (its run of S, of U) in 4 steps of
// v_mfma_f32_32x32x16_bf16: step s feeds run elements 8 s .. 8 s + 7 as the
// lane's 8 k-slots (A and B alike).
__device__ __forceinline__ void mfma_chunk(const uint4 (&a)[4],
                                           const uint4 (&b)[4], v16f& acc) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        __builtin_bit_cast(v8bf, a[s]), __builtin_bit_cast(v8bf, b[s]), acc,
        0, 0, 0);
}

// kSub 64-wide K chunks per barrier: the workgroup stages them in LDS
// together and each wave keeps the next group's S loads in flight.
constexpr int kSub = 2;

__global__ __launch_bounds__(64 * kW) void k_gemm_bf16(
    const unsigned short* __restrict__ S, const unsigned short* __restrict__ U,
    float* __restrict__ Yp, int ld, int mpad, int klen,
    const int* __restrict__ done_count, int n_sub) {
  __shared__ __align__(16) unsigned char ub[2][kSub][kR * kUPitch];
  if (*done_count >= n_sub) return;   // every right-hand side converged
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 31, h = lane >> 5;
  const int row0 = blockIdx.x * (32 * kW) + wave * 32;
  const int split = blockIdx.y;
  const int kbeg = split * klen;
  const int kend = min(ld, kbeg + klen);
  const int nch = (kend - kbeg + 63) / 64;   // klen and ld are multiples of 64
  const int ngr = (nch + kSub - 1) / kSub;   // groups of kSub chunks
  const unsigned short* srow = S + (size_t)(row0 + r) * ld + 32 * h;
  // U chunk loader: thread t < 256 moves 16 B (column t / 8, part t % 8)
  const int ucol = threadIdx.x >> 3, upart = threadIdx.x & 7;
  auto stage = [&](int g, int buf) {
    if (threadIdx.x < kR * 8) {
#pragma unroll
      for (int u = 0; u < kSub; ++u) {
        const int c = kSub * g + u;
        if (c < nch) {
          const uint4 v = *reinterpret_cast<const uint4*>(
              U + (size_t)ucol * ld + kbeg + 64 * c + 8 * upart);
          *reinterpret_cast<uint4*>(&ub[buf][u][ucol * kUPitch + 16 * upart]) =
              v;
        }
      }
    }
  };
  auto load_s = [&](int g, uint4 (&dst)[kSub][4]) {
#pragma unroll
    for (int u = 0; u < kSub; ++u) {
      const int c = min(kSub * g + u, nch - 1);   // past the end: unused
#pragma unroll
      for (int q = 0; q < 4; ++q)
        dst[u][q] = reinterpret_cast<const uint4*>(srow + kbeg + 64 * c)[q];
    }
  };
  v16f acc = {};
  uint4 a[kSub][4], an[kSub][4];
  if (nch > 0) {
    stage(0, 0);
    load_s(0, a);
  }
  for (int g = 0; g < ngr; ++g) {
    __syncthreads();   // group g staged; group g - 1's buffer is free
    if (g + 1 < ngr) {
      stage(g + 1, (g + 1) & 1);
      load_s(g + 1, an);
    }
#pragma unroll
    for (int u = 0; u < kSub; ++u) {
      if (kSub * g + u >= nch) break;   // uniform
      uint4 b[4];
      const unsigned char* bp = &ub[g & 1][u][r * kUPitch + 64 * h];
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = reinterpret_cast<const uint4*>(bp)[q];
      mfma_chunk(a[u], b, acc);
    }
#pragma unroll
    for (int u = 0; u < kSub; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) a[u][q] = an[u][q];
  }
  float* y = Yp + ((size_t)split * kR + r) * mpad + row0;
#pragma unroll
  for (int j = 0; j < 16; ++j) y[8 * (j >> 2) + 4 * h + (j & 3)] = acc[j];
}

struct FusedArgs {
  const float4* tgt;          // targets, block frames (subvolume 0)
  const int* tptr;            // subvolume 0's lists
  const double4* pos;         // batch records (w or N in .w)
  const double* d;            // batch directions (FP64)
  const unsigned short* U;    // BF16 scaled directions [R][ld]
  const float2* table;        // {f, df}
  const int* blk_nreal;
  const ds_cg_scalars* sc;
  const int* done_count;
  double* q;
  double* partials;
  double var_s;
  float tlast;
  int nblk1, ld, n_sub, nblk, table_n, form;
};


__global__ __launch_bounds__(256) void k_fused_mrhs(
    const float4* __restrict__ loc, const int* __restrict__ tidx,
    const float4* __restrict__ toff, FusedArgs a) {
  extern __shared__ float2 tab32[];
  __shared__ float part[3][2][16][64];   // waves 1..3: partial tiles
  if (*a.done_count >= a.n_sub) return;   // every right-hand side converged
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x)
    tab32[k] = a.table[k];
  __syncthreads();
  const int I = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
  const int h = lane >> 5, r = lane & 31;
  const float4 li = a.tgt[I * DS_BS + lane];
  const float tlast = a.tlast;
  v16f acc0 = {}, acc1 = {};
  const unsigned short* ucol = a.U + (size_t)r * a.ld + 8 * h;
  const int kb = __builtin_amdgcn_readfirstlane(a.tptr[I]) + wave;
  const int ke = __builtin_amdgcn_readfirstlane(a.tptr[I + 1]);
  for (int k = kb; k < ke; k += nw) {
    const int j0 = __builtin_amdgcn_readfirstlane(tidx[k]) * DS_TS;
    const float4 o = toff[k];
    const float xi = li.x + o.x, yi = li.y + o.y, zi = li.z + o.z;
    const float ri = fmaf(zi, zi, fmaf(yi, yi, xi * xi));
    const float mx = -2.0f * xi, my = -2.0f * yi, mz = -2.0f * zi;
    unsigned P[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      unsigned short gb[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const float4 pj = loc[j0 + 2 * m + e];   // {X, Y, Z, |X|^2}
        float xx = fmaf(mz, pj.z, fmaf(my, pj.y, fmaf(mx, pj.x, ri + pj.w)));
        xx = __builtin_amdgcn_fmed3f(xx, 0.0f, tlast);
        const int ix = (int)xx;
        const float t = __builtin_amdgcn_fractf(xx);
        const float2 f = tab32[ix];
        gb[e] = bf16_rne(fmaf(f.y, t, f.x));
      }
      P[m] = gb[0] | ((unsigned)gb[1] << 16);
    }
    // one K = 16 step per 32-row half: lane l of half h supplies the
    // sources 8 h .. 8 h + 7 of row 32 rb + l % 32.  A lane keeps the half
    // of its own row that its k-slots take and sends the other half to
    // lane l ^ 32, which takes it for the same row.
    unsigned own[4], rcv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      own[q] = h ? P[4 + q] : P[q];
      rcv[q] = (unsigned)__shfl_xor((int)(h ? P[q] : P[4 + q]), 32, 64);
    }
    const uint4 ow = make_uint4(own[0], own[1], own[2], own[3]);
    const uint4 rv = make_uint4(rcv[0], rcv[1], rcv[2], rcv[3]);
    const uint4 bu = *reinterpret_cast<const uint4*>(ucol + j0);
    const v8bf bv = __builtin_bit_cast(v8bf, bu);
    acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        __builtin_bit_cast(v8bf, h ? rv : ow), bv, acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        __builtin_bit_cast(v8bf, h ? ow : rv), bv, acc1, 0, 0, 0);
  }
  // combine the waves' partial tiles
  if (wave > 0) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      part[wave - 1][0][j][lane] = acc0[j];
      part[wave - 1][1][j][lane] = acc1[j];
    }
  }
  __syncthreads();
  if (wave != 0) return;
  for (int w = 1; w < nw; ++w)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      acc0[j] += part[w - 1][0][j][lane];
      acc1[j] += part[w - 1][1][j][lane];
    }
  // epilogue: lane holds rhs r, rows 32 rb + 8 (j / 4) + 4 h + j % 4
  const bool live = r < a.n_sub && !a.sc[r].done;
  const int nreal = a.blk_nreal[I];
  double dot = 0.0;
  if (live) {
    const size_t g0 = ((size_t)r * a.nblk1 + I) * DS_BS;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int i = 32 * rb + 8 * (j >> 2) + 4 * h + (j & 3);
        const size_t g = g0 + i;
        const double y = a.var_s * (double)(rb ? acc1[j] : acc0[j]);
        const double di = a.d[g];
        const double wN = a.pos[g].w;
        double qi = a.form == DS_FORM_W ? wN * y + di : y + wN * di;
        if (i >= nreal) qi = 0.0;
        a.q[g] = qi;
        dot += di * qi;
      }
  }
  dot += __shfl_xor(dot, 32, 64);
  if (live && h == 0) a.partials[(size_t)r * a.nblk1 + I] = dot;
}

// One 64-thread block per target block of the batch: right-hand side
// sv = blk_sub[I], row i = (I - sub_blk0[sv]) * 64 + lane of the shared
// geometry.  q = w .* Y + d (W) or Y + N .* d (SN); d . q -> alpha.
struct EpiArgs {
  const float* Yp;
  const double4* pos;
  const double* d;
  const int* blk_sub;
  const int* blk_nreal;
  const int* sub_blk0;
  double* q;
  double* partials;
  const ds_cg_scalars* sc;
  int mpad, nsplit, form, nblk;
};

__global__ __launch_bounds__(64) void k_dense_epi(EpiArgs a) {
  const int I = blockIdx.x;
  const int sv = a.blk_sub[I];
  const ds_cg_scalars* sc = a.sc + sv;
  if (sc->done) return;
  const int b0 = a.sub_blk0[sv];
  const int lane = threadIdx.x;
  const int i = (I - b0) * DS_BS + lane;       // row of the shared geometry
  const size_t g = (size_t)I * DS_BS + lane;   // slot in the batch vector
  double y = 0.0;
  for (int s = 0; s < a.nsplit; ++s)
    y += (double)a.Yp[((size_t)s * kR + sv) * a.mpad + i];
  const double di = a.d[g];
  const double wN = a.pos[g].w;
  double qi = a.form == DS_FORM_W ? wN * y + di : y + wN * di;
  if (lane >= a.blk_nreal[I]) qi = 0.0;
  a.q[g] = qi;
  double v[1] = {wave_sum(di * qi)};
  store_partials<1>(v, a.partials, a.nblk, I);
}

}  // namespace

hipError_t dense_build(ds_problem* p, hipStream_t st) {
  const int n1 = p->sub_n[0];
  const int ld = p->dense_ld, mpad = p->dense_mpad;
  const dim3 grid(mpad, (ld / 8 + 255) / 256);
  hipLaunchKernelGGL(k_build_s, grid, dim3(256), 0, st, p->pos64, n1, ld,
                     p->cv.var_s, p->cv.c_perp, p->cv.c_para, p->cv.dx_inv,
                     p->table_raw, p->cv.table_n, p->dense_s);
  return hipGetLastError();
}

hipError_t dense_gemm(ds_problem* p, hipStream_t st) {
  const int ld = p->dense_ld;
  const int klen = ((ld + p->dense_split - 1) / p->dense_split + 63) / 64 * 64;
  const dim3 grid(p->dense_mpad / (32 * kW), p->dense_split);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(64 * kW), 0, st, p->dense_s,
                     p->du16, p->dense_y, ld, p->dense_mpad, klen,
                     p->done_count, p->n_sub);
  return hipGetLastError();
}

hipError_t fused_mrhs(ds_problem* p, hipStream_t st) {
  FusedArgs a{};
  const int nblk1 = p->nblk / p->n_sub;
  a.tgt = p->tgt32;
  a.tptr = p->tptr;
  a.pos = p->pos64;
  a.d = p->d;
  a.U = p->du16;
  a.table = p->table32;
  a.blk_nreal = p->blk_nreal;
  a.sc = p->sc;
  a.done_count = p->done_count;
  a.q = p->q;
  a.partials = p->partials;
  a.var_s = p->sub_cv_h[0].var_s;
  a.tlast = (float)(p->cv.table_n - 1);
  a.nblk1 = nblk1;
  a.ld = p->npad / p->n_sub;
  a.n_sub = p->n_sub;
  a.nblk = p->nblk;
  a.table_n = p->cv.table_n;
  a.form = p->form;
  hipLaunchKernelGGL(k_fused_mrhs, dim3(nblk1), dim3(64 * p->waves_per_block),
                     sizeof(float2) * p->cv.table_n, st, p->loc32, p->tidx,
                     p->toff32, a);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  CgEpilogue dq;
  dq.kind = EK_CG_DQ;
  return launch_finalize(p, dq, st);
}

hipError_t dense_epilogue(ds_problem* p, hipStream_t st) {
  EpiArgs a{};
  a.Yp = p->dense_y;
  a.pos = p->pos64;
  a.d = p->d;
  a.blk_sub = p->blk_sub;
  a.blk_nreal = p->blk_nreal;
  a.sub_blk0 = p->sub_blk0;
  a.q = p->q;
  a.partials = p->partials;
  a.sc = p->sc;
  a.mpad = p->dense_mpad;
  a.nsplit = p->dense_split;
  a.form = p->form;
  a.nblk = p->nblk;
  hipLaunchKernelGGL(k_dense_epi, dim3(p->nblk), dim3(DS_BS), 0, st, a);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  CgEpilogue dq;
  dq.kind = EK_CG_DQ;
  return launch_finalize(p, dq, st);
}

}  // namespace ds
