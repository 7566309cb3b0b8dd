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
//
// k_matvec32m forms the same expanded argument on the matrix cores: per
// source tile and 16-row sub-block s of the target block, one
// v_mfma_f32_16x16x4f32 computes the 16 x 16 table arguments
//   xx_ij = |l_i|^2 + [l_i, 1] . [-2 Y_j, |Y_j|^2]
// with l_i the target in its block frame, Y_j = X_j - o the source moved
// into that frame (o = the entry's frame offset), A = [l_i, 1] and the
// accumulator input C = |l_i|^2 fixed per block, and B = one value per lane
// and tile.  Its table stores each interval's interpolant as the line
// g_k + df_k xx (g_k = f_k - k df_k, formed in FP64 on the host), so the
// VALU is left with the clamp, index, table address and two fma per pair
// (5 operations).  Lane l = 16 g + c holds A[i = c][k = g] and
// B[k = g][j = c]; the result D holds rows 4 g + r
// (r = 0..3) of column c, so each lane accumulates 16 rows x 1 source per
// tile and the 16 lanes of a group are summed once at the end.
#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

struct Matvec32Args {
  const float4* loc;      // sources: tile-frame coordinates, .w = |X|^2
  const float4* tgt;      // targets: block-frame coordinates
  const float* u32;       // scaled input (w .* a, or a) in FP32
  const double4* pos;     // w (W form) or N (SN form) in .w
  const double* wvec;     // w (W form)
  const double* in;       // FP64 input, own rows
  const int* tptr;
  const int* tidx;
  const unsigned char* tmask;   // per list entry: live 16-row sub-blocks
  const float4* toff;     // per list entry: block frame - tile frame
  const float2* table;    // {f[k], f[k+1]-f[k]}, last = {0, 0}
  const float2* table_g;  // {g, df}: the interpolant as g + df xx (MFMA)
  const int* blk_sub;
  const int* blk_nreal;
  const int* sub_blk0;
  const double4* sub_cv;
  const int* blk_order;   // launch slot -> block
  double* out;
  double* partials;
  const ds_cg_scalars* sc;
  float tlast;
  int nblk, table_n;
  int lane_zero;          // 0; lane & lane_zero makes an index lane-varying
  unsigned npad;          // slots (buffer extents of loc and u32)
  // symmetric MFMA kernel
  const int* hptr;
  float* pbuf;
  double* ydir;
  const int* gptr;
  const int* gslot;
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
  const int nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
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

typedef float v4f __attribute__((ext_vector_type(4)));

// One butterfly step over H pairs of values: the lane whose bit is clear
// keeps v[0..H) and sends v[H..2H), its partner (DPP control CTRL, a
// permutation within 16-lane rows) the reverse; v[0..H) becomes kept +
// received.
template <int H, int CTRL>
__device__ __forceinline__ void butterfly_step(float* v, int upper) {
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const float keep = upper ? v[j + H] : v[j];
    const float send = upper ? v[j] : v[j + H];
    const float recv = __int_as_float(__builtin_amdgcn_update_dpp(
        0, __float_as_int(send), CTRL, 0xF, 0xF, false));
    v[j] = keep + recv;
  }
}

// Raw buffer loads: the base and the per-entry offset stay in scalar
// registers (no per-lane 64-bit address arithmetic); loads past
// `bytes` return 0.
typedef __amdgpu_buffer_rsrc_t BufRsrc;
__device__ __forceinline__ BufRsrc buf_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0,
                                           (int)bytes, 0x00020000);
}
__device__ __forceinline__ float4 buf_f4(BufRsrc r, int voff, int soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return make_float4(__int_as_float(v[0]), __int_as_float(v[1]),
                     __int_as_float(v[2]), __int_as_float(v[3]));
}
__device__ __forceinline__ float buf_f1(BufRsrc r, int voff, int soff) {
  return __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff,
                                                             0));
}
__device__ __forceinline__ int buf_i1(BufRsrc r, int voff, int soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}

// Symmetric form: block I runs its half list (tiles T >= 4 I).  Every
// live (entry, sub-block s) 16 x 16 tile of values S_ij is evaluated once
// and used twice: y_i += S_ij u_j for the block's rows (summed in
// registers, written per row to ydir), and, when 4 I + s < T, the
// transposed y_j += S_ij u_i for the tile's rows, whose 16 column sums per
// entry go to pbuf.  For A x, k_sym_epi then adds each tile's incoming
// column sums and applies the epilogue; in a CG iteration the update kernel
// does that per row, and this kernel stores the d . q partial of its block,
// var_s (sum_i u_i y_i + sum over its transposed entries of u_j times the
// column sums) + sum_i a_i^2 (W form; SN: + N_i a_i^2), with u_i = w_i a_i
// (W) or a_i (SN) and u_j the sources' FP32 values.
template <int FORM, bool CG>
__global__ __launch_bounds__(256) void k_matvec32m(
    const float4* __restrict__ loc, const float* __restrict__ u32,
    const float4* __restrict__ toff, Matvec32Args a) {
  extern __shared__ float2 tab32[];
  __shared__ float red[4 * 64];
  __shared__ double dq_red[4];
  // the block's own vector values, then 16 zeros (the transposed factor
  // of sub-blocks without a transposed product)
  __shared__ __align__(16) float uown[64 + 16];
  const int I = a.blk_order[blockIdx.x];
  const int sv = a.blk_sub[I];
  const ds_cg_scalars* sc = a.sc + sv;
  if (CG && sc->done) return;
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x)
    tab32[k] = a.table_g[k];
  if (threadIdx.x < 64) uown[threadIdx.x] = u32[I * DS_BS + threadIdx.x];
  if (threadIdx.x < 16) uown[64 + threadIdx.x] = 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
  const int g = lane >> 4, c = lane & 15;
  const int row0 = I * DS_BS;
  const int T0 = DS_TPB * I;    // the block's first tile
  // A operands (target component g, or 1) and accumulator inputs |l|^2
  float av[4];
  v4f ci[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const float4 l = a.tgt[row0 + 16 * s + c];
    av[s] = g == 0 ? l.x : (g == 1 ? l.y : (g == 2 ? l.z : 1.0f));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float4 m = a.tgt[row0 + 16 * s + 4 * g + r];
      ci[s][r] = fmaf(m.z, m.z, fmaf(m.y, m.y, m.x * m.x));
    }
  }
  const float cx = g == 0 ? -2.0f : 0.0f, cy = g == 1 ? -2.0f : 0.0f;
  const float cz = g == 2 ? -2.0f : 0.0f, cw = g == 3 ? 1.0f : 0.0f;
  const float tlast = a.tlast;
  float acc[4][4];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[s][r] = 0.0f;
  const int kb = __builtin_amdgcn_readfirstlane(a.hptr[I]) + wave;
  const int ke = __builtin_amdgcn_readfirstlane(a.hptr[I + 1]);
  // (lz is 0 but not known to be, which keeps uniform-address loads on
  // vector loads, waited for where they are used)
  const unsigned cu = (unsigned)c;
  const unsigned lz = (unsigned)(lane & a.lane_zero);
  const BufRsrc rrec = buf_rsrc(toff, 0x7FFFFFFFu);
  const BufRsrc rslot = buf_rsrc(a.gslot, 0x7FFFFFFFu);
  const BufRsrc rloc = buf_rsrc(loc, a.npad * 16u);
  const BufRsrc ru = buf_rsrc(u32, a.npad * 4u);
  // column-sum slots: a vector store at lane offset 4 c from the slot's
  // wave-uniform base (scalar offset)
  // (problem creation keeps 64 bytes x slots within 32-bit offsets)
  const BufRsrc rpb = buf_rsrc(a.pbuf, 0xFFFFFFFFu);
  // byte offsets in uown of this lane's 4 own values for each sub-block
  // (the 16 zeros at 64 for a sub-block without a transposed product)
  unsigned uoff[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) uoff[s] = 4u * (16u * s + 4u * g);
  const int vrec = (int)(lz * 16u), vslot = (int)(lz * 4u);
  const int vloc = (int)(cu * 16u), vu = (int)(cu * 4u);
  // One entry e: its record o (frame offset; .w holds the bits of live
  // mask | tile << 4), the tile's source record X and vector value uj.
  // (The MFMA results live across entries: a sub-block an entry skips
  // keeps the last value, never read.)
  v4f d[4];
  float dqt = 0.0f;   // CG: sum of u_j times this lane's column sums
  auto entry = [&](const float4& o, const float4& X, float uj, int slot)
      __attribute__((always_inline)) {
    const unsigned bits = __builtin_amdgcn_readfirstlane(__float_as_int(o.w));
    const unsigned live = bits & 15u;
    const int dT = (int)(bits >> 4) - T0;   // >= 0 in a half list
    const unsigned tr = live & (dT >= 4 ? 15u : (1u << dT) - 1u);
    // a target in the tile frame is l + o, so the source in the block
    // frame is Y = X - o; B = -2 Y_g (g < 3) or |Y|^2 (g = 3), formed
    // without branches
    const float yx = X.x - o.x, yy = X.y - o.y, yz = X.z - o.z;
    const float y2 = fmaf(yz, yz, fmaf(yy, yy, yx * yx));
    const float b = fmaf(cx, yx, fmaf(cy, yy, fmaf(cz, yz, cw * y2)));
    // sub-blocks whose box is beyond the cutoff of the tile's box hold
    // only zero pairs (wave-uniform skips); the live sub-blocks' MFMAs are
    // issued together, ahead of the VALU work that reads them
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (live & (1u << s))
        d[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], b, ci[s], 0, 0, 0);
    float at = 0.0f;   // transposed: sum over this lane's 4 rows
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (!(live & (1u << s))) continue;
      const float4 uo = *reinterpret_cast<const float4*>(
          reinterpret_cast<const char*>(uown) +
          ((tr & (1u << s)) ? uoff[s] : 256u));
      const float uor[4] = {uo.x, uo.y, uo.z, uo.w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float xx = __builtin_amdgcn_fmed3f(d[s][r], 0.0f, tlast);
        const float2 f = tab32[(int)xx];
        const float sij = fmaf(f.y, xx, f.x);
        acc[s][r] = fmaf(sij, uj, acc[s][r]);
        at = fmaf(sij, uor[r], at);
      }
    }
    if (tr) {
      // the column sums: add the 4 lane groups of each column c (lanes
      // c + 16 g) with two half-swaps (32-lane halves, then 16-lane rows),
      // each of which leaves the pairwise sum in both halves
      const auto h32 = __builtin_amdgcn_permlane32_swap(
          __float_as_uint(at), __float_as_uint(at), false, false);
      at = __uint_as_float(h32[0]) + __uint_as_float(h32[1]);
      const auto h16 = __builtin_amdgcn_permlane16_swap(
          __float_as_uint(at), __float_as_uint(at), false, false);
      at = __uint_as_float(h16[0]) + __uint_as_float(h16[1]);
      if (g == 0) {
        __builtin_amdgcn_raw_buffer_store_b32(
            __float_as_int(at), rpb, (int)(cu * 4u),
            (int)(64u * (unsigned)__builtin_amdgcn_readfirstlane(slot)), 0);
        if (CG) dqt = fmaf(at, uj, dqt);
      }
    }
  };
  // Records are fetched two entries ahead, source records and vector
  // values one entry ahead, by vector loads from a wave-uniform base (held
  // in scalar registers) plus a fixed per-lane offset; each load is first
  // read after the next entry's arithmetic.  Indices past the list are
  // clamped to its last entry (duplicate loads, never used).
  const int kl = ke - 1;
  auto tile_of = [&](const float4& R) __attribute__((always_inline)) {
    return (__builtin_amdgcn_readfirstlane(__float_as_int(R.w)) >> 4) * DS_TS;
  };
  if (kb < ke) {
    // records and column-sum slots two entries ahead
    float4 R0 = buf_f4(rrec, vrec, 16 * kb);
    int G0 = buf_i1(rslot, vslot, 4 * kb);
    float4 R1 = buf_f4(rrec, vrec, 16 * min(kb + nw, kl));
    int G1 = buf_i1(rslot, vslot, 4 * min(kb + nw, kl));
    int j = tile_of(R0);
    float4 X0 = buf_f4(rloc, vloc, 16 * j);
    float u0 = buf_f1(ru, vu, 4 * j);
    for (int k = kb;; k += 2 * nw) {
      // entry k; prefetch the sources of k + nw, the record of k + 2 nw
      j = tile_of(R1);
      const float4 X1 = buf_f4(rloc, vloc, 16 * j);
      const float u1 = buf_f1(ru, vu, 4 * j);
      const float4 R2 = buf_f4(rrec, vrec, 16 * min(k + 2 * nw, kl));
      const int G2 = buf_i1(rslot, vslot, 4 * min(k + 2 * nw, kl));
      entry(R0, X0, u0, G0);
      if (k + nw >= ke) break;
      // entry k + nw; prefetch the sources of k + 2 nw, the record of
      // k + 3 nw
      j = tile_of(R2);
      X0 = buf_f4(rloc, vloc, 16 * j);
      u0 = buf_f1(ru, vu, 4 * j);
      const float4 R3 = buf_f4(rrec, vrec, 16 * min(k + 3 * nw, kl));
      const int G3 = buf_i1(rslot, vslot, 4 * min(k + 3 * nw, kl));
      entry(R1, X1, u1, G1);
      if (k + 2 * nw >= ke) break;
      R0 = R2;
      R1 = R3;
      G0 = G2;
      G1 = G3;
    }
  }
  // Sum over the 16 source columns (lanes c) of each lane group: a
  // butterfly that halves the 16 values at every step.  The pair (c, c')
  // of a step differs in one bit b of c (and maybe lower ones); the lane
  // with b clear keeps the lower half of its values and sends the upper,
  // its partner the reverse, and each adds what it receives.  Partners:
  // c ^ 15 (row mirror), c ^ 7 (half-row mirror), c ^ 2, c ^ 1 (quad
  // permutes); these span all 16 lanes, and lane c ends with the total of
  // value index c = 4 s + r.
  float v[16];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int r = 0; r < 4; ++r) v[4 * s + r] = acc[s][r];
  butterfly_step<8, 0x140>(v, c & 8);
  butterfly_step<4, 0x141>(v, c & 4);
  butterfly_step<2, 0x4E>(v, c & 2);
  butterfly_step<1, 0xB1>(v, c & 1);
  red[wave * 64 + 16 * (c >> 2) + 4 * g + (c & 3)] = v[0];
  if (CG) {
    const double t = wave_sum((double)dqt);
    if (lane == 0) dq_red[wave] = t;
  }
  __syncthreads();
  if (wave != 0) return;
  double y = 0.0;
  for (int w = 0; w < nw; ++w) y += (double)red[w * 64 + lane];
  const int row = row0 + lane;
  a.ydir[row] = y;
  if (CG) {
    const double var_s = a.sub_cv[sv].x;
    const double ai = a.in[row];
    const double wn = a.pos[row].w;
    const double ui = FORM == DS_FORM_W ? wn * ai : ai;
    const double ex = FORM == DS_FORM_W ? ai * ai : wn * ai * ai;
    const bool real = lane < a.blk_nreal[I];
    double tr = 0.0;
    for (int w = 0; w < nw; ++w) tr += dq_red[w];
    double v[1] = {wave_sum(real ? var_s * (ui * y) + ex : 0.0) +
                   var_s * tr};
    store_partials<1>(v, a.partials, a.nblk, I);
  }
}

// A x: rows of block I (sym_row: the block's own sums plus its tiles'
// column sums, then the epilogue), one wave per target block, four per
// workgroup.
template <int FORM>
__global__ __launch_bounds__(256) void k_sym_epi(Matvec32Args a, int nb) {
  const int slot = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (slot >= nb) return;
  const int I = a.blk_order[slot];
  const int lane = threadIdx.x & 63;
  const int row = I * DS_BS + lane;
  const SymRows rows{a.ydir, a.pbuf, a.gptr, a.pos, a.wvec, a.blk_nreal,
                     FORM};
  a.out[row] = sym_row(rows, I, lane, a.sub_cv[a.blk_sub[I]].x, a.in[row]);
}

}  // namespace

hipError_t launch_matvec32_ex(ds_problem* p, const float* u32,
                              const double* in, double* out, int cg,
                              hipStream_t st, int fold_alpha) {
  Matvec32Args a{};
  a.loc = p->loc32;
  a.tgt = p->tgt32;
  a.u32 = u32;
  a.pos = p->pos64;
  a.wvec = p->wvec;
  a.in = in;
  a.tptr = p->tptr;
  a.tidx = p->tidx;
  a.tmask = p->tmask;
  a.toff = p->toff32;
  a.table = p->table32;
  a.table_g = p->table32g;
  a.out = out;
  a.hptr = p->hptr;
  a.pbuf = p->pbuf;
  a.ydir = p->ydir;
  a.gptr = p->gptr;
  a.gslot = p->gslot;
  a.blk_sub = p->blk_sub;
  a.blk_nreal = p->blk_nreal;
  a.sub_blk0 = p->sub_blk0;
  a.sub_cv = p->sub_cv;
  a.partials = p->partials + (fold_alpha ? 2 * (size_t)p->nblk : 0);
  a.sc = p->sc;
  a.tlast = (float)(p->cv.table_n - 1);
  a.nblk = p->nblk;
  a.npad = (unsigned)p->npad;
  a.table_n = p->cv.table_n;
  const dim3 block(64 * p->waves_per_block);
  const size_t shm = sizeof(float2) * p->cv.table_n;
  const bool w = p->form == DS_FORM_W;
  // blocks of subvolumes without far points: MFMA (or expanded VALU) form;
  // the others: direct differences
  const int* ord = p->ord32 + (p->order_by_cost ? 0 : p->nblk);
  const int nf = p->n32_fast, nd = p->nblk - p->n32_fast;
#define DS_L32(F, C, D, G) \
  hipLaunchKernelGGL((k_matvec32<F, C, D>), dim3(G), block, shm, st, a.loc, \
                     a.u32, a.tidx, a.toff, a)
#define DS_L32M(F, C, G)                                                  \
  do {                                                                    \
    hipLaunchKernelGGL((k_matvec32m<F, C>), dim3(G), block, shm, st,      \
                       a.loc, a.u32, p->hrec, a);                         \
    if (!(C))                                                             \
      hipLaunchKernelGGL((k_sym_epi<F>), dim3(((G) + 3) / 4), dim3(256),  \
                         0, st, a, (G));                                  \
  } while (0)
  if (nf > 0) {
    a.blk_order = ord;
    if (p->f32_kind != DS_F32_VALU) {
      if (cg) { if (w) DS_L32M(DS_FORM_W, true, nf); else DS_L32M(DS_FORM_SN, true, nf); }
      else { if (w) DS_L32M(DS_FORM_W, false, nf); else DS_L32M(DS_FORM_SN, false, nf); }
    } else {
      if (cg) { if (w) DS_L32(DS_FORM_W, true, false, nf); else DS_L32(DS_FORM_SN, true, false, nf); }
      else { if (w) DS_L32(DS_FORM_W, false, false, nf); else DS_L32(DS_FORM_SN, false, false, nf); }
    }
  }
  if (nd > 0) {
    a.blk_order = ord + nf;
    if (cg) { if (w) DS_L32(DS_FORM_W, true, true, nd); else DS_L32(DS_FORM_SN, true, true, nd); }
    else { if (w) DS_L32(DS_FORM_W, false, true, nd); else DS_L32(DS_FORM_SN, false, true, nd); }
  }
#undef DS_L32M
#undef DS_L32
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess || !cg || fold_alpha) return err;
  CgEpilogue dq;
  dq.kind = EK_CG_DQ;
  return launch_finalize(p, dq, st);
}

}  // namespace ds
