// pairs64.hip -- FP64 pixel-pair kernels: the parity path.
//
// COMPILE WITH -ffp-contract=off.  The table argument of every pair follows
// the reference's operation order and rounding exactly
// (lib/wf_vector.cc:62-66, lib/signal_covar.h:56-58,67-73):
//     dx = xi - xj; dy = yi - yj; dz = zi - zj
//     xx = (c_perp * (dx*dx + dy*dy) + c_para * (dz*dz)) * dx_inv
// so the cutoff decision ((int)xx > table_n - 2 -> 0) and the table index are
// the reference's.  The interpolation f[i](i+1-xx) + f[i+1](xx-i) is taken as
// fma(f[i+1]-f[i], xx-i, f[i]) from an LDS table of {f[i], f[i+1]-f[i]} whose
// entry table_n-1 is {0, 0}: clamping xx to table_n-1 before the index
// conversion then yields the reference's exact 0 beyond the cutoff.  The
// per-pair value differs from the reference's by a few ulps; sums run in a
// different order.  Explicit fma() calls are fused regardless of the flag.
//
// Work decomposition: one workgroup per target block of 64 rows (lane = row),
// its waves striding over the block's list of 16-slot source tiles.  A tile's
// coordinates and vector entries are wave-uniform and are read with scalar
// loads; the only per-lane memory traffic in the loop is the table gather.
#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

enum Epi { EPI_NONE = EK_NONE, EPI_CG_DQ = EK_CG_DQ, EPI_CG_RES = EK_CG_RES };

// Sum over the listed source tiles of g(x_i, x_j) * u_j, with g the table
// interpolant (SQUARE: g^2 * u_j, for S W S).
template <bool SQUARE>
__device__ __forceinline__ double accumulate_tiles64(
    const double xi, const double yi, const double zi,
    const int* __restrict__ tidx, int k, const int ke, const int step,
    const double4* __restrict__ pos, const double* __restrict__ u,
    const double2* tab, const double cp, const double cl, const double dxi,
    const double tlast) {
  double acc = 0.0;
  for (; k < ke; k += step) {
    const int j0 = __builtin_amdgcn_readfirstlane(tidx[k]) * DS_TS;
#pragma unroll
    for (int jj = 0; jj < DS_TS; ++jj) {
      const double4 pj = pos[j0 + jj];
      const double uj = u[j0 + jj];
      const double dx = xi - pj.x;
      const double dy = yi - pj.y;
      const double dz = zi - pj.z;
      const double x_perp_2 = dx * dx + dy * dy;
      const double x_para_2 = dz * dz;
      double xx = (cp * x_perp_2 + cl * x_para_2) * dxi;
      xx = fmin(xx, tlast);
      const int ix = (int)xx;
      const double t = __builtin_amdgcn_fract(xx);   // == xx - ix, exactly
      const double2 f = tab[ix];
      const double gp = fma(f.y, t, f.x);
      acc = SQUARE ? fma(gp * gp, uj, acc) : fma(gp, uj, acc);
    }
  }
  return acc;
}

// The same sums with the target block cut into row groups of R rows
// (R = 16 or 32): lane (r, q) = (lane mod R, lane / R) holds rows g R + r of
// every group g and sources q, q + 64/R, ... of each listed tile, so a
// wave evaluates only the groups an entry's sub-block mask marks live
// (16-row sub-blocks that pass the cutoff box test; a 32-row group is live
// when either of its two is).  acc[g] holds the lane's partial sum of row
// g R + r.  Every pair's value is the R = 64 form's (same operations); the
// sums run in another order.
template <int R>
__device__ __forceinline__ void accumulate_groups64(
    const double (&px)[DS_BS / R], const double (&py)[DS_BS / R],
    const double (&pz)[DS_BS / R], double (&acc)[DS_BS / R],
    const int* __restrict__ tidx, const unsigned char* __restrict__ tmask,
    int k, const int ke, const int step, const double4* __restrict__ pos,
    const double* __restrict__ u, const double2* tab, const double cp,
    const double cl, const double dxi, const double tlast) {
  constexpr int G = DS_BS / R, NQ = 64 / R, NJ = DS_TS / NQ;
  const int q = (threadIdx.x & 63) / R;
  for (; k < ke; k += step) {
    const int j0 = __builtin_amdgcn_readfirstlane(tidx[k]) * DS_TS;
    const unsigned m = __builtin_amdgcn_readfirstlane(tmask[k]);
    double4 pj[NJ];
    double uj[NJ];
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
      pj[t] = pos[j0 + q + t * NQ];
      uj[t] = u[j0 + q + t * NQ];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const unsigned bits = R == 16 ? (m >> g) & 1u : (m >> (2 * g)) & 3u;
      if (!bits) continue;   // wave-uniform
#pragma unroll
      for (int t = 0; t < NJ; ++t) {
        const double dx = px[g] - pj[t].x;
        const double dy = py[g] - pj[t].y;
        const double dz = pz[g] - pj[t].z;
        const double x_perp_2 = dx * dx + dy * dy;
        const double x_para_2 = dz * dz;
        double xx = (cp * x_perp_2 + cl * x_para_2) * dxi;
        xx = fmin(xx, tlast);
        const int ix = (int)xx;
        const double tt = __builtin_amdgcn_fract(xx);   // == xx - ix
        const double2 f = tab[ix];
        const double gp = fma(f.y, tt, f.x);
        acc[g] = fma(gp, uj[t], acc[g]);
      }
    }
  }
}

struct MatvecArgs {
  const double4* pos;     // sorted pixels {x, y, z, w or N}
  const double* in;       // matvec input a (own rows, epilogue)
  const double* u;        // scaled input: w .* a (W form) or a (SN form)
  const double* rhs;      // wd or d (residual)
  const double* diag;     // Jacobi diagonal (CG residual epilogue)
  const float* minv;      // block-Jacobi inverses, or nullptr
  int minv_mod;
  const int* tptr;
  const int* tidx;
  const unsigned char* tmask;   // per entry: live 16-row sub-blocks
  const double2* table;
  const int* blk_sub;     // subvolume of each block
  const int* blk_nreal;   // real rows of each block
  const int* sub_blk0;    // first block of each subvolume
  const double4* sub_cv;  // {var_s, c_perp, c_para, 0} per subvolume
  const int* blk_order;   // launch slot -> block (may be null)
  double* out;            // A a, or rhs - A a
  double* out2;           // CG residual: M^-1 r
  double* partials;       // per-block sums for the finalize kernel
  const ds_cg_scalars* sc;  // per subvolume
  double dx_inv, tlast;
  int nblk, table_n;
  int residual;           // out = rhs - A a
  int need_want;          // skip unless the subvolume wants a true residual
  int single;             // one subvolume
};

template <int FORM, int EPI, int R>
__global__ __launch_bounds__(256) void k_matvec64(MatvecArgs a) {
  extern __shared__ double2 tab[];
  __shared__ double red[4 * 64];
  // a single volume's true-residual gate returns after one load when no
  // true residual is due (the usual case in a mixed solve)
  if (EPI != EPI_NONE && a.need_want && a.single &&
      (a.sc->done || !a.sc->want_true))
    return;
  const int I = a.blk_order ? a.blk_order[blockIdx.x] : blockIdx.x;
  const int sv = a.blk_sub[I];
  const ds_cg_scalars* sc = a.sc + sv;
  if (EPI != EPI_NONE) {
    if (sc->done) return;
    if (a.need_want && !sc->want_true) return;
  }
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x) tab[k] = a.table[k];
  __syncthreads();
  const double4 cvs = a.sub_cv[sv];
  const double var_s = cvs.x;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
  const int row = I * DS_BS + lane;
  const bool real = lane < a.blk_nreal[I];
  const double4 pi = a.pos[row];
  const int kb = a.tptr[I], ke = a.tptr[I + 1];
  if (R == DS_BS) {
    const double acc = accumulate_tiles64<false>(
        pi.x, pi.y, pi.z, a.tidx, kb + wave, ke, nw, a.pos, a.u, tab, cvs.y,
        cvs.z, a.dx_inv, a.tlast);
    red[wave * 64 + lane] = acc;
  } else {
    constexpr int G = DS_BS / (R < DS_BS ? R : DS_BS);
    constexpr int RR = DS_BS / G;
    const int r = lane % RR;
    double px[G], py[G], pz[G], acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const double4 pg = a.pos[I * DS_BS + g * RR + r];
      px[g] = pg.x;
      py[g] = pg.y;
      pz[g] = pg.z;
      acc[g] = 0.0;
    }
    accumulate_groups64<RR>(px, py, pz, acc, a.tidx, a.tmask, kb + wave, ke,
                            nw, a.pos, a.u, tab, cvs.y, cvs.z, a.dx_inv,
                            a.tlast);
    // sum the 64/R source columns of each row in a fixed order
#pragma unroll
    for (int g = 0; g < G; ++g)
      for (int off = RR; off < 64; off <<= 1)
        acc[g] += __shfl_xor(acc[g], off, 64);
    if (lane < RR) {
#pragma unroll
      for (int g = 0; g < G; ++g) red[wave * 64 + g * RR + r] = acc[g];
    }
  }
  __syncthreads();
  double v[2] = {0.0, 0.0};
  if (wave == 0) {
    double y = 0.0;
    for (int w = 0; w < nw; ++w) y += red[w * 64 + lane];
    const double ai = a.in[row];
    // FORM W: (w S w + I) a;  FORM SN: (S + N) a  (lib/wf_vector.cc:67,62)
    double bi = (FORM == DS_FORM_W) ? pi.w * (var_s * y) + ai
                                    : var_s * y + pi.w * ai;
    if (a.residual) bi = a.rhs[row] - bi;
    if (!real) bi = 0.0;
    a.out[row] = bi;
    if (EPI == EPI_CG_DQ) {
      v[0] = ai * bi;  // d . q
    } else if (EPI == EPI_CG_RES) {
      const double si =
          a.minv ? bj_apply(a.minv + (size_t)(I % a.minv_mod) * DS_BJ_SIZE,
                            bi)
                 : (real ? bi / a.diag[row] : 0.0);
      a.out2[row] = si;
      v[0] = bi * si;  // r . M^-1 r
      v[1] = bi * bi;  // r . r
    }
  }
  if (EPI == EPI_NONE || wave != 0) return;
  if (EPI == EPI_CG_DQ) {
    double p1[1] = {wave_sum(v[0])};
    store_partials<1>(p1, a.partials, a.nblk, I);
  } else {
    double p2[2] = {wave_sum(v[0]), wave_sum(v[1])};
    store_partials<2>(p2, a.partials, a.nblk, I);
  }
}

// Multi-RHS FP64 residual of a shared-geometry batch (every subvolume the
// same pixels, weights and covariance; subvolume r's vector at r * ld):
// out_r = rhs_r - A x_r for every subvolume that wants a true residual,
// from one pass over subvolume 0's tile lists.  Each pair's value is the
// parity kernel's (same operations, FP64); the products with the 32
// right-hand sides run on v_mfma_f64_16x16x4f64: per list entry (64 target
// rows x 16 sources) and k-step of 4 sources, lane l = 16 g + m evaluates
// the pair (row 16 rg + m, source 4 ks + g) of each 16-row group rg (the
// MFMA's A element [m][g]) and passes U[source 4 ks + g][rhs 16 ct + m]
// (B[g][m]) from the tile's 16 x 32 values staged in LDS; D element q of
// lane l is row g + 4 q, right-hand side 16 ct + m.  The row sums are FP64
// in another order than the per-subvolume kernel's (the matrix cores sum
// four sources at a time), so the residuals agree with it to FP64
// rounding, not bit for bit.  (With a lane-per-row layout every product
// needs a broadcast LDS read of the right-hand side's value: the pass was
// LDS-bound at ~0.69 ms for 100k pixels x 32.)
constexpr int kMR = DS_DENSE_R;   // right-hand sides per pass

struct Res64Args {
  const double4* pos;
  const double* x;        // x of every subvolume
  const double* xu;       // w .* x (W) or x (SN)
  const double* rhs;
  const int* tptr;        // subvolume 0's lists
  const double2* table;
  const int* blk_nreal;
  const ds_cg_scalars* sc;
  double* out;
  double var_s, c_perp, c_para, dx_inv, tlast;
  int nblk1, ld, n_sub, table_n;
  int all;                // every live subvolume (initial residual)
  // mixed solver's refinement: once any subvolume asks for a true
  // residual, every live one gets it (the pass evaluates all right-hand
  // sides anyway), so their refinements stay in step
  int sync;
  ds_cg_scalars* sc_w;
};

typedef double v4d __attribute__((ext_vector_type(4)));

template <int FORM>
__global__ __launch_bounds__(256) void k_res64_mrhs(
    const double4* __restrict__ pos, const int* __restrict__ tidx,
    Res64Args a) {
  extern __shared__ double2 tab[];
  __shared__ double us[4][DS_TS][kMR];
  __shared__ double red[2 * 4 * 4 * 64];   // one wave's 32 sums per lane
  // uniform: no subvolume asked for a true residual.  Every subvolume's
  // flags are loaded in one batch (a short-circuit test would wait for
  // each load in turn: ~16 us for 32 subvolumes).
  int want = 0;
#pragma unroll
  for (int r = 0; r < kMR; ++r) {
    const int rr = r < a.n_sub ? r : 0;
    const int w = (a.sc[rr].done == 0) & (a.all | (a.sc[rr].want_true != 0));
    want |= r < a.n_sub ? w : 0;
  }
  if (!want) return;
  if (a.sync && blockIdx.x == 0 && threadIdx.x == 0)
    for (int r = 0; r < a.n_sub; ++r)
      if (!a.sc[r].done) a.sc_w[r].want_true = 1;
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x) tab[k] = a.table[k];
  __syncthreads();
  const int I = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int m = lane & 15, g = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
  const int row0 = I * DS_BS;   // subvolume 0's slots
  // targets of this lane's A elements: row 16 rg + m of each group
  double px[4], py[4], pz[4];
#pragma unroll
  for (int rg = 0; rg < 4; ++rg) {
    const double4 t = pos[row0 + 16 * rg + m];
    px[rg] = t.x;
    py[rg] = t.y;
    pz[rg] = t.z;
  }
  v4d acc[4][2];
#pragma unroll
  for (int rg = 0; rg < 4; ++rg)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) acc[rg][ct] = v4d{0.0, 0.0, 0.0, 0.0};
  const int kb = __builtin_amdgcn_readfirstlane(a.tptr[I]) + wave;
  const int ke = __builtin_amdgcn_readfirstlane(a.tptr[I + 1]);
  for (int k = kb; k < ke; k += nw) {
    const int j0 = __builtin_amdgcn_readfirstlane(tidx[k]) * DS_TS;
    // stage u_r[j0 + jj] (lane: rhs lane % 32, sources lane / 32 + 2q)
#pragma unroll
    for (int q = 0; q < DS_TS / 2; ++q) {
      const int jj = (lane >> 5) + 2 * q, r = lane & 31;
      us[wave][jj][r] =
          r < a.n_sub ? a.xu[(size_t)r * a.ld + j0 + jj] : 0.0;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int ks = 0; ks < DS_TS / 4; ++ks) {
      const int src = 4 * ks + g;
      const double4 pj = pos[j0 + src];
      const double b0 = us[wave][src][m], b1 = us[wave][src][16 + m];
#pragma unroll
      for (int rg = 0; rg < 4; ++rg) {
        const double dx = px[rg] - pj.x;
        const double dy = py[rg] - pj.y;
        const double dz = pz[rg] - pj.z;
        const double x_perp_2 = dx * dx + dy * dy;
        const double x_para_2 = dz * dz;
        double xx = (a.c_perp * x_perp_2 + a.c_para * x_para_2) * a.dx_inv;
        xx = fmin(xx, a.tlast);
        const int ix = (int)xx;
        const double t = __builtin_amdgcn_fract(xx);   // == xx - ix, exactly
        const double2 f = tab[ix];
        const double gp = fma(f.y, t, f.x);
        acc[rg][0] =
            __builtin_amdgcn_mfma_f64_16x16x4f64(gp, b0, acc[rg][0], 0, 0, 0);
        acc[rg][1] =
            __builtin_amdgcn_mfma_f64_16x16x4f64(gp, b1, acc[rg][1], 0, 0, 0);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  // the waves' sums, added in wave order by wave 0
  for (int w = 1; w < nw; ++w) {
    __syncthreads();
    if (wave == w)
#pragma unroll
      for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            red[((rg * 2 + ct) * 4 + q) * 64 + lane] = acc[rg][ct][q];
    __syncthreads();
    if (wave == 0)
#pragma unroll
      for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            acc[rg][ct][q] += red[((rg * 2 + ct) * 4 + q) * 64 + lane];
  }
  if (wave != 0) return;
  const int nreal = a.blk_nreal[I];
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    const int r = 16 * ct + m;   // this lane's right-hand side
    if (r >= a.n_sub) continue;
    if (a.sc[r].done || !(a.all || a.sync || a.sc[r].want_true)) continue;
#pragma unroll
    for (int rg = 0; rg < 4; ++rg)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int rw = 16 * rg + g + 4 * q;   // row within the block
        const size_t gi = (size_t)r * a.ld + row0 + rw;
        const double y = acc[rg][ct][q];
        const double ai = a.x[gi];
        const double wn = pos[row0 + rw].w;
        double bi = (FORM == DS_FORM_W) ? wn * (a.var_s * y) + ai
                                        : a.var_s * y + wn * ai;
        bi = a.rhs[gi] - bi;
        if (rw >= nreal) bi = 0.0;
        a.out[gi] = bi;
      }
  }
}

// Map targets: the uniform grid of lib/util.cc:25-46, cut into 4x4x4 tiles;
// lane -> cell (lane>>4, (lane>>2)&3, lane&3) of the tile.
struct MapArgs {
  const double4* pos;
  const double* u;
  const int* tptr;
  const int* tidx;
  const double2* table;
  double* m;
  double lx, ly, lz;
  int nx, ny, nz, tx, ty, tz;
  double var_s, c_perp, c_para, dx_inv, tlast;
  int table_n;
  int t0;       // first map tile of the launch
};

template <bool SQUARE>
__global__ __launch_bounds__(256) void k_map64(MapArgs a) {
  extern __shared__ double2 tab[];
  __shared__ double red[4 * 64];
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x) tab[k] = a.table[k];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
  const int T = a.t0 + (int)blockIdx.x;
  const int tiz = T % a.tz, tiy = (T / a.tz) % a.ty, tix = T / (a.tz * a.ty);
  const int ix = tix * DS_MAP_T + (lane >> 4);
  const int iy = tiy * DS_MAP_T + ((lane >> 2) & 3);
  const int iz = tiz * DS_MAP_T + (lane & 3);
  // lib/util.cc:30-39
  const double dx = a.lx / a.nx, dy = a.ly / a.ny, dz = a.lz / a.nz;
  const double xi = dx * (ix + 0.5);
  const double yi = dy * (iy + 0.5);
  const double zi = dz * (iz + 0.5);
  const double acc = accumulate_tiles64<SQUARE>(
      xi, yi, zi, a.tidx, a.tptr[T] + wave, a.tptr[T + 1], nw, a.pos, a.u,
      tab, a.c_perp, a.c_para, a.dx_inv, a.tlast);
  red[wave * 64 + lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  double y = 0.0;
  for (int w = 0; w < nw; ++w) y += red[w * 64 + lane];
  if (ix < a.nx && iy < a.ny && iz < a.nz) {
    const size_t i = ((size_t)ix * a.ny + iy) * a.nz + iz;
    a.m[i] = SQUARE ? (a.var_s * a.var_s) * y : a.var_s * y;
  }
}

// Source-tile lists of the map grid's 4x4x4 tiles, built on the device by
// a full scan: one wavefront per 4 grid tiles tests every source tile box with
// the cutoff test of build_tile_lists (same operations, no contraction), so
// the lists, ascending in tile index, equal the host's.  Pass 0 counts each
// tile's entries, pass 1 writes them at the offsets the host scanned.
struct GridListArgs {
  const double* slo;      // source tile boxes, 3 per tile (empty: lo > hi)
  const double* shi;
  int* count;
  const int* ptr;
  int* idx;
  double lx, ly, lz;
  int nx, ny, nz, tx, ty, tz;
  double c_perp, c_para, dx_inv, limit;
  int ntile;
};

__device__ inline double axis_gap(double alo, double ahi, double blo,
                                  double bhi) {
  if (blo > ahi) return blo - ahi;
  if (alo > bhi) return alo - bhi;
  return 0.0;
}

constexpr int kGridTilesPerWave = 4;

template <bool FILL>
__global__ __launch_bounds__(256) void k_grid_lists(GridListArgs a) {
  constexpr int kT = kGridTilesPerWave;
  const int lane = threadIdx.x & 63;
  const int nt = a.tx * a.ty * a.tz;
  const int T0 = __builtin_amdgcn_readfirstlane(
      (blockIdx.x * 4 + (threadIdx.x >> 6)) * kT);
  if (T0 >= nt) return;
  // the tiles' cell-centre boxes (lib/util.cc:30-39 for the centres)
  const double dx = a.lx / a.nx, dy = a.ly / a.ny, dz = a.lz / a.nz;
  double alo[kT][3], ahi[kT][3];
  int base[kT], cnt[kT];
#pragma unroll
  for (int t = 0; t < kT; ++t) {
    const int T = min(T0 + t, nt - 1);
    const int tiz = T % a.tz, tiy = (T / a.tz) % a.ty, tix = T / (a.tz * a.ty);
    const int x0 = tix * DS_MAP_T, x1 = min(a.nx - 1, x0 + DS_MAP_T - 1);
    const int y0 = tiy * DS_MAP_T, y1 = min(a.ny - 1, y0 + DS_MAP_T - 1);
    const int z0 = tiz * DS_MAP_T, z1 = min(a.nz - 1, z0 + DS_MAP_T - 1);
    alo[t][0] = dx * (x0 + 0.5);
    ahi[t][0] = dx * (x1 + 0.5);
    alo[t][1] = dy * (y0 + 0.5);
    ahi[t][1] = dy * (y1 + 0.5);
    alo[t][2] = dz * (z0 + 0.5);
    ahi[t][2] = dz * (z1 + 0.5);
    base[t] = (FILL && T0 + t < nt) ? a.ptr[T0 + t] : 0;
    cnt[t] = 0;
  }
  for (int J0 = 0; J0 < a.ntile; J0 += 64) {
    const int J = J0 + lane;
    bool live = false;
    double l[3] = {0.0, 0.0, 0.0}, h[3] = {0.0, 0.0, 0.0};
    if (J < a.ntile) {
      const double* lo = a.slo + 3 * (size_t)J;
      const double* hi = a.shi + 3 * (size_t)J;
      l[0] = lo[0]; l[1] = lo[1]; l[2] = lo[2];
      h[0] = hi[0]; h[1] = hi[1]; h[2] = hi[2];
      live = l[0] <= h[0];
    }
#pragma unroll
    for (int t = 0; t < kT; ++t) {
      bool keep = false;
      if (live && T0 + t < nt) {
        const double g0 = axis_gap(alo[t][0], ahi[t][0], l[0], h[0]);
        const double g1 = axis_gap(alo[t][1], ahi[t][1], l[1], h[1]);
        const double g2 = axis_gap(alo[t][2], ahi[t][2], l[2], h[2]);
        const double x_perp_2 = (g0 * g0 + g1 * g1);
        const double x_para_2 = g2 * g2;
        const double ga_perp = a.c_perp * x_perp_2;
        const double ga_para = a.c_para * x_para_2;
        keep = (ga_perp + ga_para) * a.dx_inv < a.limit;
      }
      const unsigned long long b = __ballot(keep);
      if (FILL && keep) {
        const int r = __builtin_amdgcn_mbcnt_hi(
            (unsigned)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
        a.idx[base[t] + cnt[t] + r] = J;
      }
      cnt[t] += __popcll(b);
    }
  }
  if (!FILL && lane == 0) {
#pragma unroll
    for (int t = 0; t < kT; ++t)
      if (T0 + t < nt) a.count[T0 + t] = cnt[t];
  }
}

// Arbitrary target points, sorted into blocks of 64 by their own blocking.
struct PointsArgs {
  const double4* tgt;     // sorted targets (npad_t)
  const int* tgt_perm;    // sorted slot -> output index
  const double4* pos;
  const double* u;
  const int* tptr;
  const int* tidx;
  const double2* table;
  double* m;
  int n_tgt;
  double var_s, c_perp, c_para, dx_inv, tlast;
  int table_n;
};

template <bool SQUARE>
__global__ __launch_bounds__(256) void k_points64(PointsArgs a) {
  extern __shared__ double2 tab[];
  __shared__ double red[4 * 64];
  for (int k = threadIdx.x; k < a.table_n; k += blockDim.x) tab[k] = a.table[k];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
  const int I = blockIdx.x;
  const int slot = I * DS_BS + lane;
  const double4 ti = a.tgt[slot];
  const double acc = accumulate_tiles64<SQUARE>(
      ti.x, ti.y, ti.z, a.tidx, a.tptr[I] + wave, a.tptr[I + 1], nw, a.pos,
      a.u, tab, a.c_perp, a.c_para, a.dx_inv, a.tlast);
  red[wave * 64 + lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  double y = 0.0;
  for (int w = 0; w < nw; ++w) y += red[w * 64 + lane];
  if (slot < a.n_tgt)
    a.m[a.tgt_perm[slot]] = SQUARE ? (a.var_s * a.var_s) * y : a.var_s * y;
}

}  // namespace

// ---- launchers ------------------------------------------------------------

hipError_t launch_matvec64_ex(ds_problem* p, int residual, const double* in,
                              const double* u, double* out,
                              const CgEpilogue& e, hipStream_t st) {
  MatvecArgs a{};
  a.pos = p->pos64;
  a.in = in;
  a.u = u;
  a.rhs = p->rhs64;
  a.diag = p->diag;
  a.minv = p->use_minv() ? p->minv : nullptr;
  a.minv_mod = p->minv_mod;
  a.tptr = p->tptr;
  a.tidx = p->tidx;
  a.tmask = p->tmask;
  a.table = p->table64;
  a.blk_sub = p->blk_sub;
  a.blk_nreal = p->blk_nreal;
  a.sub_blk0 = p->sub_blk0;
  a.sub_cv = p->sub_cv;
  a.blk_order = p->order_by_cost ? p->blk_order : nullptr;
  a.out = out;
  a.out2 = e.out2;
  a.partials = p->partials;
  a.sc = p->sc;
  a.dx_inv = p->cv.dx_inv;
  a.tlast = (double)(p->cv.table_n - 1);
  a.nblk = p->nblk;
  a.table_n = p->cv.table_n;
  a.residual = residual;
  a.need_want = e.need_want;
  a.single = p->n_sub == 1;
  const dim3 grid(p->nblk), block(64 * p->waves_per_block);
  const size_t shm = sizeof(double2) * p->cv.table_n;
  const bool w = p->form == DS_FORM_W;
#define DS_L(F, E)                                                         \
  do {                                                                     \
    if (p->f64_rows == 16)                                                 \
      hipLaunchKernelGGL((k_matvec64<F, E, 16>), grid, block, shm, st, a); \
    else if (p->f64_rows == 32)                                            \
      hipLaunchKernelGGL((k_matvec64<F, E, 32>), grid, block, shm, st, a); \
    else                                                                   \
      hipLaunchKernelGGL((k_matvec64<F, E, 64>), grid, block, shm, st, a); \
  } while (0)
  if (e.kind == EPI_NONE) {
    if (w) DS_L(DS_FORM_W, EPI_NONE); else DS_L(DS_FORM_SN, EPI_NONE);
  } else if (e.kind == EPI_CG_DQ) {
    if (w) DS_L(DS_FORM_W, EPI_CG_DQ); else DS_L(DS_FORM_SN, EPI_CG_DQ);
  } else {
    if (w) DS_L(DS_FORM_W, EPI_CG_RES); else DS_L(DS_FORM_SN, EPI_CG_RES);
  }
#undef DS_L
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess || e.kind == EPI_NONE || e.fold) return err;
  return launch_finalize(p, e, st);
}

hipError_t launch_res64_mrhs(ds_problem* p, const double* in,
                             const double* u, double* out, int all,
                             hipStream_t st, int sync) {
  Res64Args a{};
  a.sync = sync;
  a.sc_w = p->sc;
  a.pos = p->pos64;
  a.x = in;
  a.xu = u;
  a.rhs = p->rhs64;
  a.tptr = p->tptr;
  a.table = p->table64;
  a.blk_nreal = p->blk_nreal;
  a.sc = p->sc;
  a.out = out;
  a.var_s = p->sub_cv_h[0].var_s;
  a.c_perp = p->sub_cv_h[0].c_perp;
  a.c_para = p->sub_cv_h[0].c_para;
  a.dx_inv = p->cv.dx_inv;
  a.tlast = (double)(p->cv.table_n - 1);
  a.nblk1 = p->nblk / p->n_sub;
  a.ld = p->npad / p->n_sub;
  a.n_sub = p->n_sub;
  a.table_n = p->cv.table_n;
  a.all = all;
  const dim3 grid(a.nblk1), block(64 * p->waves_per_block);
  const size_t shm = sizeof(double2) * p->cv.table_n;
  if (p->form == DS_FORM_W)
    hipLaunchKernelGGL((k_res64_mrhs<DS_FORM_W>), grid, block, shm, st,
                       p->pos64, p->tidx, a);
  else
    hipLaunchKernelGGL((k_res64_mrhs<DS_FORM_SN>), grid, block, shm, st,
                       p->pos64, p->tidx, a);
  return hipGetLastError();
}

hipError_t launch_map64(int mode, const MapGrid& g, const double4* pos,
                        const double* u, const int* tptr, const int* tidx,
                        const double2* table64, const CovarDev& cv,
                        double* m, hipStream_t st, int x0, int x1) {
  if (x1 < 0) x1 = g.tx;
  if (x1 <= x0) return hipSuccess;
  MapArgs a{};
  a.pos = pos;
  a.u = u;
  a.tptr = tptr;
  a.tidx = tidx;
  a.table = table64;
  a.m = m;
  a.lx = g.lx; a.ly = g.ly; a.lz = g.lz;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.tx = g.tx; a.ty = g.ty; a.tz = g.tz;
  a.var_s = cv.var_s;
  a.c_perp = cv.c_perp;
  a.c_para = cv.c_para;
  a.dx_inv = cv.dx_inv;
  a.tlast = (double)(cv.table_n - 1);
  a.table_n = cv.table_n;
  const size_t per_x = (size_t)g.ty * g.tz;   // tiles per x tile index
  a.t0 = (int)(per_x * x0);
  const dim3 grid((unsigned)(per_x * (x1 - x0))), block(256);
  const size_t shm = sizeof(double2) * cv.table_n;
  if (mode == MM_SIGMA)
    hipLaunchKernelGGL(k_map64<true>, grid, block, shm, st, a);
  else
    hipLaunchKernelGGL(k_map64<false>, grid, block, shm, st, a);
  return hipGetLastError();
}

hipError_t launch_grid_lists(const MapGrid& g, const double* slo,
                             const double* shi, int ntile,
                             const CovarDev& cv, int* count, const int* ptr,
                             int* idx, hipStream_t st) {
  GridListArgs a{};
  a.slo = slo;
  a.shi = shi;
  a.count = count;
  a.ptr = ptr;
  a.idx = idx;
  a.lx = g.lx; a.ly = g.ly; a.lz = g.lz;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.tx = g.tx; a.ty = g.ty; a.tz = g.tz;
  a.c_perp = cv.c_perp;
  a.c_para = cv.c_para;
  a.dx_inv = cv.dx_inv;
  a.limit = cull_limit(cv);
  a.ntile = ntile;
  const size_t nt = (size_t)g.tx * g.ty * g.tz;
  const size_t waves = (nt + kGridTilesPerWave - 1) / kGridTilesPerWave;
  const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
  if (idx)
    hipLaunchKernelGGL(k_grid_lists<true>, grid, block, 0, st, a);
  else
    hipLaunchKernelGGL(k_grid_lists<false>, grid, block, 0, st, a);
  return hipGetLastError();
}

hipError_t launch_points64(int mode, int n_tgt, const double4* tgt,
                           const int* tgt_perm, const double4* pos,
                           const double* u, const int* tptr, const int* tidx,
                           int n_tblk, const double2* table64,
                           const CovarDev& cv, double* m, hipStream_t st) {
  PointsArgs a{};
  a.tgt = tgt;
  a.tgt_perm = tgt_perm;
  a.pos = pos;
  a.u = u;
  a.tptr = tptr;
  a.tidx = tidx;
  a.table = table64;
  a.m = m;
  a.n_tgt = n_tgt;
  a.var_s = cv.var_s;
  a.c_perp = cv.c_perp;
  a.c_para = cv.c_para;
  a.dx_inv = cv.dx_inv;
  a.tlast = (double)(cv.table_n - 1);
  a.table_n = cv.table_n;
  const dim3 grid(n_tblk), block(256);
  const size_t shm = sizeof(double2) * cv.table_n;
  if (mode == MM_SIGMA)
    hipLaunchKernelGGL(k_points64<true>, grid, block, shm, st, a);
  else
    hipLaunchKernelGGL(k_points64<false>, grid, block, shm, st, a);
  return hipGetLastError();
}

}  // namespace ds
