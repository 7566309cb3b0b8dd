// warmstart.hip -- the learned initial guess of the Wiener-filter solve.
//
// A point network predicts each pixel's solution value x_i = [(S+N)^-1 d]_i
// (the quantity the W-form solve returns, lib/wf_vector.cc:268-271) from its
// neighbourhood; the prediction enters PCG as the initial guess
// y0_i = x0_i / w_i, the reference's x in/out argument (lib/pcg.h:98).  The
// solve's stopping rule is unchanged, so the prediction only changes how
// many iterations are needed.
//
// Features of pixel i (FEAT = 5 K + 2, K = 32 nearest neighbours under the
// covariance metric, nearest first; absent neighbours are zeros):
//   per neighbour j: (x_j - x_i)/l_perp, (y_j - y_i)/l_perp, (z_j - z_i)/l_para,
//                    n_j / sqrt(var_s), d_j / sqrt(var_s)
//   own:             n_i / sqrt(var_s), d_i / sqrt(var_s)
// stored as BF16.  Network: FEAT -> 64 -> 64 -> 64 -> 1, ReLU, FP32 weights
// (layout [in][out]), activations rounded to BF16 (optional), output
// t = x_i sqrt(var_s).
//
// Kernels: k_knn_ring and k_knn (neighbour search over the problem's
// culled tile lists, once per problem: the ring search, then the general
// search on the tiles it leaves), k_ws_features, k_fwd (inference), k_fwd_bwd (training forward + per-pixel
// back-propagated deltas), k_wgrad (deterministic chunked weight-gradient
// sums), k_reduce_chunks, k_adam, k_gather_pred.
#include <cmath>
#include <cstring>
#include <vector>

#include "cg_device.h"
#include "ds_internal.h"
#include "warmstart.h"

namespace ds {

namespace {

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float(((unsigned)h) << 16);
}

// ---- neighbour search and features --------------------------------------

struct KnnArgs {
  const double4* pos;     // {x, y, z, w or N}
  const double* rhs;      // wd (W) or d (SN)
  const float4* tgt;
  const int* tptr;
  const int* tidx;
  int* nbr;               // npad x K neighbour rows, -1 = none
  const float4* tbox;     // per tile: min, max of its records (tile frame)
  unsigned short* feat;   // npad x WS_FEAT
  // tiles the ring search could not finish: fb[0] = count, fb[1..] = tiles
  // (k_knn takes tile fb[1 + blockIdx.x] when fb is set)
  int* fb;
  int n, form;
  float inv_lp, inv_ll, inv_sqrt_vs, cutoff;
  float r0;               // first ring radius (sqrt of d2 units)
  int ntile;
};

constexpr int KNN_BINS = 64;    // histogram bins per refinement level
constexpr int KNN_LEVELS = 3;   // refinement levels
constexpr int KNN_CAP = 48;     // candidates kept per row for the selection

// d2 of row (frame-relative li) and source slot j of list entry e: the
// table argument in FP32 with direct differences in the tile frame, the
// same operations in every pass.
__device__ __forceinline__ float knn_d2(const float4& li, const float4& o,
                                        const float4& pj) {
  const float xi = li.x + o.x, yi = li.y + o.y, zi = li.z + o.z;
  const float dX = xi - pj.x, dY = yi - pj.y, dZ = zi - pj.z;
  return fmaf(dZ, dZ, fmaf(dY, dY, dX * dX));
}

// Rows per wavefront of the neighbour search: the 16 rows of one tile, each
// row's candidates split over 4 lane groups (group g takes slots 4 g .. 4 g
// + 3 of every visited source tile).
constexpr int KNN_RW = DS_TS;
constexpr int KNN_G = 64 / KNN_RW;

// Visit every slot of the listed source tiles of the rows' block whose
// entry passes the distance filter: visit(d2, j, ok) with ok = the slot is
// a candidate neighbour of target `row` (below n, not row itself, d2 inside
// the cutoff, row itself below n).  The visitor runs for every slot of a
// visited entry so that it can be written without branches.
//
// Filter: an entry's pairs are all at least lb apart, the gap between the
// rows' box (bmin, bmax: the 16 rows in the block frame, shifted into the
// tile frame by o = block frame minus tile frame) and the tile's box (its
// 16 records in the tile frame), in the scaled units where distance^2 =
// d2, less a margin that covers FP32 rounding; entries with lb in [lo, hi)
// are visited.  With hi = r, every pair with computed d2 < r^2 lies in a
// visited entry; with lo = r, none of the visited pairs has d2 < r^2.
// A visited tile's 16 records are staged in LDS (trec, two buffers) by
// lanes 0-15 while the previous tile is computed.
template <class F>
__device__ __forceinline__ void knn_scan(const KnnArgs& a,
                                         const float4* __restrict__ loc,
                                         const float4* __restrict__ toff,
                                         float4* trec, int I, int row,
                                         const float4& li,
                                         const float4& bmin,
                                         const float4& bmax, float lo,
                                         float hi, F&& visit) {
  const int kb = __builtin_amdgcn_readfirstlane(a.tptr[I]);
  const int ke = __builtin_amdgcn_readfirstlane(a.tptr[I + 1]);
  const bool live = row < a.n;
  const bool all = lo == -__builtin_inff() && hi == __builtin_inff();
  const int lane = threadIdx.x & 63;
  const int g = lane / KNN_RW;
  int buf = 0;
  // 64 entries at a time: lane k reads entry base + k and applies the
  // filter, so the entries' loads overlap instead of forming a chain; the
  // passing entries then run in ascending order
  for (int base = kb; base < ke; base += 64) {
    const int el = base + lane;
    int Tl = 0;
    float4 ol = make_float4(0.f, 0.f, 0.f, 0.f);
    bool pass = false;
    if (el < ke) {
      Tl = a.tidx[el];
      ol = toff[el];
      pass = true;
      if (!all) {
        const float4 tmin = a.tbox[2 * Tl], tmax = a.tbox[2 * Tl + 1];
        const float gx = fmaxf(0.f, fmaxf(tmin.x - (bmax.x + ol.x),
                                          (bmin.x + ol.x) - tmax.x));
        const float gy = fmaxf(0.f, fmaxf(tmin.y - (bmax.y + ol.y),
                                          (bmin.y + ol.y) - tmax.y));
        const float gz = fmaxf(0.f, fmaxf(tmin.z - (bmax.z + ol.z),
                                          (bmin.z + ol.z) - tmax.z));
        const float scale = fabsf(ol.x) + fabsf(ol.y) + fabsf(ol.z) +
                            (bmax.x - bmin.x) + (bmax.y - bmin.y) +
                            (bmax.z - bmin.z);
        const float lb = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx))) -
                         (1e-3f + 1e-5f * scale);
        pass = lb >= lo && lb < hi;
      }
    }
    unsigned long long m = __ballot(pass);
    const int sl = lane & (DS_TS - 1);
    int k = m ? __builtin_ctzll(m) : 0;
    float4 nxt = m ? loc[__builtin_amdgcn_readlane(Tl, k) * DS_TS + sl]
                   : make_float4(0.f, 0.f, 0.f, 0.f);
    while (m) {
      m &= m - 1;
      const int T = __builtin_amdgcn_readlane(Tl, k);
      float4 o;
      o.x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ol.x), k));
      o.y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ol.y), k));
      o.z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ol.z), k));
      o.w = 0.f;
      if (lane < DS_TS) trec[buf * DS_TS + lane] = nxt;
      if (m) {
        k = __builtin_ctzll(m);
        nxt = loc[__builtin_amdgcn_readlane(Tl, k) * DS_TS + sl];
      }
      __syncthreads();
      const int j0 = T * DS_TS + KNN_G * g;
      const float4* rb = trec + buf * DS_TS + KNN_G * g;
      const bool full = T * DS_TS + DS_TS <= a.n;   // no padding slot
#pragma unroll
      for (int it = 0; it < KNN_G; ++it) {
        const int j = j0 + it;
        const float d2 = knn_d2(li, o, rb[it]);
        visit(d2, j, live && (full || j < a.n) && j != row &&
                         d2 < a.cutoff);
      }
      buf ^= 1;
    }
  }
}

// Level 0 bins d2 logarithmically, eight bins per octave over [1, 256)
// (the cutoff is 199): the bin is the float's exponent and top three
// mantissa bits; d2 < 1 falls in bin 0.  Levels 1 and 2 split the crossing
// bin of the level above into KNN_BINS equal bins.
__device__ __forceinline__ int knn_bin0(float d2) {
  const int b = (int)(__float_as_uint(d2) >> 20) - (127 << 3);
  return min(max(b, 0), KNN_BINS - 1);
}
__device__ __forceinline__ float knn_bin0_lower(int b) {
  return b <= 0 ? 0.0f : __uint_as_float((unsigned)(b + (127 << 3)) << 20);
}
__device__ __forceinline__ int knn_bin(float d2, float lo, float scale) {
  const int b = (int)((d2 - lo) * scale);
  return min(max(b, 0), KNN_BINS - 1);
}

// (d, j) < (e, k) in (distance, index) order.
__device__ __forceinline__ bool knn_less(float d, int j, float e, int k) {
  return d < e || (d == e && (unsigned)j < (unsigned)k);
}

// The box of each tile's 16 records in its frame (padding slots included:
// the box only grows).  One lane per record (coalesced loads), reduced over
// the tile's 16 lanes.
// (Also clears the ring search's unfinished-tile count.)
__global__ __launch_bounds__(256) void k_tile_boxes(
    const float4* __restrict__ loc, int ntile, float4* tbox, int* fb) {
  const int g = blockIdx.x * 256 + threadIdx.x;   // record
  if (g == 0) fb[0] = 0;
  const int T = g / DS_TS;
  const float4 v = loc[min(g, ntile * DS_TS - 1)];
  float lx = v.x, ly = v.y, lz = v.z, hx = v.x, hy = v.y, hz = v.z;
#pragma unroll
  for (int o = DS_TS / 2; o >= 1; o >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, o));
    ly = fminf(ly, __shfl_xor(ly, o));
    lz = fminf(lz, __shfl_xor(lz, o));
    hx = fmaxf(hx, __shfl_xor(hx, o));
    hy = fmaxf(hy, __shfl_xor(hy, o));
    hz = fmaxf(hz, __shfl_xor(hz, o));
  }
  if ((g & (DS_TS - 1)) == 0 && T < ntile) {
    tbox[2 * T] = make_float4(lx, ly, lz, v.w);
    tbox[2 * T + 1] = make_float4(hx, hy, hz, v.w);
  }
}

// The K nearest neighbours of each row in (d2, index) order, one wave per
// 16-row tile (4 lanes per row).  A row never inserts into a top-K list
// while scanning (the whole wave would run the insertion whenever any lane
// accepts a candidate).  Histogram passes find a threshold instead: level 0
// bins d2 logarithmically; the bin where the cumulative count reaches K is
// split into KNN_BINS bins at the next level while the candidates up to it
// would exceed KNN_CAP.  The accepted set {bin path <= crossing path} is
// closed downward in d2 and holds at least K candidates (or all of them),
// ties at the threshold included.  An append pass keeps it in LDS as
// {d2, j}; each kept candidate's rank in (d2, j) order among the row's kept
// ones is its position in the neighbour list (ranks >= K are dropped).  The
// row's state (bins, crossing path) is computed alike by its 4 lanes.  More
// than KNN_CAP survive only when more than about KNN_CAP - K candidates
// share one FP32 distance after three refinements; then the first KNN_CAP
// appended are kept.
template <int K>
__global__ __launch_bounds__(64) void k_knn(
    const float4* __restrict__ loc, const float4* __restrict__ toff,
    KnnArgs a) {
  __shared__ unsigned hist[KNN_BINS * KNN_RW];   // [bin][row]
  __shared__ uint2 cand[KNN_CAP * KNN_RW];       // [slot][row]
  __shared__ int ccnt[KNN_RW];
  __shared__ float4 trec[2 * DS_TS];
  const int lane = threadIdx.x;
  const int r = lane & (KNN_RW - 1), g = lane / KNN_RW;
  // the rows' tile (with a.fb: the ring search's unfinished tiles only)
  int Tt = blockIdx.x;
  if (a.fb) {
    if ((int)blockIdx.x >= a.fb[0]) return;
    Tt = a.fb[1 + blockIdx.x];
  }
  const int I = Tt / DS_TPB;              // and its block
  const int row = Tt * DS_TS + r;
  const float4 li = a.tgt[row];

  float lo[KNN_LEVELS], sc[KNN_LEVELS];
  int cb[KNN_LEVELS];
#pragma unroll
  for (int L = 0; L < KNN_LEVELS; ++L) {
    lo[L] = 0.0f;
    sc[L] = 0.0f;
    cb[L] = 0;
  }
  int nlev = 0;              // levels whose crossing bin is fixed
  bool settled = row >= a.n; // padding rows keep no neighbours
  int below = 0;             // candidates strictly below the current domain

  auto bin_at = [&](float d2, int L) {
    return L == 0 ? knn_bin0(d2) : knn_bin(d2, lo[L], sc[L]);
  };
  // -1: below the domain of level p (accepted), 0: inside it, 1: above.
  auto classify = [&](float d2, int p) {
    int c = 0;
#pragma unroll
    for (int L = 0; L < KNN_LEVELS; ++L) {
      if (L < p && c == 0) {
        const int b = bin_at(d2, L);
        c = b < cb[L] ? -1 : (b > cb[L] ? 1 : 0);
      }
    }
    return c;
  };
  auto hbin = [&](int b) -> unsigned& { return hist[b * KNN_RW + r]; };

  const float kInf = __builtin_inff();
  // the rows' box (block frame, scaled units); the 4 lane groups hold the
  // same 16 rows
  float4 bmin = li, bmax = li;
#pragma unroll
  for (int off = 1; off < KNN_RW; off <<= 1) {
    bmin.x = fminf(bmin.x, __shfl_xor(bmin.x, off, 64));
    bmin.y = fminf(bmin.y, __shfl_xor(bmin.y, off, 64));
    bmin.z = fminf(bmin.z, __shfl_xor(bmin.z, off, 64));
    bmax.x = fmaxf(bmax.x, __shfl_xor(bmax.x, off, 64));
    bmax.y = fmaxf(bmax.y, __shfl_xor(bmax.y, off, 64));
    bmax.z = fmaxf(bmax.z, __shfl_xor(bmax.z, off, 64));
  }

  for (int p = 0; p < KNN_LEVELS; ++p) {
    if (!__any(!settled)) break;
    for (int b = g; b < KNN_BINS; b += KNN_G) hbin(b) = 0u;
    __syncthreads();
    if (p == 0) {
      // Level 0 in two parts: first the entries that can hold d2 < R0^2;
      // a row that already finds K candidates there has its crossing bin
      // below kNearBins, where the counts are complete.  The other entries
      // are scanned only when some row did not.
      constexpr float R0 = 8.0f;        // d2 < 64: bins below kNearBins
      constexpr int kNearBins = 48;
      auto add = [&](float d2, int, bool ok) {
        atomicAdd(&hbin(knn_bin0(d2)), ok ? 1u : 0u);
      };
      knn_scan(a, loc, toff, trec, I, row, li, bmin, bmax, -kInf, R0, add);
      __syncthreads();
      int near = 0;
      for (int b = 0; b < kNearBins; ++b) near += (int)hbin(b);
      if (__any(!settled && near < K))
        knn_scan(a, loc, toff, trec, I, row, li, bmin, bmax, R0, kInf, add);
    } else {
      knn_scan(a, loc, toff, trec, I, row, li, bmin, bmax, -kInf, kInf,
               [&](float d2, int, bool ok) {
                 if (ok && !settled && classify(d2, p) == 0)
                   atomicAdd(&hbin(bin_at(d2, p)), 1u);
               });
    }
    __syncthreads();
    if (settled) continue;
    int cum = below, cross = -1, before = below;
    for (int b = 0; b < KNN_BINS; ++b) {
      const int h = (int)hbin(b);
      if (cross < 0 && cum + h >= K) {
        cross = b;
        before = cum;
      }
      cum += h;
    }
    if (cross < 0) {   // fewer than K candidates in all: take the domain
      nlev = p;
      settled = true;
      continue;
    }
    cb[p] = cross;
    nlev = p + 1;
    const int take = before + (int)hbin(cross);
    if (take <= KNN_CAP || p + 1 == KNN_LEVELS) {
      settled = true;
      continue;
    }
    below = before;
    float l0, l1;
    if (p == 0) {
      l0 = knn_bin0_lower(cross);
      l1 = knn_bin0_lower(cross + 1);
    } else {
      const float w = 1.0f / sc[p];
      l0 = lo[p] + (float)cross * w;
      l1 = l0 + w;
    }
    lo[p + 1] = l0;
    sc[p + 1] = (float)KNN_BINS / (l1 - l0);
  }

  // append: below the crossing path, or inside the final crossing bin (or
  // the whole final domain when it holds fewer than K)
  if (lane < KNN_RW) ccnt[lane] = 0;
  __syncthreads();
  auto append = [&](float d2, int j) {
    const int s = atomicAdd(&ccnt[r], 1);
    if (s < KNN_CAP)
      cand[s * KNN_RW + r] = make_uint2(__float_as_uint(d2), (unsigned)j);
  };
  if (__all(nlev <= 1)) {
    const int tb = nlev == 0 ? KNN_BINS - 1 : cb[0];
    // kept candidates have d2 below the upper edge of bin tb
    float rk = row >= a.n ? 0.0f
                          : (nlev == 0 ? kInf : sqrtf(knn_bin0_lower(tb + 1)));
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
      rk = fmaxf(rk, __shfl_xor(rk, off, 64));
    knn_scan(a, loc, toff, trec, I, row, li, bmin, bmax, -kInf, rk,
             [&](float d2, int j, bool ok) {
      if (ok && knn_bin0(d2) <= tb) append(d2, j);
    });
  } else {
    knn_scan(a, loc, toff, trec, I, row, li, bmin, bmax, -kInf, kInf,
             [&](float d2, int j, bool ok) {
      if (ok && classify(d2, nlev) <= 0) append(d2, j);
    });
  }
  __syncthreads();

  // rank of each kept candidate among the row's kept ones; the row's 4
  // lanes take candidates g, g + 4, ...
  const int cnt = min(ccnt[r], KNN_CAP);
  int* o = a.nbr + (size_t)row * K;
  for (int u = g; u < cnt; u += KNN_G) {
    const uint2 cu = cand[u * KNN_RW + r];
    const float du = __uint_as_float(cu.x);
    const int ju = (int)cu.y;
    int rk = 0;
    for (int t = 0; t < cnt; ++t) {
      const uint2 c = cand[t * KNN_RW + r];
      rk += knn_less(__uint_as_float(c.x), (int)c.y, du, ju) ? 1 : 0;
    }
    if (rk < K) o[rk] = ju;
  }
  for (int k = cnt + g; k < K; k += KNN_G) o[k] = -1;
}

// ---- ring search: the common case in one pass over the near entries ----
//
// The same K nearest neighbours in (d2, index) order as k_knn, found by
// collecting every candidate of a row inside a radius and selecting from
// LDS.  Ring m covers d2 in [R_{m-1}^2, R_m^2): R_0 = a.r0 (the radius that
// holds about 64 points at the problem's mean density), R_{m+1} = 1.26 R_m
// (twice the volume), the last ring ending at the cutoff.  A row takes part
// in ring m only while it has fewer than K candidates, so a row's
// candidates are all its points with d2 below the end of its last ring:
// every point nearer than its K-th (or all of its neighbours when it has
// fewer than K).  A ring visits the entries whose box gap is below its
// outer radius (knn_scan's filter), 4 entries per step: lane (s, g) loads
// slot s of entry 4 m + g (the next two steps' records in flight), the 64
// records go through LDS, and lane (r, g) computes row r against the 16
// slots of entry g.  Candidates are appended per row (positions from the
// lanes' accept masks); a tile with a row beyond KNN_NCAP candidates is
// left to k_knn (fb).
// Selection: a KNN_NBIN-bin histogram of the row's d2 finds the bin where
// the count reaches K, the candidates up to that bin are compacted to the
// front of the row's list (16 per round, so no slot is overwritten before
// it is read), and each kept candidate's rank in (d2, j) order among the
// kept ones is its position.
constexpr int KNN_NCAP = 128;
constexpr int KNN_NBIN = 32;
constexpr float KNN_RING = 1.26f;

// A candidate as one 64-bit key: d2's bits (d2 >= 0, so they order like
// the values) above the index, so (d2, j) order is unsigned key order.
__device__ __forceinline__ unsigned long long knn_key(float d2, int j) {
  return ((unsigned long long)__float_as_uint(d2) << 32) | (unsigned)j;
}

struct KnnRingSmem {
  unsigned long long cand[KNN_NCAP * KNN_RW];   // [slot][row]: knn_key
  int cross[KNN_RW];
  union {
    struct {
      float4 rec[2][4 * 17];       // a step's 4 entries x 16 records
                                   // (stride 17: the 4 entries' records
                                   // start in different LDS banks)
      float4 po[64];               // passing entries: frame offset
      int pT[64];                  // and tile
    } s;
    unsigned hist[KNN_NBIN * KNN_RW];   // [bin][row]
  } u;
};

// One ring: append the candidates of rows with `want` whose d2 is in
// [lo2, hi2), from the block's entries whose box gap is below `hi`.  A
// row's count lives in `cnt` (the same in its 4 lanes): each step's
// accepted slots form a 16-bit mask per lane, and the positions come from
// the masks' counts, summed over the row's 4 lanes (one per 16-lane row of
// the wave) with two permlane swaps, so no LDS atomic is waited for.  The
// next chunk's entry indices and offsets are loaded during this chunk's
// steps, and its tile boxes at this chunk's end.
template <bool FIRST>
__device__ __forceinline__ void knn_ring(
    const KnnArgs& a, const float4* __restrict__ loc,
    const float4* __restrict__ toff, KnnRingSmem& sm, int I, int row,
    const float4& li, const float4& bmin, const float4& bmax, float hi,
    float lo2, float hi2, bool want, int& cnt) {
  const int lane = threadIdx.x;
  const int r = lane & (KNN_RW - 1), g = lane / KNN_RW;
  const int kb = __builtin_amdgcn_readfirstlane(a.tptr[I]);
  const int ke = __builtin_amdgcn_readfirstlane(a.tptr[I + 1]);
  // chunk metadata, one entry per lane, one chunk ahead
  int Tn = 0;
  float4 on = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 bn0 = on, bn1 = on;
  if (kb + lane < ke) {
    Tn = a.tidx[kb + lane];
    on = toff[kb + lane];
    bn0 = a.tbox[2 * Tn];
    bn1 = a.tbox[2 * Tn + 1];
  }
  for (int base = kb; base < ke; base += 64) {
    const bool inb = base + lane < ke;
    const int Tl = Tn;
    float4 ol = on;
    ol.w = 0.f;
    const float4 tmin = bn0, tmax = bn1;
    const bool more = base + 64 + lane < ke;
    if (more) {
      Tn = a.tidx[base + 64 + lane];
      on = toff[base + 64 + lane];
    }
    bool pass = false;
    if (inb) {
      const float gx = fmaxf(0.f, fmaxf(tmin.x - (bmax.x + ol.x),
                                        (bmin.x + ol.x) - tmax.x));
      const float gy = fmaxf(0.f, fmaxf(tmin.y - (bmax.y + ol.y),
                                        (bmin.y + ol.y) - tmax.y));
      const float gz = fmaxf(0.f, fmaxf(tmin.z - (bmax.z + ol.z),
                                        (bmin.z + ol.z) - tmax.z));
      const float scale = fabsf(ol.x) + fabsf(ol.y) + fabsf(ol.z) +
                          (bmax.x - bmin.x) + (bmax.y - bmin.y) +
                          (bmax.z - bmin.z);
      const float lb = sqrtf(fmaf(gz, gz, fmaf(gy, gy, gx * gx))) -
                       (1e-3f + 1e-5f * scale);
      pass = lb < hi;
    }
    const unsigned long long m = __ballot(pass);
    const int np = __popcll(m);
    if (np > 0) {
      const int pr = __builtin_amdgcn_mbcnt_hi(
          (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      if (pass) {
        sm.u.s.pT[pr] = Tl;
        sm.u.s.po[pr] = ol;
      }
      __syncthreads();
      const int steps = (np + 3) >> 2;
      const int slot = lane & (DS_TS - 1);
      auto load = [&](int b) {
        const int e = min(4 * b + g, np - 1);
        return loc[sm.u.s.pT[e] * DS_TS + slot];
      };
      float4 n0 = load(0);
      float4 n1 = steps > 1 ? load(1) : n0;
      for (int b = 0; b < steps; ++b) {
        const float4 cur = n0;
        n0 = n1;
        if (b + 2 < steps) n1 = load(b + 2);
        sm.u.s.rec[b & 1][17 * g + slot] = cur;
        __syncthreads();
        const int e = 4 * b + g;
        const int ec = min(e, np - 1);
        const int T = sm.u.s.pT[ec];
        const float4 o = sm.u.s.po[ec];
        const int j0 = T * DS_TS;
        const bool on_e = want && e < np;
        const bool full = j0 + DS_TS <= a.n;   // no padding slot
        const float4* rb = sm.u.s.rec[b & 1] + 17 * g;
        // the entry's slots that can hold a candidate of this row: below
        // n, not the row itself
        unsigned valid = full ? 0xFFFFu : (a.n > j0 ? (1u << (a.n - j0)) - 1u
                                                    : 0u);
        const unsigned self = (unsigned)(row - j0);
        if (self < (unsigned)DS_TS) valid &= ~(1u << self);
        if (!on_e) valid = 0u;
        float d2v[DS_TS];
        unsigned acc = 0u;
#pragma unroll
        for (int t = 0; t < DS_TS; ++t) {
          const float d2 = knn_d2(li, o, rb[t]);
          d2v[t] = d2;
          const bool in = FIRST ? d2 < hi2 : (d2 >= lo2 && d2 < hi2);
          acc |= in ? 1u << t : 0u;
        }
        acc &= valid;
        // positions: this lane's accepted slots follow those of the row's
        // lanes in lower 16-lane rows (g), after the row's count so far
        const unsigned c = (unsigned)__popc(acc);
        const auto p16 = __builtin_amdgcn_permlane16_swap(c, c, false, false);
        const unsigned x1 = (g & 1) ? p16[0] : p16[1];
        const unsigned s2 = p16[0] + p16[1];
        const auto p32 =
            __builtin_amdgcn_permlane32_swap(s2, s2, false, false);
        const unsigned x2 = (g & 2) ? p32[0] : p32[1];
        int pos = cnt + (int)((g & 1) ? x1 : 0u) + (int)((g & 2) ? x2 : 0u);
        cnt += (int)(s2 + x2);
        if (acc) {
#pragma unroll
          for (int t = 0; t < DS_TS; ++t) {
            if (acc & (1u << t)) {
              if (pos < KNN_NCAP)
                sm.cand[pos * KNN_RW + r] =
                    knn_key(d2v[t], j0 + t);
              ++pos;
            }
          }
        }
      }
    }
    // the next chunk's tile boxes (its indices have arrived by now)
    if (more) {
      bn0 = a.tbox[2 * Tn];
      bn1 = a.tbox[2 * Tn + 1];
    }
    __syncthreads();
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_knn_ring(
    const float4* __restrict__ loc, const float4* __restrict__ toff,
    KnnArgs a) {
  __shared__ KnnRingSmem sm;
  const int lane = threadIdx.x;
  const int r = lane & (KNN_RW - 1), g = lane / KNN_RW;
  // Workgroups go to the 8 XCDs round-robin by index: XCD x = i mod 8
  // takes the contiguous tile range x * span .. (x + 1) * span - 1, so
  // spatially adjacent tiles (which read the same source records) share an
  // L2.  The grid is 8 * span workgroups.
  const int span = (a.ntile + 7) / 8;
  const int Tt = (int)(blockIdx.x % 8u) * span + (int)(blockIdx.x / 8u);
  if (Tt >= a.ntile) return;
  const int I = Tt / DS_TPB;
  const int row = Tt * DS_TS + r;
  const bool live = row < a.n;
  const float4 li = a.tgt[row];
  float4 bmin = li, bmax = li;
#pragma unroll
  for (int off = 1; off < KNN_RW; off <<= 1) {
    bmin.x = fminf(bmin.x, __shfl_xor(bmin.x, off, 64));
    bmin.y = fminf(bmin.y, __shfl_xor(bmin.y, off, 64));
    bmin.z = fminf(bmin.z, __shfl_xor(bmin.z, off, 64));
    bmax.x = fmaxf(bmax.x, __shfl_xor(bmax.x, off, 64));
    bmax.y = fmaxf(bmax.y, __shfl_xor(bmax.y, off, 64));
    bmax.z = fmaxf(bmax.z, __shfl_xor(bmax.z, off, 64));
  }
  int cnt = 0;   // the row's candidates so far (same in its 4 lanes)
  const float rmax2 = a.cutoff;
  float rlo = 0.0f, rhi = a.r0;
  float row_hi2 = 0.0f;   // end of the last ring this row took part in
  bool want = live;
  for (;;) {
    const bool last = rhi * rhi >= rmax2;
    const float hi2 = last ? rmax2 : rhi * rhi;
    if (rlo == 0.0f)
      knn_ring<true>(a, loc, toff, sm, I, row, li, bmin, bmax,
                     last ? __builtin_inff() : rhi, 0.0f, hi2, want, cnt);
    else
      knn_ring<false>(a, loc, toff, sm, I, row, li, bmin, bmax,
                      last ? __builtin_inff() : rhi, rlo * rlo, hi2, want,
                      cnt);
    if (want) row_hi2 = hi2;
    const int c = cnt;
    if (__any(c > KNN_NCAP)) {
      if (lane == 0) {
        const int k = atomicAdd(&a.fb[0], 1);
        a.fb[1 + k] = Tt;
      }
      return;
    }
    want = live && c < K;
    if (last || !__any(want)) break;
    rlo = rhi;
    rhi *= KNN_RING;
  }

  // selection: the bin where the row's count reaches K
  const float inv_w = row_hi2 > 0.0f ? (float)KNN_NBIN / row_hi2 : 0.0f;
  auto bin_of = [&](unsigned long long key) {
    return min((int)(__uint_as_float((unsigned)(key >> 32)) * inv_w),
               KNN_NBIN - 1);
  };
  for (int b = g; b < KNN_NBIN; b += KNN_G) sm.u.hist[b * KNN_RW + r] = 0u;
  __syncthreads();
  if (cnt > K)
    for (int u = g; u < cnt; u += KNN_G)
      atomicAdd(&sm.u.hist[bin_of(sm.cand[u * KNN_RW + r]) * KNN_RW + r],
                1u);
  __syncthreads();
  constexpr int BPL = KNN_NBIN / KNN_G;   // bins per lane
  int hs = 0;
#pragma unroll
  for (int b = 0; b < BPL; ++b) hs += (int)sm.u.hist[(BPL * g + b) * KNN_RW + r];
  const int h0 = __shfl(hs, r, 64), h1 = __shfl(hs, r + 16, 64);
  const int h2 = __shfl(hs, r + 32, 64);
  int cum = (g > 0 ? h0 : 0) + (g > 1 ? h1 : 0) + (g > 2 ? h2 : 0);
  if (cnt > K) {
#pragma unroll
    for (int b = 0; b < BPL; ++b) {
      const int h = (int)sm.u.hist[(BPL * g + b) * KNN_RW + r];
      if (cum < K && cum + h >= K) sm.cross[r] = BPL * g + b;
      cum += h;
    }
  } else if (g == 0) {
    sm.cross[r] = KNN_NBIN - 1;   // fewer than K + 1: keep them all
  }
  __syncthreads();
  const int cross = sm.cross[r];
  // compact the kept candidates (bin <= cross) to the front, 16 per round
  int rounds = (cnt + 15) >> 4;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1)
    rounds = max(rounds, __shfl_xor(rounds, off, 64));
  int kept = 0;
  for (int q = 0; q < rounds; ++q) {
    unsigned long long c4[4];
    bool kp[4];
    int kc = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int u = 16 * q + 4 * g + i;
      kp[i] = u < cnt;
      c4[i] = kp[i] ? sm.cand[u * KNN_RW + r] : 0ull;
      kp[i] = kp[i] && bin_of(c4[i]) <= cross;
      kc += kp[i] ? 1 : 0;
    }
    const int k0 = __shfl(kc, r, 64), k1 = __shfl(kc, r + 16, 64);
    const int k2 = __shfl(kc, r + 32, 64), k3 = __shfl(kc, r + 48, 64);
    int pos = kept + (g > 0 ? k0 : 0) + (g > 1 ? k1 : 0) + (g > 2 ? k2 : 0);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (kp[i]) sm.cand[(pos++) * KNN_RW + r] = c4[i];
    kept += k0 + k1 + k2 + k3;
    __syncthreads();
  }
  // rank of each kept candidate among the kept ones (the others read
  // eight at a time, so their LDS latencies overlap)
  int* o = a.nbr + (size_t)row * K;
  for (int u = g; u < kept; u += KNN_G) {
    const unsigned long long cu = sm.cand[u * KNN_RW + r];
    int rk = 0;
    int t0 = 0;
    for (; t0 + 8 <= kept; t0 += 8) {
      unsigned long long c[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) c[i] = sm.cand[(t0 + i) * KNN_RW + r];
#pragma unroll
      for (int i = 0; i < 8; ++i) rk += c[i] < cu ? 1 : 0;
    }
    for (; t0 < kept; ++t0) rk += sm.cand[t0 * KNN_RW + r] < cu ? 1 : 0;
    if (rk < K) o[rk] = (int)(unsigned)cu;
  }
  for (int k = kept + g; k < K; k += KNN_G) o[k] = -1;
}

// The WS_FEAT features of one row from its neighbour list and the current
// data, as BF16 into f[0, WS_FEAT).
template <int K>
__device__ __forceinline__ void ws_feature_row(const KnnArgs& a, int row,
                                               unsigned short* f) {
  if (row >= a.n) {
    for (int c = 0; c < WS_FEAT; ++c) f[c] = 0;
  } else {
    const double4 pi = a.pos[row];
    auto noise = [&](double wN) {
      return a.form == DS_FORM_W ? 1.0 / wN : sqrt(wN);
    };
    auto datum = [&](int j, double wN) {
      return a.form == DS_FORM_W ? a.rhs[j] / wN : a.rhs[j];
    };
    // the row's K neighbour indices in 16-byte loads, then their records
    // gathered eight neighbours at a time (absent neighbours read the row
    // itself and give zeros)
    static_assert(K % 8 == 0, "neighbour batches of 8");
    const int4* nb4 = reinterpret_cast<const int4*>(a.nbr + (size_t)row * K);
    int nb[K];
#pragma unroll
    for (int q = 0; q < K / 4; ++q) {
      const int4 t = nb4[q];
      nb[4 * q] = t.x;
      nb[4 * q + 1] = t.y;
      nb[4 * q + 2] = t.z;
      nb[4 * q + 3] = t.w;
    }
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 8) {
      double4 pj[8];
      double dj[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int j = nb[k0 + u] >= 0 ? nb[k0 + u] : row;
        pj[u] = a.pos[j];
        dj[u] = a.rhs[j];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + u;
        float v[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        if (nb[k] >= 0) {
          v[0] = (float)(pj[u].x - pi.x) * a.inv_lp;
          v[1] = (float)(pj[u].y - pi.y) * a.inv_lp;
          v[2] = (float)(pj[u].z - pi.z) * a.inv_ll;
          v[3] = (float)noise(pj[u].w) * a.inv_sqrt_vs;
          v[4] = (float)(a.form == DS_FORM_W ? dj[u] / pj[u].w : dj[u]) *
                 a.inv_sqrt_vs;
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) f[5 * k + c] = f2bf(v[c]);
      }
    }
    f[5 * K] = f2bf((float)noise(pi.w) * a.inv_sqrt_vs);
    f[5 * K + 1] = f2bf((float)datum(row, pi.w) * a.inv_sqrt_vs);
  }
}

// Features of each row, staged in LDS so that the block's 64 contiguous
// feature rows are written out in coalesced dwords.
template <int K>
__global__ __launch_bounds__(64) void k_ws_features(KnnArgs a) {
  __shared__ unsigned short stage[DS_BS * WS_FEAT];   // 20.25 KB
  const int lane = threadIdx.x;
  const int I = blockIdx.x;
  ws_feature_row<K>(a, I * DS_BS + lane, stage + lane * WS_FEAT);
  __syncthreads();
  const unsigned* src = reinterpret_cast<const unsigned*>(stage);
  unsigned* dst = reinterpret_cast<unsigned*>(a.feat + (size_t)I * DS_BS *
                                                           WS_FEAT);
  for (int k = lane; k < DS_BS * WS_FEAT / 2; k += 64) dst[k] = src[k];
}

// ---- network ------------------------------------------------------------

constexpr int H = WS_HIDDEN;

struct NetPtrs {
  const float* W1;  // [FEAT][H]
  const float* b1;
  const float* W2;  // [H][H]
  const float* b2;
  const float* W3;
  const float* b3;
  const float* W4;  // [H]
  const float* b4;
};

__device__ __forceinline__ NetPtrs net_ptrs(const float* P) {
  NetPtrs n;
  n.W1 = P + WS_OFF_W1;
  n.b1 = P + WS_OFF_B1;
  n.W2 = P + WS_OFF_W2;
  n.b2 = P + WS_OFF_B2;
  n.W3 = P + WS_OFF_W3;
  n.b3 = P + WS_OFF_B3;
  n.W4 = P + WS_OFF_W4;
  n.b4 = P + WS_OFF_B4;
  return n;
}

template <bool Q>
__device__ __forceinline__ float act(float z) {
  const float r = fmaxf(z, 0.0f);
  return Q ? bf2f(f2bf(r)) : r;
}

// h_out = act(b + h_in W) for a 64 -> 64 layer (weights [in][out]).
template <bool Q>
__device__ __forceinline__ void dense64(const float* __restrict__ W,
                                        const float* __restrict__ b,
                                        const float (&hin)[H], float (&hout)[H]) {
#pragma unroll
  for (int o = 0; o < H; ++o) hout[o] = b[o];
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const float x = hin[i];
#pragma unroll
    for (int o = 0; o < H; ++o) hout[o] = fmaf(x, W[i * H + o], hout[o]);
  }
#pragma unroll
  for (int o = 0; o < H; ++o) hout[o] = act<Q>(hout[o]);
}

// Forward pass of one pixel; fills h1..h3 and returns the output.
template <bool Q>
__device__ __forceinline__ float forward(const NetPtrs& n,
                                         const unsigned short* __restrict__ f,
                                         float (&h1)[H], float (&h2)[H],
                                         float (&h3)[H]) {
#pragma unroll
  for (int o = 0; o < H; ++o) h1[o] = n.b1[o];
  for (int c = 0; c < WS_FEAT; ++c) {
    const float x = bf2f(f[c]);
#pragma unroll
    for (int o = 0; o < H; ++o) h1[o] = fmaf(x, n.W1[c * H + o], h1[o]);
  }
#pragma unroll
  for (int o = 0; o < H; ++o) h1[o] = act<Q>(h1[o]);
  dense64<Q>(n.W2, n.b2, h1, h2);
  dense64<Q>(n.W3, n.b3, h2, h3);
  float y = n.b4[0];
#pragma unroll
  for (int i = 0; i < H; ++i) y = fmaf(h3[i], n.W4[i], y);
  return y;
}

struct FwdArgs {
  const unsigned short* feat;
  const float* P;
  const double4* pos;     // prediction -> W-space initial guess
  double* x_out;          // y0 = x0 / w (W) or x0 (SN); may be null
  float* t_out;           // raw outputs; may be null
  int n, npad, form, quant;
  double inv_sqrt_vs;     // x0 = t / sqrt(var_s)
};

// One wave per 64 pixels: the block's feature rows (contiguous) are copied
// to LDS in coalesced dwords, then each lane runs its pixel through the
// network with the weights fetched by scalar loads (P is read-only).
template <bool Q>
__global__ __launch_bounds__(64) void k_fwd(const float* __restrict__ P,
                                            FwdArgs a) {
  __shared__ unsigned short stage[64 * WS_FEAT];
  const int lane = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  const int rows = min(64, a.npad - r0);
  const unsigned* src =
      reinterpret_cast<const unsigned*>(a.feat + (size_t)r0 * WS_FEAT);
  unsigned* dst = reinterpret_cast<unsigned*>(stage);
  for (int k = lane; k < rows * (WS_FEAT / 2); k += 64) dst[k] = src[k];
  __syncthreads();
  const int i = r0 + lane;
  if (i >= a.npad) return;
  if (i >= a.n) {
    if (a.x_out) a.x_out[i] = 0.0;
    if (a.t_out) a.t_out[i] = 0.0f;
    return;
  }
  const NetPtrs n = net_ptrs(P);
  float h1[H], h2[H], h3[H];
  const float t = forward<Q>(n, stage + lane * WS_FEAT, h1, h2, h3);
  if (a.t_out) a.t_out[i] = t;
  if (a.x_out) {
    const double x0 = (double)t * a.inv_sqrt_vs;
    const double w = a.pos[i].w;
    a.x_out[i] = a.form == DS_FORM_W ? x0 / w : x0;
  }
}

// ---- inference on the matrix cores ----------------------------------------
//
// Each FP32 weight is split into three BF16 parts, w = hi + mid + lo (8
// mantissa bits each, so the sum carries the FP32 value); activations are
// BF16 already (features, and every hidden layer with quant), so the
// products x * part are exact in FP32 and v_mfma_f32_32x32x16_bf16 (gfx950's
// BF16 MFMA) sums them with FP32 accumulation: the same arithmetic class as
// the VALU forward, in another summation order.  One wave runs 64 pixels:
// A = the activation rows (LDS), B = 8 consecutive inputs of one unit of one
// part (global, L2-resident), D = 32 pixels x 32 units per MFMA; the epilogue
// adds the bias, applies ReLU and rounds to BF16 into LDS for the next
// layer.  The 64 -> 1 output layer is a VALU dot product.

typedef __bf16 v8bf_ws __attribute__((ext_vector_type(8)));
typedef float v16f_ws __attribute__((ext_vector_type(16)));

// Wsplit layout per layer L (inputs K_L padded to Kp_L): [part][unit][Kp_L].
__global__ __launch_bounds__(256) void k_split_weights(const float* P,
                                                       unsigned short* ws) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n1 = H * WS_K1P, n2 = H * H;
  if (t >= n1 + 2 * n2) return;
  int layer, rem;
  if (t < n1) { layer = 0; rem = t; }
  else if (t < n1 + n2) { layer = 1; rem = t - n1; }
  else { layer = 2; rem = t - n1 - n2; }
  const int kp = layer == 0 ? WS_K1P : H;
  const int kin = layer == 0 ? WS_FEAT : H;
  const int j = rem / kp, k = rem % kp;
  const float* W = P + (layer == 0 ? WS_OFF_W1 : layer == 1 ? WS_OFF_W2
                                                            : WS_OFF_W3);
  const float w = k < kin ? W[k * H + j] : 0.0f;
  const unsigned short hi = f2bf(w);
  const float r1 = w - bf2f(hi);
  const unsigned short mid = f2bf(r1);
  const float r2 = r1 - bf2f(mid);
  const unsigned short lo = f2bf(r2);
  const size_t base = layer == 0 ? 0 : (size_t)3 * n1 + (size_t)3 * n2 *
                                                           (layer - 1);
  const size_t plane = (size_t)H * kp;
  ws[base + 0 * plane + rem] = hi;
  ws[base + 1 * plane + rem] = mid;
  ws[base + 2 * plane + rem] = lo;
}

// 8 BF16 values (16 bytes, 16-byte aligned).
__device__ __forceinline__ v8bf_ws ld8(const unsigned short* p) {
  return __builtin_bit_cast(v8bf_ws, *reinterpret_cast<const uint4*>(p));
}

// acc[mh][nh] += act[32 mh + r][k] * W[k][32 nh + r] over NKS K-steps of 16
// (lane half h takes k-slots 8 h .. 8 h + 7 of each step, in A and B).
template <int NKS>
__device__ __forceinline__ void mfma_layer(const unsigned short* act,
                                           int pitch,
                                           const unsigned short* __restrict__ w,
                                           int kp, v16f_ws (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
  const size_t plane = (size_t)H * kp;
  // The weights (the same for every wave, cache-resident) arrive in
  // batches of up to 8 k-steps, every load of a batch issued before its
  // first MFMA; the MFMAs run in k-step order as before.
  constexpr int CH = NKS <= 4 ? NKS : (NKS + 2) / 3;
#pragma unroll
  for (int c0 = 0; c0 < NKS; c0 += CH) {
    v8bf_ws b[CH][3][2];
#pragma unroll
    for (int u = 0; u < CH; ++u)
      if (c0 + u < NKS)
#pragma unroll
        for (int part = 0; part < 3; ++part)
#pragma unroll
          for (int nh = 0; nh < 2; ++nh)
            b[u][part][nh] = ld8(w + part * plane + (size_t)(32 * nh + r) * kp +
                                 16 * (c0 + u) + 8 * h);
    // keep the batch together (the scheduler would otherwise move each
    // load next to its MFMA and expose its latency)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (c0 + u >= NKS) break;
      const int ks = c0 + u;
      const v8bf_ws a0 = ld8(act + r * pitch + 16 * ks + 8 * h);
      const v8bf_ws a1 = ld8(act + (32 + r) * pitch + 16 * ks + 8 * h);
#pragma unroll
      for (int part = 0; part < 3; ++part)
#pragma unroll
        for (int nh = 0; nh < 2; ++nh) {
          acc[0][nh] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              a0, b[u][part][nh], acc[0][nh], 0, 0, 0);
          acc[1][nh] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
              a1, b[u][part][nh], acc[1][nh], 0, 0, 0);
        }
    }
  }
}

// bias + ReLU + BF16 into act[pixel][unit] (pitch), D layout of the MFMA:
// register j of lane l holds row 8 (j / 4) + 4 (l / 32) + j % 4, column
// l % 32 (plus 32 mh / 32 nh).
__device__ __forceinline__ void mfma_epilogue(const v16f_ws (&acc)[2][2],
                                              const float* __restrict__ b,
                                              unsigned short* act, int pitch) {
  const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
#pragma unroll
  for (int mh = 0; mh < 2; ++mh)
#pragma unroll
    for (int nh = 0; nh < 2; ++nh) {
      const float bias = b[32 * nh + r];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int px = 32 * mh + 8 * (j >> 2) + 4 * h + (j & 3);
        act[px * pitch + 32 * nh + r] =
            f2bf(fmaxf(acc[mh][nh][j] + bias, 0.0f));
      }
    }
}

// LDS pitches (BF16) of the feature rows and the hidden activations: whole
// 16-byte units (the MFMA operands are 16-byte loads), at least the padded
// input widths, rows spread over the LDS banks
constexpr int kFP = 184;
constexpr int kAP = 72;
static_assert(kFP >= WS_K1P && kFP % 8 == 0 && kAP >= WS_HIDDEN &&
              kAP % 8 == 0, "MFMA operand rows");

// The network on a block's 64 feature rows staged in sf (pitch kFP, pad
// columns zero): layers on the matrix cores, activations through LDS, the
// output layer on the VALU, the prediction stored for rows r0 + lane.
__device__ __forceinline__ void fwd_mfma_rows(
    const float* __restrict__ P, const unsigned short* __restrict__ wsplit,
    const FwdArgs& a, unsigned short* sf, unsigned short* sa, int r0) {
  const int lane = threadIdx.x;
  const size_t n1 = (size_t)H * WS_K1P, n2 = (size_t)H * H;
  v16f_ws acc[2][2];
  auto zero = [&]() {
#pragma unroll
    for (int mh = 0; mh < 2; ++mh)
#pragma unroll
      for (int nh = 0; nh < 2; ++nh) acc[mh][nh] = (v16f_ws){};
  };
  zero();
  mfma_layer<WS_K1P / 16>(sf, kFP, wsplit, WS_K1P, acc);
  mfma_epilogue(acc, P + WS_OFF_B1, sa, kAP);
  __syncthreads();
  zero();
  mfma_layer<H / 16>(sa, kAP, wsplit + 3 * n1, H, acc);
  mfma_epilogue(acc, P + WS_OFF_B2, sf, kAP);
  __syncthreads();
  zero();
  mfma_layer<H / 16>(sf, kAP, wsplit + 3 * n1 + 3 * n2, H, acc);
  mfma_epilogue(acc, P + WS_OFF_B3, sa, kAP);
  __syncthreads();
  const int i = r0 + lane;
  if (i >= a.npad) return;
  if (i >= a.n) {
    if (a.x_out) a.x_out[i] = 0.0;
    if (a.t_out) a.t_out[i] = 0.0f;
    return;
  }
  const float* W4 = P + WS_OFF_W4;
  float t = P[WS_OFF_B4];
#pragma unroll 8
  for (int u = 0; u < H; ++u) t = fmaf(bf2f(sa[lane * kAP + u]), W4[u], t);
  if (a.t_out) a.t_out[i] = t;
  if (a.x_out) {
    const double x0 = (double)t * a.inv_sqrt_vs;
    const double w = a.pos[i].w;
    a.x_out[i] = a.form == DS_FORM_W ? x0 / w : x0;
  }
}

__global__ __launch_bounds__(64) void k_fwd_mfma(
    const float* __restrict__ P, const unsigned short* __restrict__ wsplit,
    FwdArgs a) {
  __shared__ __align__(16) unsigned short sf[64 * kFP];   // features / h2
  __shared__ __align__(16) unsigned short sa[64 * kAP];   // h1 / h3
  const int lane = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  // The block's 64 feature rows (162 BF16 = 81 dwords each; the feature
  // buffers hold whole 64-row blocks, rows past npad only feed outputs
  // that are not stored) are one contiguous, 16-byte aligned run: every
  // 16-byte load is issued before the first is used, then the dwords go
  // to pitch-kFP rows whose pad columns are zero.
  constexpr int kRowD = WS_FEAT / 2, kPitchD = kFP / 2;
  constexpr int kQ = 64 * kRowD / 4, kIt = (kQ + 63) / 64;
  static_assert(64 * kRowD % 4 == 0, "feature block is whole uint4s");
  const uint4* src4 =
      reinterpret_cast<const uint4*>(a.feat + (size_t)r0 * WS_FEAT);
  unsigned* dst = reinterpret_cast<unsigned*>(sf);
  uint4 q[kIt];
#pragma unroll
  for (int t = 0; t < kIt; ++t) {
    const int k = lane + 64 * t;
    q[t] = src4[min(k, kQ - 1)];
  }
#pragma unroll
  for (int t = 0; t < kIt; ++t) {
    const int k = lane + 64 * t;
    if (k < kQ) {
      const unsigned v[4] = {q[t].x, q[t].y, q[t].z, q[t].w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = 4 * k + e, row = idx / kRowD;
        dst[row * kPitchD + idx - row * kRowD] = v[e];
      }
    }
  }
  for (int k = lane; k < 64 * (kPitchD - kRowD); k += 64) {
    const int row = k / (kPitchD - kRowD);
    dst[row * kPitchD + kRowD + k - row * (kPitchD - kRowD)] = 0u;
  }
  __syncthreads();
  fwd_mfma_rows(P, wsplit, a, sf, sa, r0);
}

// Prediction in one pass: the block's features go straight into the
// network's LDS rows (the same values as k_ws_features writes), then
// fwd_mfma_rows.  No feature array in HBM.
template <int K>
__global__ __launch_bounds__(64) void k_predict_mfma(
    const float* __restrict__ P, const unsigned short* __restrict__ wsplit,
    KnnArgs ka, FwdArgs a) {
  __shared__ __align__(16) unsigned short sf[64 * kFP];   // features / h2
  __shared__ __align__(16) unsigned short sa[64 * kAP];   // h1 / h3
  const int lane = threadIdx.x;
  const int r0 = blockIdx.x * 64;
  unsigned short* f = sf + lane * kFP;
  ws_feature_row<K>(ka, r0 + lane, f);
  for (int c = WS_FEAT; c < kFP; ++c) f[c] = 0;
  __syncthreads();
  fwd_mfma_rows(P, wsplit, a, sf, sa, r0);
}

// Training: forward + back-propagation of the squared error of sample
// idx[b] (b < B) into per-sample activations (BF16) and deltas (FP32).
struct TrainArgs {
  const unsigned short* feat;   // store of all samples
  const float* target;
  const int* idx;               // minibatch -> sample
  const float* P;
  unsigned short* act;          // B x 3H: h1, h2, h3
  float* delta;                 // B x 3H: d1, d2, d3
  float* dy;                    // B
  float* loss_part;             // gridDim.x
  int B;
  float scale;                  // d loss / d y = scale * (y - t)
};

template <bool Q>
__global__ __launch_bounds__(256) void k_fwd_bwd(TrainArgs a) {
  __shared__ double red[4];
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (b < a.B) {
    const int s = a.idx[b];
    const NetPtrs n = net_ptrs(a.P);
    float h1[H], h2[H], h3[H];
    const float y = forward<Q>(n, a.feat + (size_t)s * WS_FEAT, h1, h2, h3);
    const float e = y - a.target[s];
    v[0] = (double)e * e;
    const float g = a.scale * e;
    a.dy[b] = g;
    unsigned short* A = a.act + (size_t)b * 3 * H;
    float* D = a.delta + (size_t)b * 3 * H;
    float d3[H], d2[H];
#pragma unroll
    for (int o = 0; o < H; ++o) {
      d3[o] = h3[o] > 0.0f ? g * n.W4[o] : 0.0f;
      D[2 * H + o] = d3[o];
      A[2 * H + o] = f2bf(h3[o]);
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
      float s3 = 0.0f;
#pragma unroll
      for (int o = 0; o < H; ++o) s3 = fmaf(n.W3[i * H + o], d3[o], s3);
      d2[i] = h2[i] > 0.0f ? s3 : 0.0f;
    }
#pragma unroll
    for (int o = 0; o < H; ++o) {
      D[H + o] = d2[o];
      A[H + o] = f2bf(h2[o]);
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
      float s2 = 0.0f;
#pragma unroll
      for (int o = 0; o < H; ++o) s2 = fmaf(n.W2[i * H + o], d2[o], s2);
      D[i] = h1[i] > 0.0f ? s2 : 0.0f;
      A[i] = f2bf(h1[i]);
    }
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) a.loss_part[blockIdx.x] = (float)v[0];
}

// Weight gradients G[r][o] = sum_b In[b][r] * D[b][o] over a chunk of the
// minibatch, one 16 x 64 tile of G per block (r = input row, o = output).
// layer 0: In = features (BF16, FEAT cols), D = delta1
// layer 1: In = h1 (act cols 0..63),         D = delta2
// layer 2: In = h2,                          D = delta3
// layer 3: In = h3, D = dy (one column)
// A final row r == nin holds the bias gradient (In = 1).
struct WgradArgs {
  const unsigned short* feat;
  const int* idx;
  const unsigned short* act;
  const float* delta;
  const float* dy;
  float* part;          // [nchunk][WS_NPARAM]
  int B, chunk;
};

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a, int layer) {
  const int nin = layer == 0 ? WS_FEAT : H;
  const int nout = layer == 3 ? 1 : H;
  const int rtiles = (nin + 1 + 15) / 16;
  const int rt = blockIdx.x % rtiles;
  const int ch = blockIdx.x / rtiles;
  const int r = rt * 16 + (threadIdx.x >> 4);
  const int o4 = (threadIdx.x & 15) * 4;
  const int b0 = ch * a.chunk, b1 = min(a.B, b0 + a.chunk);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  if (r <= nin && o4 < nout) {
    for (int b = b0; b < b1; ++b) {
      float x;
      if (r == nin) x = 1.0f;
      else if (layer == 0) x = bf2f(a.feat[(size_t)a.idx[b] * WS_FEAT + r]);
      else x = bf2f(a.act[(size_t)b * 3 * H + (layer - 1) * H + r]);
      if (layer == 3) {
        acc[0] = fmaf(x, a.dy[b], acc[0]);
      } else {
        const float4 d = *reinterpret_cast<const float4*>(
            a.delta + (size_t)b * 3 * H + layer * H + o4);
        acc[0] = fmaf(x, d.x, acc[0]);
        acc[1] = fmaf(x, d.y, acc[1]);
        acc[2] = fmaf(x, d.z, acc[2]);
        acc[3] = fmaf(x, d.w, acc[3]);
      }
    }
  }
  if (r > nin || o4 >= nout) return;
  float* G = a.part + (size_t)ch * WS_NPARAM;
  const int offW = layer == 0 ? WS_OFF_W1 : layer == 1 ? WS_OFF_W2
                   : layer == 2 ? WS_OFF_W3 : WS_OFF_W4;
  const int offB = layer == 0 ? WS_OFF_B1 : layer == 1 ? WS_OFF_B2
                   : layer == 2 ? WS_OFF_B3 : WS_OFF_B4;
  float* dst = r == nin ? G + offB : G + offW + (size_t)r * nout;
  const int m = nout == 1 ? 1 : 4;
  for (int k = 0; k < m; ++k) dst[o4 + k] = acc[k];
}

__global__ __launch_bounds__(256) void k_reduce_chunks(const float* part,
                                                       int nchunk,
                                                       float* grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= WS_NPARAM) return;
  float s = 0.0f;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * WS_NPARAM + p];
  grad[p] = s;
}

__global__ __launch_bounds__(256) void k_adam(float* P, float* m, float* v,
                                              const float* g, float lr,
                                              float b1, float b2, float eps,
                                              float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= WS_NPARAM) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1.0f - b1) * gi;
  const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  P[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// training targets: t = x sqrt(var_s) from the solved x (sorted order)
__global__ __launch_bounds__(256) void k_targets(const double* x, int n,
                                                 double sqrt_vs, float* t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = (float)(x[i] * sqrt_vs);
}

}  // namespace

// ---- host side ----------------------------------------------------------

#define WS_TRY(expr)                  \
  do {                                \
    hipError_t e_ = (expr);           \
    if (e_ != hipSuccess) return e_;  \
  } while (0)

namespace {

// The feature kernels' arguments; builds the problem's neighbour lists on
// first use.
hipError_t knn_args(ds_problem* p, unsigned short* feat, hipStream_t st,
                    KnnArgs& a) {
  a = KnnArgs{};
  a.pos = p->pos64;
  a.rhs = p->rhs64;
  a.tgt = p->tgt32;
  a.tptr = p->tptr;
  a.tidx = p->tidx;
  a.feat = feat;
  a.n = p->n;
  a.form = p->form;
  a.inv_lp = (float)std::sqrt(2.0 * p->cv.c_perp);   // 1 / l_perp
  a.inv_ll = (float)std::sqrt(2.0 * p->cv.c_para);
  a.inv_sqrt_vs = (float)(1.0 / std::sqrt(p->cv.var_s));
  a.cutoff = (float)(p->cv.table_n - 1);
  if (!p->knn_nbr) {
    // The neighbour lists depend on the geometry only: built once per
    // problem, like its tile lists.  Candidates are kept as list entry * 16
    // + tile slot in 32 bits.
    if (p->ntile_pairs >= (1LL << 28)) return hipErrorInvalidValue;
    // one allocation: the lists, the tiles' boxes, then the unfinished-tile
    // list of the ring search (count + tiles)
    int* nbr = nullptr;
    const size_t nb = sizeof(int) * (size_t)p->npad * WS_K;
    const size_t nbox = 2 * sizeof(float4) * p->ntile;
    const size_t bytes = nb + nbox + sizeof(int) * (p->ntile + 1);
    hipError_t e = ds::pool_alloc(p, (void**)&nbr, bytes);
    if (e != hipSuccess) return e;
    a.nbr = nbr;
    float4* tbox = reinterpret_cast<float4*>(reinterpret_cast<char*>(nbr) + nb);
    a.tbox = tbox;
    a.fb = reinterpret_cast<int*>(reinterpret_cast<char*>(nbr) + nb + nbox);
    a.r0 = p->knn_r0;
    hipLaunchKernelGGL(k_tile_boxes,
                       dim3((unsigned)(((size_t)p->ntile * DS_TS + 255) / 256)),
                       dim3(256), 0, st, p->loc32, p->ntile, tbox, a.fb);
    // the ring search, then the general search on the tiles it left
    a.ntile = p->ntile;
    hipLaunchKernelGGL(k_knn_ring<WS_K>, dim3(8 * ((p->ntile + 7) / 8)),
                       dim3(64), 0, st, p->loc32, p->toff32, a);
    hipLaunchKernelGGL(k_knn<WS_K>, dim3(p->ntile), dim3(64), 0, st,
                       p->loc32, p->toff32, a);
    a.fb = nullptr;
    e = hipGetLastError();
    if (e != hipSuccess) {
      ds::pfree(p, nbr);
      return e;
    }
    p->knn_nbr = nbr;
  }
  a.nbr = p->knn_nbr;
  return hipSuccess;
}

FwdArgs fwd_args(const float* P, const unsigned short* feat, int n, int npad,
                 int quant, const ds_problem* p, double* x_out,
                 float* t_out) {
  FwdArgs a{};
  a.feat = feat;
  a.P = P;
  a.pos = p ? p->pos64 : nullptr;
  a.x_out = x_out;
  a.t_out = t_out;
  a.n = n;
  a.npad = npad;
  a.form = p ? p->form : DS_FORM_SN;
  a.quant = quant;
  a.inv_sqrt_vs = p ? 1.0 / std::sqrt(p->cv.var_s) : 1.0;
  return a;
}

hipError_t split_weights(const float* P, unsigned short* wsplit,
                         hipStream_t st) {
  const int ns = H * (WS_K1P + 2 * H);
  hipLaunchKernelGGL(k_split_weights, dim3((ns + 255) / 256), dim3(256), 0,
                     st, P, wsplit);
  return hipGetLastError();
}

}  // namespace

hipError_t ws_features(ds_problem* p, unsigned short* feat,
                       hipStream_t st) {
  KnnArgs a;
  const hipError_t e = knn_args(p, feat, st, a);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ws_features<WS_K>, dim3(p->nblk), dim3(64), 0, st, a);
  return hipGetLastError();
}

hipError_t ws_predict(const float* P, unsigned short* wsplit, int quant,
                      ds_problem* p, unsigned short* feat, hipStream_t st) {
  if (!(quant && wsplit)) {
    // VALU inference reads the feature array
    const hipError_t e = ws_features(p, feat, st);
    if (e != hipSuccess) return e;
    return ws_forward(P, wsplit, feat, p->n, p->npad, quant, p, p->x,
                      nullptr, st);
  }
  KnnArgs ka;
  hipError_t e = knn_args(p, nullptr, st, ka);
  if (e == hipSuccess) e = split_weights(P, wsplit, st);
  if (e != hipSuccess) return e;
  const FwdArgs a = fwd_args(P, nullptr, p->n, p->npad, quant, p, p->x,
                             nullptr);
  hipLaunchKernelGGL(k_predict_mfma<WS_K>, dim3(p->nblk), dim3(64), 0, st, P,
                     wsplit, ka, a);
  return hipGetLastError();
}

hipError_t ws_forward(const float* P, unsigned short* wsplit,
                      const unsigned short* feat, int n, int npad, int quant,
                      const ds_problem* p, double* x_out, float* t_out,
                      hipStream_t st) {
  const FwdArgs a = fwd_args(P, feat, n, npad, quant, p, x_out, t_out);
  const dim3 g((npad + 63) / 64), b(64);
  if (quant && wsplit) {
    const hipError_t e = split_weights(P, wsplit, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fwd_mfma, g, b, 0, st, P, wsplit, a);
  } else if (quant) {
    hipLaunchKernelGGL(k_fwd<true>, g, b, 0, st, P, a);
  } else {
    hipLaunchKernelGGL(k_fwd<false>, g, b, 0, st, P, a);
  }
  return hipGetLastError();
}

hipError_t ws_targets(const ds_problem* p, float* t, hipStream_t st) {
  hipLaunchKernelGGL(k_targets, dim3((p->n + 255) / 256), dim3(256), 0, st,
                     p->x, p->n, std::sqrt(p->cv.var_s), t);
  return hipGetLastError();
}

// One minibatch: forward/backward, gradient into `grad` (and, with lr > 0,
// an Adam step).  loss_part receives ceil(B/256) partial sums.
hipError_t ws_batch(WsTrainer& T, const unsigned short* feat,
                    const float* target, const int* idx, int B, int quant,
                    float* loss_part, hipStream_t st) {
  TrainArgs a{};
  a.feat = feat;
  a.target = target;
  a.idx = idx;
  a.P = T.P;
  a.act = T.act;
  a.delta = T.delta;
  a.dy = T.dy;
  a.loss_part = loss_part;
  a.B = B;
  a.scale = 2.0f / (float)B;
  const dim3 g((B + 255) / 256), blk(256);
  if (quant) hipLaunchKernelGGL(k_fwd_bwd<true>, g, blk, 0, st, a);
  else hipLaunchKernelGGL(k_fwd_bwd<false>, g, blk, 0, st, a);
  WS_TRY(hipGetLastError());
  WgradArgs w{};
  w.feat = feat;
  w.idx = idx;
  w.act = T.act;
  w.delta = T.delta;
  w.dy = T.dy;
  w.part = T.part;
  w.B = B;
  w.chunk = WS_CHUNK;
  const int nchunk = (B + WS_CHUNK - 1) / WS_CHUNK;
  for (int layer = 0; layer < 4; ++layer) {
    const int nin = layer == 0 ? WS_FEAT : WS_HIDDEN;
    const int rtiles = (nin + 1 + 15) / 16;
    hipLaunchKernelGGL(k_wgrad, dim3(rtiles * nchunk), dim3(256), 0, st, w,
                       layer);
    WS_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_reduce_chunks, dim3((WS_NPARAM + 255) / 256),
                     dim3(256), 0, st, T.part, nchunk, T.grad);
  return hipGetLastError();
}

hipError_t ws_adam(WsTrainer& T, float lr, hipStream_t st) {
  T.step += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)T.step);
  const float c2 = 1.0f - std::pow(b2, (float)T.step);
  hipLaunchKernelGGL(k_adam, dim3((WS_NPARAM + 255) / 256), dim3(256), 0, st,
                     T.P, T.m, T.v, T.grad, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace ds

namespace ds {
namespace {
__global__ __launch_bounds__(256) void k_gather_rows(
    const unsigned short* src, const int* perm, int n, unsigned short* dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)n * WS_FEAT;
  if (t >= total) return;
  const size_t r = t / WS_FEAT, c = t % WS_FEAT;
  dst[t] = src[(size_t)perm[r] * WS_FEAT + c];
}
__global__ __launch_bounds__(256) void k_gather_f(const float* src,
                                                  const int* perm, int n,
                                                  float* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
}  // namespace

hipError_t ws_gather(const unsigned short* fsrc, const float* tsrc,
                     const int* perm, int n, unsigned short* fdst,
                     float* tdst, hipStream_t st) {
  const size_t total = (size_t)n * WS_FEAT;
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((total + 255) / 256)),
                     dim3(256), 0, st, fsrc, perm, n, fdst);
  hipLaunchKernelGGL(k_gather_f, dim3((n + 255) / 256), dim3(256), 0, st,
                     tsrc, perm, n, tdst);
  return hipGetLastError();
}
}  // namespace ds
