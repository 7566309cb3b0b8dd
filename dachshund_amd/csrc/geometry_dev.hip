// geometry_dev.hip -- a problem's ordering, culled tile lists and per-slot
// arrays, built on the device.
//
// The reference evaluates all n^2 pairs (lib/wf_vector.cc:54-73); the pair
// kernels here visit only (64-row block, 16-slot tile) pairs whose boxes
// pass the exact cutoff test.  This file builds everything those kernels
// read from the caller's pixel records, with the host threads doing only
// the O(n) checks, the tree schedule (a function of the pixel counts) and
// the launch orders:
//
//   1. k-d ordering.  A node splits along its largest weighted quantized
//      extent; its left child takes the `left` points of smallest key
//      (quantized coordinate, point index) and both children keep the
//      node's order.  Nodes above 1024 points are split level by level by
//      all workgroups (radix select of the split key over 8-bit digits,
//      each launch taking the previous digit from its counts, a per-chunk
//      count, a scan and a stable scatter); each node of at most 1024
//      points is then finished, down to its 16-slot tiles, by one
//      wavefront in LDS.  A select stops at the first digit whose bucket
//      starts at the split's rank: that prefix, lower digits zero, puts
//      the same keys below it.  The splits are those of build_blocking.
//   2. One wavefront per block: per-slot records (positions, right-hand
//      side, diagonal, weights), the FP32 frames (per-axis medians of the
//      block and of each tile), tile and block boxes, the far-point test.
//   3. Culled lists: per target block, groups of 64 source tiles are tested
//      first (one box per group), then the tiles of the groups that pass,
//      with the 16-row sub-block masks: counts, a scan, then the entries
//      (ascending tile order, the host builder's lists).
//   4. The symmetric kernel's half lists and their column-sum slots: the
//      blocks whose records feed tile T are read from T's own block's list
//      (the sub-block test is symmetric), so no sort is needed.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ds_internal.h"

namespace ds {

namespace {

constexpr int kChunk = 2048;      // big-node chunk: 256 threads x 8 points
constexpr int kEpt = kChunk / 256;
constexpr int kWaveMax = 1024;    // nodes up to this size: one wavefront
constexpr int kStack = 64;

struct BigNode {
  int b, e, left, sub, c0, nc;    // slots [b, e), chunks [c0, c0 + nc)
};
struct Chunk {
  int node, b, e;
};

__device__ inline unsigned qax(const uint4& r, int a) {
  return a == 0 ? r.y : (a == 1 ? r.z : r.w);
}

__device__ inline unsigned wmin_u(unsigned v) {
  for (int o = 32; o >= 1; o >>= 1) v = min(v, (unsigned)__shfl_xor((int)v, o));
  return v;
}
__device__ inline unsigned wmax_u(unsigned v) {
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o));
  return v;
}
__device__ inline int wsum_i(int v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline int wscan_incl(int v) {
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}
__device__ inline int mbcnt64(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}
// min / max within aligned groups of `width` lanes
template <int W>
__device__ inline double gmin_d(double v) {
  for (int o = W / 2; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
template <int W>
__device__ inline double gmax_d(double v) {
  for (int o = W / 2; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// The culling test of build_tile_lists: box_gap_xx(a, b) < limit with the
// same operations (this file is compiled without contraction).
__device__ inline bool box_pass(const double alo[3], const double ahi[3],
                                const double blo[3], const double bhi[3],
                                double c_perp, double c_para, double dx_inv,
                                double limit) {
  double g[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (blo[a] > ahi[a]) g[a] = blo[a] - ahi[a];
    else if (alo[a] > bhi[a]) g[a] = alo[a] - bhi[a];
    else g[a] = 0.0;
  }
  const double x_perp_2 = (g[0] * g[0] + g[1] * g[1]);
  const double x_para_2 = g[2] * g[2];
  const double ga_perp = c_perp * x_perp_2;
  const double ga_para = c_para * x_para_2;
  return (ga_perp + ga_para) * dx_inv < limit;
}

struct KdArgs {
  const double* pix;        // n x 5, caller order
  uint4* rec;               // npad records {index, q0, q1, q2}
  uint4* tmp;
  const int* blk_sub;
  const int* sub_blk0;
  const int* sub_n;
  const int* sub_off;
  const double* quant;      // per subvolume: lo[3], inv[3], w[3]
  int npad;
};

__global__ __launch_bounds__(256) void k_kd_init(KdArgs a) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= a.npad) return;
  const int k = a.blk_sub[g / DS_BS];
  const int loc = g - a.sub_blk0[k] * DS_BS;
  uint4 r = make_uint4(0xFFFFFFFFu, 0u, 0u, 0u);
  if (loc < a.sub_n[k]) {
    const int idx = a.sub_off[k] + loc;
    const double* P = a.pix + 5 * (size_t)idx;
    const double* Q = a.quant + 9 * (size_t)k;
    r = make_uint4((unsigned)idx, kd_q(P[0], Q[0], Q[3]),
                   kd_q(P[1], Q[1], Q[4]), kd_q(P[2], Q[2], Q[5]));
  }
  a.rec[g] = r;
}

// ---- nodes above kWaveMax points: level-synchronous, all workgroups -----

struct BigArgs {
  KdArgs kd;
  const BigNode* nodes;
  const Chunk* chunks;
  unsigned* nq;               // per node: qlo[3], qhi[3]
  // The split key's prefix and the split's rank within it after an even
  // (first nn entries) or odd (next nn) number of digits; rank 0: found.
  // After the last digit the first nn prefixes hold the split keys.
  unsigned long long* pre;
  int* rank;
  int* hist;                  // 3 x max_nn x 256 digit counts, in turn
  unsigned* nq_next;          // the next level's boxes (the other half of
                              // a double buffer), set up by k_big_count
  int* ccount;                // per chunk: keys below the split
  int* coff;                  // per chunk: left keys of earlier chunks
  int* err;                   // bit 8: a split count differs from `left`,
                              // bit 16: a digit histogram does not add up
  int nn, nc;
  int max_nn;                 // node capacity of hist and of each nq half
  int nn_next;                // the next level's node count
};

__device__ inline int node_axis(const BigArgs& a, int i) {
  unsigned lo[3], hi[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    lo[t] = a.nq[6 * i + t];
    hi[t] = a.nq[6 * i + 3 + t];
  }
  return kd_axis(lo, hi, a.kd.quant + 9 * (size_t)a.nodes[i].sub + 6);
}

// Adds the lanes with m set to their digit's bin; a wavefront whose
// counted lanes share one digit (the common case near a node's edges and
// in the prefix's digits) adds them with one atomic.
__device__ inline void lds_count(int* lh, bool m, unsigned dg) {
  const unsigned long long am = __ballot(m);
  if (!am) return;
  const int leader = __builtin_ctzll(am);
  const unsigned dl = (unsigned)__shfl((int)dg, leader);
  if (__ballot(m && dg == dl) == am) {
    if ((int)(threadIdx.x & 63) == leader) atomicAdd(&lh[dl], __popcll(am));
  } else if (m) {
    atomicAdd(&lh[dg], 1);
  }
}

// The next digit of a split key: h counts that digit over the node's keys
// that carry the prefix; the digit holding `rank` joins the prefix and the
// rank becomes the rank within it.  At rank 0 the split is the first key
// of the prefix's bucket, so the prefix with every lower digit zero splits
// the keys the same way, and nothing changes.  All 256 threads call this
// with the same prefix and rank.
__device__ inline void big_select(const int* h, int shift,
                                  unsigned long long& pre, int& rank,
                                  int* sel, int* err) {
  if (rank == 0) return;
  const int t = threadIdx.x, wave = t >> 6;
  const int c = h[t];
  const int iw = wscan_incl(c);
  if ((t & 63) == 63) sel[wave] = iw;
  if (t == 0) sel[4] = -1;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += sel[w];
  const int incl = before + iw, excl = incl - c;
  if (excl <= rank && rank < incl) {
    sel[4] = t;
    sel[5] = excl;
  }
  __syncthreads();
  const int d = sel[4], below = sel[5];
  __syncthreads();
  if (d < 0) {
    if (t == 0) atomicOr(err, 16);
    rank = 0;
    return;
  }
  pre |= (unsigned long long)d << shift;
  rank -= below;
}

__global__ __launch_bounds__(256) void k_big_init(BigArgs a) {
  const int i = blockIdx.x, t = threadIdx.x;
  a.hist[256 * (size_t)i + t] = 0;
  if (t < 3) {
    a.nq[6 * i + t] = ~0u;
    a.nq[6 * i + 3 + t] = 0u;
  }
}

__global__ __launch_bounds__(256) void k_big_bbox(BigArgs a) {
  __shared__ unsigned red[4][6];
  const Chunk c = a.chunks[blockIdx.x];
  const int t = threadIdx.x;
  uint4 v[kEpt];
#pragma unroll
  for (int u = 0; u < kEpt; ++u) {
    const int j = c.b + t + 256 * u;
    v[u] = a.kd.rec[min(j, c.e - 1)];
  }
  unsigned lo[3] = {~0u, ~0u, ~0u}, hi[3] = {0u, 0u, 0u};
#pragma unroll
  for (int u = 0; u < kEpt; ++u) {
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      lo[x] = min(lo[x], qax(v[u], x));
      hi[x] = max(hi[x], qax(v[u], x));
    }
  }
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    lo[x] = wmin_u(lo[x]);
    hi[x] = wmax_u(hi[x]);
  }
  if ((t & 63) == 0) {
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      red[t >> 6][x] = lo[x];
      red[t >> 6][3 + x] = hi[x];
    }
  }
  __syncthreads();
  if (t < 6) {
    unsigned r = red[0][t];
    for (int w = 1; w < 4; ++w)
      r = t < 3 ? min(r, red[w][t]) : max(r, red[w][t]);
    if (t < 3) atomicMin(&a.nq[6 * c.node + t], r);
    else atomicMax(&a.nq[6 * c.node + t], r);
  }
}

// Digit `pass` of the split keys (8 bits, most significant first).  Each
// workgroup first takes the previous digit from its counts (every
// workgroup of a node the same way; the node's first chunk records the
// result and clears the counts the next launch adds to), then counts its
// chunk's keys that carry the prefix, unless the split is already found.
__global__ __launch_bounds__(256) void k_big_hist(BigArgs a, int pass) {
  __shared__ int lh[256];
  __shared__ int sel[8];
  const int t = threadIdx.x;
  lh[t] = 0;
  const Chunk c = a.chunks[blockIdx.x];
  const int i = c.node;
  const BigNode nd = a.nodes[i];
  const size_t hs = 256 * (size_t)a.max_nn;
  unsigned long long pre = 0ull;
  int rank = nd.left;
  if (pass > 0) {
    const int s = (pass - 1) & 1;
    pre = a.pre[s * a.nn + i];
    rank = a.rank[s * a.nn + i];
    big_select(a.hist + ((pass - 1) % 3) * hs + 256 * (size_t)i,
               64 - 8 * pass, pre, rank, sel, a.err);
  }
  if ((int)blockIdx.x == nd.c0) {
    if (t == 0) {
      a.pre[(pass & 1) * a.nn + i] = pre;
      a.rank[(pass & 1) * a.nn + i] = rank;
    }
    a.hist[((pass + 1) % 3) * hs + 256 * (size_t)i + t] = 0;
  }
  if (rank == 0) return;
  const int axis = node_axis(a, i);
  const int shift = 56 - 8 * pass;
  uint4 v[kEpt];
#pragma unroll
  for (int u = 0; u < kEpt; ++u) {
    const int j = c.b + t + 256 * u;
    v[u] = a.kd.rec[min(j, c.e - 1)];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kEpt; ++u) {
    const int j = c.b + t + 256 * u;
    const unsigned long long key = kd_key(qax(v[u], axis), (int)v[u].x);
    const bool m = j < c.e && (pass == 0 ||
                               (key >> (shift + 8)) == (pre >> (shift + 8)));
    lds_count(lh, m, (unsigned)(key >> shift) & 255u);
  }
  __syncthreads();
  const int cnt = lh[t];
  if (cnt) atomicAdd(&a.hist[(pass % 3) * hs + 256 * (size_t)i + t], cnt);
}

// The last digit (as in k_big_hist), then the chunk's keys below the
// split key; the node's first chunk records the key for the scatter.
__global__ __launch_bounds__(256) void k_big_count(BigArgs a) {
  __shared__ int part[4];
  __shared__ int sel[8];
  const int t = threadIdx.x;
  const Chunk c = a.chunks[blockIdx.x];
  const int i = c.node;
  unsigned long long K = a.pre[a.nn + i];
  int rank = a.rank[a.nn + i];
  big_select(a.hist + 256 * ((size_t)a.max_nn + i), 0, K, rank, sel, a.err);
  if ((int)blockIdx.x == a.nodes[i].c0 && t == 0) {
    a.pre[i] = K;
    if (rank != 0) atomicOr(a.err, 16);
  }
  // the next level's first counts (slot 0, last read by this level's
  // seventh digit launch) and boxes
  for (int m = blockIdx.x; m < a.nn_next; m += gridDim.x) {
    a.hist[256 * (size_t)m + t] = 0;
    if (t < 6) a.nq_next[6 * m + t] = t < 3 ? ~0u : 0u;
  }
  const int axis = node_axis(a, i);
  uint4 v[kEpt];
#pragma unroll
  for (int u = 0; u < kEpt; ++u) {
    const int j = c.b + t + 256 * u;
    v[u] = a.kd.rec[min(j, c.e - 1)];
  }
  int n = 0;
#pragma unroll
  for (int u = 0; u < kEpt; ++u)
    n += c.b + t + 256 * u < c.e && kd_key(qax(v[u], axis), (int)v[u].x) < K;
  n = wsum_i(n);
  if ((t & 63) == 0) part[t >> 6] = n;
  __syncthreads();
  if (t == 0) a.ccount[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// One wavefront per node: the chunks' left-key offsets.
__global__ __launch_bounds__(64) void k_big_scan(BigArgs a) {
  const int lane = threadIdx.x;
  const BigNode nd = a.nodes[blockIdx.x];
  int run = 0;
  for (int c0 = 0; c0 < nd.nc; c0 += 64) {
    const int c = c0 + lane;
    const int v = c < nd.nc ? a.ccount[nd.c0 + c] : 0;
    const int incl = wscan_incl(v);
    if (c < nd.nc) a.coff[nd.c0 + c] = run + incl - v;
    run += __shfl(incl, 63);
  }
  if (lane == 0 && run != nd.left) atomicOr(a.err, 8);
}

// Stable partition of each chunk into tmp: keys below the split to the
// left child (after the earlier chunks' left keys), the rest to the right.
__global__ __launch_bounds__(256) void k_big_scatter(BigArgs a) {
  __shared__ int wsum[4];
  const Chunk c = a.chunks[blockIdx.x];
  const BigNode nd = a.nodes[c.node];
  const int axis = node_axis(a, c.node);
  const unsigned long long K = a.pre[c.node];
  const int j0 = c.b + (int)threadIdx.x * kEpt;
  uint4 v[kEpt];
  unsigned fl = 0;
  int nl = 0;
#pragma unroll
  for (int t = 0; t < kEpt; ++t) {
    const int j = j0 + t;
    v[t] = j < c.e ? a.kd.rec[j] : make_uint4(0u, 0u, 0u, 0u);
    const bool f = j < c.e && kd_key(qax(v[t], axis), (int)v[t].x) < K;
    fl |= (unsigned)f << t;
    nl += f;
  }
  // exclusive scan of the left counts over the 256 threads
  const int incl = wscan_incl(nl);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wsum[w];
  int li = before + incl - nl;                 // left keys before me
  int ri = (j0 - c.b) - li;                    // right keys before me
  const int lbase = nd.b + a.coff[blockIdx.x];
  const int rbase = nd.b + nd.left + (c.b - nd.b) - a.coff[blockIdx.x];
#pragma unroll
  for (int t = 0; t < kEpt; ++t) {
    const int j = j0 + t;
    const bool f = (fl >> t) & 1u;
    const int dst = f ? lbase + li : rbase + ri;
    if (j < c.e && dst >= nd.b && dst < nd.e) a.kd.tmp[dst] = v[t];
    li += f;
    ri += !f;
  }
}

// ---- nodes of at most kWaveMax points: one wavefront each, in LDS -------

// Node (b, e) of subvolume `sub` (more than 64 points) in block phase:
// every descendant split down to leaves of 64 made in place (the splits of
// each block into its tiles are k_kd_blocks').  The node's records come
// from the record buffer its last big-node level wrote (wn.w: 1 = tmp), and
// the finished order goes to rec.  The stack keeps the leaf size with each
// node for the split rule, always 64 here.
__global__ __launch_bounds__(64) void k_kd_wave(KdArgs a, const int4* wn) {
  __shared__ uint4 buf[kWaveMax], buf2[kWaveMax];
  __shared__ int hist[256];
  __shared__ int stk[3 * kStack];
  const int lane = threadIdx.x;
  const int4 node = wn[blockIdx.x];
  const int b = node.x, cnt = node.y - node.x;
  const double* W = a.quant + 9 * (size_t)node.z + 6;
  const uint4* src = node.w ? a.tmp : a.rec;
  for (int j = lane; j < cnt; j += 64) buf[j] = src[b + j];
  int sp = 0;
  if (lane == 0) {
    stk[0] = 0;
    stk[1] = cnt;
    stk[2] = DS_BS;
  }
  sp = cnt > DS_BS ? 1 : 0;
  __syncthreads();
  while (sp > 0) {
    --sp;
    const int lb = stk[3 * sp], le = stk[3 * sp + 1], L = stk[3 * sp + 2];
    const int m = le - lb;
    unsigned lo[3] = {~0u, ~0u, ~0u}, hi[3] = {0u, 0u, 0u};
    for (int j = lb + lane; j < le; j += 64) {
      const uint4 r = buf[j];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        lo[t] = min(lo[t], qax(r, t));
        hi[t] = max(hi[t], qax(r, t));
      }
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      lo[t] = wmin_u(lo[t]);
      hi[t] = wmax_u(hi[t]);
    }
    const int axis = kd_axis(lo, hi, W);
    const int nleaf = (m + L - 1) / L;
    const int left = L * (nleaf / 2);
    // radix select of the key of rank `left`.  The digits above the first
    // one in which the node's coordinates on this axis differ are the
    // box's; the search ends once the rank falls on the first key of a
    // bucket, whose prefix (lower digits zero) splits the node the same way
    const unsigned qd = lo[axis] ^ hi[axis];
    int pass = qd ? (int)(__builtin_clz(qd) >> 3) : 4;
    unsigned long long pre =
        pass ? (unsigned long long)(lo[axis] >> (32 - 8 * pass))
                   << (64 - 8 * pass)
             : 0ull;
    int rk = left;
    for (; pass < 8 && rk > 0; ++pass) {
      const int shift = 56 - 8 * pass;
#pragma unroll
      for (int t = 0; t < 4; ++t) hist[4 * lane + t] = 0;
      __syncthreads();
      for (int j0 = lb; j0 < le; j0 += 64) {
        const int j = j0 + lane;
        const uint4 r = buf[min(j, le - 1)];
        const unsigned long long key = kd_key(qax(r, axis), (int)r.x);
        const bool m = j < le && (pass == 0 || (key >> (shift + 8)) ==
                                                   (pre >> (shift + 8)));
        lds_count(hist, m, (unsigned)(key >> shift) & 255u);
      }
      __syncthreads();
      int c[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) c[t] = hist[4 * lane + t];
      const int s = c[0] + c[1] + c[2] + c[3];
      const int incl = wscan_incl(s), excl = incl - s;
      const bool mine = excl <= rk && rk < incl;
      int d = 0, below = excl;
      if (mine) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (rk < below + c[t]) {
            d = 4 * lane + t;
            break;
          }
          below += c[t];
        }
      }
      const unsigned long long mm = __ballot(mine);
      const int src = mm ? __builtin_ctzll(mm) : 0;
      d = __shfl(d, src);
      below = __shfl(below, src);
      pre |= (unsigned long long)d << shift;
      rk -= below;
      __syncthreads();
    }
    // stable partition into buf2, then back
    int nl = 0, nr = 0;
    for (int base = lb; base < le; base += 64) {
      const int j = base + lane;
      const bool valid = j < le;
      uint4 r = make_uint4(0u, 0u, 0u, 0u);
      bool f = false;
      if (valid) {
        r = buf[j];
        f = kd_key(qax(r, axis), (int)r.x) < pre;
      }
      const unsigned long long bl = __ballot(f), bv = __ballot(valid);
      const int rl = mbcnt64(bl), rv = mbcnt64(bv);
      const int dst = f ? lb + nl + rl : lb + left + nr + (rv - rl);
      if (valid && dst < le) buf2[dst] = r;
      nl += __popcll(bl);
      nr += __popcll(bv) - __popcll(bl);
    }
    __syncthreads();
    for (int j = lb + lane; j < le; j += 64) buf[j] = buf2[j];
    __syncthreads();
    // children that hold more than one block
    const int cb[2] = {lb, lb + left}, ce[2] = {lb + left, le};
    int push = 0;
#pragma unroll
    for (int h = 1; h >= 0; --h) {
      const int cn = ce[h] - cb[h];
      const int Lc = cn > DS_BS ? DS_BS : 0;
      if (Lc) {
        if (lane == 0) {
          stk[3 * (sp + push)] = cb[h];
          stk[3 * (sp + push) + 1] = ce[h];
          stk[3 * (sp + push) + 2] = Lc;
        }
        ++push;
      }
    }
    sp += push;
    __syncthreads();
  }
  for (int j = lane; j < cnt; j += 64) a.rec[b + j] = buf[j];
}

// ---- per block: slots, frames, boxes ------------------------------------

struct BlkArgs {
  const uint4* rec;
  const double* pix;
  const double* quant;      // per subvolume: lo[3], inv[3], w[3]
  const int* blk_sub;
  const int* sub_blk0;
  const int* sub_n;
  const double4* subcv;     // {var_s, c_perp, c_para, 0}
  const double* s3;         // per subvolume: s_perp, s_para
  int form;
  int* perm;
  double4* pos;
  double* rhs;
  double* diag;
  double* wvec;
  float4* loc;
  float4* tgt;
  double* tlo;
  double* thi;
  int* tcount;
  double* cen;
  double* tcen;
  unsigned* msrc;           // per subvolume: max |source coordinate| bits
  int* err;                 // bit 4: a slot without a valid record
  int n;
};

// The value of rank cnt / 2 among the real lanes of a group (median_frame
// of the host builder: nth_element at size / 2).
template <int W>
__device__ inline double group_median(double v, bool real, int cnt) {
  const int lane = threadIdx.x & 63, g0 = lane & ~(W - 1);
  int rank = 0;
  for (int j = 0; j < W; ++j) {
    const double u = __shfl(v, g0 + j);
    const bool rj = j < cnt;
    rank += rj && (u < v || (u == v && g0 + j < lane));
  }
  const bool hit = real && rank == cnt / 2;
  const unsigned long long m = __ballot(hit);
  const unsigned long long gm = W == 64 ? m : (m >> g0) & ((1ull << W) - 1);
  const int src = gm ? g0 + __builtin_ctzll(gm) : lane;
  return __shfl(v, src);
}

// The splits of one 64-slot block into its 16-slot tiles (build_blocking
// with leaf 16: the node's box picks the axis, its left child takes the
// `left` smallest keys, both children keep the node's order).  Lane i
// holds the record of slot i (lanes from nreal on hold none).  Every node
// is a range of lanes made of whole 16-lane groups (16 divides each split
// point), and the nodes of one level split together: the box from the
// groups' boxes, each key's rank within its node by a pass over the
// node's lanes, then every record pushed to its new lane.
__device__ inline uint4 tile_order(uint4 r, int nreal, const double w[3]) {
  const int lane = threadIdx.x & 63;
  const bool real = lane < nreal;
  int nb = real ? 0 : lane, ne = real ? nreal : lane;   // this lane's node
  for (int level = 0; level < 3; ++level) {
    const bool act = ne - nb > DS_TS;
    if (!__ballot(act)) break;
    unsigned lo[3], hi[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      unsigned gl = real ? qax(r, ax) : ~0u, gh = real ? qax(r, ax) : 0u;
#pragma unroll
      for (int o = 8; o >= 1; o >>= 1) {
        gl = min(gl, (unsigned)__shfl_xor((int)gl, o));
        gh = max(gh, (unsigned)__shfl_xor((int)gh, o));
      }
      lo[ax] = ~0u;
      hi[ax] = 0u;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const unsigned sl = (unsigned)__builtin_amdgcn_readlane((int)gl, 16 * g);
        const unsigned sh = (unsigned)__builtin_amdgcn_readlane((int)gh, 16 * g);
        if (16 * g >= nb && 16 * g < ne) {
          lo[ax] = min(lo[ax], sl);
          hi[ax] = max(hi[ax], sh);
        }
      }
    }
    const int axis = kd_axis(lo, hi, w);
    const unsigned q = qax(r, axis), id = r.x;
    int rank = 0;
    for (int j = 0; j < nreal; ++j) {
      const unsigned qj = (unsigned)__builtin_amdgcn_readlane((int)q, j);
      const unsigned ij = (unsigned)__builtin_amdgcn_readlane((int)id, j);
      const bool less = qj < q || (qj == q && ij < id);
      rank += (j >= nb && j < ne && less) ? 1 : 0;
    }
    const int m = ne - nb;
    const int left = DS_TS * (((m + DS_TS - 1) / DS_TS) / 2);
    const bool isl = act && rank < left;
    const unsigned long long bl = __ballot(isl);
    int dst = lane;
    if (act) {
      const unsigned long long below_nb = (1ull << nb) - 1ull;
      const int lb4 = mbcnt64(bl) - __popcll(bl & below_nb);
      dst = isl ? nb + lb4 : nb + left + (lane - nb - lb4);
    }
    r.x = (unsigned)__builtin_amdgcn_ds_permute(4 * dst, (int)r.x);
    r.y = (unsigned)__builtin_amdgcn_ds_permute(4 * dst, (int)r.y);
    r.z = (unsigned)__builtin_amdgcn_ds_permute(4 * dst, (int)r.z);
    r.w = (unsigned)__builtin_amdgcn_ds_permute(4 * dst, (int)r.w);
    if (act) {
      if (lane < nb + left) ne = nb + left;
      else nb += left;
    }
  }
  return r;
}

__global__ __launch_bounds__(256) void k_kd_blocks(BlkArgs a, int nblk) {
  const int B = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (B >= nblk) return;
  const int lane = threadIdx.x & 63;
  const int k = a.blk_sub[B];
  const int nreal = min(DS_BS, a.sub_n[k] - (B - a.sub_blk0[k]) * DS_BS);
  const bool real = lane < nreal;
  const size_t g = (size_t)B * DS_BS + lane;
  uint4 r = a.rec[(size_t)B * DS_BS + (real ? lane : nreal - 1)];
  {
    const double* Q = a.quant + 9 * (size_t)k + 6;
    const double w[3] = {Q[0], Q[1], Q[2]};
    r = tile_order(r, nreal, w);
    // the slots after the last real one repeat its (ordered) record
    const int src = real ? lane : nreal - 1;
    r.x = (unsigned)__shfl((int)r.x, src);
    r.y = (unsigned)__shfl((int)r.y, src);
    r.z = (unsigned)__shfl((int)r.z, src);
    r.w = (unsigned)__shfl((int)r.w, src);
  }
  int idx = (int)r.x;
  if (idx < 0 || idx >= a.n) {
    atomicOr(a.err, 4);
    idx = 0;
  }
  const double* P = a.pix + 5 * (size_t)idx;
  const double p[3] = {P[0], P[1], P[2]};
  const double w0 = P[3], dval = P[4];
  const double4 cv = a.subcv[k];
  const double var_s = cv.x;
  a.perm[g] = real ? idx : -1;
  const double wN = real ? w0 : (a.form == DS_FORM_W ? 0.0 : 1.0);
  a.pos[g] = make_double4(p[0], p[1], p[2], wN);
  // lib/wf_vector.cc:190-192 (W) and :44-47 (SN)
  a.rhs[g] = real ? dval : 0.0;
  a.diag[g] = real ? (a.form == DS_FORM_W ? wN * wN * var_s + 1.0
                                          : var_s + wN)
                   : 1.0;
  a.wvec[g] = real ? (a.form == DS_FORM_W ? wN : 1.0) : 0.0;
  const int t = lane >> 4;
  const int tcnt = max(0, min(DS_TS, nreal - DS_TS * t));
  const int T = B * DS_TPB + t;
  const double inf = __builtin_huge_val();
  double c[3], tc[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const double tl = gmin_d<16>(real ? p[ax] : inf);
    const double th = gmax_d<16>(real ? p[ax] : -inf);
    if ((lane & 15) == 0) {
      a.tlo[3 * (size_t)T + ax] = tl;
      a.thi[3 * (size_t)T + ax] = th;
    }
    c[ax] = group_median<64>(p[ax], real, nreal);
    const double tm = group_median<16>(p[ax], real, tcnt);
    tc[ax] = tcnt > 0 ? tm : c[ax];   // an empty tile takes its block's
    if (lane == 0) a.cen[3 * (size_t)B + ax] = c[ax];
    if ((lane & 15) == 0) a.tcen[3 * (size_t)T + ax] = tc[ax];
  }
  if ((lane & 15) == 0) a.tcount[T] = tcnt;
  const double s3[3] = {a.s3[2 * k], a.s3[2 * k], a.s3[2 * k + 1]};
  float l[3], tg[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    l[ax] = (float)((p[ax] - tc[ax]) * s3[ax]);
    tg[ax] = (float)((p[ax] - c[ax]) * s3[ax]);
  }
  a.loc[g] = make_float4(l[0], l[1], l[2],
                         l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
  a.tgt[g] = make_float4(tg[0], tg[1], tg[2], 0.0f);
  float ms = real ? fmaxf(fabsf(l[0]), fmaxf(fabsf(l[1]), fabsf(l[2]))) : 0.0f;
  for (int o = 32; o >= 1; o >>= 1) ms = fmaxf(ms, __shfl_xor(ms, o));
  if (lane == 0) atomicMax(&a.msrc[k], __float_as_uint(ms));
}

// Box of each group of 64 consecutive tiles (live tiles only).
__global__ __launch_bounds__(256) void k_kd_groups(const double* tlo,
                                                   const double* thi,
                                                   const int* tcount,
                                                   int ntile, int ngroup,
                                                   double* glo, double* ghi) {
  const int G = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (G >= ngroup) return;
  const int T = G * 64 + (threadIdx.x & 63);
  const bool live = T < ntile && tcount[T] > 0;
  const double inf = __builtin_huge_val();
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const double lo = gmin_d<64>(live ? tlo[3 * (size_t)T + ax] : inf);
    const double hi = gmax_d<64>(live ? thi[3 * (size_t)T + ax] : -inf);
    if ((threadIdx.x & 63) == 0) {
      glo[3 * (size_t)G + ax] = lo;
      ghi[3 * (size_t)G + ax] = hi;
    }
  }
}

struct ListArgs {
  const double* tlo;
  const double* thi;
  const int* tcount;
  const double* glo;
  const double* ghi;
  const int* blk_sub;
  const int* sub_blk0;
  const double4* subcv;
  const double* s3;
  const double* cen;
  const double* tcen;
  double dx_inv, limit;
  int nblk;
  int* bstat;               // count pass: 5 per block
  const int* tptr;          // fill pass
  int* tidx;
  unsigned char* tmask;
  float4* toff;
};

// Per target block: the source tiles of its subvolume that pass the cutoff
// test, with the mask of its 16-row sub-blocks (its own tiles) that do.
// Count pass: list length, half-list length and the pair counts; fill
// pass: the entries.
template <bool FILL>
__global__ __launch_bounds__(256) void k_kd_lists(ListArgs a) {
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (I >= a.nblk) return;
  const int lane = threadIdx.x & 63;
  const int k = a.blk_sub[I];
  const int T0 = a.sub_blk0[k] * DS_TPB, T1 = a.sub_blk0[k + 1] * DS_TPB;
  const double c_perp = a.subcv[k].y, c_para = a.subcv[k].z;
  const double inf = __builtin_huge_val();
  double slo[DS_TPB][3], shi[DS_TPB][3], alo[3], ahi[3];
  bool slive[DS_TPB];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    alo[ax] = inf;
    ahi[ax] = -inf;
  }
#pragma unroll
  for (int s = 0; s < DS_TPB; ++s) {
    const size_t Ts = (size_t)I * DS_TPB + s;
    slive[s] = a.tcount[Ts] > 0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      slo[s][ax] = a.tlo[3 * Ts + ax];
      shi[s][ax] = a.thi[3 * Ts + ax];
      if (slive[s]) {
        alo[ax] = fmin(alo[ax], slo[s][ax]);
        ahi[ax] = fmax(ahi[ax], shi[s][ax]);
      }
    }
  }
  float o3[3] = {0.f, 0.f, 0.f};
  double cI[3] = {0.0, 0.0, 0.0}, s3[3] = {0.0, 0.0, 0.0};
  if (FILL) {
    s3[0] = s3[1] = a.s3[2 * k];
    s3[2] = a.s3[2 * k + 1];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) cI[ax] = a.cen[3 * (size_t)I + ax];
  }
  const int base = FILL ? a.tptr[I] : 0;
  int cnt = 0, hc = 0, npop = 0, nhalf = 0, neval = 0;
  const int G0 = T0 >> 6, G1 = (T1 - 1) >> 6;
  for (int Gb = G0; Gb <= G1; Gb += 64) {
    const int Gl = Gb + lane;
    bool okg = false;
    if (Gl <= G1) {
      double gl[3], gh[3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        gl[ax] = a.glo[3 * (size_t)Gl + ax];
        gh[ax] = a.ghi[3 * (size_t)Gl + ax];
      }
      okg = gl[0] <= gh[0] &&
            box_pass(alo, ahi, gl, gh, c_perp, c_para, a.dx_inv, a.limit);
    }
    unsigned long long bm = __ballot(okg);
    while (bm) {
      const int G = Gb + __builtin_ctzll(bm);
      bm &= bm - 1;
      const int T = G * 64 + lane;
      bool ok = T >= T0 && T < T1 && a.tcount[T] > 0;
      unsigned msk = 0;
      if (ok) {
        double jl[3], jh[3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          jl[ax] = a.tlo[3 * (size_t)T + ax];
          jh[ax] = a.thi[3 * (size_t)T + ax];
        }
        ok = box_pass(alo, ahi, jl, jh, c_perp, c_para, a.dx_inv, a.limit);
        if (ok) {
#pragma unroll
          for (int s = 0; s < DS_TPB; ++s)
            if (slive[s] && box_pass(slo[s], shi[s], jl, jh, c_perp, c_para,
                                     a.dx_inv, a.limit))
              msk |= 1u << s;
          ok = msk != 0;
        }
      }
      const unsigned long long bb = __ballot(ok);
      if (ok) {
        const int e = base + cnt + mbcnt64(bb);
        if (FILL && e < a.tptr[I + 1]) {
          a.tidx[e] = T;
          a.tmask[e] = (unsigned char)msk;
#pragma unroll
          for (int ax = 0; ax < 3; ++ax)
            o3[ax] = (float)((cI[ax] - a.tcen[3 * (size_t)T + ax]) * s3[ax]);
          const unsigned bits = msk | ((unsigned)T << 4);
          a.toff[e] = make_float4(o3[0], o3[1], o3[2], __uint_as_float(bits));
        } else if (!FILL) {
          npop += __popc(msk);
          nhalf += ((msk & 3u) != 0) + ((msk & 12u) != 0);
          const int dT = T - DS_TPB * I;
          if (dT >= 0) {
            const unsigned keep = dT >= DS_TPB - 1 ? 15u : (2u << dT) - 1u;
            const unsigned mm = msk & keep;
            if (mm) {
              ++hc;
              neval += __popc(mm);
            }
          }
        }
      }
      cnt += __popcll(bb);
    }
  }
  if (!FILL) {
    hc = wsum_i(hc);
    npop = wsum_i(npop);
    nhalf = wsum_i(nhalf);
    neval = wsum_i(neval);
    if (lane == 0) {
      int* st = a.bstat + 5 * (size_t)I;
      st[0] = cnt;
      st[1] = hc;
      st[2] = npop;
      st[3] = nhalf;
      st[4] = neval;
    }
  }
}

// Exclusive scan of in[i * stride] (n values) into out[0 .. n]; out[n] = -1
// when the total exceeds INT_MAX.  One workgroup.
__global__ __launch_bounds__(1024) void k_kd_exscan(const int* in, int stride,
                                                    int n, int* out) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const long long per = ((long long)n + 1023) / 1024;
  const long long b = per * t, e = std::min<long long>(n, b + per);
  long long s = 0;
  for (long long i = b; i < e; ++i) s += in[i * stride];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const long long v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - s;
  const long long total = part[1023];
  const bool ok = total <= 2147483647LL;
  for (long long i = b; i < e; ++i) {
    out[i] = ok ? (int)run : 0;
    run += in[i * stride];
  }
  if (t == 1023) out[n] = ok ? (int)total : -1;
}

struct HalfArgs {
  const int* tptr;
  const int* tidx;
  const unsigned char* tmask;
  const float4* toff;
  const int* hptr;
  float4* hrec;
  int* gslot;
  int* gcnt;
  const int* gptr;
  int* err;
  int nblk;
};

// Block B's half list: its entries for tiles T >= 4 B, live bits kept for
// the sub-blocks s with 4 B + s <= T.
__global__ __launch_bounds__(256) void k_kd_half(HalfArgs a) {
  const int B = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (B >= a.nblk) return;
  const int lane = threadIdx.x & 63;
  const int e0 = a.tptr[B], e1 = a.tptr[B + 1], base = a.hptr[B];
  int c = 0;
  for (int eb = e0; eb < e1; eb += 64) {
    const int e = eb + lane;
    bool keep_rec = false;
    unsigned mm = 0;
    int T = 0;
    if (e < e1) {
      T = a.tidx[e];
      const int dT = T - DS_TPB * B;
      if (dT >= 0) {
        const unsigned keep = dT >= DS_TPB - 1 ? 15u : (2u << dT) - 1u;
        mm = a.tmask[e] & keep;
        keep_rec = mm != 0;
      }
    }
    const unsigned long long bb = __ballot(keep_rec);
    const int r = base + c + mbcnt64(bb);
    if (keep_rec && r < a.hptr[B + 1]) {
      const float4 o = a.toff[e];
      const unsigned bits = mm | ((unsigned)T << 4);
      a.hrec[r] = make_float4(o.x, o.y, o.z, __uint_as_float(bits));
      a.gslot[r] = -1;
    }
    c += __popcll(bb);
  }
}

// Column-sum slots: tile T is fed by the records (B', T) of the blocks B'
// with a sub-block s, 4 B' + s < T, live against T; by the symmetry of the
// box test these are the tiles T' = 4 B' + s < T in the list of T's own
// block with T's bit set, in ascending order.  Count pass: slots per tile;
// fill pass: each record's slot (tile's first slot + the block's rank).
template <bool FILL>
__global__ __launch_bounds__(256) void k_kd_gather(HalfArgs a) {
  const int Bt = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (Bt >= a.nblk) return;
  const int lane = threadIdx.x & 63;
  const int e0 = a.tptr[Bt], e1 = a.tptr[Bt + 1];
  int carry[DS_TPB], run[DS_TPB];
#pragma unroll
  for (int t = 0; t < DS_TPB; ++t) {
    carry[t] = -1;
    run[t] = 0;
  }
  for (int eb = e0; eb < e1; eb += 64) {
    const int e = eb + lane;
    const bool valid = e < e1;
    const int Tp = valid ? a.tidx[e] : 0;
    const unsigned m = valid ? a.tmask[e] : 0u;
    const int Bp = Tp >> 2;
#pragma unroll
    for (int t = 0; t < DS_TPB; ++t) {
      const int T = DS_TPB * Bt + t;
      const bool q = valid && Tp < T && ((m >> t) & 1u);
      const unsigned long long bq = __ballot(q);
      const unsigned long long below =
          lane ? bq & ((1ull << lane) - 1ull) : 0ull;
      const int src = below ? 63 - __builtin_clzll(below) : lane;
      const int prevB = __shfl(Bp, src);
      const bool first = q && Bp != (below ? prevB : carry[t]);
      const unsigned long long bf = __ballot(first);
      if (FILL && first) {
        const int rnk = run[t] + mbcnt64(bf);
        int lo = a.hptr[Bp], hi = a.hptr[Bp + 1];
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          const int Tm = (int)(__float_as_uint(a.hrec[mid].w) >> 4);
          if (Tm < T) lo = mid + 1;
          else hi = mid;
        }
        if (lo < a.hptr[Bp + 1] &&
            (int)(__float_as_uint(a.hrec[lo].w) >> 4) == T)
          a.gslot[lo] = a.gptr[T] + rnk;
        else
          atomicOr(a.err, 1);
      }
      run[t] += __popcll(bf);
      const int last = bq ? 63 - __builtin_clzll(bq) : 0;
      const int lastB = __shfl(Bp, last);
      if (bq) carry[t] = lastB;
    }
  }
  if (!FILL && lane < DS_TPB) {
    int v = run[0];
#pragma unroll
    for (int t = 1; t < DS_TPB; ++t)
      if (lane == t) v = run[t];
    a.gcnt[DS_TPB * Bt + lane] = v;
  }
}

// A temporary of the builder, taken from the context's buffer cache and
// returned to it (the stream has been synchronized by then: see SyncGuard).
template <class T>
struct Tmp {
  T* p = nullptr;
  ds_ctx* c = nullptr;
  size_t cap = 0;
  hipError_t alloc(ds_ctx* ctx, size_t n) {
    c = ctx;
    return ds::ctx_buf_get(c, (void**)&p, sizeof(T) * std::max<size_t>(n, 1),
                           &cap);
  }
  ~Tmp() {
    if (p) ds::ctx_buf_put(c, (void*)p, cap);
  }
};

// Waits for the stream when the builder returns, before the temporaries
// (declared earlier, so destroyed later) go back to the cache.
struct SyncGuard {
  hipStream_t st;
  ~SyncGuard() { (void)hipStreamSynchronize(st); }
};

#define GTRY(x)                                   \
  do {                                            \
    const hipError_t e_ = (x);                    \
    if (e_ != hipSuccess) return e_;              \
  } while (0)

}  // namespace

hipError_t build_geometry_device(ds_problem* p, const double* dpix,
                                 const std::vector<KdQuant>& quant,
                                 std::vector<float>& max_src,
                                 std::vector<int>& bstat, int* error) {
  *error = 0;
  hipStream_t st = p->ctx->stream;
  const int n = p->n, n_sub = p->n_sub, nblk = p->nblk, ntile = p->ntile,
            npad = p->npad;
  // ---- the tree schedule, from the pixel counts alone ----
  std::vector<std::vector<BigNode>> lv_nodes;
  std::vector<std::vector<Chunk>> lv_chunks;
  std::vector<int4> wave_nodes;
  struct Pend { int b, e, sub, depth; };
  std::vector<Pend> todo;
  for (int k = 0; k < n_sub; ++k) {
    const int b = p->sub_blk0_h[k] * DS_BS;
    todo.push_back({b, b + p->sub_n[k], k, 0});
  }
  while (!todo.empty()) {
    const Pend d = todo.back();
    todo.pop_back();
    const int cnt = d.e - d.b;
    if (cnt <= DS_BS) continue;   // one block: k_kd_blocks makes its tiles
    if (cnt <= kWaveMax) {
      // depth d >= 1: written by level d - 1's scatter, which writes tmp
      // at even levels and rec at odd ones (depth 0: rec, from k_kd_init)
      wave_nodes.push_back(make_int4(d.b, d.e, d.sub, d.depth & 1));
      continue;
    }
    if ((int)lv_nodes.size() <= d.depth) {
      lv_nodes.resize(d.depth + 1);
      lv_chunks.resize(d.depth + 1);
    }
    const int nleaf = (cnt + DS_BS - 1) / DS_BS;
    const int left = DS_BS * (nleaf / 2);
    std::vector<Chunk>& ch = lv_chunks[d.depth];
    BigNode nd{d.b, d.e, left, d.sub, (int)ch.size(), 0};
    const int id = (int)lv_nodes[d.depth].size();
    for (int c = d.b; c < d.e; c += kChunk)
      ch.push_back({id, c, std::min(d.e, c + kChunk)});
    nd.nc = (int)ch.size() - nd.c0;
    lv_nodes[d.depth].push_back(nd);
    todo.push_back({d.b + left, d.e, d.sub, d.depth + 1});
    todo.push_back({d.b, d.b + left, d.sub, d.depth + 1});
  }
  size_t max_nn = 1, max_nc = 1, tot_nn = 0, tot_nc = 0;
  for (size_t l = 0; l < lv_nodes.size(); ++l) {
    max_nn = std::max(max_nn, lv_nodes[l].size());
    max_nc = std::max(max_nc, lv_chunks[l].size());
    tot_nn += lv_nodes[l].size();
    tot_nc += lv_chunks[l].size();
  }
  // ---- inputs ----
  std::vector<double> hq(9 * (size_t)n_sub), hs3(2 * (size_t)n_sub);
  std::vector<int> hoff(n_sub), hblk_sub(nblk), hnreal(nblk);
  std::vector<double4> hcv(n_sub);
  for (int k = 0; k < n_sub; ++k) {
    for (int a = 0; a < 3; ++a) {
      hq[9 * k + a] = quant[k].lo[a];
      hq[9 * k + 3 + a] = quant[k].inv[a];
      hq[9 * k + 6 + a] = quant[k].w[a];
    }
    const CovarDev& cv = p->sub_cv_h[k];
    hs3[2 * k] = cv.s_perp;
    hs3[2 * k + 1] = cv.s_para;
    hcv[k] = make_double4(cv.var_s, cv.c_perp, cv.c_para, 0.0);
    hoff[k] = p->sub_off[k];
    for (int B = p->sub_blk0_h[k]; B < p->sub_blk0_h[k + 1]; ++B) {
      hblk_sub[B] = k;
      hnreal[B] = std::min(DS_BS, p->sub_n[k] - (B - p->sub_blk0_h[k]) * DS_BS);
    }
  }
  std::vector<BigNode> all_nodes;
  std::vector<Chunk> all_chunks;
  all_nodes.reserve(tot_nn);
  all_chunks.reserve(tot_nc);
  for (size_t l = 0; l < lv_nodes.size(); ++l) {
    all_nodes.insert(all_nodes.end(), lv_nodes[l].begin(), lv_nodes[l].end());
    all_chunks.insert(all_chunks.end(), lv_chunks[l].begin(),
                      lv_chunks[l].end());
  }
  Tmp<double> dq, ds3, dcen, dtcen, dglo, dghi;
  Tmp<int> dsub_n, dsub_off, dcc, dco, drank, dhist, dgcnt, derr, dstat;
  Tmp<double4> dcv;
  Tmp<uint4> drec, dtmp;
  Tmp<BigNode> dnodes;
  Tmp<Chunk> dchunks;
  Tmp<unsigned> dnq, dmsrc;
  Tmp<unsigned long long> dpre;
  Tmp<int4> dwave;
  SyncGuard guard{st};
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st)
                 : hipSuccess;
  };
  GTRY(dq.alloc(p->ctx, hq.size()));
  GTRY(up(dq.p, hq.data(), sizeof(double) * hq.size()));
  GTRY(ds3.alloc(p->ctx, hs3.size()));
  GTRY(up(ds3.p, hs3.data(), sizeof(double) * hs3.size()));
  GTRY(dcv.alloc(p->ctx, n_sub));
  GTRY(up(dcv.p, hcv.data(), sizeof(double4) * n_sub));
  GTRY(dsub_n.alloc(p->ctx, n_sub));
  GTRY(up(dsub_n.p, p->sub_n.data(), sizeof(int) * n_sub));
  GTRY(dsub_off.alloc(p->ctx, n_sub));
  GTRY(up(dsub_off.p, hoff.data(), sizeof(int) * n_sub));
  GTRY(ds::palloc(p, &p->blk_sub, nblk));
  GTRY(up(p->blk_sub, hblk_sub.data(), sizeof(int) * nblk));
  GTRY(ds::palloc(p, &p->blk_nreal, nblk));
  GTRY(up(p->blk_nreal, hnreal.data(), sizeof(int) * nblk));
  GTRY(ds::palloc(p, &p->sub_blk0, n_sub + 1));
  GTRY(up(p->sub_blk0, p->sub_blk0_h.data(), sizeof(int) * (n_sub + 1)));
  GTRY(drec.alloc(p->ctx, npad));
  GTRY(derr.alloc(p->ctx, 1));
  GTRY(hipMemsetAsync(derr.p, 0, sizeof(int), st));
  if (tot_nn) {
    GTRY(dtmp.alloc(p->ctx, npad));
    GTRY(dnodes.alloc(p->ctx, tot_nn));
    GTRY(up(dnodes.p, all_nodes.data(), sizeof(BigNode) * tot_nn));
    GTRY(dchunks.alloc(p->ctx, tot_nc));
    GTRY(up(dchunks.p, all_chunks.data(), sizeof(Chunk) * tot_nc));
    GTRY(dnq.alloc(p->ctx, 2 * 6 * max_nn));
    GTRY(dpre.alloc(p->ctx, 2 * max_nn));
    GTRY(drank.alloc(p->ctx, 2 * max_nn));
    GTRY(dhist.alloc(p->ctx, 3 * 256 * max_nn));
    GTRY(dcc.alloc(p->ctx, max_nc));
    GTRY(dco.alloc(p->ctx, max_nc));
  }
  KdArgs kd{};
  kd.pix = dpix;
  kd.rec = drec.p;
  kd.tmp = dtmp.p;
  kd.blk_sub = p->blk_sub;
  kd.sub_blk0 = p->sub_blk0;
  kd.sub_n = dsub_n.p;
  kd.sub_off = dsub_off.p;
  kd.quant = dq.p;
  kd.npad = npad;
  hipLaunchKernelGGL(k_kd_init, dim3((npad + 255) / 256), dim3(256), 0, st,
                     kd);
  GTRY(hipGetLastError());
  // ---- big nodes, level by level; each level reads one record buffer
  // and scatters into the other ----
  size_t no = 0, co = 0;
  for (size_t l = 0; l < lv_nodes.size(); ++l) {
    BigArgs b{};
    b.kd = kd;
    b.kd.rec = (l & 1) ? dtmp.p : drec.p;
    b.kd.tmp = (l & 1) ? drec.p : dtmp.p;
    b.nodes = dnodes.p + no;
    b.chunks = dchunks.p + co;
    b.nq = dnq.p + (l & 1) * 6 * max_nn;
    b.nq_next = dnq.p + ((l + 1) & 1) * 6 * max_nn;
    b.max_nn = (int)max_nn;
    b.nn_next = l + 1 < lv_nodes.size() ? (int)lv_nodes[l + 1].size() : 0;
    b.pre = dpre.p;
    b.rank = drank.p;
    b.hist = dhist.p;
    b.ccount = dcc.p;
    b.coff = dco.p;
    b.err = derr.p;
    b.nn = (int)lv_nodes[l].size();
    b.nc = (int)lv_chunks[l].size();
    const dim3 gn(b.nn), gc(b.nc);
    if (l == 0) hipLaunchKernelGGL(k_big_init, gn, dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_big_bbox, gc, dim3(256), 0, st, b);
    for (int pass = 0; pass < 8; ++pass)
      hipLaunchKernelGGL(k_big_hist, gc, dim3(256), 0, st, b, pass);
    hipLaunchKernelGGL(k_big_count, gc, dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_big_scan, gn, dim3(64), 0, st, b);
    hipLaunchKernelGGL(k_big_scatter, gc, dim3(256), 0, st, b);
    GTRY(hipGetLastError());
    no += lv_nodes[l].size();
    co += lv_chunks[l].size();
  }
  // ---- the rest of each tree, one wavefront per node ----
  if (!wave_nodes.empty()) {
    GTRY(dwave.alloc(p->ctx, wave_nodes.size()));
    GTRY(up(dwave.p, wave_nodes.data(), sizeof(int4) * wave_nodes.size()));
    hipLaunchKernelGGL(k_kd_wave, dim3((unsigned)wave_nodes.size()), dim3(64),
                       0, st, kd, (const int4*)dwave.p);
    GTRY(hipGetLastError());
  }
  // ---- per block ----
  GTRY(ds::palloc(p, &p->pos64, npad));
  GTRY(ds::palloc(p, &p->rhs64, npad));
  GTRY(ds::palloc(p, &p->diag, npad));
  GTRY(ds::palloc(p, &p->wvec, npad));
  GTRY(ds::palloc(p, &p->loc32, npad));
  GTRY(ds::palloc(p, &p->tgt32, npad));
  GTRY(ds::palloc(p, &p->perm_dev, npad));
  GTRY(ds::palloc(p, &p->tbox_lo, 3 * (size_t)ntile));
  GTRY(ds::palloc(p, &p->tbox_hi, 3 * (size_t)ntile));
  GTRY(ds::palloc(p, &p->tcount, ntile));
  GTRY(dcen.alloc(p->ctx, 3 * (size_t)nblk));
  GTRY(dtcen.alloc(p->ctx, 3 * (size_t)ntile));
  GTRY(dmsrc.alloc(p->ctx, n_sub));
  GTRY(hipMemsetAsync(dmsrc.p, 0, sizeof(unsigned) * n_sub, st));
  BlkArgs ba{};
  ba.rec = drec.p;
  ba.pix = dpix;
  ba.quant = dq.p;
  ba.blk_sub = p->blk_sub;
  ba.sub_blk0 = p->sub_blk0;
  ba.sub_n = dsub_n.p;
  ba.subcv = dcv.p;
  ba.s3 = ds3.p;
  ba.form = p->form;
  ba.perm = p->perm_dev;
  ba.pos = p->pos64;
  ba.rhs = p->rhs64;
  ba.diag = p->diag;
  ba.wvec = p->wvec;
  ba.loc = p->loc32;
  ba.tgt = p->tgt32;
  ba.tlo = p->tbox_lo;
  ba.thi = p->tbox_hi;
  ba.tcount = p->tcount;
  ba.cen = dcen.p;
  ba.tcen = dtcen.p;
  ba.msrc = dmsrc.p;
  ba.err = derr.p;
  ba.n = n;
  const dim3 gb((nblk + 3) / 4);
  hipLaunchKernelGGL(k_kd_blocks, gb, dim3(256), 0, st, ba, nblk);
  GTRY(hipGetLastError());
  const int ngroup = (ntile + 63) / 64;
  GTRY(dglo.alloc(p->ctx, 3 * (size_t)ngroup));
  GTRY(dghi.alloc(p->ctx, 3 * (size_t)ngroup));
  hipLaunchKernelGGL(k_kd_groups, dim3((ngroup + 3) / 4), dim3(256), 0, st,
                     p->tbox_lo, p->tbox_hi, p->tcount, ntile, ngroup, dglo.p,
                     dghi.p);
  GTRY(hipGetLastError());
  // ---- culled lists: count, scan, fill ----
  GTRY(dstat.alloc(p->ctx, 5 * (size_t)nblk));
  ListArgs la{};
  la.tlo = p->tbox_lo;
  la.thi = p->tbox_hi;
  la.tcount = p->tcount;
  la.glo = dglo.p;
  la.ghi = dghi.p;
  la.blk_sub = p->blk_sub;
  la.sub_blk0 = p->sub_blk0;
  la.subcv = dcv.p;
  la.s3 = ds3.p;
  la.cen = dcen.p;
  la.tcen = dtcen.p;
  la.dx_inv = p->cv.dx_inv;
  la.limit = cull_limit(p->cv);
  la.nblk = nblk;
  la.bstat = dstat.p;
  hipLaunchKernelGGL(k_kd_lists<false>, gb, dim3(256), 0, st, la);
  GTRY(hipGetLastError());
  GTRY(ds::palloc(p, &p->tptr, nblk + 1));
  GTRY(ds::palloc(p, &p->hptr, nblk + 1));
  hipLaunchKernelGGL(k_kd_exscan, dim3(1), dim3(1024), 0, st, dstat.p, 5,
                     nblk, p->tptr);
  hipLaunchKernelGGL(k_kd_exscan, dim3(1), dim3(1024), 0, st, dstat.p + 1, 5,
                     nblk, p->hptr);
  GTRY(hipGetLastError());
  int tot[2] = {0, 0};
  GTRY(hipMemcpyAsync(&tot[0], p->tptr + nblk, sizeof(int),
                      hipMemcpyDeviceToHost, st));
  GTRY(hipMemcpyAsync(&tot[1], p->hptr + nblk, sizeof(int),
                      hipMemcpyDeviceToHost, st));
  GTRY(hipStreamSynchronize(st));
  if (tot[0] < 0 || tot[1] < 0) {
    *error = 1;   // more than INT_MAX list entries
    return hipSuccess;
  }
  GTRY(ds::palloc(p, &p->tidx, tot[0]));
  GTRY(ds::palloc(p, &p->tmask, tot[0]));
  GTRY(ds::palloc(p, &p->toff32, tot[0]));
  la.tptr = p->tptr;
  la.tidx = p->tidx;
  la.tmask = p->tmask;
  la.toff = p->toff32;
  hipLaunchKernelGGL(k_kd_lists<true>, gb, dim3(256), 0, st, la);
  GTRY(hipGetLastError());
  // ---- half lists and column-sum slots ----
  GTRY(ds::palloc(p, &p->hrec, tot[1]));
  GTRY(ds::palloc(p, &p->gslot, tot[1]));
  GTRY(ds::palloc(p, &p->gptr, ntile + 1));
  GTRY(dgcnt.alloc(p->ctx, ntile));
  HalfArgs ha{};
  ha.tptr = p->tptr;
  ha.tidx = p->tidx;
  ha.tmask = p->tmask;
  ha.toff = p->toff32;
  ha.hptr = p->hptr;
  ha.hrec = p->hrec;
  ha.gslot = p->gslot;
  ha.gcnt = dgcnt.p;
  ha.gptr = p->gptr;
  ha.err = derr.p;
  ha.nblk = nblk;
  hipLaunchKernelGGL(k_kd_half, gb, dim3(256), 0, st, ha);
  hipLaunchKernelGGL(k_kd_gather<false>, gb, dim3(256), 0, st, ha);
  hipLaunchKernelGGL(k_kd_exscan, dim3(1), dim3(1024), 0, st, dgcnt.p, 1,
                     ntile, p->gptr);
  hipLaunchKernelGGL(k_kd_gather<true>, gb, dim3(256), 0, st, ha);
  GTRY(hipGetLastError());
  // ---- what the host needs: per-block counts, far points, slot order ----
  // (one pinned buffer from the context's pool, copied back in one go)
  bstat.resize(5 * (size_t)nblk);
  std::vector<unsigned> ms(n_sub);
  int nslot = 0, gerr = 0;
  p->perm.resize(npad);
  const size_t o_stat = 0, o_ms = o_stat + sizeof(int) * bstat.size(),
               o_slot = o_ms + sizeof(unsigned) * n_sub,
               o_err = o_slot + sizeof(int), o_perm = o_err + sizeof(int),
               o_end = o_perm + sizeof(int) * (size_t)npad;
  char* hb = nullptr;
  GTRY(ds::pool_alloc(p, (void**)&hb, o_end, true));
  auto down = [&](size_t off, const void* src, size_t bytes) {
    return hipMemcpyAsync(hb + off, src, bytes, hipMemcpyDeviceToHost, st);
  };
  hipError_t ce = down(o_stat, dstat.p, o_ms - o_stat);
  if (ce == hipSuccess) ce = down(o_ms, dmsrc.p, o_slot - o_ms);
  if (ce == hipSuccess) ce = down(o_slot, p->gptr + ntile, sizeof(int));
  if (ce == hipSuccess) ce = down(o_err, derr.p, sizeof(int));
  if (ce == hipSuccess) ce = down(o_perm, p->perm_dev, o_end - o_perm);
  if (ce == hipSuccess) ce = hipStreamSynchronize(st);
  if (ce == hipSuccess) {
    std::memcpy(bstat.data(), hb + o_stat, o_ms - o_stat);
    std::memcpy(ms.data(), hb + o_ms, o_slot - o_ms);
    std::memcpy(&nslot, hb + o_slot, sizeof(int));
    std::memcpy(&gerr, hb + o_err, sizeof(int));
    std::memcpy(p->perm.data(), hb + o_perm, o_end - o_perm);
  } else {
    (void)hipStreamSynchronize(st);
  }
  ds::pool_free(p, hb);
  GTRY(ce);
  if (nslot < 0 || gerr) {
    *error = nslot < 0 ? 1 : 2 + gerr;
    return hipSuccess;
  }
  max_src.resize(n_sub);
  for (int k = 0; k < n_sub; ++k) {
    float f;
    std::memcpy(&f, &ms[k], sizeof(float));
    max_src[k] = f;
  }
  p->nslot = nslot;
  return hipSuccess;
}

}  // namespace ds
