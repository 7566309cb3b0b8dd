// geometry.cpp -- spatial blocking and exact cutoff culling (host side).
//
// The reference evaluates every one of the n^2 pairs (lib/wf_vector.cc:54-73)
// but its covariance is exactly zero once the table index reaches
// table_n - 1 (lib/signal_covar.h:57-62).  Pixels are sorted into compact
// blocks of 64 (one wavefront of target rows) with a k-d tree, each block is
// split into 4 tiles of 16 source slots, and a (target block, source tile)
// pair is dropped only when the smallest separation between the two bounding
// boxes already gives an index >= table_n - 1.  Rounding is monotone, so
// evaluating the reference's own expression on the box gaps bounds every
// pair's computed index from below: the dropped terms are the reference's
// exact +0.0.
#include <algorithm>
#include <cstdint>
#include <cmath>
#include <limits>
#include <numeric>
#include <thread>

#include "ds_internal.h"

namespace ds {

namespace {

unsigned n_threads() {
  unsigned t = std::thread::hardware_concurrency();
  return t == 0 ? 1u : std::min(t, 16u);
}

template <class F>
void parallel_for(int n, F f) {
  const unsigned nt = std::min<unsigned>(n_threads(), (unsigned)std::max(n, 1));
  if (nt <= 1 || n < 64) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (int i = (int)t; i < n; i += (int)nt) f(i);
    });
  for (auto& x : th) x.join();
}

// Recursively split perm[b, e) into leaves of `leaf` points; every left child
// holds a multiple of `leaf` points so only the last leaf can be short.  A
// node splits along the axis of its largest quantized extent (weighted by
// the covariance metric); its left child takes the `left` points of
// smallest key (quantized coordinate on that axis, then point index) and
// both children keep the node's order (a stable partition), so the
// ordering is a function of the points alone: the device builder makes the
// same splits with a radix select and a scan.
void kd_order(const double* pts, int stride, const KdQuant& Q, int* perm,
              int b, int e, int leaf) {
  struct Node { int b, e; };
  std::vector<Node> stack;
  std::vector<uint64_t> keys;
  if (e > b) stack.push_back({b, e});
  while (!stack.empty()) {
    Node nd = stack.back();
    stack.pop_back();
    const int cnt = nd.e - nd.b;
    if (cnt <= leaf) continue;
    unsigned lo[3] = {~0u, ~0u, ~0u}, hi[3] = {0u, 0u, 0u};
    for (int k = nd.b; k < nd.e; ++k) {
      const double* p = pts + (size_t)stride * perm[k];
      for (int a = 0; a < 3; ++a) {
        const unsigned q = kd_q(p[a], Q.lo[a], Q.inv[a]);
        lo[a] = std::min(lo[a], q);
        hi[a] = std::max(hi[a], q);
      }
    }
    const int axis = kd_axis(lo, hi, Q.w);
    const int nleaf = (cnt + leaf - 1) / leaf;
    const int left = leaf * (nleaf / 2);
    auto key = [&](int i) {
      return kd_key(kd_q(pts[(size_t)stride * i + axis], Q.lo[axis],
                         Q.inv[axis]), i);
    };
    keys.resize(cnt);
    for (int k = 0; k < cnt; ++k) keys[k] = key(perm[nd.b + k]);
    std::nth_element(keys.begin(), keys.begin() + left, keys.end());
    const uint64_t split = keys[left];
    std::stable_partition(perm + nd.b, perm + nd.e,
                          [&](int i) { return key(i) < split; });
    stack.push_back({nd.b + left, nd.e});
    stack.push_back({nd.b, nd.b + left});
  }
}

void box_of(const double* pts, int stride, const int* perm, int b, int e,
            double* lo, double* hi) {
  const double inf = std::numeric_limits<double>::infinity();
  for (int a = 0; a < 3; ++a) { lo[a] = inf; hi[a] = -inf; }
  for (int k = b; k < e; ++k) {
    const double* p = pts + (size_t)stride * perm[k];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[a]);
      hi[a] = std::max(hi[a], p[a]);
    }
  }
}

}  // namespace

KdQuant kd_quant(const double* pts, int stride, int n, const double scale[3]) {
  KdQuant Q;
  for (int a = 0; a < 3; ++a) {
    double lo = std::numeric_limits<double>::infinity(), hi = -lo;
    for (int i = 0; i < n; ++i) {
      const double x = pts[(size_t)stride * i + a];
      lo = std::min(lo, x);
      hi = std::max(hi, x);
    }
    kd_quant_axis(lo, hi, scale[a], Q, a);
  }
  return Q;
}

void kd_quant_axis(double lo, double hi, double scale, KdQuant& Q, int a) {
  const double span = hi - lo;
  Q.lo[a] = std::isfinite(lo) ? lo : 0.0;
  if (span > 0.0 && std::isfinite(span)) {
    Q.inv[a] = 4294967295.0 / span;
    Q.w[a] = scale * (span / 4294967295.0);
  } else {
    Q.inv[a] = 0.0;
    Q.w[a] = 0.0;
  }
}

CovarDev make_covar(const ds_covar* c) {
  CovarDev cv;
  cv.var_s = c->var_s;
  cv.c_perp = c->c_perp;
  cv.c_para = c->c_para;
  cv.dx_inv = c->dx_inv;
  cv.table_n = c->table_n;
  cv.s_perp = std::sqrt(c->c_perp * c->dx_inv);
  cv.s_para = std::sqrt(c->c_para * c->dx_inv);
  return cv;
}

void build_blocking(const double* pts, int stride, int n, const double scale[3],
                    Blocking& out) {
  out.n = n;
  out.nblk = (n + DS_BS - 1) / DS_BS;
  out.ntile = out.nblk * DS_TPB;
  out.npad = out.nblk * DS_BS;
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  const KdQuant Q = kd_quant(pts, stride, n, scale);
  kd_order(pts, stride, Q, perm.data(), 0, n, DS_BS);
  // split each block into tiles
  parallel_for(out.nblk, [&](int B) {
    const int b = B * DS_BS, e = std::min(n, b + DS_BS);
    kd_order(pts, stride, Q, perm.data(), b, e, DS_TS);
  });
  out.perm.assign(out.npad, n > 0 ? perm[n - 1] : 0);
  std::copy(perm.begin(), perm.end(), out.perm.begin());
  out.blo.assign(3 * (size_t)out.nblk, 0.0);
  out.bhi.assign(3 * (size_t)out.nblk, 0.0);
  out.tlo.assign(3 * (size_t)out.ntile, 0.0);
  out.thi.assign(3 * (size_t)out.ntile, 0.0);
  out.tcount.assign(out.ntile, 0);
  for (int B = 0; B < out.nblk; ++B)
    box_of(pts, stride, out.perm.data(), B * DS_BS,
           std::min(n, (B + 1) * DS_BS), &out.blo[3 * (size_t)B],
           &out.bhi[3 * (size_t)B]);
  for (int T = 0; T < out.ntile; ++T) {
    const int b = T * DS_TS, e = std::max(b, std::min(n, b + DS_TS));
    out.tcount[T] = e - b;
    box_of(pts, stride, out.perm.data(), b, e, &out.tlo[3 * (size_t)T],
           &out.thi[3 * (size_t)T]);
  }
}

double box_gap_xx(const double* alo, const double* ahi, const double* blo,
                  const double* bhi, const CovarDev& cv) {
  double g[3];
  for (int a = 0; a < 3; ++a) {
    if (blo[a] > ahi[a]) g[a] = blo[a] - ahi[a];
    else if (alo[a] > bhi[a]) g[a] = alo[a] - bhi[a];
    else g[a] = 0.0;
  }
  // lib/wf_vector.cc:65-66 then lib/signal_covar.h:69-72, :56
  const double x_perp_2 = (g[0] * g[0] + g[1] * g[1]);
  const double x_para_2 = g[2] * g[2];
  const double ga_perp = cv.c_perp * x_perp_2;
  const double ga_para = cv.c_para * x_para_2;
  return (ga_perp + ga_para) * cv.dx_inv;
}

double cull_limit(const CovarDev& cv) {
  // The reference drops a pair when (int)xx > table_n - 2, i.e. when
  // xx >= table_n - 1.  The box gap is a lower bound of every pair's
  // coordinate differences; a relative margin of 1e-12 covers the few ulps
  // by which the bound's own rounding may exceed a pair's.
  return (double)(cv.table_n - 1) * (1.0 + 1e-12);
}

namespace {

// Source tiles binned on a uniform grid by box centre, so that a target
// box visits only the tiles whose centres can lie within the cutoff of it
// instead of all of them.  The search is a conservative prefilter: the
// exact box test below still decides every entry, and the entries kept are
// put in ascending tile order, so the lists equal those of a full scan.
struct TileBins {
  bool ok = false;
  double cmin[3], inv[3], reach[3];
  int nb[3];
  std::vector<int> start, tiles;

  TileBins(const Blocking& src, const CovarDev& cv, double limit) {
    // A pair passes the cutoff test only if each axis' own term does:
    // c_perp g^2 dx_inv < limit for x and y, c_para g^2 dx_inv < limit for z.
    const double kperp = cv.c_perp * cv.dx_inv, kpara = cv.c_para * cv.dx_inv;
    if (!(kperp > 0.0) || !(kpara > 0.0)) return;
    const double r[3] = {std::sqrt(limit / kperp), std::sqrt(limit / kperp),
                         std::sqrt(limit / kpara)};
    // The bins span the 1st to 99th percentile of the tile centres per
    // axis (outliers would stretch a min-max grid until the bulk shares a
    // few bins); centres outside fall into the edge bins, since the bin
    // index is clamped, and a clamped monotone index keeps every query a
    // superset of the tiles within reach.
    double cmax[3], h[3] = {0.0, 0.0, 0.0};
    std::vector<double> cs[3];
    for (int J = 0; J < src.ntile; ++J) {
      if (src.tcount[J] == 0) continue;
      for (int a = 0; a < 3; ++a) {
        const double lo = src.tlo[3 * (size_t)J + a];
        const double hi = src.thi[3 * (size_t)J + a];
        cs[a].push_back(0.5 * (lo + hi));
        h[a] = std::max(h[a], 0.5 * (hi - lo));
      }
    }
    const int live = (int)cs[0].size();
    if (live == 0) return;
    for (int a = 0; a < 3; ++a) {
      std::vector<double>& v = cs[a];
      const size_t q0 = (size_t)(0.01 * (live - 1)), q1 = live - 1 - q0;
      std::nth_element(v.begin(), v.begin() + q0, v.end());
      cmin[a] = v[q0];
      std::nth_element(v.begin(), v.begin() + q1, v.end());
      cmax[a] = v[q1];
    }
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(r[a]) || !std::isfinite(cmax[a] - cmin[a])) return;
      // centres farther than the cutoff plus the widest tile's half-extent
      // from a target box hold no pair within the cutoff; the relative
      // margin covers the rounding of centres and half-extents
      reach[a] = (r[a] + h[a]) * (1.0 + 1e-9) +
                 1e-12 * (std::fabs(cmin[a]) + std::fabs(cmax[a]));
      const double span = cmax[a] - cmin[a];
      const double w = std::max(reach[a], span / 1024.0);
      nb[a] = (w > 0.0) ? std::max(1, (int)std::min(1024.0, span / w) + 1) : 1;
    }
    // any bin grid gives the same candidates' superset; keep the grid
    // within a few bins per tile (outliers stretch the span, not the work)
    const size_t cap = std::max<size_t>(4096, 2 * (size_t)live);
    auto count = [&]() { return (size_t)nb[0] * nb[1] * nb[2]; };
    while (count() > cap) {
      int a = 0;
      for (int q = 1; q < 3; ++q)
        if (nb[q] > nb[a]) a = q;
      nb[a] = (nb[a] + 1) / 2;
    }
    const size_t total = count();
    for (int a = 0; a < 3; ++a) {
      const double span = cmax[a] - cmin[a];
      inv[a] = (span > 0.0) ? nb[a] / (span * (1.0 + 1e-12)) : 0.0;
    }
    start.assign(total + 1, 0);
    std::vector<int> bin(src.ntile, -1);
    for (int J = 0; J < src.ntile; ++J) {
      if (src.tcount[J] == 0) continue;
      int b[3];
      for (int a = 0; a < 3; ++a)
        b[a] = index(a, 0.5 * (src.tlo[3 * (size_t)J + a] +
                               src.thi[3 * (size_t)J + a]));
      bin[J] = (b[0] * nb[1] + b[1]) * nb[2] + b[2];
      ++start[bin[J] + 1];
    }
    for (size_t k = 0; k < total; ++k) start[k + 1] += start[k];
    tiles.resize(live);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int J = 0; J < src.ntile; ++J)
      if (bin[J] >= 0) tiles[fill[bin[J]]++] = J;
    ok = true;
  }

  // bin of a coordinate: monotone in x, clamped to the grid
  int index(int a, double x) const {
    const double t = (x - cmin[a]) * inv[a];
    if (!(t > 0.0)) return 0;
    return (int)std::min<double>(nb[a] - 1, std::floor(t));
  }

  // tiles whose centres may lie within reach of [lo, hi], bin by bin
  void candidates(const double* lo, const double* hi,
                  std::vector<int>& out) const {
    out.clear();
    int b0[3], b1[3];
    for (int a = 0; a < 3; ++a) {
      b0[a] = index(a, lo[a] - reach[a]);
      b1[a] = index(a, hi[a] + reach[a]);
    }
    for (int x = b0[0]; x <= b1[0]; ++x)
      for (int y = b0[1]; y <= b1[1]; ++y)
        for (int z = b0[2]; z <= b1[2]; ++z) {
          const int k = (x * nb[1] + y) * nb[2] + z;
          out.insert(out.end(), tiles.begin() + start[k],
                     tiles.begin() + start[k + 1]);
        }
  }
};

}  // namespace

void build_tile_lists(int ntgt, const double* tlo, const double* thi,
                      const Blocking& src, const CovarDev& cv,
                      std::vector<int>& ptr, std::vector<int>& idx,
                      const double* sublo, const double* subhi,
                      std::vector<unsigned char>* mask) {
  const double limit = cull_limit(cv);
  const bool sub = sublo && subhi && mask;
  const TileBins bins(src, cv, limit);
  std::vector<std::vector<int>> lists(ntgt);
  std::vector<std::vector<unsigned char>> masks(sub ? ntgt : 0);
  std::vector<int> all;
  if (!bins.ok) {
    all.resize(src.ntile);
    std::iota(all.begin(), all.end(), 0);
  }
  parallel_for(ntgt, [&](int I) {
    std::vector<int>& L = lists[I];
    const double* alo = tlo + 3 * (size_t)I;
    const double* ahi = thi + 3 * (size_t)I;
    if (!(alo[0] <= ahi[0])) return;  // empty target
    std::vector<int> cand;
    if (bins.ok) bins.candidates(alo, ahi, cand);
    const std::vector<int>& C = bins.ok ? cand : all;
    for (int J : C) {
      if (src.tcount[J] == 0) continue;
      const double* jlo = &src.tlo[3 * (size_t)J];
      const double* jhi = &src.thi[3 * (size_t)J];
      const double xx = box_gap_xx(alo, ahi, jlo, jhi, cv);
      if (!(xx < limit)) continue;
      if (sub) {
        // the same test for each 16-row sub-block; an entry none of them
        // passes holds only pairs beyond the cutoff
        unsigned m = 0;
        for (int s = 0; s < DS_TPB; ++s) {
          const double* slo = sublo + 3 * ((size_t)I * DS_TPB + s);
          const double* shi = subhi + 3 * ((size_t)I * DS_TPB + s);
          if (!(slo[0] <= shi[0])) continue;   // empty sub-block
          if (box_gap_xx(slo, shi, jlo, jhi, cv) < limit) m |= 1u << s;
        }
        if (!m) continue;
        masks[I].push_back((unsigned char)m);
      }
      L.push_back(J);
    }
    if (bins.ok && !std::is_sorted(L.begin(), L.end())) {
      // ascending tile order, as a full scan would give
      std::vector<int> o(L.size());
      std::iota(o.begin(), o.end(), 0);
      std::sort(o.begin(), o.end(), [&](int a, int b) { return L[a] < L[b]; });
      std::vector<int> L2(L.size());
      for (size_t k = 0; k < o.size(); ++k) L2[k] = L[o[k]];
      L.swap(L2);
      if (sub) {
        std::vector<unsigned char> M2(o.size());
        for (size_t k = 0; k < o.size(); ++k) M2[k] = masks[I][o[k]];
        masks[I].swap(M2);
      }
    }
  });
  ptr.assign(ntgt + 1, 0);
  for (int I = 0; I < ntgt; ++I) ptr[I + 1] = ptr[I] + (int)lists[I].size();
  idx.resize(ptr[ntgt]);
  for (int I = 0; I < ntgt; ++I)
    std::copy(lists[I].begin(), lists[I].end(), idx.begin() + ptr[I]);
  if (sub) {
    mask->resize(ptr[ntgt]);
    for (int I = 0; I < ntgt; ++I)
      std::copy(masks[I].begin(), masks[I].end(), mask->begin() + ptr[I]);
  }
}

}  // namespace ds
