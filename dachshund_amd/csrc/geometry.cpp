// geometry.cpp -- spatial blocking and exact cutoff culling (host side).
//
// The reference evaluates every one of the n^2 pairs (lib/wf_vector.cc:54-73)
// but its covariance is exactly zero once the table index reaches
// table_n - 1 (lib/signal_covar.h:57-62).  Pixels are sorted into compact
// blocks of 64 with a k-d tree, and a (target, source) block pair is dropped
// only when the smallest separation between the two bounding boxes already
// gives an index >= table_n - 1.  Rounding is monotone, so evaluating the
// reference's own expression on the box gaps bounds every pair's computed
// index from below: the dropped terms are the reference's exact +0.0.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>

#include "ds_internal.h"

namespace ds {

namespace {

struct Node {
  int b, e;
};

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

}  // namespace

void build_blocking(const double* pts, int stride, int n, const double scale[3],
                    Blocking& out) {
  out.n = n;
  out.nblk = (n + DS_BS - 1) / DS_BS;
  out.npad = out.nblk * DS_BS;
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  // Every left child holds a multiple of DS_BS points, so each leaf starts
  // at a multiple of DS_BS and only the last leaf can be short.
  std::vector<Node> stack;
  if (n > 0) stack.push_back({0, n});
  while (!stack.empty()) {
    Node nd = stack.back();
    stack.pop_back();
    const int cnt = nd.e - nd.b;
    if (cnt <= DS_BS) continue;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = nd.b; k < nd.e; ++k) {
      const double* p = pts + (size_t)stride * perm[k];
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], p[a]);
        hi[a] = std::max(hi[a], p[a]);
      }
    }
    int axis = 0;
    double best = -1.0;
    for (int a = 0; a < 3; ++a) {
      const double ext = (hi[a] - lo[a]) * scale[a];
      if (ext > best) { best = ext; axis = a; }
    }
    const int nleaf = (cnt + DS_BS - 1) / DS_BS;
    const int left = DS_BS * (nleaf / 2);
    auto mid = perm.begin() + nd.b + left;
    std::nth_element(perm.begin() + nd.b, mid, perm.begin() + nd.e,
                     [&](int i, int j) {
                       const double a = pts[(size_t)stride * i + axis];
                       const double b = pts[(size_t)stride * j + axis];
                       return a < b || (a == b && i < j);
                     });
    stack.push_back({nd.b + left, nd.e});
    stack.push_back({nd.b, nd.b + left});
  }
  out.perm.assign(out.npad, n > 0 ? perm[n - 1] : 0);
  std::copy(perm.begin(), perm.end(), out.perm.begin());
  out.lo.assign(3 * (size_t)out.nblk, 0.0);
  out.hi.assign(3 * (size_t)out.nblk, 0.0);
  for (int B = 0; B < out.nblk; ++B) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    const int e = std::min(n, (B + 1) * DS_BS);
    for (int k = B * DS_BS; k < e; ++k) {
      const double* p = pts + (size_t)stride * out.perm[k];
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], p[a]);
        hi[a] = std::max(hi[a], p[a]);
      }
    }
    for (int a = 0; a < 3; ++a) {
      out.lo[3 * (size_t)B + a] = lo[a];
      out.hi[3 * (size_t)B + a] = hi[a];
    }
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

static void lists_from(int ntgt, const double* tlo, const double* thi,
                       const Blocking& src, const CovarDev& cv,
                       std::vector<int>& ptr, std::vector<int>& idx) {
  const double limit = (double)(cv.table_n - 1) * (1.0 + 1e-12);
  std::vector<std::vector<int>> lists(ntgt);
  parallel_for(ntgt, [&](int I) {
    std::vector<int>& L = lists[I];
    for (int J = 0; J < src.nblk; ++J) {
      const double xx = box_gap_xx(tlo + 3 * (size_t)I, thi + 3 * (size_t)I,
                                   &src.lo[3 * (size_t)J],
                                   &src.hi[3 * (size_t)J], cv);
      if (xx < limit) L.push_back(J);
    }
  });
  ptr.assign(ntgt + 1, 0);
  for (int I = 0; I < ntgt; ++I) ptr[I + 1] = ptr[I] + (int)lists[I].size();
  idx.resize(ptr[ntgt]);
  for (int I = 0; I < ntgt; ++I)
    std::copy(lists[I].begin(), lists[I].end(), idx.begin() + ptr[I]);
}

void build_pair_lists(const Blocking& tgt, const Blocking& src,
                      const CovarDev& cv, std::vector<int>& ptr,
                      std::vector<int>& idx) {
  lists_from(tgt.nblk, tgt.lo.data(), tgt.hi.data(), src, cv, ptr, idx);
}

void build_pair_lists_boxes(int ntgt, const double* tlo, const double* thi,
                            const Blocking& src, const CovarDev& cv,
                            std::vector<int>& ptr, std::vector<int>& idx) {
  lists_from(ntgt, tlo, thi, src, cv, ptr, idx);
}

}  // namespace ds
