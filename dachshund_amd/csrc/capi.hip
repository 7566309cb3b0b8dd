// capi.hip -- the extern "C" surface of libdachshund_gpu.so
// (declared in include/dachshund_gpu.h).
//
// Host side only: validation, spatial ordering and culling (geometry.cpp),
// uploads to HBM in sorted order, launches, and the permutation back to the
// caller's order.  Errors are returned as ds_status codes with a message in
// the context; nothing here exits or falls back to host arithmetic.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "ds_internal.h"

using ds::Blocking;
using ds::CovarDev;

namespace {

int fail(ds_ctx* ctx, int code, const char* fmt, ...)
    __attribute__((format(printf, 3, 4)));

int fail(ds_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIP_OR_FAIL(ctx, expr)                                            \
  do {                                                                    \
    hipError_t e_ = (expr);                                               \
    if (e_ != hipSuccess)                                                 \
      return fail((ctx), e_ == hipErrorOutOfMemory ? DS_E_NOMEM : DS_E_HIP, \
                  "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),         \
                  __FILE__, __LINE__);                                    \
  } while (0)

template <class T>
hipError_t dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, sizeof(T) * count);
}

template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// RAII device buffer for per-call temporaries.
template <class T>
struct DevBuf {
  T* p = nullptr;
  hipError_t alloc(size_t n) { return dev_alloc(&p, n); }
  ~DevBuf() { dev_free(p); }
};

double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int check_covar(ds_ctx* ctx, const ds_covar* c) {
  if (!c) return fail(ctx, DS_E_ARG, "covar is NULL");
  if (c->table_n < 2 || c->table_n > DS_MAX_TABLE)
    return fail(ctx, DS_E_ARG, "table_n %d outside [2, %d]", c->table_n,
                DS_MAX_TABLE);
  if (!c->table) return fail(ctx, DS_E_ARG, "covar table is NULL");
  if (!(std::isfinite(c->var_s) && std::isfinite(c->c_perp) &&
        std::isfinite(c->c_para) && std::isfinite(c->dx_inv)) ||
      c->c_perp <= 0.0 || c->c_para <= 0.0 || c->dx_inv <= 0.0)
    return fail(ctx, DS_E_ARG, "covariance constants must be finite, > 0");
  for (int k = 0; k < c->table_n; ++k)
    if (!std::isfinite(c->table[k]))
      return fail(ctx, DS_E_ARG, "table[%d] is not finite", k);
  return DS_OK;
}

int check_points(ds_ctx* ctx, const double* pts, int stride, int n,
                 const char* what) {
  if (n < 0) return fail(ctx, DS_E_ARG, "%s: negative count", what);
  if (n > 0 && !pts) return fail(ctx, DS_E_ARG, "%s: NULL array", what);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(pts[(size_t)stride * i + a]))
        return fail(ctx, DS_E_ARG, "%s: coordinate %d of record %d is not "
                    "finite", what, a, i);
  return DS_OK;
}

// {T[k], T[k+1]-T[k]} with a final {0, 0}: the interpolant of
// lib/signal_covar.h:53-65 as fma(df, t, f), exactly 0 at index table_n-1.
void make_tables(const ds_covar* c, std::vector<double2>& t64,
                 std::vector<float2>& t32) {
  const int n = c->table_n;
  t64.resize(n);
  t32.resize(n);
  for (int k = 0; k < n - 1; ++k) {
    t64[k] = make_double2(c->table[k], c->table[k + 1] - c->table[k]);
    t32[k] = make_float2((float)c->table[k],
                         (float)(c->table[k + 1] - c->table[k]));
  }
  t64[n - 1] = make_double2(0.0, 0.0);
  t32[n - 1] = make_float2(0.0f, 0.0f);
}

void axis_scale(const CovarDev& cv, double s[3]) {
  s[0] = std::sqrt(cv.c_perp);
  s[1] = std::sqrt(cv.c_perp);
  s[2] = std::sqrt(cv.c_para);
}

// A blocked set of source points on the device: sorted coordinates (with a
// per-point weight in .w) and the source vector u in sorted order.
struct SrcSet {
  Blocking blk;
  DevBuf<double4> pos;
  DevBuf<double> u;
};

// u_mode 0: u = vals[j] (smp); 1: u = 1 / pts[5j+3] (sigma, NPixel N).
int make_srcset(ds_ctx* ctx, const double* pts, int stride, int n,
                const double* vals, int u_mode, const CovarDev& cv,
                SrcSet& s) {
  double sc[3];
  axis_scale(cv, sc);
  ds::build_blocking(pts, stride, n, sc, s.blk);
  const int npad = s.blk.npad;
  std::vector<double4> hp(npad);
  std::vector<double> hu(npad, 0.0);
  for (int k = 0; k < npad; ++k) {
    const int j = s.blk.perm[k];
    const double* P = pts + (size_t)stride * j;
    hp[k] = make_double4(P[0], P[1], P[2], 0.0);
    if (k < n) hu[k] = u_mode == 0 ? vals[j] : 1.0 / P[3];
  }
  HIP_OR_FAIL(ctx, s.pos.alloc(npad));
  HIP_OR_FAIL(ctx, s.u.alloc(npad));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(s.pos.p, hp.data(), sizeof(double4) * npad,
                                  hipMemcpyHostToDevice, ctx->stream));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(s.u.p, hu.data(), sizeof(double) * npad,
                                  hipMemcpyHostToDevice, ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  return DS_OK;
}

// Map tiles of 4x4x4 cells; AABB of the cell centres of each tile, with the
// expression of lib/util.cc:30-39 (the kernels generate the same values).
void grid_tiles(const ds::MapGrid& g, std::vector<double>& lo,
                std::vector<double>& hi) {
  const size_t nt = (size_t)g.tx * g.ty * g.tz;
  lo.resize(3 * nt);
  hi.resize(3 * nt);
  const double dx = g.lx / g.nx, dy = g.ly / g.ny, dz = g.lz / g.nz;
  for (int a = 0; a < g.tx; ++a)
    for (int b = 0; b < g.ty; ++b)
      for (int c = 0; c < g.tz; ++c) {
        const size_t T = ((size_t)a * g.ty + b) * g.tz + c;
        const int x0 = a * DS_MAP_T, x1 = std::min(g.nx - 1, x0 + DS_MAP_T - 1);
        const int y0 = b * DS_MAP_T, y1 = std::min(g.ny - 1, y0 + DS_MAP_T - 1);
        const int z0 = c * DS_MAP_T, z1 = std::min(g.nz - 1, z0 + DS_MAP_T - 1);
        lo[3 * T + 0] = dx * (x0 + 0.5);
        hi[3 * T + 0] = dx * (x1 + 0.5);
        lo[3 * T + 1] = dy * (y0 + 0.5);
        hi[3 * T + 1] = dy * (y1 + 0.5);
        lo[3 * T + 2] = dz * (z0 + 0.5);
        hi[3 * T + 2] = dz * (z1 + 0.5);
      }
}

int check_grid(ds_ctx* ctx, double lx, double ly, double lz, int nx, int ny,
               int nz, ds::MapGrid& g) {
  if (nx <= 0 || ny <= 0 || nz <= 0)
    return fail(ctx, DS_E_ARG, "map grid dimensions must be positive");
  if ((double)nx * ny * nz > 2.0e9)
    return fail(ctx, DS_E_ARG, "map grid too large");
  if (!(std::isfinite(lx) && std::isfinite(ly) && std::isfinite(lz)))
    return fail(ctx, DS_E_ARG, "map box must be finite");
  g.lx = lx; g.ly = ly; g.lz = lz;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.tx = (nx + DS_MAP_T - 1) / DS_MAP_T;
  g.ty = (ny + DS_MAP_T - 1) / DS_MAP_T;
  g.tz = (nz + DS_MAP_T - 1) / DS_MAP_T;
  return DS_OK;
}

// Exclusive offsets of the per-map-tile list counts: ptr[T] = sum of
// cnt[t < T], ptr[nt] = the total, or -1 when the total exceeds cap
// entries (then no offset is written).  One workgroup of 1024 threads, each
// summing a contiguous run of counts.
__global__ __launch_bounds__(1024) void k_list_offsets(const int* cnt, int nt,
                                                       long long cap,
                                                       int* ptr) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const int per = (nt + 1023) / 1024;
  const int b = min(nt, t * per), e = min(nt, b + per);
  long long sum = 0;
  for (int i = b; i < e; ++i) sum += cnt[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {   // inclusive scan
    const long long v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  const long long total = part[1023];
  const bool ok = total <= cap;
  long long run = t > 0 ? part[t - 1] : 0;
  if (ok)
    for (int i = b; i < e; ++i) {
      ptr[i] = (int)run;
      run += cnt[i];
    }
  if (t == 0) ptr[nt] = ok ? (int)total : -1;
}

// Map onto a grid from a blocked source set (u in sorted order on device)
// whose ntile tile boxes are on the device (dlo / dhi, 3 per tile); hblk,
// when given, is the host blocking of the same set.
int map_grid_run(ds_ctx* ctx, int mode, const ds::MapGrid& g, int ntile,
                 const double* dlo, const double* dhi, const Blocking* hblk,
                 const double4* pos, const double* u, const double2* table64,
                 const CovarDev& cv, double* m_host) {
  const int nt = g.tx * g.ty * g.tz;
  std::vector<int> ptr, idx;
  ds::PoolTmp<int> dptr, didx;
  ds::PoolTmp<double> dm;
  const size_t n_map = (size_t)g.nx * g.ny * g.nz;
  hipStream_t st = ctx->stream;
  HIP_OR_FAIL(ctx, dptr.alloc(ctx, (size_t)nt + 1));
  const char* host_lists = std::getenv("DS_HOST_GRID_LISTS");
  if (hblk && host_lists && host_lists[0] == '1') {
    const Blocking& blk = *hblk;
    // diagnostic: the host's binned builder (the lists are identical)
    std::vector<double> lo, hi;
    grid_tiles(g, lo, hi);
    ds::build_tile_lists(nt, lo.data(), hi.data(), blk, cv, ptr, idx);
    HIP_OR_FAIL(ctx, didx.alloc(ctx, std::max<size_t>(idx.size(), 1)));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(dptr.p, ptr.data(),
                                    sizeof(int) * ptr.size(),
                                    hipMemcpyHostToDevice, st));
    if (!idx.empty())
      HIP_OR_FAIL(ctx, hipMemcpyAsync(didx.p, idx.data(),
                                      sizeof(int) * idx.size(),
                                      hipMemcpyHostToDevice, st));
  } else {
    // full scan on the device: count, offsets on the device, fill
    ds::PoolTmp<int> dcnt;
    HIP_OR_FAIL(ctx, dcnt.alloc(ctx, nt));
    HIP_OR_FAIL(ctx, ds::launch_grid_lists(g, dlo, dhi, ntile, cv, dcnt.p,
                                           nullptr, nullptr, st));
    hipLaunchKernelGGL(k_list_offsets, dim3(1), dim3(1024), 0, st, dcnt.p,
                       nt, (long long)INT_MAX, dptr.p);
    HIP_OR_FAIL(ctx, hipGetLastError());
    int total = 0;
    HIP_OR_FAIL(ctx, hipMemcpyAsync(&total, dptr.p + nt, sizeof(int),
                                    hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
    if (total < 0)
      return fail(ctx, DS_E_ARG, "map grid: too many tile-list entries");
    HIP_OR_FAIL(ctx, didx.alloc(ctx, std::max(total, 1)));
    HIP_OR_FAIL(ctx, ds::launch_grid_lists(g, dlo, dhi, ntile, cv, nullptr,
                                           dptr.p, didx.p, st));
  }
  HIP_OR_FAIL(ctx, dm.alloc(ctx, n_map));
  // the map tiles in up to 4 slabs along x: each slab's cells are one
  // contiguous run of the map, copied out while the next slab is computed
  ds::HostCopy out;
  HIP_OR_FAIL(ctx, out.begin(ctx, m_host, sizeof(double) * n_map, st));
  const int parts = std::min(4, g.tx);
  const size_t slab = (size_t)DS_MAP_T * g.ny * g.nz;   // cells per tile row
  for (int q = 0; q < parts; ++q) {
    const int x0 = (int)((long long)g.tx * q / parts);
    const int x1 = (int)((long long)g.tx * (q + 1) / parts);
    HIP_OR_FAIL(ctx, ds::launch_map64(mode, g, pos, u, dptr.p, didx.p,
                                      table64, cv, dm.p, st, x0, x1));
    const size_t c0 = std::min(n_map, (size_t)x0 * slab);
    const size_t c1 = std::min(n_map, (size_t)x1 * slab);
    HIP_OR_FAIL(ctx, out.enqueue(dm.p, sizeof(double) * c0,
                                 sizeof(double) * (c1 - c0)));
  }
  HIP_OR_FAIL(ctx, out.finish());
  return DS_OK;
}

// Map onto arbitrary points from a blocked source set.
int map_points_run(ds_ctx* ctx, int mode, int n_map, const double* map_pts,
                   const SrcSet& src, const double2* table64,
                   const CovarDev& cv, double* m_host) {
  if (n_map == 0) return DS_OK;
  double sc[3];
  axis_scale(cv, sc);
  Blocking tb;
  ds::build_blocking(map_pts, 3, n_map, sc, tb);
  std::vector<double4> ht(tb.npad);
  for (int k = 0; k < tb.npad; ++k) {
    const double* P = map_pts + 3 * (size_t)tb.perm[k];
    ht[k] = make_double4(P[0], P[1], P[2], 0.0);
  }
  std::vector<int> ptr, idx;
  ds::build_tile_lists(tb.nblk, tb.blo.data(), tb.bhi.data(), src.blk, cv, ptr,
                       idx);
  DevBuf<double4> dt;
  DevBuf<int> dperm, dptr, didx;
  DevBuf<double> dm;
  HIP_OR_FAIL(ctx, dt.alloc(tb.npad));
  HIP_OR_FAIL(ctx, dperm.alloc(tb.npad));
  HIP_OR_FAIL(ctx, dptr.alloc(ptr.size()));
  HIP_OR_FAIL(ctx, didx.alloc(idx.size()));
  HIP_OR_FAIL(ctx, dm.alloc(n_map));
  hipStream_t st = ctx->stream;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(dt.p, ht.data(), sizeof(double4) * tb.npad,
                                  hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(dperm.p, tb.perm.data(),
                                  sizeof(int) * tb.npad, hipMemcpyHostToDevice,
                                  st));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(dptr.p, ptr.data(), sizeof(int) * ptr.size(),
                                  hipMemcpyHostToDevice, st));
  if (!idx.empty())
    HIP_OR_FAIL(ctx, hipMemcpyAsync(didx.p, idx.data(),
                                    sizeof(int) * idx.size(),
                                    hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(ctx, ds::launch_points64(mode, n_map, dt.p, dperm.p,
                                       src.pos.p, src.u.p, dptr.p, didx.p,
                                       tb.nblk, table64, cv, dm.p, st));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(m_host, dm.p, sizeof(double) * n_map,
                                  hipMemcpyDeviceToHost, st));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
  return DS_OK;
}

int upload_table64(ds_ctx* ctx, const ds_covar* c, DevBuf<double2>& t) {
  std::vector<double2> t64;
  std::vector<float2> t32;
  make_tables(c, t64, t32);
  HIP_OR_FAIL(ctx, t.alloc(t64.size()));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(t.p, t64.data(), sizeof(double2) * t64.size(),
                                  hipMemcpyHostToDevice, ctx->stream));
  return DS_OK;
}

// Launch order of `blocks` (ascending block indices, spatially ordered):
// workgroups go to the 8 XCDs round-robin by launch slot (slot i to XCD
// i mod 8), so slot 8k + x gets the k-th block of XCD x's contiguous share
// of the list, each share taken longest (cost) first.  Neighbouring
// blocks, which read the same source tiles, then run on one XCD and share
// its L2 instead of each XCD fetching every tile.
template <class Cost>
std::vector<int> xcd_order(const std::vector<int>& blocks, Cost cost) {
  constexpr int kXcd = 8;
  const int n = (int)blocks.size();
  std::vector<int> out(n);
  int start = 0;
  for (int x = 0; x < kXcd; ++x) {
    const int cnt = n / kXcd + (x < n % kXcd ? 1 : 0);
    std::vector<int> share(blocks.begin() + start,
                           blocks.begin() + start + cnt);
    std::stable_sort(share.begin(), share.end(),
                     [&](int a, int b) { return cost(a) > cost(b); });
    for (int k = 0; k < cnt; ++k) out[(size_t)k * kXcd + x] = share[k];
    start += cnt;
  }
  return out;
}

void problem_free(ds_problem* p) {
  // the stream work that used them has completed (ds_problem_destroy and
  // the failed-creation path synchronize first)
  ds::pfree(p, p->knn_nbr);
  ds::pfree(p, p->minv);
  ds::pfree(p, p->perm_dev);
  ds::pfree(p, p->tbox_lo);
  ds::pfree(p, p->tbox_hi);
  ds::pfree(p, p->pos64);
  ds::pfree(p, p->rhs64);
  ds::pfree(p, p->diag);
  ds::pfree(p, p->wvec);
  ds::pfree(p, p->loc32);
  ds::pfree(p, p->tgt32);
  ds::pfree(p, p->tptr);
  ds::pfree(p, p->tidx);
  ds::pfree(p, p->tmask);
  ds::pfree(p, p->toff32);
  ds::pfree(p, p->table64);
  ds::pfree(p, p->table32);
  ds::pfree(p, p->table32g);
  ds::pfree(p, p->table_raw);
  ds::pfree(p, p->blk_sub);
  ds::pfree(p, p->blk_nreal);
  ds::pfree(p, p->blk_order);
  ds::pfree(p, p->ord32);
  ds::pfree(p, p->hptr);
  ds::pfree(p, p->hrec);
  ds::pfree(p, p->pbuf);
  ds::pfree(p, p->ydir);
  ds::pfree(p, p->gptr);
  ds::pfree(p, p->sub_sym);
  ds::pfree(p, p->gslot);
  ds::pfree(p, p->sub_blk0);
  ds::pfree(p, p->sub_cv);
  ds::pfree(p, p->tcount);
  ds::pfree(p, p->done_count);
  ds::pfree(p, p->dense_s);
  ds::pfree(p, p->dense_y);
  ds::pfree(p, p->du16);
  if (p->xu != p->x) ds::pfree(p, p->xu);
  if (p->du != p->d) ds::pfree(p, p->du);
  p->xu = p->du = nullptr;
  ds::pfree(p, p->x);
  ds::pfree(p, p->r);
  ds::pfree(p, p->d);
  ds::pfree(p, p->s);
  ds::pfree(p, p->q);
  ds::pfree(p, p->tmp);
  ds::pfree(p, p->du32);
  ds::pfree(p, p->partials);
  ds::pfree(p, p->sc);
  ds::pfree(p, p->norms);
  ds::pfree(p, p->sc_host);
  ds::pfree(p, p->done_host);
  ds::pfree(p, p->norm_host);
  ds::pfree(p, p->host_tmp);
  // anything else taken from the pool goes back too
  while (!p->pooled.empty()) ds::pool_free(p, p->pooled.back().ptr);
}

// Host-order vector -> sorted device vector (pads 0).
int upload_sorted(ds_problem* p, const double* v, double* dev) {
  for (int k = 0; k < p->npad; ++k)
    p->host_tmp[k] = p->perm[k] >= 0 ? v[p->perm[k]] : 0.0;
  HIP_OR_FAIL(p->ctx, hipMemcpyAsync(dev, p->host_tmp,
                                     sizeof(double) * p->npad,
                                     hipMemcpyHostToDevice, p->ctx->stream));
  HIP_OR_FAIL(p->ctx, hipStreamSynchronize(p->ctx->stream));
  return DS_OK;
}

int download_sorted(ds_problem* p, const double* dev, double* v) {
  HIP_OR_FAIL(p->ctx, hipMemcpyAsync(p->host_tmp, dev,
                                     sizeof(double) * p->npad,
                                     hipMemcpyDeviceToHost, p->ctx->stream));
  HIP_OR_FAIL(p->ctx, hipStreamSynchronize(p->ctx->stream));
  for (int k = 0; k < p->npad; ++k)
    if (p->perm[k] >= 0) v[p->perm[k]] = p->host_tmp[k];
  return DS_OK;
}

int check_pcg_params(ds_problem* p, const ds_pcg_params* prm) {
  if (!prm) return fail(p->ctx, DS_E_ARG, "params is NULL");
  if (prm->n != p->n)
    return fail(p->ctx, DS_E_ARG, "params.n = %d but the problem has %d "
                "pixels", prm->n, p->n);
  if (prm->step_r < 1)
    return fail(p->ctx, DS_E_ARG, "step_r must be >= 1 (the reference "
                "divides by it, lib/pcg.h:119)");
  if (!std::isfinite(prm->tol))
    return fail(p->ctx, DS_E_ARG, "tol is not finite");
  if (prm->max_iter > 100000000)
    return fail(p->ctx, DS_E_ARG, "max_iter too large");
  return DS_OK;
}

}  // namespace

namespace ds {

void ctx_retain(ds_ctx* c) { c->refs.fetch_add(1); }

void ctx_release(ds_ctx* c) {
  if (c->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(c->device);
  for (auto& e : c->buf_cache) (void)hipFree(e.second);
  c->buf_cache.clear();
  for (auto& e : c->host_cache) (void)hipHostFree(e.second);
  c->host_cache.clear();
  if (c->aux) (void)hipStreamDestroy(c->aux);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

namespace {

// Smallest cached buffer of at least `bytes` and at most bytes + bytes / 8
// + 64 KiB, taken out of the cache; false when there is none.
bool cache_take(std::vector<std::pair<size_t, void*>>& cache, size_t& held,
                size_t bytes, void** p, size_t* cap) {
  const size_t hi = bytes + bytes / 8 + (size_t(64) << 10);
  size_t best = cache.size();
  for (size_t k = 0; k < cache.size(); ++k) {
    const size_t b = cache[k].first;
    if (b >= bytes && b <= hi && (best == cache.size() || b < cache[best].first))
      best = k;
  }
  if (best == cache.size()) return false;
  *p = cache[best].second;
  *cap = cache[best].first;
  cache.erase(cache.begin() + (long)best);
  held -= *cap;
  return true;
}

}  // namespace

hipError_t ctx_buf_get(ds_ctx* c, void** p, size_t bytes, size_t* cap) {
  {
    std::lock_guard<std::mutex> lk(c->buf_mu);
    if (cache_take(c->buf_cache, c->buf_cache_bytes, bytes, p, cap))
      return hipSuccess;
  }
  *cap = bytes;
  return hipMalloc(p, bytes);
}

void ctx_buf_put(ds_ctx* c, void* p, size_t cap) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(c->buf_mu);
  if (cap > kBufCacheBytes) {
    (void)hipFree(p);
    return;
  }
  if (c->pool_poison) (void)hipMemset(p, 0xFF, cap);
  // oldest out first
  while (!c->buf_cache.empty() && c->buf_cache_bytes + cap > kBufCacheBytes) {
    (void)hipFree(c->buf_cache.front().second);
    c->buf_cache_bytes -= c->buf_cache.front().first;
    c->buf_cache.erase(c->buf_cache.begin());
  }
  c->buf_cache.emplace_back(cap, p);
  c->buf_cache_bytes += cap;
}

namespace {

hipError_t host_get(ds_ctx* c, void** p, size_t bytes, size_t* cap) {
  {
    std::lock_guard<std::mutex> lk(c->buf_mu);
    if (cache_take(c->host_cache, c->host_cache_bytes, bytes, p, cap))
      return hipSuccess;
  }
  *cap = bytes;
  return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

void host_put(ds_ctx* c, void* p, size_t cap) {
  std::lock_guard<std::mutex> lk(c->buf_mu);
  if (cap > kHostCacheBytes) {
    (void)hipHostFree(p);
    return;
  }
  if (c->pool_poison) std::memset(p, 0xFF, cap);
  while (!c->host_cache.empty() &&
         c->host_cache_bytes + cap > kHostCacheBytes) {
    (void)hipHostFree(c->host_cache.front().second);
    c->host_cache_bytes -= c->host_cache.front().first;
    c->host_cache.erase(c->host_cache.begin());
  }
  c->host_cache.emplace_back(cap, p);
  c->host_cache_bytes += cap;
}

}  // namespace

hipError_t HostCopy::begin(ds_ctx* ctx, void* d, size_t n, hipStream_t s) {
  c = ctx;
  dst = d;
  bytes = n;
  st = s;
  constexpr size_t kMin = size_t(4) << 20;
  direct = bytes < kMin || bytes > kHostCacheBytes ||
           host_get(c, &stage, bytes, &cap) != hipSuccess;
  if (direct) {
    stage = nullptr;
    (void)hipGetLastError();
  }
  return hipSuccess;
}

hipError_t HostCopy::enqueue(const void* src, size_t off, size_t len) {
  if (len == 0) return hipSuccess;
  if (direct)
    return hipMemcpyAsync(static_cast<char*>(dst) + off,
                          static_cast<const char*>(src) + off, len,
                          hipMemcpyDeviceToHost, st);
  // the part's copy runs on the context's second stream once the work
  // enqueued so far on st has produced it, so st goes on with the next part
  hipEvent_t ready, e;
  hipError_t err = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
  if (err != hipSuccess) return err;
  err = hipEventCreateWithFlags(&e, hipEventDisableTiming);
  if (err != hipSuccess) {
    (void)hipEventDestroy(ready);
    return err;
  }
  parts.push_back({off, len});
  ev.push_back(e);
  ev.push_back(ready);
  err = hipEventRecord(ready, st);
  if (err == hipSuccess) err = hipStreamWaitEvent(c->aux, ready, 0);
  if (err == hipSuccess)
    err = hipMemcpyAsync(static_cast<char*>(stage) + off,
                         static_cast<const char*>(src) + off, len,
                         hipMemcpyDeviceToHost, c->aux);
  return err != hipSuccess ? err : hipEventRecord(e, c->aux);
}

hipError_t HostCopy::finish() {
  hipError_t err = hipSuccess;
  if (!direct && !parts.empty()) {
    // each host thread copies its share of a part as soon as it arrives
    constexpr int kThreads = 4;
    std::atomic<int> bad{0};
    auto work = [&](int t) {
      if (t > 0 && hipSetDevice(c->device) != hipSuccess) {
        bad = 1;
        return;
      }
      for (size_t k = 0; k < parts.size(); ++k) {
        if (hipEventSynchronize(ev[2 * k]) != hipSuccess) {
          bad = 1;
          return;
        }
        const size_t off = parts[k].first, len = parts[k].second;
        const size_t per = (len + kThreads - 1) / kThreads;
        const size_t b = std::min(len, (size_t)t * per);
        const size_t e = std::min(len, b + per);
        if (e > b)
          std::memcpy(static_cast<char*>(dst) + off + b,
                      static_cast<const char*>(stage) + off + b, e - b);
      }
    };
    std::thread th[kThreads - 1];
    for (int t = 1; t < kThreads; ++t) th[t - 1] = std::thread(work, t);
    work(0);
    for (auto& x : th) x.join();
    if (bad) err = hipErrorUnknown;
  }
  // every copy has completed (or failed) before the buffer goes back
  hipError_t e2 = hipStreamSynchronize(st);
  if (err == hipSuccess) err = e2;
  e2 = hipStreamSynchronize(c->aux);
  if (err == hipSuccess) err = e2;
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  ev.clear();
  parts.clear();
  if (stage) host_put(c, stage, cap);
  stage = nullptr;
  return err;
}

HostCopy::~HostCopy() {
  if (stage || !ev.empty()) (void)finish();
}

hipError_t pool_alloc(ds_problem* p, void** ptr, size_t bytes, bool host) {
  size_t cap = 0;
  const hipError_t e = host ? host_get(p->ctx, ptr, bytes, &cap)
                            : ctx_buf_get(p->ctx, ptr, bytes, &cap);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return e;
  }
  p->pooled.push_back({*ptr, cap, host});
  return hipSuccess;
}

void pool_free(ds_problem* p, void* ptr) {
  for (size_t k = 0; k < p->pooled.size(); ++k) {
    if (p->pooled[k].ptr == ptr) {
      const ds_problem::Pooled e = p->pooled[k];
      p->pooled.erase(p->pooled.begin() + (long)k);
      if (e.host) host_put(p->ctx, ptr, e.cap);
      else ctx_buf_put(p->ctx, ptr, e.cap);
      return;
    }
  }
  (void)hipFree(ptr);
}

}  // namespace ds

// ---- static scratch of the no-alloc entry point -------------------------
// Every GPU the library runs on holds one copy of this array until the
// process ends.  In host code its name is only a stand-in, never a device
// address: noalloc_setup() looks the address up once per device (with that
// device current, at context creation) and host code passes that pointer
// to kernels and copies.
static __device__ __align__(256) unsigned char scratch[DS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current.
hipError_t noalloc_setup(int device, unsigned char** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = static_cast<unsigned char*>(ptr);
  }
  *out = g_noalloc_ptr[device];
  return hipSuccess;
}


}  // namespace

extern "C" {

int ds_plan_tiles(int n, const double* pts, int stride,
                  const ds_covar* covar, int* perm_out, int* tile_ptr,
                  int* tile_idx, int idx_cap, int* n_blocks, int* n_idx) {
  if (n <= 0 || stride < 3 || !n_blocks || !n_idx) return DS_E_ARG;
  if (check_covar(nullptr, covar)) return DS_E_ARG;
  if (check_points(nullptr, pts, stride, n, "points")) return DS_E_ARG;
  const CovarDev cv = ds::make_covar(covar);
  double sc[3];
  axis_scale(cv, sc);
  Blocking b;
  ds::build_blocking(pts, stride, n, sc, b);
  std::vector<int> ptr, idx;
  ds::build_tile_lists(b.nblk, b.blo.data(), b.bhi.data(), b, cv, ptr, idx);
  *n_blocks = b.nblk;
  *n_idx = (int)idx.size();
  if (idx_cap <= 0) return DS_OK;
  if (idx_cap < (int)idx.size() || !perm_out || !tile_ptr || !tile_idx)
    return DS_E_ARG;
  std::copy(b.perm.begin(), b.perm.end(), perm_out);
  std::copy(ptr.begin(), ptr.end(), tile_ptr);
  std::copy(idx.begin(), idx.end(), tile_idx);
  return DS_OK;
}

int ds_ctx_create(int device, ds_ctx** out) {
  if (!out) return DS_E_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return DS_E_HIP;
  if (device < 0 || device >= count) return DS_E_ARG;
  std::unique_ptr<ds_ctx> c(new ds_ctx);
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) return DS_E_HIP;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return DS_E_HIP;
  if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess)
    return DS_E_HIP;
  if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount,
                            device) != hipSuccess)
    return DS_E_HIP;
  if (noalloc_setup(device, &c->noalloc) != hipSuccess) return DS_E_HIP;
  const char* poison = std::getenv("DS_POOL_POISON");
  c->pool_poison = poison && poison[0] == '1';
  *out = c.release();
  return DS_OK;
}

int ds_ctx_destroy(ds_ctx* ctx) {
  if (!ctx) return DS_E_ARG;
  ds::ctx_release(ctx);
  return DS_OK;
}

const char* ds_ctx_last_error(const ds_ctx* ctx) {
  return ctx ? ctx->err : "NULL context";
}

int ds_ctx_synchronize(ds_ctx* ctx) {
  if (!ctx) return DS_E_ARG;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->aux));
  return DS_OK;
}

int ds_ctx_num_cu(const ds_ctx* ctx) { return ctx ? ctx->num_cu : 0; }

int ds_ctx_set_host_geometry(ds_ctx* ctx, int on) {
  if (!ctx) return DS_E_ARG;
  ctx->host_geometry = on ? 1 : 0;
  return DS_OK;
}

}  // extern "C"

namespace {

// Build a problem of n_sub independent subvolumes (sub_n[k] pixels each,
// records consecutive in `pixels`, covariance covars[k]).  Each subvolume is
// ordered and culled on its own; its blocks are concatenated after the
// previous subvolume's.
// The host builder of a problem's geometry: the ordering, culled lists,
// per-slot arrays and column-sum slots that build_geometry_device makes,
// made here by the host threads and uploaded (the reference for the device
// builder, selected with ds_ctx_set_host_geometry).  Fills bstat (5 counts
// per block, as the device builder) and the per-subvolume far-point test.
//
// FP32 frames: every target block and every source tile gets the per-axis
// median of its real pixels as origin (an outlier cannot move it).  Sources
// are stored relative to their tile's frame, targets relative to their
// block's frame, and each list entry carries the difference of the two
// frames; for any pair within the cutoff both points are then near the
// source frame, which keeps the FP32 expansion |Xi|^2 + |Xj|^2 - 2 Xi.Xj
// well conditioned.
int geometry_host(ds_problem* p, const double* pixels,
                  std::vector<float>& max_src_out, std::vector<int>& bstat) {
  ds_ctx* ctx = p->ctx;
  const int n_sub = p->n_sub, form = p->form;
  std::vector<Blocking> blks(n_sub);
  std::vector<std::vector<int>> ptrs(n_sub), idxs(n_sub);
  std::vector<std::vector<unsigned char>> msks(n_sub);
  auto build = [&](int k) {
    const ds::CovarDev& cv = p->sub_cv_h[k];
    double sc[3];
    axis_scale(cv, sc);
    const double* pk = pixels + 5 * (size_t)p->sub_off[k];
    ds::build_blocking(pk, 5, p->sub_n[k], sc, blks[k]);
    // a block's 16-row sub-blocks are its own tiles (targets = sources)
    ds::build_tile_lists(blks[k].nblk, blks[k].blo.data(), blks[k].bhi.data(),
                         blks[k], cv, ptrs[k], idxs[k], blks[k].tlo.data(),
                         blks[k].thi.data(), &msks[k]);
  };
  if (n_sub == 1) {
    build(0);
  } else {
    // subvolumes in parallel (each ordering is itself serial)
    std::vector<std::thread> th;
    const int nt = std::max(1, std::min(n_sub, 16));
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&, t]() {
        for (int k = t; k < n_sub; k += nt) build(k);
      });
    for (auto& x : th) x.join();
  }
  const int nblk = p->nblk, npad = p->npad, ntile = p->ntile;
  // concatenated tile lists
  std::vector<int> ptr(nblk + 1, 0), idx;
  for (int k = 0; k < n_sub; ++k) {
    const int B0 = p->sub_blk0_h[k];
    for (int I = 0; I < blks[k].nblk; ++I)
      ptr[B0 + I + 1] = ptr[B0 + I] + (ptrs[k][I + 1] - ptrs[k][I]);
  }
  idx.resize(ptr[nblk]);
  std::vector<unsigned char> msk(ptr[nblk]);
  for (int k = 0; k < n_sub; ++k) {
    const int B0 = p->sub_blk0_h[k];
    for (int I = 0; I < blks[k].nblk; ++I)
      for (int e = ptrs[k][I]; e < ptrs[k][I + 1]; ++e) {
        idx[ptr[B0 + I] + (e - ptrs[k][I])] = idxs[k][e] + B0 * DS_TPB;
        msk[ptr[B0 + I] + (e - ptrs[k][I])] = msks[k][e];
      }
  }
  bstat.assign(5 * (size_t)nblk, 0);
  for (int B = 0; B < nblk; ++B) {
    int* c = &bstat[5 * (size_t)B];
    c[0] = ptr[B + 1] - ptr[B];
    for (int e = ptr[B]; e < ptr[B + 1]; ++e) {
      c[2] += __builtin_popcount(msk[e]);
      c[3] += ((msk[e] & 3u) != 0) + ((msk[e] & 12u) != 0);
    }
  }
  // per-slot, per-block and per-tile host arrays
  std::vector<double4> hpos(npad);
  std::vector<double> hrhs(npad, 0.0), hdiag(npad, 1.0), hw(npad, 0.0);
  std::vector<float4> hloc(npad), htgt(npad);
  std::vector<int> hblk_sub(nblk), hnreal(nblk), htcount(ntile);
  std::vector<double> tlo(3 * (size_t)ntile), thi(3 * (size_t)ntile);
  p->perm.assign(npad, -1);
  std::vector<double> cen(3 * (size_t)nblk), tcen(3 * (size_t)ntile);
  std::vector<double> max_src(n_sub, 0.0);
  for (int k = 0; k < n_sub; ++k) {
    const Blocking& b = blks[k];
    const ds::CovarDev& cv = p->sub_cv_h[k];
    const int B0 = p->sub_blk0_h[k];
    const int off = p->sub_off[k];
    const double* pk = pixels + 5 * (size_t)off;
    const double var_s = cv.var_s;
    auto median_frame = [&](int k0, int k1, double* c) {
      std::vector<double> v(k1 - k0);
      for (int a = 0; a < 3; ++a) {
        for (int q = k0; q < k1; ++q) v[q - k0] = pk[5 * (size_t)b.perm[q] + a];
        std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
        c[a] = v[v.size() / 2];
      }
    };
    for (int B = 0; B < b.nblk; ++B) {
      hblk_sub[B0 + B] = k;
      hnreal[B0 + B] = std::min(DS_BS, b.n - B * DS_BS);
      median_frame(B * DS_BS, std::min(b.n, (B + 1) * DS_BS),
                   &cen[3 * (size_t)(B0 + B)]);
    }
    for (int T = 0; T < b.ntile; ++T) {
      const int TG = B0 * DS_TPB + T;
      htcount[TG] = b.tcount[T];
      for (int a = 0; a < 3; ++a) {
        tlo[3 * (size_t)TG + a] = b.tlo[3 * (size_t)T + a];
        thi[3 * (size_t)TG + a] = b.thi[3 * (size_t)T + a];
      }
      if (b.tcount[T] > 0)
        median_frame(T * DS_TS, T * DS_TS + b.tcount[T], &tcen[3 * (size_t)TG]);
      else
        for (int a = 0; a < 3; ++a)
          tcen[3 * (size_t)TG + a] = cen[3 * (size_t)(B0 + T / DS_TPB) + a];
    }
    const double s3[3] = {cv.s_perp, cv.s_perp, cv.s_para};
    for (int q = 0; q < b.npad; ++q) {
      const size_t g = (size_t)B0 * DS_BS + q;
      const int j = b.perm[q];
      const double* Pj = pk + 5 * (size_t)j;
      const bool real = q < b.n;
      if (real) p->perm[g] = off + j;
      const double wN = real ? Pj[3] : (form == DS_FORM_W ? 0.0 : 1.0);
      hpos[g] = make_double4(Pj[0], Pj[1], Pj[2], wN);
      if (real) {
        hrhs[g] = Pj[4];
        // lib/wf_vector.cc:190-192 (W) and :44-47 (SN)
        hdiag[g] = form == DS_FORM_W ? wN * wN * var_s + 1.0 : var_s + wN;
        hw[g] = form == DS_FORM_W ? wN : 1.0;
      }
      const size_t B = (size_t)B0 + q / DS_BS, T = (size_t)B0 * DS_TPB + q / DS_TS;
      float l[3], t[3];
      for (int a = 0; a < 3; ++a) {
        l[a] = (float)((Pj[a] - tcen[3 * T + a]) * s3[a]);
        t[a] = (float)((Pj[a] - cen[3 * B + a]) * s3[a]);
        if (real) max_src[k] = std::max(max_src[k], std::fabs((double)l[a]));
      }
      hloc[g] = make_float4(l[0], l[1], l[2],
                            l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
      htgt[g] = make_float4(t[0], t[1], t[2], 0.0f);
    }
  }
  // per list entry: target block frame minus source tile frame (pre-scaled)
  std::vector<float4> hoff(idx.size());
  for (int I = 0; I < nblk; ++I) {
    const ds::CovarDev& cv = p->sub_cv_h[hblk_sub[I]];
    const double s3[3] = {cv.s_perp, cv.s_perp, cv.s_para};
    for (int e = ptr[I]; e < ptr[I + 1]; ++e) {
      const size_t T = (size_t)idx[e];
      float o[3];
      for (int a = 0; a < 3; ++a)
        o[a] = (float)((cen[3 * (size_t)I + a] - tcen[3 * T + a]) * s3[a]);
      // .w carries the bits of (live sub-block mask | tile << 4), so the
      // MFMA kernel reads an entry in one 16-byte load
      const unsigned bits = (unsigned)msk[e] | ((unsigned)T << 4);
      float wbits;
      std::memcpy(&wbits, &bits, sizeof(float));
      hoff[e] = make_float4(o[0], o[1], o[2], wbits);
    }
  }
  // Half lists of the symmetric MFMA kernel.  The 16-row sub-blocks of a
  // block are its own tiles (tile 4 B + s), and the sub-block test is the
  // tile-pair test, so the pair {P, Q} of tiles is live in P's block's
  // entry for Q exactly when it is live in Q's block's entry for P: keep
  // it once, in the block of the smaller tile (both products from one
  // evaluation), and a tile with itself once (its own product only).
  std::vector<int> hptr(nblk + 1, 0);
  std::vector<float4> hrec;
  std::vector<int> gcnt(ntile + 1, 0), gsrc;   // (entry, tile) pairs
  hrec.reserve(idx.size() / 2 + nblk);
  for (int B = 0; B < nblk; ++B) {
    for (int e = ptr[B]; e < ptr[B + 1]; ++e) {
      const int T = idx[e], dT = T - DS_TPB * B;
      if (dT < 0) continue;
      // sub-blocks s <= dT evaluate, s < dT also give the transposed part
      const unsigned keep = dT >= DS_TPB - 1 ? 15u : (2u << dT) - 1u;
      const unsigned m = msk[e] & keep;
      if (!m) continue;
      const unsigned tr = m & (dT >= DS_TPB ? 15u : (1u << dT) - 1u);
      const unsigned bits = m | ((unsigned)T << 4);
      float wbits;
      std::memcpy(&wbits, &bits, sizeof(float));
      if (tr) {
        gsrc.push_back((int)hrec.size());
        gsrc.push_back(T);
        ++gcnt[T + 1];
      }
      hrec.push_back(make_float4(hoff[e].x, hoff[e].y, hoff[e].z, wbits));
      bstat[5 * (size_t)B + 4] += __builtin_popcount(m);
    }
    hptr[B + 1] = (int)hrec.size();
    bstat[5 * (size_t)B + 1] = hptr[B + 1] - hptr[B];
  }
  std::vector<int> gptr(ntile + 1, 0), gslot(hrec.size(), -1);
  for (int T = 0; T < ntile; ++T) gptr[T + 1] = gptr[T] + gcnt[T + 1];
  {
    // slots in block order within each tile (gsrc is in block order)
    std::vector<int> fill(gptr.begin(), gptr.end() - 1);
    for (size_t k = 0; k < gsrc.size(); k += 2)
      gslot[gsrc[k]] = fill[gsrc[k + 1]]++;
  }
  p->nslot = (long long)(gsrc.size() / 2);
  max_src_out.resize(n_sub);
  for (int k = 0; k < n_sub; ++k) max_src_out[k] = (float)max_src[k];
  hipStream_t st = ctx->stream;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st)
                 : hipSuccess;
  };
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->pos64, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->rhs64, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->diag, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->wvec, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->loc32, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->tgt32, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->perm_dev, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->tptr, ptr.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->tidx, idx.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->tmask, msk.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->toff32, hoff.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->blk_sub, nblk));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->blk_nreal, nblk));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->hptr, nblk + 1));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->hrec, hrec.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->gslot, gslot.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->gptr, ntile + 1));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->sub_blk0, n_sub + 1));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->tcount, ntile));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->tbox_lo, tlo.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->tbox_hi, thi.size()));
  HIP_OR_FAIL(ctx, up(p->pos64, hpos.data(), sizeof(double4) * npad));
  HIP_OR_FAIL(ctx, up(p->rhs64, hrhs.data(), sizeof(double) * npad));
  HIP_OR_FAIL(ctx, up(p->diag, hdiag.data(), sizeof(double) * npad));
  HIP_OR_FAIL(ctx, up(p->wvec, hw.data(), sizeof(double) * npad));
  HIP_OR_FAIL(ctx, up(p->loc32, hloc.data(), sizeof(float4) * npad));
  HIP_OR_FAIL(ctx, up(p->tgt32, htgt.data(), sizeof(float4) * npad));
  HIP_OR_FAIL(ctx, up(p->perm_dev, p->perm.data(), sizeof(int) * npad));
  HIP_OR_FAIL(ctx, up(p->tptr, ptr.data(), sizeof(int) * ptr.size()));
  HIP_OR_FAIL(ctx, up(p->tidx, idx.data(), sizeof(int) * idx.size()));
  HIP_OR_FAIL(ctx, up(p->tmask, msk.data(), msk.size()));
  HIP_OR_FAIL(ctx, up(p->toff32, hoff.data(), sizeof(float4) * hoff.size()));
  HIP_OR_FAIL(ctx, up(p->blk_sub, hblk_sub.data(), sizeof(int) * nblk));
  HIP_OR_FAIL(ctx, up(p->blk_nreal, hnreal.data(), sizeof(int) * nblk));
  HIP_OR_FAIL(ctx, up(p->hptr, hptr.data(), sizeof(int) * (nblk + 1)));
  HIP_OR_FAIL(ctx, up(p->hrec, hrec.data(), sizeof(float4) * hrec.size()));
  HIP_OR_FAIL(ctx, up(p->gslot, gslot.data(), sizeof(int) * gslot.size()));
  HIP_OR_FAIL(ctx, up(p->gptr, gptr.data(), sizeof(int) * (ntile + 1)));
  HIP_OR_FAIL(ctx, up(p->sub_blk0, p->sub_blk0_h.data(),
                      sizeof(int) * (n_sub + 1)));
  HIP_OR_FAIL(ctx, up(p->tcount, htcount.data(), sizeof(int) * ntile));
  HIP_OR_FAIL(ctx, up(p->tbox_lo, tlo.data(), sizeof(double) * tlo.size()));
  HIP_OR_FAIL(ctx, up(p->tbox_hi, thi.data(), sizeof(double) * thi.size()));
  // the host arrays are freed on return
  HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
  return DS_OK;
}

// The device builder's copy of the caller's pixel records, sent from a
// host thread of its own so that it overlaps the host's checks of the same
// records.  The copy only reads them, and nothing on the device uses it
// before join() (the stream keeps the order from there on).
struct PixelUpload {
  ds_ctx* c = nullptr;
  double* d = nullptr;
  size_t cap = 0;
  std::thread th;
  hipError_t err = hipSuccess;
  hipError_t start(ds_ctx* ctx, const double* src, size_t count) {
    c = ctx;
    const hipError_t e =
        ds::ctx_buf_get(ctx, (void**)&d, sizeof(double) * count, &cap);
    if (e != hipSuccess) return e;
    th = std::thread([this, src, count]() {
      err = hipSetDevice(c->device);
      if (err == hipSuccess)
        err = hipMemcpyAsync(d, src, sizeof(double) * count,
                             hipMemcpyHostToDevice, c->stream);
    });
    return hipSuccess;
  }
  hipError_t join() {
    if (th.joinable()) th.join();
    return err;
  }
  // after the stream has finished with the copy
  void release() {
    (void)join();
    if (d) ds::ctx_buf_put(c, d, cap);
    d = nullptr;
  }
  ~PixelUpload() {
    (void)join();
    if (d) {
      (void)hipStreamSynchronize(c->stream);
      release();
    }
  }
};

// One pass over the pixel records on host threads: the finite checks of
// the coordinates (check_points) and of the weights and data, reporting the
// first failing record in index order; each subvolume's coordinate bounds
// (kd_quant: min and max, exact in any order) and |b|, each subvolume's sum
// in index order as lib/wf_vector.cc:105-112 forms it, on one thread.
int scan_pixels(ds_ctx* ctx, const double* pixels, int n_sub,
                const int* sub_n, const std::vector<int>& sub_off,
                const std::vector<ds::CovarDev>& cvs,
                std::vector<ds::KdQuant>& quant, std::vector<double>& norm_b) {
  constexpr int kRun = 32768;
  struct Run {
    int k, b, e;
    double lo[3], hi[3];
    long long bad_x = -1;   // 3 * record + coordinate
    int bad_wd = -1;
  };
  std::vector<Run> runs;
  for (int k = 0; k < n_sub; ++k)
    for (int b = sub_off[k]; b < sub_off[k + 1]; b += kRun)
      runs.push_back({k, b, std::min(sub_off[k + 1], b + kRun), {}, {}});
  norm_b.assign(n_sub, 0.0);
  const int ntask = n_sub + (int)runs.size();
  std::atomic<int> next(0);
  auto work = [&]() {
    for (int t = next++; t < ntask; t = next++) {
      if (t < n_sub) {
        const double* pk = pixels + 5 * (size_t)sub_off[t];
        double sum_b2 = 0.0;
        for (int i = 0; i < sub_n[t]; ++i) {
          const double bi = pk[5 * (size_t)i + 4];
          sum_b2 += bi * bi;
        }
        norm_b[t] = std::sqrt(sum_b2);
        continue;
      }
      Run& r = runs[t - n_sub];
      for (int a = 0; a < 3; ++a) {
        r.lo[a] = std::numeric_limits<double>::infinity();
        r.hi[a] = -r.lo[a];
      }
      for (int i = r.b; i < r.e; ++i) {
        const double* q = pixels + 5 * (size_t)i;
        for (int a = 0; a < 3; ++a) {
          if (r.bad_x < 0 && !std::isfinite(q[a])) r.bad_x = 3LL * i + a;
          r.lo[a] = std::min(r.lo[a], q[a]);
          r.hi[a] = std::max(r.hi[a], q[a]);
        }
        if (r.bad_wd < 0 && !(std::isfinite(q[3]) && std::isfinite(q[4])))
          r.bad_wd = i;
      }
    }
  };
  const long long nt = std::min<long long>(
      16, std::max<long long>(1, std::thread::hardware_concurrency()));
  const int nth = (int)std::min<long long>(nt, ntask);
  std::vector<std::thread> th;
  for (int t = 1; t < nth; ++t) th.emplace_back(work);
  work();
  for (auto& x : th) x.join();
  for (const Run& r : runs)
    if (r.bad_x >= 0)
      return fail(ctx, DS_E_ARG, "pixels: coordinate %d of record %lld is "
                  "not finite", (int)(r.bad_x % 3), r.bad_x / 3);
  for (const Run& r : runs)
    if (r.bad_wd >= 0)
      return fail(ctx, DS_E_ARG, "pixel %d: weight or data not finite",
                  r.bad_wd);
  quant.assign(n_sub, ds::KdQuant());
  size_t ri = 0;
  for (int k = 0; k < n_sub; ++k) {
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<double>::infinity();
      hi[a] = -lo[a];
    }
    for (; ri < runs.size() && runs[ri].k == k; ++ri)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], runs[ri].lo[a]);
        hi[a] = std::max(hi[a], runs[ri].hi[a]);
      }
    double sc[3];
    axis_scale(cvs[k], sc);
    for (int a = 0; a < 3; ++a) ds::kd_quant_axis(lo[a], hi[a], sc[a], quant[k], a);
  }
  return DS_OK;
}

int create_problem(ds_ctx* ctx, int form, int n_sub, const int* sub_n,
                   const double* pixels, const ds_covar* covars,
                   ds_problem** out) {
  if (!ctx || !out || n_sub < 1 || !sub_n || !covars) return DS_E_ARG;
  *out = nullptr;
  if (form != DS_FORM_W && form != DS_FORM_SN)
    return fail(ctx, DS_E_ARG, "form must be DS_FORM_W or DS_FORM_SN");
  long long ntot = 0;
  for (int k = 0; k < n_sub; ++k) {
    if (sub_n[k] <= 0) return fail(ctx, DS_E_ARG, "subvolume %d is empty", k);
    if (int rc = check_covar(ctx, &covars[k])) return rc;
    const ds_covar& c0 = covars[0];
    if (covars[k].table_n != c0.table_n || covars[k].dx_inv != c0.dx_inv ||
        !std::equal(c0.table, c0.table + c0.table_n, covars[k].table))
      return fail(ctx, DS_E_ARG, "subvolume %d: every subvolume must use the "
                  "same covariance table", k);
    ntot += sub_n[k];
  }
  if (ntot > 500000000LL) return fail(ctx, DS_E_ARG, "too many pixels");
  const int n = (int)ntot;
  if (!pixels) return fail(ctx, DS_E_ARG, "pixels: NULL array");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device));
  const double t0 = now_s();
  PixelUpload upload;
  if (!ctx->host_geometry)
    HIP_OR_FAIL(ctx, upload.start(ctx, pixels, 5 * (size_t)n));
  std::unique_ptr<ds_problem, void (*)(ds_problem*)> P(
      new ds_problem, [](ds_problem* q) {
        // a failed creation: its enqueued work ends before the buffers go
        // back to the context's cache
        if (q->ctx) (void)hipStreamSynchronize(q->ctx->stream);
        problem_free(q);
        delete q;
      });
  ds_problem* p = P.get();
  p->ctx = ctx;
  p->form = form;
  p->n = n;
  p->n_sub = n_sub;
  p->cv = ds::make_covar(&covars[0]);
  p->table.assign(covars[0].table, covars[0].table + covars[0].table_n);
  p->sub_n.assign(sub_n, sub_n + n_sub);
  p->sub_off.assign(n_sub + 1, 0);
  for (int k = 0; k < n_sub; ++k) p->sub_off[k + 1] = p->sub_off[k] + sub_n[k];
  p->sub_cv_h.resize(n_sub);
  for (int k = 0; k < n_sub; ++k) p->sub_cv_h[k] = ds::make_covar(&covars[k]);
  // per subvolume: the ordering's quantization and |b| in the reference's
  // order (lib/wf_vector.cc:105-112, :253-260)
  std::vector<ds::KdQuant> quant;
  if (int rc = scan_pixels(ctx, pixels, n_sub, sub_n, p->sub_off,
                           p->sub_cv_h, quant, p->sub_norm_b))
    return rc;
  HIP_OR_FAIL(ctx, upload.join());
  p->norm_b = p->sub_norm_b[0];
  // The neighbour search's first ring radius (square root of table-argument
  // units): the ball that holds about 64 points at subvolume 0's mean
  // density over its bounding box (a flat axis counts as one unit wide).
  {
    const double sd = std::sqrt(p->cv.dx_inv);
    double vol = 1.0;
    for (int ax = 0; ax < 3; ++ax)
      vol *= std::max(1.0, quant[0].w[ax] * 4294967295.0 * sd);
    const double rho = sub_n[0] / vol;
    const double r0 = std::cbrt(64.0 / (rho * (4.0 / 3.0) * M_PI));
    p->knn_r0 = (float)std::min(std::max(r0, 1.0),
                                std::sqrt((double)(p->cv.table_n - 1)));
  }
  // do all subvolumes share one geometry (coordinates and weights)?
  p->shared_geometry = true;
  for (int k = 1; k < n_sub && p->shared_geometry; ++k) {
    if (sub_n[k] != sub_n[0] || covars[k].var_s != covars[0].var_s ||
        covars[k].c_perp != covars[0].c_perp ||
        covars[k].c_para != covars[0].c_para) {
      p->shared_geometry = false;
      break;
    }
    const double* a = pixels + 5 * (size_t)p->sub_off[k];
    for (int i = 0; i < sub_n[0] && p->shared_geometry; ++i)
      for (int c = 0; c < 4; ++c)
        if (a[5 * (size_t)i + c] != pixels[5 * (size_t)i + c]) {
          p->shared_geometry = false;
          break;
        }
  }
  p->sub_blk0_h.assign(n_sub + 1, 0);
  for (int k = 0; k < n_sub; ++k)
    p->sub_blk0_h[k + 1] = p->sub_blk0_h[k] + (sub_n[k] + DS_BS - 1) / DS_BS;
  p->nblk = p->sub_blk0_h[n_sub];
  p->ntile = p->nblk * DS_TPB;
  p->npad = p->nblk * DS_BS;
  const int nblk = p->nblk, npad = p->npad;
  // ordering, culled lists, per-slot arrays and column-sum slots: on the
  // device, or with the host builder (the same arrays)
  std::vector<float> max_src;
  std::vector<int> bstat;
  if (ctx->host_geometry) {
    if (int rc = geometry_host(p, pixels, max_src, bstat)) return rc;
  } else {
    int gerr = 0;
    HIP_OR_FAIL(ctx, ds::build_geometry_device(p, upload.d, quant, max_src,
                                               bstat, &gerr));
    upload.release();
    if (gerr == 1)
      return fail(ctx, DS_E_ARG, "more than INT_MAX tile-list entries");
    if (gerr)
      return fail(ctx, DS_E_STATE, "device geometry builder: inconsistent "
                  "result (code %d)", gerr);
  }
  // pair counts from the per-block counts {list, half list, live
  // sub-block entries, live 32-row halves, evaluated sub-block entries}
  std::vector<int> hblk_sub(nblk);
  for (int k = 0; k < n_sub; ++k)
    for (int B = p->sub_blk0_h[k]; B < p->sub_blk0_h[k + 1]; ++B)
      hblk_sub[B] = k;
  p->blk_len_h.resize(nblk);
  p->ntile_pairs = p->nsub_pairs = p->nhalf_pairs = p->nsub_eval = 0;
  p->sub_pairs64.assign(n_sub, 0);
  p->sub_pairs32.assign(n_sub, 0);
  p->sub_half64.assign(n_sub, 0);
  for (int B = 0; B < nblk; ++B) {
    const int* c = &bstat[5 * (size_t)B];
    const int k = hblk_sub[B];
    p->blk_len_h[B] = c[0];
    p->ntile_pairs += c[0];
    p->nsub_pairs += c[2];
    p->nhalf_pairs += c[3];
    p->nsub_eval += c[4];
    p->sub_pairs64[k] += c[0];
    p->sub_pairs32[k] += c[2];
    p->sub_half64[k] += c[3];
  }
  for (int k = 0; k < n_sub; ++k) {
    p->sub_pairs64[k] *= DS_BS * DS_TS;
    p->sub_pairs32[k] *= DS_TS * DS_TS;
    p->sub_half64[k] *= 32 * DS_TS;
  }
  // the CG update kernel addresses the slots by 32-bit byte offsets
  if ((size_t)p->nslot > (size_t)UINT32_MAX / (16 * sizeof(float)))
    return fail(ctx, DS_E_ARG, "%lld column-sum slots exceed the 32-bit "
                "offsets of the update kernel", (long long)p->nslot);
  // launch order: longest tile list first (for the autotuner), grouped by
  // XCD (xcd_order)
  std::vector<int> order(nblk);
  std::iota(order.begin(), order.end(), 0);
  order = xcd_order(order, [&](int I) { return p->blk_len_h[I]; });
  // FP32 launch lists: longest list first and block order, each split into
  // the blocks of subvolumes without far points (lengths of their half
  // lists, which the MFMA kernel runs), then the others.  A tile holding a
  // point far from its median (isolated outliers) would make the expansion
  // lose its accuracy: that subvolume's blocks use direct differences
  // (decided per subvolume, so a batch runs each subvolume's blocks with
  // the kernel it would get alone).
  std::vector<int> ord32(2 * (size_t)nblk);
  {
    auto direct = [&](int I) { return max_src[hblk_sub[I]] > 16.0f; };
    int nf = 0;
    for (int I : order) nf += !direct(I);
    p->n32_fast = nf;
    // the MFMA kernel runs the half lists: longest half list first; each
    // part is launched on its own, so each is grouped by XCD on its own
    std::vector<int> fast, slow;
    for (int I = 0; I < nblk; ++I) (direct(I) ? slow : fast).push_back(I);
    const std::vector<int> f0 = xcd_order(fast, [&](int I) {
      return bstat[5 * (size_t)I + 1];
    });
    const std::vector<int> s0 = xcd_order(slow, [&](int I) {
      return p->blk_len_h[I];
    });
    std::copy(f0.begin(), f0.end(), ord32.begin());
    std::copy(s0.begin(), s0.end(), ord32.begin() + nf);
    int f = 0, d = nf;
    for (int k = 0; k < nblk; ++k) {
      int* dst = ord32.data() + nblk;
      if (direct(k)) dst[d++] = k; else dst[f++] = k;
    }
  }
  std::vector<double4> hsubcv(n_sub);
  for (int k = 0; k < n_sub; ++k) {
    const ds::CovarDev& cv = p->sub_cv_h[k];
    hsubcv[k] = make_double4(cv.var_s, cv.c_perp, cv.c_para, 0.0);
  }
  std::vector<double2> t64;
  std::vector<float2> t32;
  make_tables(&covars[0], t64, t32);
  // The FP32 MFMA kernel stores each interval's interpolant
  // f_k + df_k (xx - k) as the line g_k + df_k xx (g_k = f_k - k df_k,
  // formed in FP64), one fma per pair with no fraction.
  std::vector<float2> t32g(t64.size());
  for (size_t k = 0; k < t64.size(); ++k)
    t32g[k] = make_float2((float)(t64[k].x - (double)k * t64[k].y),
                          (float)t64[k].y);

  hipStream_t st = ctx->stream;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st)
                 : hipSuccess;
  };
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->table64, t64.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->table32, t32.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->table32g, t32g.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->table_raw, p->table.size()));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->blk_order, nblk));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->ord32, 2 * (size_t)nblk));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->pbuf, std::max<size_t>(16 * (size_t)p->nslot,
                                                        1)));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->ydir, npad));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->sub_sym, n_sub));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->sub_cv, n_sub));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->done_count, 1));
  double** vecs[] = {&p->x, &p->r, &p->d, &p->s, &p->q, &p->tmp};
  for (double** v : vecs) HIP_OR_FAIL(ctx, ds::palloc(p, v, npad));
  if (form == DS_FORM_W) {
    HIP_OR_FAIL(ctx, ds::palloc(p, &p->xu, npad));
    HIP_OR_FAIL(ctx, ds::palloc(p, &p->du, npad));
  } else {
    p->xu = p->x;
    p->du = p->d;
  }
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->du32, npad));
  // two partial vectors, then the d . q partials of a single volume's
  // mixed iterations (summed by the update kernel)
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->partials, 3 * (size_t)nblk));
  HIP_OR_FAIL(ctx, ds::palloc(p, &p->sc, n_sub));
  HIP_OR_FAIL(ctx, ds::pool_alloc(p, (void**)&p->sc_host,
                                  sizeof(ds_cg_scalars) * n_sub, true));
  HIP_OR_FAIL(ctx, ds::pool_alloc(p, (void**)&p->done_host, 2 * sizeof(int),
                                  true));
  HIP_OR_FAIL(ctx, ds::pool_alloc(p, (void**)&p->norm_host,
                                  2 * sizeof(double), true));
  HIP_OR_FAIL(ctx, ds::pool_alloc(p, (void**)&p->host_tmp,
                                  sizeof(double) * npad, true));
  std::memset(p->sc_host, 0, sizeof(ds_cg_scalars) * n_sub);
  HIP_OR_FAIL(ctx, up(p->table64, t64.data(), sizeof(double2) * t64.size()));
  HIP_OR_FAIL(ctx, up(p->table32, t32.data(), sizeof(float2) * t32.size()));
  HIP_OR_FAIL(ctx, up(p->table32g, t32g.data(), sizeof(float2) * t32g.size()));
  HIP_OR_FAIL(ctx, up(p->table_raw, p->table.data(),
                      sizeof(double) * p->table.size()));
  HIP_OR_FAIL(ctx, up(p->blk_order, order.data(), sizeof(int) * nblk));
  HIP_OR_FAIL(ctx, up(p->ord32, ord32.data(), sizeof(int) * 2 * (size_t)nblk));
  {
    std::vector<int> sym(n_sub);
    for (int k = 0; k < n_sub; ++k) sym[k] = max_src[k] > 16.0f ? 0 : 1;
    HIP_OR_FAIL(ctx, up(p->sub_sym, sym.data(), sizeof(int) * n_sub));
  }
  HIP_OR_FAIL(ctx, up(p->sub_cv, hsubcv.data(), sizeof(double4) * n_sub));
  HIP_OR_FAIL(ctx, ds::launch_fill(p->x, 0.0, npad, st));
  HIP_OR_FAIL(ctx, up(p->sc, p->sc_host, sizeof(ds_cg_scalars) * n_sub));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
  p->setup_seconds = now_s() - t0;
  ds::ctx_retain(ctx);
  *out = P.release();
  return DS_OK;
}

}  // namespace

extern "C" {

int ds_problem_create(ds_ctx* ctx, int form, int n, const double* pixels,
                      const ds_covar* covar, ds_problem** out) {
  if (!ctx || !out) return DS_E_ARG;
  if (n <= 0) return fail(ctx, DS_E_ARG, "n must be positive");
  if (int rc = check_covar(ctx, covar)) return rc;
  return create_problem(ctx, form, 1, &n, pixels, covar, out);
}

int ds_batch_create(ds_ctx* ctx, int form, int n_sub, const int* sub_n,
                    const double* pixels, const ds_covar* covars,
                    ds_problem** out) {
  if (!ctx || !out) return DS_E_ARG;
  return create_problem(ctx, form, n_sub, sub_n, pixels, covars, out);
}

int ds_problem_destroy(ds_problem* p) {
  if (!p) return DS_E_ARG;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  ds_ctx* ctx = p->ctx;
  if (p->minv_ev) {
    (void)hipEventSynchronize(p->minv_ev);   // an inverse build in flight
    (void)hipEventDestroy(p->minv_ev);
    p->minv_ev = nullptr;
  }
  problem_free(p);
  delete p;
  ds::ctx_release(ctx);
  return DS_OK;
}

int ds_problem_get_stats(const ds_problem* p, ds_problem_stats* out) {
  if (!p || !out) return DS_E_ARG;
  out->n = p->n;
  out->n_blocks = p->nblk;
  out->n_tiles = p->ntile;
  out->tile_pairs = p->ntile_pairs;
  out->pairs_per_matvec = p->ntile_pairs * (long long)(DS_BS * DS_TS);
  out->sub_pairs_per_matvec = p->nsub_pairs * (long long)(DS_TS * DS_TS);
  out->setup_seconds = p->setup_seconds;
  out->sub_pairs_evaluated = p->nsub_eval * (long long)(DS_TS * DS_TS);
  out->f64_pairs_per_matvec =
      p->f64_rows == 16 ? out->sub_pairs_per_matvec
      : p->f64_rows == 32 ? p->nhalf_pairs * (long long)(32 * DS_TS)
                          : out->pairs_per_matvec;
  return DS_OK;
}

int ds_problem_get_geometry(const ds_problem* p, int which, void* out,
                            long long cap_bytes, long long* bytes) {
  if (!p || !bytes) return DS_E_ARG;
  const size_t npad = p->npad, nblk = p->nblk, ntile = p->ntile;
  int tot[2] = {0, 0};   // list entries, half-list records
  HIP_OR_FAIL(p->ctx, hipMemcpy(&tot[0], p->tptr + nblk, sizeof(int),
                                hipMemcpyDeviceToHost));
  HIP_OR_FAIL(p->ctx, hipMemcpy(&tot[1], p->hptr + nblk, sizeof(int),
                                hipMemcpyDeviceToHost));
  const void* src = nullptr;
  size_t nb = 0;
  switch (which) {
    case DS_GEOM_PERM: src = p->perm_dev; nb = 4 * npad; break;
    case DS_GEOM_POS: src = p->pos64; nb = 32 * npad; break;
    case DS_GEOM_RHS: src = p->rhs64; nb = 8 * npad; break;
    case DS_GEOM_DIAG: src = p->diag; nb = 8 * npad; break;
    case DS_GEOM_WVEC: src = p->wvec; nb = 8 * npad; break;
    case DS_GEOM_LOC32: src = p->loc32; nb = 16 * npad; break;
    case DS_GEOM_TGT32: src = p->tgt32; nb = 16 * npad; break;
    case DS_GEOM_TCOUNT: src = p->tcount; nb = 4 * ntile; break;
    case DS_GEOM_TBOX_LO: src = p->tbox_lo; nb = 24 * ntile; break;
    case DS_GEOM_TBOX_HI: src = p->tbox_hi; nb = 24 * ntile; break;
    case DS_GEOM_TPTR: src = p->tptr; nb = 4 * (nblk + 1); break;
    case DS_GEOM_TIDX: src = p->tidx; nb = 4 * (size_t)tot[0]; break;
    case DS_GEOM_TMASK: src = p->tmask; nb = (size_t)tot[0]; break;
    case DS_GEOM_TOFF32: src = p->toff32; nb = 16 * (size_t)tot[0]; break;
    case DS_GEOM_HPTR: src = p->hptr; nb = 4 * (nblk + 1); break;
    case DS_GEOM_HREC: src = p->hrec; nb = 16 * (size_t)tot[1]; break;
    case DS_GEOM_GPTR: src = p->gptr; nb = 4 * (ntile + 1); break;
    case DS_GEOM_GSLOT: src = p->gslot; nb = 4 * (size_t)tot[1]; break;
    case DS_GEOM_BLK_ORDER: src = p->blk_order; nb = 4 * nblk; break;
    case DS_GEOM_ORD32: src = p->ord32; nb = 8 * nblk; break;
    case DS_GEOM_MINV:
      if (!p->use_minv())
        return fail(p->ctx, DS_E_STATE, "block-Jacobi preconditioner not "
                    "selected");
      if (p->minv_ev) HIP_OR_FAIL(p->ctx, hipEventSynchronize(p->minv_ev));
      src = p->minv;
      nb = p->minv_bytes;
      break;
    case DS_GEOM_KNN:
      if (!p->knn_nbr)
        return fail(p->ctx, DS_E_STATE, "no neighbour lists yet (they are "
                    "built by the first warm-start prediction)");
      src = p->knn_nbr;
      nb = 4 * 32 * npad;
      break;
    default:
      return fail(p->ctx, DS_E_ARG, "unknown geometry array %d", which);
  }
  *bytes = (long long)nb;
  if (!out) return DS_OK;
  if (cap_bytes < (long long)nb)
    return fail(p->ctx, DS_E_ARG, "geometry array %d needs %zu bytes", which,
                nb);
  HIP_OR_FAIL(p->ctx, hipStreamSynchronize(p->ctx->stream));
  if (nb)
    HIP_OR_FAIL(p->ctx, hipMemcpy(out, src, nb, hipMemcpyDeviceToHost));
  return DS_OK;
}

int ds_problem_set_waves(ds_problem* p, int waves) {
  if (!p) return DS_E_ARG;
  if (waves != 1 && waves != 2 && waves != 4)
    return fail(p->ctx, DS_E_ARG, "waves_per_block must be 1, 2 or 4");
  p->waves_per_block = waves;
  return DS_OK;
}

int ds_A_product(ds_problem* p, int prec, const double* a, double* b) {
  if (!p || !a || !b) return DS_E_ARG;
  if (prec != DS_PREC_F64 && prec != DS_PREC_F32)
    return fail(p->ctx, DS_E_ARG, "unknown precision %d", prec);
  if (int rc = upload_sorted(p, a, p->tmp)) return rc;
  hipStream_t st = p->ctx->stream;
  const double* w = p->form == DS_FORM_W ? p->wvec : nullptr;
  if (prec == DS_PREC_F64) {
    HIP_OR_FAIL(p->ctx, ds::launch_scale(p->tmp, w, p->r, nullptr, p->npad,
                                         st));
    HIP_OR_FAIL(p->ctx, ds::launch_matvec64_ex(p, 0, p->tmp, p->r, p->q,
                                               ds::CgEpilogue(), st));
  } else {
    HIP_OR_FAIL(p->ctx, ds::launch_scale(p->tmp, w, nullptr, p->du32,
                                         p->npad, st));
    HIP_OR_FAIL(p->ctx, ds::launch_matvec32_ex(p, p->du32, p->tmp, p->q, 0,
                                               st));
  }
  return download_sorted(p, p->q, b);
}

int ds_A_residual(ds_problem* p, const double* x, double* r) {
  if (!p || !x || !r) return DS_E_ARG;
  if (int rc = upload_sorted(p, x, p->tmp)) return rc;
  hipStream_t st = p->ctx->stream;
  const double* w = p->form == DS_FORM_W ? p->wvec : nullptr;
  HIP_OR_FAIL(p->ctx, ds::launch_scale(p->tmp, w, p->r, nullptr, p->npad, st));
  if (p->shared_geometry && p->n_sub > 1 && (p->use_fused || p->use_dense)) {
    // the multi-RHS residual pass the mixed solver uses on such batches;
    // it acts on the subvolumes that want a true residual: all of them
    std::vector<ds_cg_scalars> h(p->n_sub);
    std::memset(h.data(), 0, sizeof(ds_cg_scalars) * p->n_sub);
    for (auto& c : h) c.want_true = 1;
    HIP_OR_FAIL(p->ctx, hipMemcpyAsync(p->sc, h.data(),
                                       sizeof(ds_cg_scalars) * p->n_sub,
                                       hipMemcpyHostToDevice, st));
    HIP_OR_FAIL(p->ctx, ds::launch_res64_mrhs(p, p->tmp, p->r, p->q, 1, st));
  } else {
    HIP_OR_FAIL(p->ctx, ds::launch_matvec64_ex(p, 1, p->tmp, p->r, p->q,
                                               ds::CgEpilogue(), st));
  }
  return download_sorted(p, p->q, r);
}

int ds_precond(ds_problem* p, const double* b, double* x) {
  if (!p || !b || !x) return DS_E_ARG;
  if (int rc = upload_sorted(p, b, p->tmp)) return rc;
  HIP_OR_FAIL(p->ctx, ds::launch_precond_p(p, p->tmp, p->q, p->ctx->stream));
  return download_sorted(p, p->q, x);
}

int ds_precond_r32(ds_problem* p, const double* b, double* x) {
  if (!p || !b || !x) return DS_E_ARG;
  if (!p->use_minv())
    return fail(p->ctx, DS_E_STATE, "block-Jacobi preconditioner not "
                "selected");
  if (int rc = upload_sorted(p, b, p->tmp)) return rc;
  HIP_OR_FAIL(p->ctx, ds::launch_precond_r32(p, p->tmp, p->q,
                                             p->ctx->stream));
  return download_sorted(p, p->q, x);
}

int ds_norm_b(ds_problem* p, double* out) {
  if (!p || !out) return DS_E_ARG;
  *out = p->norm_b;
  return DS_OK;
}

int ds_problem_set_x(ds_problem* p, const double* x) {
  if (!p || !x) return DS_E_ARG;
  return upload_sorted(p, x, p->x);
}

int ds_problem_get_x(ds_problem* p, double* x) {
  if (!p || !x) return DS_E_ARG;
  return download_sorted(p, p->x, x);
}

int ds_solve_pcg_resident(ds_problem* p, const ds_pcg_params* prm,
                          ds_pcg_result* res) {
  if (!p || !res) return DS_E_ARG;
  if (int rc = check_pcg_params(p, prm)) return rc;
  if (p->n_sub != 1)
    return fail(p->ctx, DS_E_STATE, "batch problem: use ds_batch_solve");
  HIP_OR_FAIL(p->ctx, ds::run_solver(p, prm, 0, 0.0, res, 0));
  return DS_OK;
}

int ds_solve_mixed_resident(ds_problem* p, const ds_pcg_params* prm,
                            double drop, ds_pcg_result* res) {
  if (!p || !res) return DS_E_ARG;
  if (int rc = check_pcg_params(p, prm)) return rc;
  if (!(drop > 0.0 && drop < 1.0))
    return fail(p->ctx, DS_E_ARG, "refine_drop must lie in (0, 1)");
  if (p->n_sub != 1)
    return fail(p->ctx, DS_E_STATE, "batch problem: use ds_batch_solve");
  // Plain launches: a single volume's iterations keep the GPU busy while
  // the host enqueues, and capturing plus instantiating the chunk's graph
  // on every call cost more than its replay saved (7.78 vs 7.38 ms at 300k
  // pixels).  ds_batch_solve keeps the choice.
  HIP_OR_FAIL(p->ctx, ds::run_solver(p, prm, 1, drop, res, 0));
  return DS_OK;
}

int ds_problem_get_norms(ds_problem* p, double* norms, int cap, int* count) {
  if (!p || (!norms && cap > 0)) return DS_E_ARG;
  const int nn = p->sc_host ? p->sc_host[0].n_norms : 0;
  const int m = std::min(nn, std::max(cap, 0));
  if (m > 0 && p->norms)
    HIP_OR_FAIL(p->ctx, hipMemcpy(norms, p->norms, sizeof(double) * m,
                                  hipMemcpyDeviceToHost));
  if (count) *count = nn;
  return DS_OK;
}

int ds_problem_get_deltas(ds_problem* p, double* deltas, int cap,
                          int* count) {
  if (!p || (!deltas && cap > 0)) return DS_E_ARG;
  const int nn = p->sc_host ? p->sc_host[0].n_norms : 0;
  const int m = std::min(nn, std::max(cap, 0));
  if (m > 0 && p->deltas)
    HIP_OR_FAIL(p->ctx, hipMemcpy(deltas, p->deltas, sizeof(double) * m,
                                  hipMemcpyDeviceToHost));
  if (count) *count = nn;
  return DS_OK;
}

static int solve_common(ds_problem* p, const ds_pcg_params* prm, int mixed,
                        double drop, double* x, double* norms, int cap,
                        ds_pcg_result* res) {
  if (!p || !x || !res) return DS_E_ARG;
  if (int rc = ds_problem_set_x(p, x)) return rc;
  const int rc = mixed ? ds_solve_mixed_resident(p, prm, drop, res)
                       : ds_solve_pcg_resident(p, prm, res);
  if (rc) return rc;
  if (int rc2 = ds_problem_get_x(p, x)) return rc2;
  if (norms && cap > 0) return ds_problem_get_norms(p, norms, cap, nullptr);
  return DS_OK;
}

int ds_solve_pcg(ds_problem* p, const ds_pcg_params* prm, double* x,
                 double* norms, int cap, ds_pcg_result* res) {
  return solve_common(p, prm, 0, 0.0, x, norms, cap, res);
}

int ds_solve_mixed(ds_problem* p, const ds_pcg_params* prm, double drop,
                   double* x, double* norms, int cap, ds_pcg_result* res) {
  return solve_common(p, prm, 1, drop, x, norms, cap, res);
}

int ds_problem_smp_grid(ds_problem* p, double lx, double ly, double lz,
                        int nx, int ny, int nz, double* m) {
  if (!p || !m) return DS_E_ARG;
  ds::MapGrid g;
  if (int rc = check_grid(p->ctx, lx, ly, lz, nx, ny, nz, g)) return rc;
  if (p->n_sub != 1)
    return fail(p->ctx, DS_E_STATE, "map projection needs a single volume");
  return map_grid_run(p->ctx, ds::MM_SMP, g, p->ntile, p->tbox_lo,
                      p->tbox_hi, nullptr, p->pos64, p->x, p->table64, p->cv,
                      m);
}

// ds_problem_smp_grid without device allocations: the source-tile boxes,
// the map tiles' list counts, offsets and entries and the map itself live
// in the static scratch array (layout below, 256-byte aligned parts); the
// offsets are formed on the device, so the host allocates nothing either.
// The lists and the projection are ds_problem_smp_grid's (same kernels,
// same lists), so the maps are equal bit for bit.
int ds_problem_smp_grid_noalloc(ds_problem* p, double lx, double ly,
                                double lz, int nx, int ny, int nz,
                                double* m) {
  if (!p || !m) return DS_E_ARG;
  ds_ctx* ctx = p->ctx;
  ds::MapGrid g;
  if (int rc = check_grid(ctx, lx, ly, lz, nx, ny, nz, g)) return rc;
  if (p->n_sub != 1)
    return fail(ctx, DS_E_STATE, "map projection needs a single volume");
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device));
  if (!ctx->noalloc)
    return fail(ctx, DS_E_STATE, "no-alloc scratch not set up");
  const size_t nt = (size_t)g.tx * g.ty * g.tz;
  const size_t n_map = (size_t)g.nx * g.ny * g.nz;
  auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_cnt = 0;
  const size_t o_ptr = align(o_cnt + sizeof(int) * nt);
  const size_t o_map = align(o_ptr + sizeof(int) * (nt + 1));
  const size_t o_idx = align(o_map + sizeof(double) * n_map);
  if (o_idx >= (size_t)DS_NOALLOC_SCRATCH_BYTES)
    return fail(ctx, DS_E_ARG,
                "map grid of %zu cells needs %zu bytes before its lists, "
                "more than the %u-byte no-alloc scratch",
                n_map, o_idx, (unsigned)DS_NOALLOC_SCRATCH_BYTES);
  const long long cap = std::min<long long>(
      ((size_t)DS_NOALLOC_SCRATCH_BYTES - o_idx) / sizeof(int), INT_MAX);
  unsigned char* base = ctx->noalloc;
  const double* dlo = p->tbox_lo;
  const double* dhi = p->tbox_hi;
  int* dcnt = reinterpret_cast<int*>(base + o_cnt);
  int* dptr = reinterpret_cast<int*>(base + o_ptr);
  double* dmap = reinterpret_cast<double*>(base + o_map);
  int* didx = reinterpret_cast<int*>(base + o_idx);
  hipStream_t st = ctx->stream;
  HIP_OR_FAIL(ctx, ds::launch_grid_lists(g, dlo, dhi, p->ntile, p->cv, dcnt,
                                         nullptr, nullptr, st));
  hipLaunchKernelGGL(k_list_offsets, dim3(1), dim3(1024), 0, st, dcnt,
                     (int)nt, cap, dptr);
  HIP_OR_FAIL(ctx, hipGetLastError());
  int total = 0;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(&total, dptr + nt, sizeof(int),
                                  hipMemcpyDeviceToHost, st));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
  if (total < 0)
    return fail(ctx, DS_E_ARG,
                "map grid lists need more than the %lld entries the no-alloc "
                "scratch holds beside this %zu-cell map", cap, n_map);
  HIP_OR_FAIL(ctx, ds::launch_grid_lists(g, dlo, dhi, p->ntile, p->cv,
                                         nullptr, dptr, didx, st));
  HIP_OR_FAIL(ctx, ds::launch_map64(ds::MM_SMP, g, p->pos64, p->x, dptr, didx,
                                    p->table64, p->cv, dmap, st));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(m, dmap, sizeof(double) * n_map,
                                  hipMemcpyDeviceToHost, st));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
  return DS_OK;
}

int ds_smp_product(ds_ctx* ctx, int n_pix, const double* pix_pts, int n_map,
                   const double* map_pts, const ds_covar* covar,
                   const double* x, double* m) {
  if (!ctx || (n_map > 0 && !m) || (n_pix > 0 && !x)) return DS_E_ARG;
  if (int rc = check_covar(ctx, covar)) return rc;
  if (int rc = check_points(ctx, pix_pts, 3, n_pix, "pixel points")) return rc;
  if (int rc = check_points(ctx, map_pts, 3, n_map, "map points")) return rc;
  if (n_map == 0) return DS_OK;
  if (n_pix == 0) {
    std::fill(m, m + n_map, 0.0);
    return DS_OK;
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device));
  const CovarDev cv = ds::make_covar(covar);
  SrcSet src;
  if (int rc = make_srcset(ctx, pix_pts, 3, n_pix, x, 0, cv, src)) return rc;
  DevBuf<double2> t64;
  if (int rc = upload_table64(ctx, covar, t64)) return rc;
  return map_points_run(ctx, ds::MM_SMP, n_map, map_pts, src, t64.p, cv, m);
}

int ds_smp_product_grid(ds_ctx* ctx, int n_pix, const double* pix_pts,
                        double lx, double ly, double lz, int nx, int ny,
                        int nz, const ds_covar* covar, const double* x,
                        double* m) {
  if (!ctx || !m || (n_pix > 0 && !x)) return DS_E_ARG;
  if (int rc = check_covar(ctx, covar)) return rc;
  if (int rc = check_points(ctx, pix_pts, 3, n_pix, "pixel points")) return rc;
  ds::MapGrid g;
  if (int rc = check_grid(ctx, lx, ly, lz, nx, ny, nz, g)) return rc;
  if (n_pix == 0) {
    std::fill(m, m + (size_t)nx * ny * nz, 0.0);
    return DS_OK;
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device));
  const CovarDev cv = ds::make_covar(covar);
  SrcSet src;
  if (int rc = make_srcset(ctx, pix_pts, 3, n_pix, x, 0, cv, src)) return rc;
  DevBuf<double2> t64;
  if (int rc = upload_table64(ctx, covar, t64)) return rc;
  DevBuf<double> dlo, dhi;
  const size_t nb = 3 * (size_t)src.blk.ntile;
  HIP_OR_FAIL(ctx, dlo.alloc(nb));
  HIP_OR_FAIL(ctx, dhi.alloc(nb));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(dlo.p, src.blk.tlo.data(),
                                  sizeof(double) * nb, hipMemcpyHostToDevice,
                                  ctx->stream));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(dhi.p, src.blk.thi.data(),
                                  sizeof(double) * nb, hipMemcpyHostToDevice,
                                  ctx->stream));
  return map_grid_run(ctx, ds::MM_SMP, g, src.blk.ntile, dlo.p, dhi.p,
                      &src.blk, src.pos.p, src.u.p, t64.p, cv, m);
}

int ds_sigma_m_noise_dom(ds_ctx* ctx, int n_pix, const double* npix,
                         int n_map, const double* map_pts,
                         const ds_covar* covar, double* sigma) {
  if (!ctx || (n_map > 0 && !sigma)) return DS_E_ARG;
  if (int rc = check_covar(ctx, covar)) return rc;
  if (int rc = check_points(ctx, npix, 5, n_pix, "pixels")) return rc;
  if (int rc = check_points(ctx, map_pts, 3, n_map, "map points")) return rc;
  if (n_map == 0) return DS_OK;
  if (n_pix == 0) {
    std::fill(sigma, sigma + n_map, 0.0);
    return DS_OK;
  }
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device));
  const CovarDev cv = ds::make_covar(covar);
  SrcSet src;
  if (int rc = make_srcset(ctx, npix, 5, n_pix, nullptr, 1, cv, src))
    return rc;
  DevBuf<double2> t64;
  if (int rc = upload_table64(ctx, covar, t64)) return rc;
  return map_points_run(ctx, ds::MM_SIGMA, n_map, map_pts, src, t64.p, cv,
                        sigma);
}

int ds_solve_cf(ds_problem* p, int exact, double* x) {
  if (!p || !x) return DS_E_ARG;
  if (p->n_sub != 1)
    return fail(p->ctx, DS_E_STATE, "dense solve needs a single volume");
  if ((double)p->n * p->n * 8.0 > 1.5e11)
    return fail(p->ctx, DS_E_ARG, "dense solve of %d pixels needs %.0f GB",
                p->n, (double)p->n * p->n * 8e-9);
  hipStream_t st = p->ctx->stream;
  HIP_OR_FAIL(p->ctx, ds::launch_fill(p->x, 0.0, p->npad, st));
  int info = 0;
  const hipError_t e = ds::dense_solve(p, exact ? 1 : 0, p->table_raw, &info);
  if (e == hipErrorInvalidValue && info > 0)
    return fail(p->ctx, DS_E_STATE, "A is not positive definite (leading "
                "minor %d)", info);
  HIP_OR_FAIL(p->ctx, e);
  // lib/wf_vector.cc:316-318: x = w .* y
  if (p->form == DS_FORM_W)
    HIP_OR_FAIL(p->ctx, ds::launch_scale(p->x, p->wvec, p->x, nullptr,
                                         p->npad, st));
  return download_sorted(p, p->x, x);
}

int ds_map_covar_diag(ds_ctx* ctx, int n_pix, const double* npix, int n_map,
                      const double* map_pts, const ds_covar* covar,
                      const ds_pcg_params* prm, double* sigma,
                      int* num_iters, int* success) {
  if (!ctx || !prm || (n_map > 0 && (!sigma || !map_pts))) return DS_E_ARG;
  if (int rc = check_points(ctx, map_pts, 3, n_map, "map points")) return rc;
  if (int rc = check_covar(ctx, covar)) return rc;
  if (n_pix <= 0) return fail(ctx, DS_E_ARG, "n must be positive");
  if (int rc = check_points(ctx, npix, 5, n_pix, "pixels")) return rc;
  if (n_map == 0) return DS_OK;
  // The cells' solves share one operator (S + N): they run kMR at a time
  // as the subvolumes of a shared-geometry batch (R copies of the pixels,
  // one cell's column as each copy's data), so every kernel of a step
  // serves all of them; each subvolume's PCG is the single-volume solve bit
  // for bit (same blocking, same reduction order).
  constexpr int kMR = DS_DENSE_R;
  const int R = std::min(kMR, n_map);
  const size_t n = (size_t)n_pix;
  std::vector<double> copies(5 * n * R);
  for (int k = 0; k < R; ++k)
    for (size_t i = 0; i < n; ++i) {
      for (int c = 0; c < 4; ++c)
        copies[5 * (n * k + i) + c] = npix[5 * i + c];
      copies[5 * (n * k + i) + 4] = 0.0;   // replaced by the columns
    }
  std::vector<int> sub_n(R, n_pix);
  std::vector<ds_covar> covs(R, *covar);
  ds_problem* p = nullptr;
  if (int rc = ds_batch_create(ctx, DS_FORM_SN, R, sub_n.data(),
                               copies.data(), covs.data(), &p))
    return rc;
  std::unique_ptr<ds_problem, int (*)(ds_problem*)> guard(p,
                                                         ds_problem_destroy);
  if (prm->n != n_pix)
    return fail(ctx, DS_E_ARG, "params.n = %d but the problem has %d pixels",
                prm->n, n_pix);
  ds_pcg_params bprm = *prm;   // the batch holds R copies of the pixels
  bprm.n = p->n;
  if (int rc = check_pcg_params(p, &bprm)) return rc;
  if (!p->shared_geometry)
    return fail(ctx, DS_E_STATE, "map covariance: copies not shared");
  const int npad1 = p->npad / R;
  // slot of pixel j in subvolume 0 (every copy has the same blocking)
  std::vector<int> slot0(n);
  for (int s = 0; s < npad1; ++s)
    if (p->perm[s] >= 0) slot0[p->perm[s]] = s;
  hipStream_t st = ctx->stream;
  DevBuf<double> dots;
  HIP_OR_FAIL(ctx, dots.alloc(R));
  std::vector<double> hcol((size_t)p->npad), hdot(R);
  std::vector<ds_pcg_result> res(R);
  for (int i0 = 0; i0 < n_map; i0 += R) {
    const int r = std::min(R, n_map - i0);
    // lib/map_covar.cc:40-71: rhs = S(map_i, pix_k), x = 0, SN-form PCG;
    // the copies past the last cell get rhs 0 (|b| = 0: no solve)
    for (int k = 0; k < R; ++k) {
      double* col = p->rhs64 + (size_t)k * npad1;
      if (k < r)
        HIP_OR_FAIL(ctx, ds::launch_column_sub(
                             p, k, p->table_raw,
                             map_pts + 3 * (size_t)(i0 + k), col, st));
      else
        HIP_OR_FAIL(ctx, ds::launch_fill(col, 0.0, npad1, st));
    }
    HIP_OR_FAIL(ctx, hipMemcpyAsync(hcol.data(), p->rhs64,
                                    sizeof(double) * p->npad,
                                    hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
    for (int k = 0; k < R; ++k) {
      double sum_b2 = 0.0;   // |b| in the reference's order
      const double* col = hcol.data() + (size_t)k * npad1;
      for (size_t j = 0; j < n; ++j) {
        const double bj = col[slot0[j]];
        sum_b2 += bj * bj;
      }
      p->sub_norm_b[k] = std::sqrt(sum_b2);
    }
    p->norm_b = p->sub_norm_b[0];
    HIP_OR_FAIL(ctx, ds::launch_fill(p->x, 0.0, p->npad, st));
    HIP_OR_FAIL(ctx, ds::run_solver(p, &bprm, 0, 0.0, res.data(), 0));
    for (int k = 0; k < r; ++k)
      if (res[k].success)
        // lib/map_covar.cc:95-109: sigma_i = sum_j S_ij x_j
        HIP_OR_FAIL(ctx, ds::launch_dot1(p->rhs64 + (size_t)k * npad1,
                                         p->x + (size_t)k * npad1, npad1,
                                         dots.p + k, st));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(hdot.data(), dots.p, sizeof(double) * R,
                                    hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
    for (int k = 0; k < r; ++k) {
      sigma[i0 + k] = res[k].success ? hdot[k] : 0.0;
      if (num_iters) num_iters[i0 + k] = res[k].num_iters;
      if (success) success[i0 + k] = res[k].success;
    }
  }
  return DS_OK;
}

int ds_problem_fill_x(ds_problem* p, double value) {
  if (!p) return DS_E_ARG;
  hipStream_t st = p->ctx->stream;
  if (value == 0.0) {
    HIP_OR_FAIL(p->ctx, ds::launch_fill(p->x, 0.0, p->npad, st));
    HIP_OR_FAIL(p->ctx, hipStreamSynchronize(st));
    return DS_OK;
  }
  // padding slots must stay 0
  std::vector<double> v(p->n, value);
  return upload_sorted(p, v.data(), p->x);
}

int ds_batch_num_sub(const ds_problem* p) { return p ? p->n_sub : 0; }

int ds_batch_sub_pairs(const ds_problem* p, long long* pairs64,
                       long long* pairs32) {
  if (!p || !pairs64 || !pairs32) return DS_E_ARG;
  for (int k = 0; k < p->n_sub; ++k) {
    pairs64[k] = p->f64_rows == 16   ? p->sub_pairs32[k]
                 : p->f64_rows == 32 ? p->sub_half64[k]
                                     : p->sub_pairs64[k];
    pairs32[k] = p->sub_pairs32[k];
  }
  return DS_OK;
}

int ds_batch_solve(ds_problem* p, const ds_pcg_params* prm, int mixed,
                   double drop, int use_graph, ds_pcg_result* res) {
  if (!p || !res) return DS_E_ARG;
  if (int rc = check_pcg_params(p, prm)) return rc;
  if (mixed && !(drop > 0.0 && drop < 1.0))
    return fail(p->ctx, DS_E_ARG, "refine_drop must lie in (0, 1)");
  HIP_OR_FAIL(p->ctx, ds::run_solver(p, prm, mixed ? 1 : 0, drop, res,
                                     use_graph ? 1 : 0));
  return DS_OK;
}

int ds_problem_set_launch(ds_problem* p, int waves, int order_by_cost) {
  if (!p) return DS_E_ARG;
  if (waves != 1 && waves != 2 && waves != 4)
    return fail(p->ctx, DS_E_ARG, "waves_per_block must be 1, 2 or 4");
  p->waves_per_block = waves;
  p->order_by_cost = order_by_cost ? 1 : 0;
  return DS_OK;
}

int ds_problem_set_precond(ds_problem* p, int kind) {
  if (!p) return DS_E_ARG;
  if (kind != DS_PRECOND_JACOBI && kind != DS_PRECOND_BLOCK)
    return fail(p->ctx, DS_E_ARG, "unknown preconditioner %d", kind);
  if (kind == DS_PRECOND_BLOCK && !p->minv) {
    // built on the context's second stream, after the work already on the
    // main stream, so that it overlaps what the caller enqueues next (the
    // warm start's neighbour search); the first reader of the inverses on
    // the main stream waits for it (ds::minv_join)
    ds_ctx* ctx = p->ctx;
    // subvolumes sharing one geometry share one set of block inverses
    p->minv_mod = p->shared_geometry ? p->nblk / p->n_sub : p->nblk;
    p->minv_bytes = sizeof(float) * (size_t)p->minv_mod * DS_BJ_SIZE;
    HIP_OR_FAIL(ctx, ds::palloc(p, &p->minv, p->minv_bytes / sizeof(float)));
    if (!p->minv_ev)
      HIP_OR_FAIL(ctx, hipEventCreateWithFlags(&p->minv_ev,
                                               hipEventDisableTiming));
    HIP_OR_FAIL(ctx, hipEventRecord(p->minv_ev, ctx->stream));
    HIP_OR_FAIL(ctx, hipStreamWaitEvent(ctx->aux, p->minv_ev, 0));
    HIP_OR_FAIL(ctx, ds::bj_build(p, ctx->aux));
    HIP_OR_FAIL(ctx, hipEventRecord(p->minv_ev, ctx->aux));
    p->minv_pending = true;
  }
  p->precond_kind = kind;
  return DS_OK;
}

int ds_problem_set_f32_kernel(ds_problem* p, int kind) {
  if (!p) return DS_E_ARG;
  if (kind != DS_F32_VALU && kind != DS_F32_MFMA && kind != DS_F32_MFMA_CLAMPED)
    return fail(p->ctx, DS_E_ARG, "unknown FP32 kernel kind %d", kind);
  p->f32_kind = kind;
  return DS_OK;
}

int ds_problem_set_tile(ds_problem* p, int f64_rows) {
  if (!p) return DS_E_ARG;
  if (f64_rows != 16 && f64_rows != 32 && f64_rows != 64)
    return fail(p->ctx, DS_E_ARG, "FP64 tile rows must be 16, 32 or 64");
  p->f64_rows = f64_rows;
  return DS_OK;
}

int ds_problem_get_launch(const ds_problem* p, int* waves, int* order,
                          int* f64_rows, int* f32_kind) {
  if (!p) return DS_E_ARG;
  if (waves) *waves = p->waves_per_block;
  if (order) *order = p->order_by_cost;
  if (f64_rows) *f64_rows = p->f64_rows;
  if (f32_kind) *f32_kind = p->f32_kind;
  return DS_OK;
}

int ds_problem_autotune(ds_problem* p, int prec, int reps, int* waves_out,
                        int* order_out, double* ms_out) {
  if (!p || reps < 1) return DS_E_ARG;
  if (prec != DS_PREC_F64 && prec != DS_PREC_F32)
    return fail(p->ctx, DS_E_ARG, "unknown precision %d", prec);
  // the candidates: FP32 kernel form (VALU 64 x 16 or MFMA 16 x 16) or FP64
  // tile rows (64, 32, 16 x 16 slots), x waves per block x launch order
  const int w0 = p->waves_per_block, o0 = p->order_by_cost;
  const int k0 = p->f32_kind, r0 = p->f64_rows;
  double best = 1e300;
  int bw = w0, bo = o0, bk = k0, br = r0;
  const bool f32 = prec == DS_PREC_F32;
  const int kinds[2] = {DS_F32_VALU, DS_F32_MFMA};
  const int rows[3] = {64, 32, 16};
  const int nshape = f32 ? 2 : 3;
  for (int sh = 0; sh < nshape; ++sh)
    for (int order = 0; order < 2; ++order)
      for (int waves : {1, 2, 4}) {
        p->waves_per_block = waves;
        p->order_by_cost = order;
        if (f32) p->f32_kind = kinds[sh];
        else p->f64_rows = rows[sh];
        double ms = 0.0;
        if (int rc = ds_problem_time_matvec(p, prec, reps, &ms)) {
          // leave the launch settings as they were before the call
          p->waves_per_block = w0;
          p->order_by_cost = o0;
          p->f32_kind = k0;
          p->f64_rows = r0;
          return rc;
        }
        if (ms < best) {
          best = ms;
          bw = waves;
          bo = order;
          bk = p->f32_kind;
          br = p->f64_rows;
        }
      }
  p->waves_per_block = bw;
  p->order_by_cost = bo;
  p->f32_kind = bk;
  p->f64_rows = br;
  if (waves_out) *waves_out = bw;
  if (order_out) *order_out = bo;
  if (ms_out) *ms_out = best;
  return DS_OK;
}

int ds_problem_time_matvec(ds_problem* p, int prec, int reps,
                           double* ms_per_launch) {
  if (!p || !ms_per_launch || reps < 1) return DS_E_ARG;
  if (prec != DS_PREC_F64 && prec != DS_PREC_F32)
    return fail(p->ctx, DS_E_ARG, "unknown precision %d", prec);
  hipStream_t st = p->ctx->stream;
  const double* w = p->form == DS_FORM_W ? p->wvec : nullptr;
  // input: the resident x (any finite vector will do for timing)
  HIP_OR_FAIL(p->ctx, ds::launch_scale(p->x, w, p->r, p->du32, p->npad, st));
  hipEvent_t e0, e1;
  HIP_OR_FAIL(p->ctx, hipEventCreate(&e0));
  HIP_OR_FAIL(p->ctx, hipEventCreate(&e1));
  hipError_t e = hipSuccess;
  for (int warm = 0; warm < 2 && e == hipSuccess; ++warm) {
    e = hipEventRecord(e0, st);
    const int k = warm ? reps : 1;
    for (int i = 0; i < k && e == hipSuccess; ++i)
      e = prec == DS_PREC_F64
              ? ds::launch_matvec64_ex(p, 0, p->x, p->r, p->q,
                                       ds::CgEpilogue(), st)
              : ds::launch_matvec32_ex(p, p->du32, p->x, p->q, 0, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
  }
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_OR_FAIL(p->ctx, e);
  *ms_per_launch = (double)ms / reps;
  return DS_OK;
}

int ds_problem_build_dense_bf16(ds_problem* p) {
  if (!p) return DS_E_ARG;
  if (!p->shared_geometry || p->n_sub > DS_DENSE_R)
    return fail(p->ctx, DS_E_STATE, "the explicit block path needs at most "
                "%d subvolumes sharing one geometry", DS_DENSE_R);
  const int ld = p->npad / p->n_sub;
  const int mpad = (ld + 255) / 256 * 256;   // GEMM workgroups: 256 rows
  const double bytes = 2.0 * (double)mpad * ld;
  if (bytes > 2.4e11)
    return fail(p->ctx, DS_E_NOMEM, "explicit S would need %.1f GB",
                bytes * 1e-9);
  hipStream_t st = p->ctx->stream;
  // a rebuild: the work that read the old block has completed before its
  // buffers go back to the context's cache
  HIP_OR_FAIL(p->ctx, hipStreamSynchronize(st));
  ds::pfree(p, p->dense_s);
  ds::pfree(p, p->dense_y);
  ds::pfree(p, p->du16);
  p->dense_ld = ld;
  p->dense_mpad = mpad;
  HIP_OR_FAIL(p->ctx, ds::palloc(p, &p->dense_s, (size_t)mpad * ld));
  HIP_OR_FAIL(p->ctx, ds::palloc(p, &p->dense_y,
                                (size_t)DS_DENSE_SPLIT_MAX * DS_DENSE_R * mpad));
  HIP_OR_FAIL(p->ctx, ds::palloc(p, &p->du16, (size_t)DS_DENSE_R * ld));
  // unused right-hand-side columns stay zero
  HIP_OR_FAIL(p->ctx, hipMemsetAsync(p->du16, 0, sizeof(unsigned short) *
                                     (size_t)DS_DENSE_R * ld, st));
  HIP_OR_FAIL(p->ctx, ds::dense_build(p, st));
  // Split-K count of the GEMM: how evenly its workgroups fall on the CUs
  // decides the rate, so time each candidate on the problem itself (the
  // rate does not depend on the data) and keep the fastest.
  const int zero = 0;
  HIP_OR_FAIL(p->ctx, hipMemcpyAsync(p->done_count, &zero, sizeof(int),
                                     hipMemcpyHostToDevice, st));
  hipEvent_t e0, e1;
  HIP_OR_FAIL(p->ctx, hipEventCreate(&e0));
  HIP_OR_FAIL(p->ctx, hipEventCreate(&e1));
  hipError_t e = hipSuccess;
  float best = 1e30f;
  int best_split = p->dense_split;
  for (int split : {2, 3, 4, 5, 6, 8}) {
    static_assert(DS_DENSE_SPLIT_MAX >= 8, "dense_y holds 8 splits");
    p->dense_split = split;
    e = ds::dense_gemm(p, st);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = ds::dense_gemm(p, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) break;
    if (ms < best) {
      best = ms;
      best_split = split;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_OR_FAIL(p->ctx, e);
  p->dense_split = best_split;
  HIP_OR_FAIL(p->ctx, hipStreamSynchronize(st));
  p->use_dense = true;
  return DS_OK;
}

int ds_problem_set_dense(ds_problem* p, int on) {
  if (!p) return DS_E_ARG;
  if (on && !p->dense_s)
    return fail(p->ctx, DS_E_STATE, "no explicit block built");
  p->use_dense = on != 0;
  if (on) p->use_fused = false;
  return DS_OK;
}

int ds_problem_set_fused(ds_problem* p, int on) {
  if (!p) return DS_E_ARG;
  if (on) {
    if (!p->shared_geometry || p->n_sub > DS_DENSE_R)
      return fail(p->ctx, DS_E_STATE, "the fused multi-RHS operator needs "
                  "at most %d subvolumes sharing one geometry", DS_DENSE_R);
    if (!p->du16) {
      const size_t ld = (size_t)p->npad / p->n_sub;
      HIP_OR_FAIL(p->ctx, ds::palloc(p, &p->du16, (size_t)DS_DENSE_R * ld));
      // unused right-hand-side columns stay zero
      HIP_OR_FAIL(p->ctx, hipMemsetAsync(p->du16, 0, sizeof(unsigned short) *
                                         DS_DENSE_R * ld, p->ctx->stream));
      HIP_OR_FAIL(p->ctx, hipStreamSynchronize(p->ctx->stream));
    }
    p->use_dense = false;
  }
  p->use_fused = on != 0;
  return DS_OK;
}

int ds_problem_time_dense(ds_problem* p, int reps, double* ms_per_launch) {
  if (!p || !ms_per_launch || reps < 1) return DS_E_ARG;
  if (!p->dense_s && !p->use_fused)
    return fail(p->ctx, DS_E_STATE, "no multi-RHS operator selected");
  hipStream_t st = p->ctx->stream;
  // the timing input: bf16 copy of the resident x (any vector will do);
  // done_count must read "not finished" for the kernel to run
  HIP_OR_FAIL(p->ctx, ds::launch_scale(p->x, nullptr, nullptr, p->du32,
                                       p->npad, st));
  HIP_OR_FAIL(p->ctx, ds::launch_to_bf16(p->du32, p->du16, p->npad, st));
  const int zero = 0;
  HIP_OR_FAIL(p->ctx, hipMemcpyAsync(p->done_count, &zero, sizeof(int),
                                     hipMemcpyHostToDevice, st));
  hipEvent_t e0, e1;
  HIP_OR_FAIL(p->ctx, hipEventCreate(&e0));
  HIP_OR_FAIL(p->ctx, hipEventCreate(&e1));
  auto op = [&]() {
    return p->use_fused ? ds::fused_mrhs(p, st) : ds::dense_gemm(p, st);
  };
  hipError_t e = op();
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int i = 0; i < reps && e == hipSuccess; ++i) e = op();
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_OR_FAIL(p->ctx, e);
  *ms_per_launch = (double)ms / reps;
  return DS_OK;
}

int ds_problem_dense_product(ds_problem* p, const double* a, double* b) {
  // b = A a through the explicit BF16 block (every subvolume at once);
  // the pair of kernels the mixed solver uses, without the CG epilogue's
  // scalar updates.
  if (!p || !a || !b) return DS_E_ARG;
  if (!p->dense_s && !p->use_fused)
    return fail(p->ctx, DS_E_STATE, "no multi-RHS operator selected");
  hipStream_t st = p->ctx->stream;
  if (int rc = upload_sorted(p, a, p->d)) return rc;
  const double* w = p->form == DS_FORM_W ? p->wvec : nullptr;
  HIP_OR_FAIL(p->ctx, ds::launch_scale(p->d, w, nullptr, p->du32, p->npad,
                                       st));
  HIP_OR_FAIL(p->ctx, ds::launch_to_bf16(p->du32, p->du16, p->npad, st));
  // scalars: not done, zero tickets (the epilogue also computes alpha)
  std::vector<ds_cg_scalars> h(p->n_sub);
  std::memset(h.data(), 0, sizeof(ds_cg_scalars) * p->n_sub);
  const int zero = 0;
  HIP_OR_FAIL(p->ctx, hipMemcpyAsync(p->sc, h.data(),
                                     sizeof(ds_cg_scalars) * p->n_sub,
                                     hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(p->ctx, hipMemcpyAsync(p->done_count, &zero, sizeof(int),
                                     hipMemcpyHostToDevice, st));
  if (p->use_fused && !p->use_dense) {
    HIP_OR_FAIL(p->ctx, ds::fused_mrhs(p, st));
  } else {
    HIP_OR_FAIL(p->ctx, ds::dense_gemm(p, st));
    HIP_OR_FAIL(p->ctx, ds::dense_epilogue(p, st));
  }
  return download_sorted(p, p->q, b);
}

}  // extern "C"
