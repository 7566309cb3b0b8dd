// capi.hip -- the extern "C" surface of libdachshund_gpu.so
// (declared in include/dachshund_gpu.h).
//
// Host side only: validation, spatial ordering and culling (geometry.cpp),
// uploads to HBM in sorted order, launches, and the permutation back to the
// caller's order.  Errors are returned as ds_status codes with a message in
// the context; nothing here exits or falls back to host arithmetic.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "ds_internal.h"

using ds::Blocking;
using ds::CovarDev;

namespace {

int fail(ds_ctx* ctx, int code, const char* fmt, ...)
    __attribute__((format(printf, 3, 4)));

int fail(ds_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
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

// Map onto a grid from a blocked source set (u in sorted order on device).
int map_grid_run(ds_ctx* ctx, int mode, const ds::MapGrid& g,
                 const Blocking& blk, const double4* pos, const double* u,
                 const double2* table64, const CovarDev& cv, double* m_host) {
  std::vector<double> lo, hi;
  grid_tiles(g, lo, hi);
  const int nt = g.tx * g.ty * g.tz;
  std::vector<int> ptr, idx;
  ds::build_tile_lists(nt, lo.data(), hi.data(), blk, cv, ptr, idx);
  DevBuf<int> dptr, didx;
  DevBuf<double> dm;
  const size_t n_map = (size_t)g.nx * g.ny * g.nz;
  HIP_OR_FAIL(ctx, dptr.alloc(ptr.size()));
  HIP_OR_FAIL(ctx, didx.alloc(idx.size()));
  HIP_OR_FAIL(ctx, dm.alloc(n_map));
  hipStream_t st = ctx->stream;
  HIP_OR_FAIL(ctx, hipMemcpyAsync(dptr.p, ptr.data(), sizeof(int) * ptr.size(),
                                  hipMemcpyHostToDevice, st));
  if (!idx.empty())
    HIP_OR_FAIL(ctx, hipMemcpyAsync(didx.p, idx.data(),
                                    sizeof(int) * idx.size(),
                                    hipMemcpyHostToDevice, st));
  HIP_OR_FAIL(ctx, ds::launch_map64(mode, g, pos, u, dptr.p, didx.p, table64,
                                    cv, dm.p, st));
  HIP_OR_FAIL(ctx, hipMemcpyAsync(m_host, dm.p, sizeof(double) * n_map,
                                  hipMemcpyDeviceToHost, st));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
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

void problem_free(ds_problem* p) {
  dev_free(p->pos64);
  dev_free(p->rhs64);
  dev_free(p->diag);
  dev_free(p->wvec);
  dev_free(p->loc32);
  dev_free(p->blk_c);
  dev_free(p->tptr);
  dev_free(p->tidx);
  dev_free(p->toff32);
  dev_free(p->table64);
  dev_free(p->table32);
  dev_free(p->table_raw);
  dev_free(p->perm);
  if (p->xu != p->x) dev_free(p->xu);
  if (p->du != p->d) dev_free(p->du);
  p->xu = p->du = nullptr;
  dev_free(p->x);
  dev_free(p->r);
  dev_free(p->d);
  dev_free(p->s);
  dev_free(p->q);
  dev_free(p->tmp);
  dev_free(p->du32);
  dev_free(p->partials);
  dev_free(p->sc);
  dev_free(p->norms);
  if (p->sc_host) (void)hipHostFree(p->sc_host);
  if (p->host_tmp) (void)hipHostFree(p->host_tmp);
  p->sc_host = nullptr;
  p->host_tmp = nullptr;
}

// Host-order vector -> sorted device vector (pads 0).
int upload_sorted(ds_problem* p, const double* v, double* dev) {
  for (int k = 0; k < p->npad; ++k)
    p->host_tmp[k] = k < p->n ? v[p->blk.perm[k]] : 0.0;
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
  for (int k = 0; k < p->n; ++k) v[p->blk.perm[k]] = p->host_tmp[k];
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
  if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount,
                            device) != hipSuccess)
    return DS_E_HIP;
  *out = c.release();
  return DS_OK;
}

int ds_ctx_destroy(ds_ctx* ctx) {
  if (!ctx) return DS_E_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return DS_OK;
}

const char* ds_ctx_last_error(const ds_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "NULL context";
}

int ds_ctx_synchronize(ds_ctx* ctx) {
  if (!ctx) return DS_E_ARG;
  HIP_OR_FAIL(ctx, hipStreamSynchronize(ctx->stream));
  return DS_OK;
}

int ds_ctx_num_cu(const ds_ctx* ctx) { return ctx ? ctx->num_cu : 0; }

int ds_problem_create(ds_ctx* ctx, int form, int n, const double* pixels,
                      const ds_covar* covar, ds_problem** out) {
  if (!ctx || !out) return DS_E_ARG;
  *out = nullptr;
  if (form != DS_FORM_W && form != DS_FORM_SN)
    return fail(ctx, DS_E_ARG, "form must be DS_FORM_W or DS_FORM_SN");
  if (n <= 0) return fail(ctx, DS_E_ARG, "n must be positive");
  if (int rc = check_covar(ctx, covar)) return rc;
  if (int rc = check_points(ctx, pixels, 5, n, "pixels")) return rc;
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(pixels[5 * (size_t)i + 3]) ||
        !std::isfinite(pixels[5 * (size_t)i + 4]))
      return fail(ctx, DS_E_ARG, "pixel %d: weight or data not finite", i);
  HIP_OR_FAIL(ctx, hipSetDevice(ctx->device));
  const double t0 = now_s();
  std::unique_ptr<ds_problem, void (*)(ds_problem*)> P(
      new ds_problem, [](ds_problem* q) { problem_free(q); delete q; });
  ds_problem* p = P.get();
  p->ctx = ctx;
  p->form = form;
  p->n = n;
  p->cv = ds::make_covar(covar);
  p->table.assign(covar->table, covar->table + covar->table_n);
  double sc[3];
  axis_scale(p->cv, sc);
  ds::build_blocking(pixels, 5, n, sc, p->blk);
  const Blocking& b = p->blk;
  p->nblk = b.nblk;
  p->ntile = b.ntile;
  p->npad = b.npad;
  std::vector<int> ptr, idx;
  ds::build_tile_lists(b.nblk, b.blo.data(), b.bhi.data(), b, p->cv, ptr, idx);
  p->ntile_pairs = (long long)idx.size();
  // |b| in the reference's order (lib/wf_vector.cc:105-112, :253-260)
  double sum_b2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double bi = pixels[5 * (size_t)i + 4];
    sum_b2 += bi * bi;
  }
  p->norm_b = std::sqrt(sum_b2);

  const int npad = p->npad;
  const double var_s = p->cv.var_s;
  std::vector<double4> hpos(npad);
  std::vector<double> hrhs(npad, 0.0), hdiag(npad, 1.0), hw(npad, 0.0);
  std::vector<float4> hloc(npad);
  std::vector<double> cen(3 * (size_t)b.nblk);
  // FP32 frame of each block: the per-axis median of its real pixels, so an
  // outlier in a block cannot move the frame away from the others.
  for (int B = 0; B < b.nblk; ++B) {
    const int k0 = B * DS_BS, k1 = std::min(n, k0 + DS_BS);
    std::vector<double> v(k1 - k0);
    for (int a = 0; a < 3; ++a) {
      for (int k = k0; k < k1; ++k) v[k - k0] = pixels[5 * (size_t)b.perm[k] + a];
      std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
      cen[3 * B + a] = v[v.size() / 2];
    }
  }
  const double s3[3] = {p->cv.s_perp, p->cv.s_perp, p->cv.s_para};
  for (int k = 0; k < npad; ++k) {
    const int j = b.perm[k];
    const double* Pj = pixels + 5 * (size_t)j;
    const bool real = k < n;
    const double wN = real ? Pj[3] : (form == DS_FORM_W ? 0.0 : 1.0);
    hpos[k] = make_double4(Pj[0], Pj[1], Pj[2], wN);
    if (real) {
      hrhs[k] = Pj[4];
      // lib/wf_vector.cc:190-192 (W) and :44-47 (SN)
      hdiag[k] = form == DS_FORM_W ? wN * wN * var_s + 1.0 : var_s + wN;
      hw[k] = form == DS_FORM_W ? wN : 1.0;
    }
    const int B = k / DS_BS;
    float l[3];
    for (int a = 0; a < 3; ++a)
      l[a] = (float)((Pj[a] - cen[3 * B + a]) * s3[a]);
    hloc[k] = make_float4(l[0], l[1], l[2], 0.0f);
  }
  // per list entry: centre of the target block minus centre of the source
  // tile's block, in pre-scaled units
  std::vector<float4> hoff(idx.size());
  for (int I = 0; I < b.nblk; ++I)
    for (int e = ptr[I]; e < ptr[I + 1]; ++e) {
      const int B = idx[e] / DS_TPB;
      float o[3];
      for (int a = 0; a < 3; ++a)
        o[a] = (float)((cen[3 * I + a] - cen[3 * B + a]) * s3[a]);
      hoff[e] = make_float4(o[0], o[1], o[2], 0.0f);
    }
  std::vector<double2> t64;
  std::vector<float2> t32;
  make_tables(covar, t64, t32);

  hipStream_t st = ctx->stream;
  HIP_OR_FAIL(ctx, dev_alloc(&p->pos64, npad));
  HIP_OR_FAIL(ctx, dev_alloc(&p->rhs64, npad));
  HIP_OR_FAIL(ctx, dev_alloc(&p->diag, npad));
  HIP_OR_FAIL(ctx, dev_alloc(&p->wvec, npad));
  HIP_OR_FAIL(ctx, dev_alloc(&p->loc32, npad));
  HIP_OR_FAIL(ctx, dev_alloc(&p->tptr, ptr.size()));
  HIP_OR_FAIL(ctx, dev_alloc(&p->tidx, idx.size()));
  HIP_OR_FAIL(ctx, dev_alloc(&p->toff32, hoff.size()));
  HIP_OR_FAIL(ctx, dev_alloc(&p->table64, t64.size()));
  HIP_OR_FAIL(ctx, dev_alloc(&p->table32, t32.size()));
  HIP_OR_FAIL(ctx, dev_alloc(&p->table_raw, p->table.size()));
  double** vecs[] = {&p->x, &p->r, &p->d, &p->s, &p->q, &p->tmp};
  for (double** v : vecs) HIP_OR_FAIL(ctx, dev_alloc(v, npad));
  if (form == DS_FORM_W) {
    HIP_OR_FAIL(ctx, dev_alloc(&p->xu, npad));
    HIP_OR_FAIL(ctx, dev_alloc(&p->du, npad));
  } else {
    p->xu = p->x;
    p->du = p->d;
  }
  HIP_OR_FAIL(ctx, dev_alloc(&p->du32, npad));
  const size_t nparts = 2 * (size_t)std::max(p->nblk, npad / 256 + 1);
  HIP_OR_FAIL(ctx, dev_alloc(&p->partials, nparts));
  HIP_OR_FAIL(ctx, dev_alloc(&p->sc, 1));
  HIP_OR_FAIL(ctx, hipHostMalloc((void**)&p->sc_host, 3 * sizeof(ds_cg_scalars),
                                 hipHostMallocDefault));
  HIP_OR_FAIL(ctx, hipHostMalloc((void**)&p->host_tmp, sizeof(double) * npad,
                                 hipHostMallocDefault));
  std::memset(p->sc_host, 0, 3 * sizeof(ds_cg_scalars));
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st)
                 : hipSuccess;
  };
  HIP_OR_FAIL(ctx, up(p->pos64, hpos.data(), sizeof(double4) * npad));
  HIP_OR_FAIL(ctx, up(p->rhs64, hrhs.data(), sizeof(double) * npad));
  HIP_OR_FAIL(ctx, up(p->diag, hdiag.data(), sizeof(double) * npad));
  HIP_OR_FAIL(ctx, up(p->wvec, hw.data(), sizeof(double) * npad));
  HIP_OR_FAIL(ctx, up(p->loc32, hloc.data(), sizeof(float4) * npad));
  HIP_OR_FAIL(ctx, up(p->tptr, ptr.data(), sizeof(int) * ptr.size()));
  HIP_OR_FAIL(ctx, up(p->tidx, idx.data(), sizeof(int) * idx.size()));
  HIP_OR_FAIL(ctx, up(p->toff32, hoff.data(), sizeof(float4) * hoff.size()));
  HIP_OR_FAIL(ctx, up(p->table64, t64.data(), sizeof(double2) * t64.size()));
  HIP_OR_FAIL(ctx, up(p->table32, t32.data(), sizeof(float2) * t32.size()));
  HIP_OR_FAIL(ctx, up(p->table_raw, p->table.data(),
                      sizeof(double) * p->table.size()));
  std::vector<double> zero(npad, 0.0);
  HIP_OR_FAIL(ctx, up(p->x, zero.data(), sizeof(double) * npad));
  HIP_OR_FAIL(ctx, up(p->sc, p->sc_host, sizeof(ds_cg_scalars)));
  HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
  p->setup_seconds = now_s() - t0;
  *out = P.release();
  return DS_OK;
}

int ds_problem_destroy(ds_problem* p) {
  if (!p) return DS_E_ARG;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  problem_free(p);
  delete p;
  return DS_OK;
}

int ds_problem_get_stats(const ds_problem* p, ds_problem_stats* out) {
  if (!p || !out) return DS_E_ARG;
  out->n = p->n;
  out->n_blocks = p->nblk;
  out->n_tiles = p->ntile;
  out->tile_pairs = p->ntile_pairs;
  out->pairs_per_matvec = p->ntile_pairs * (long long)(DS_BS * DS_TS);
  out->setup_seconds = p->setup_seconds;
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
  HIP_OR_FAIL(p->ctx, ds::launch_matvec64_ex(p, 1, p->tmp, p->r, p->q,
                                             ds::CgEpilogue(), st));
  return download_sorted(p, p->q, r);
}

int ds_precond(ds_problem* p, const double* b, double* x) {
  if (!p || !b || !x) return DS_E_ARG;
  if (int rc = upload_sorted(p, b, p->tmp)) return rc;
  HIP_OR_FAIL(p->ctx, ds::launch_precond(p->tmp, p->diag, p->q, p->npad,
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
  HIP_OR_FAIL(p->ctx, ds::run_solver(p, prm, 0, 0.0, res));
  return DS_OK;
}

int ds_solve_mixed_resident(ds_problem* p, const ds_pcg_params* prm,
                            double drop, ds_pcg_result* res) {
  if (!p || !res) return DS_E_ARG;
  if (int rc = check_pcg_params(p, prm)) return rc;
  if (!(drop > 0.0 && drop < 1.0))
    return fail(p->ctx, DS_E_ARG, "refine_drop must lie in (0, 1)");
  HIP_OR_FAIL(p->ctx, ds::run_solver(p, prm, 1, drop, res));
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
  return map_grid_run(p->ctx, ds::MM_SMP, g, p->blk, p->pos64, p->x,
                      p->table64, p->cv, m);
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
  return map_grid_run(ctx, ds::MM_SMP, g, src.blk, src.pos.p, src.u.p, t64.p,
                      cv, m);
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
  ds_problem* p = nullptr;
  if (int rc = ds_problem_create(ctx, DS_FORM_SN, n_pix, npix, covar, &p))
    return rc;
  std::unique_ptr<ds_problem, int (*)(ds_problem*)> guard(p,
                                                         ds_problem_destroy);
  if (int rc = check_pcg_params(p, prm)) return rc;
  hipStream_t st = ctx->stream;
  DevBuf<double> col, dot;
  HIP_OR_FAIL(ctx, col.alloc(p->npad));
  HIP_OR_FAIL(ctx, dot.alloc(1));
  std::vector<int> slot(p->n);
  for (int k = 0; k < p->n; ++k) slot[p->blk.perm[k]] = k;
  double* rhs_saved = p->rhs64;
  std::vector<double> hcol(p->npad);
  for (int i = 0; i < n_map; ++i) {
    // lib/map_covar.cc:40-71: rhs = S(map_i, pix_k), x = 0, SN-form PCG
    HIP_OR_FAIL(ctx, ds::launch_column(p, p->table_raw, map_pts + 3 * (size_t)i,
                                       col.p, st));
    HIP_OR_FAIL(ctx, hipMemcpyAsync(hcol.data(), col.p,
                                    sizeof(double) * p->npad,
                                    hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
    double sum_b2 = 0.0;   // |b| in the reference's order
    for (int j = 0; j < p->n; ++j) {
      const double bj = hcol[slot[j]];
      sum_b2 += bj * bj;
    }
    p->norm_b = std::sqrt(sum_b2);
    p->rhs64 = col.p;
    HIP_OR_FAIL(ctx, ds::launch_fill(p->x, 0.0, p->npad, st));
    ds_pcg_result res;
    const hipError_t e = ds::run_solver(p, prm, 0, 0.0, &res);
    p->rhs64 = rhs_saved;
    HIP_OR_FAIL(ctx, e);
    double s = 0.0;
    if (res.success) {
      // lib/map_covar.cc:95-109: sigma_i = sum_j S_ij x_j
      HIP_OR_FAIL(ctx, ds::launch_dot1(col.p, p->x, p->npad, dot.p, st));
      HIP_OR_FAIL(ctx, hipMemcpyAsync(&s, dot.p, sizeof(double),
                                      hipMemcpyDeviceToHost, st));
      HIP_OR_FAIL(ctx, hipStreamSynchronize(st));
    }
    sigma[i] = s;
    if (num_iters) num_iters[i] = res.num_iters;
    if (success) success[i] = res.success;
  }
  return DS_OK;
}

int ds_problem_fill_x(ds_problem* p, double value) {
  if (!p) return DS_E_ARG;
  hipStream_t st = p->ctx->stream;
  HIP_OR_FAIL(p->ctx, ds::launch_fill(p->x, value, p->n, st));
  if (p->npad > p->n)
    HIP_OR_FAIL(p->ctx, ds::launch_fill(p->x + p->n, 0.0, p->npad - p->n, st));
  HIP_OR_FAIL(p->ctx, hipStreamSynchronize(st));
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

}  // extern "C"
