// cg.hip -- Jacobi-preconditioned CG on the device (lib/pcg.h:27-160).
//
// COMPILE WITH -ffp-contract=off: the elementwise updates below are the
// reference's own rounded operations (x += alpha d, r -= alpha q,
// s = r / diag, d = s + beta d; lib/pcg.h:39,48,52,62-64, lib/linalg.cc:32-49).
//
// All CG scalars live in device memory (ds_cg_scalars); every kernel first
// checks `done` and returns at once after convergence, so the host can
// enqueue iterations in chunks and poll the flag without a round trip per
// iteration.  Kernels per iteration:
//   plain FP64:  matvec64+dq/alpha | update(+precond, dots, stop) | dir
//                (every step_r-th: x-update | residual64+precond+dots+stop)
//   mixed:       matvec32+dq/alpha | update(+dots, refine test) |
//                residual64 if requested (true residual) | dir
#include <algorithm>
#include <chrono>
#include <cstring>

#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

constexpr int kVecThreads = 256;

struct UpdArgs {
  const double* d;
  const double* q;
  const double* diag;
  const double* wvec;     // w (W form) or nullptr
  double* x;
  double* xu;             // w .* x (W form); may alias x
  double* r;
  double* s;
  double* partials;
  ds_cg_scalars* sc;
  double* norms;
  double drop;
  int npad;
  int only_x, mixed;
};

// lib/pcg.h:37-58: x += alpha d; r -= alpha q; s = M^-1 r; r.s and r.r.
__global__ __launch_bounds__(kVecThreads) void k_cg_update(UpdArgs a) {
  __shared__ double red[2 * (kVecThreads / 64)];
  __shared__ bool is_last;
  if (a.sc->done) return;
  const double alpha = a.sc->alpha;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < a.npad) {
    const double xi = a.x[i] + alpha * a.d[i];
    a.x[i] = xi;
    if (a.wvec) a.xu[i] = a.wvec[i] * xi;
    if (!a.only_x) {
      const double ri = a.r[i] - alpha * a.q[i];
      a.r[i] = ri;
      const double si = ri / a.diag[i];
      a.s[i] = si;
      v[0] = ri * si;
      v[1] = ri * ri;
    }
  }
  if (a.only_x) return;
  block_sum<2>(v, red);
  if (!publish_partials<2>(v, a.partials, &a.sc->counter, &is_last)) return;
  gather_partials<2>(v, a.partials, red);
  if (threadIdx.x == 0)
    finalize_step(a.sc, a.norms, v[0], v[1], 0, a.mixed, a.drop, 0);
  release_ticket(&a.sc->counter);
}

struct DirArgs {
  const double* s;
  const double* wvec;
  double* d;
  double* du;
  float* du32;
  ds_cg_scalars* sc;
  int npad;
};

// lib/pcg.h:61-64: d = s + beta d (and its scaled copies).
__global__ __launch_bounds__(kVecThreads) void k_cg_dir(DirArgs a) {
  if (a.sc->done) return;
  const double beta = a.sc->beta;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.npad) return;
  const double di = a.s[i] + beta * a.d[i];
  a.d[i] = di;
  const double ui = a.wvec ? a.wvec[i] * di : di;
  if (a.wvec) a.du[i] = ui;
  if (a.du32) a.du32[i] = (float)ui;
}

// u = w .* v (w may be null: u = v), optional FP32 copy.
__global__ __launch_bounds__(kVecThreads) void k_scale(const double* v,
                                                       const double* w,
                                                       double* u, float* u32,
                                                       int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const double ui = w ? w[i] * v[i] : v[i];
  if (u) u[i] = ui;
  if (u32) u32[i] = (float)ui;
}

// lib/wf_vector.cc:42-49 / :185-194: x = b / diag.
__global__ __launch_bounds__(kVecThreads) void k_precond(const double* b,
                                                         const double* diag,
                                                         double* x, int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  x[i] = b[i] / diag[i];
}

inline dim3 vec_grid(int npad) {
  return dim3((unsigned)((npad + kVecThreads - 1) / kVecThreads));
}

}  // namespace

hipError_t launch_scale(const double* v, const double* w, double* u,
                        float* u32, int npad, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, vec_grid(npad), dim3(kVecThreads), 0, st, v, w,
                     u, u32, npad);
  return hipGetLastError();
}

hipError_t launch_precond(const double* b, const double* diag, double* x,
                          int npad, hipStream_t st) {
  hipLaunchKernelGGL(k_precond, vec_grid(npad), dim3(kVecThreads), 0, st, b,
                     diag, x, npad);
  return hipGetLastError();
}

namespace {

#define DS_TRY(expr)                     \
  do {                                   \
    hipError_t e_ = (expr);              \
    if (e_ != hipSuccess) return e_;     \
  } while (0)

hipError_t launch_update(ds_problem* p, int only_x, int mixed, double drop,
                         hipStream_t st) {
  UpdArgs a{};
  a.d = p->d;
  a.q = p->q;
  a.diag = p->diag;
  a.wvec = p->form == DS_FORM_W ? p->wvec : nullptr;
  a.x = p->x;
  a.xu = p->xu;
  a.r = p->r;
  a.s = p->s;
  a.partials = p->partials;
  a.sc = p->sc;
  a.norms = p->norms;
  a.drop = drop;
  a.npad = p->npad;
  a.only_x = only_x;
  a.mixed = mixed;
  hipLaunchKernelGGL(k_cg_update, vec_grid(p->npad), dim3(kVecThreads), 0, st,
                     a);
  return hipGetLastError();
}

hipError_t launch_dir(ds_problem* p, int mixed, hipStream_t st) {
  DirArgs a{};
  a.s = p->s;
  a.wvec = p->form == DS_FORM_W ? p->wvec : nullptr;
  a.d = p->d;
  a.du = p->du;
  a.du32 = mixed ? p->du32 : nullptr;
  a.sc = p->sc;
  a.npad = p->npad;
  hipLaunchKernelGGL(k_cg_dir, vec_grid(p->npad), dim3(kVecThreads), 0, st, a);
  return hipGetLastError();
}

// One iteration's kernels, in stream order.
hipError_t enqueue_iteration(ds_problem* p, int iter, int step_r, int mixed,
                             double drop, hipStream_t st) {
  CgEpilogue dq;
  dq.kind = EK_CG_DQ;
  if (mixed) {
    DS_TRY(launch_matvec32_ex(p, p->du32, p->d, p->q, 1, st));
    DS_TRY(launch_update(p, 0, 1, drop, st));
    CgEpilogue tr;
    tr.kind = EK_CG_RES;
    tr.out2 = p->s;
    tr.mixed = 1;
    tr.is_true = 1;
    tr.need_want = 1;
    tr.drop = drop;
    DS_TRY(launch_matvec64_ex(p, 1, p->x, p->xu, p->r, tr, st));
  } else {
    DS_TRY(launch_matvec64_ex(p, 0, p->d, p->du, p->q, dq, st));
    if (iter % step_r == 0) {
      // lib/pcg.h:42-45: full residual r = b - A x after the x update.
      DS_TRY(launch_update(p, 1, 0, 0.0, st));
      CgEpilogue res;
      res.kind = EK_CG_RES;
      res.out2 = p->s;
      DS_TRY(launch_matvec64_ex(p, 1, p->x, p->xu, p->r, res, st));
    } else {
      DS_TRY(launch_update(p, 0, 0, 0.0, st));
    }
  }
  return launch_dir(p, mixed, st);
}

}  // namespace

// Solve with the problem's resident x (sorted order, W space on entry) as the
// initial guess; on return x holds the solution (x = w .* y for the W form,
// lib/wf_vector.cc:268-271).
hipError_t run_solver(ds_problem* p, const ds_pcg_params* prm, int mixed,
                      double drop, ds_pcg_result* res) {
  hipStream_t st = p->ctx->stream;
  const int max_iter = prm->max_iter;
  const int cap = std::max(max_iter, 1) + 1;
  if (cap > p->norms_cap) {
    if (p->norms) DS_TRY(hipFree(p->norms));
    p->norms = nullptr;
    DS_TRY(hipMalloc(&p->norms, sizeof(double) * (size_t)cap));
    p->norms_cap = cap;
  }
  std::memset(res, 0, sizeof(*res));
  res->norm_b = p->norm_b;
  hipEvent_t e0, e1;
  DS_TRY(hipEventCreate(&e0));
  DS_TRY(hipEventCreate(&e1));
  DS_TRY(hipEventRecord(e0, st));
  ds_cg_scalars& h = *p->sc_host;
  std::memset(&h, 0, sizeof(h));
  h.tol_norm_b = prm->tol * p->norm_b;
  h.max_iter = max_iter;
  const double* w = p->form == DS_FORM_W ? p->wvec : nullptr;
  if (p->norm_b <= 0.0) {
    // lib/pcg.h:82-86: early return, x untouched (then x *= w).
    h.done = 1;
  } else {
    DS_TRY(hipMemcpyAsync(p->sc, &h, sizeof(h), hipMemcpyHostToDevice, st));
    // lib/pcg.h:97-110: r = b - A x; d = M^-1 r; delta = r . d; |r|.
    DS_TRY(launch_scale(p->x, w, w ? p->xu : nullptr, nullptr, p->npad, st));
    CgEpilogue init;
    init.kind = EK_CG_RES;
    init.out2 = p->d;
    init.init = 1;
    init.mixed = mixed;
    DS_TRY(launch_matvec64_ex(p, 1, p->x, w ? p->xu : p->x, p->r, init, st));
    DS_TRY(launch_scale(p->d, w, w ? p->du : nullptr,
                        mixed ? p->du32 : nullptr, p->npad, st));
    // Enqueue iterations in chunks; after each chunk a snapshot of the
    // scalars is copied to pinned memory (two slots, alternating) and the
    // host inspects the previous chunk's snapshot while this one runs.
    const int chunk = 8;
    hipEvent_t ev[2];
    DS_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    DS_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    int iter = 1, k = 0;
    while (iter < max_iter) {
      const int end = std::min(max_iter, iter + chunk);
      for (; iter < end; ++iter)
        DS_TRY(enqueue_iteration(p, iter, prm->step_r, mixed, drop, st));
      DS_TRY(hipMemcpyAsync(&p->sc_host[1 + (k & 1)], p->sc, sizeof(h),
                            hipMemcpyDeviceToHost, st));
      DS_TRY(hipEventRecord(ev[k & 1], st));
      if (k > 0) {
        DS_TRY(hipEventSynchronize(ev[(k - 1) & 1]));
        if (p->sc_host[1 + ((k - 1) & 1)].done) break;
      }
      ++k;
    }
    DS_TRY(hipMemcpyAsync(&h, p->sc, sizeof(h), hipMemcpyDeviceToHost, st));
    DS_TRY(hipStreamSynchronize(st));
    DS_TRY(hipEventDestroy(ev[0]));
    DS_TRY(hipEventDestroy(ev[1]));
    if (max_iter <= 1) {
      // lib/pcg.h:118,152-158: no step taken, iter == 1.
      h.num_iters = 1;
      h.success = (max_iter != 1);
      h.n_norms = 1;
      h.done = 1;
    }
  }
  // lib/wf_vector.cc:268-271: x = w .* y.
  if (w) DS_TRY(launch_scale(p->x, w, p->x, nullptr, p->npad, st));
  DS_TRY(hipEventRecord(e1, st));
  DS_TRY(hipEventSynchronize(e1));
  float ms = 0.f;
  DS_TRY(hipEventElapsedTime(&ms, e0, e1));
  DS_TRY(hipEventDestroy(e0));
  DS_TRY(hipEventDestroy(e1));
  res->success = h.success;
  res->num_iters = h.num_iters;
  res->n_norms = h.n_norms;
  res->final_norm = h.norm_r;
  res->true_norm = mixed ? h.true_norm : h.norm_r;
  res->n_true = h.n_true;
  res->seconds = ms * 1e-3;
  // pair-evaluation passes: init residual + one per iteration (+ refreshes)
  if (p->norm_b > 0.0) {
    const int steps = std::max(0, std::min(h.num_iters, max_iter - 1));
    if (mixed) {
      res->n_matvec32 = steps;
      res->n_matvec64 = 1 + h.n_true;
    } else {
      res->n_matvec32 = 0;
      res->n_matvec64 = 1 + steps + steps / std::max(prm->step_r, 1);
    }
  }
  return hipSuccess;
}

}  // namespace ds
