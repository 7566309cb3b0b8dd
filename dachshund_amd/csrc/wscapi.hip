// wscapi.hip -- C ABI of the learned warm start (include/dachshund_gpu.h,
// ds_ws_*): network state, training store in HBM, training loop, inference
// into a problem's resident initial guess.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

#include "ds_internal.h"
#include "warmstart.h"

struct WsSample {
  unsigned short* feat = nullptr;   // n x WS_FEAT, shuffled
  float* target = nullptr;          // n
  int n = 0;
};

struct ds_warmstart {
  ds_ctx* ctx = nullptr;
  ds::WsTrainer T;
  int quant = 1;
  std::vector<WsSample> samples;
  unsigned short* fscratch = nullptr;   // features of one problem
  size_t fscratch_rows = 0;
  float* tscratch = nullptr;
  int* iscratch = nullptr;
  size_t tscratch_n = 0;
};

namespace {

int ws_fail(ds_ctx* ctx, int code, const char* msg) {
  if (ctx) snprintf(ctx->err, sizeof(ctx->err), "%s", msg);
  return code;
}

#define WS_HIP(ws, expr)                                                     \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      char b_[384];                                                          \
      snprintf(b_, sizeof(b_), "%s: %s (%s:%d)", #expr,                      \
               hipGetErrorString(e_), __FILE__, __LINE__);                   \
      return ws_fail((ws)->ctx,                                              \
                     e_ == hipErrorOutOfMemory ? DS_E_NOMEM : DS_E_HIP, b_); \
    }                                                                        \
  } while (0)

unsigned short host_f2bf(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

void free_trainer(ds::WsTrainer& T) {
  void* ps[] = {T.P, T.m, T.v, T.grad, T.part, T.act, T.delta, T.dy,
                T.loss_part, T.iota, T.wsplit};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  T = ds::WsTrainer();
}

int ensure_scratch(ds_warmstart* ws, size_t rows) {
  if (rows > ws->fscratch_rows) {
    if (ws->fscratch) (void)hipFree(ws->fscratch);
    ws->fscratch = nullptr;
    // whole 64-row blocks: the MFMA forward reads its 64 rows at once (a
    // partial block's extra rows only feed outputs that are not stored)
    const size_t frows = (rows + 63) / 64 * 64;
    WS_HIP(ws, hipMalloc(&ws->fscratch, sizeof(unsigned short) * frows *
                                            WS_FEAT));
    ws->fscratch_rows = frows;
  }
  if (rows > ws->tscratch_n) {
    if (ws->tscratch) (void)hipFree(ws->tscratch);
    if (ws->iscratch) (void)hipFree(ws->iscratch);
    ws->tscratch = nullptr;
    ws->iscratch = nullptr;
    WS_HIP(ws, hipMalloc(&ws->tscratch, sizeof(float) * rows));
    WS_HIP(ws, hipMalloc(&ws->iscratch, sizeof(int) * rows));
    ws->tscratch_n = rows;
  }
  return DS_OK;
}

}  // namespace

extern "C" {

int ds_ws_num_params(void) { return WS_NPARAM; }
int ds_ws_feature_dim(void) { return WS_FEAT; }

int ds_ws_create(ds_ctx* ctx, unsigned long long seed, ds_warmstart** out) {
  if (!ctx || !out) return DS_E_ARG;
  *out = nullptr;
  std::unique_ptr<ds_warmstart> ws(new ds_warmstart);
  ws->ctx = ctx;
  ds::WsTrainer& T = ws->T;
  const int nchunk = WS_MAX_BATCH / WS_CHUNK;
  WS_HIP(ws, hipSetDevice(ctx->device));
  WS_HIP(ws, hipMalloc(&T.P, sizeof(float) * WS_NPARAM));
  WS_HIP(ws, hipMalloc(&T.m, sizeof(float) * WS_NPARAM));
  WS_HIP(ws, hipMalloc(&T.v, sizeof(float) * WS_NPARAM));
  WS_HIP(ws, hipMalloc(&T.grad, sizeof(float) * WS_NPARAM));
  WS_HIP(ws, hipMalloc(&T.part, sizeof(float) * WS_NPARAM * nchunk));
  WS_HIP(ws, hipMalloc(&T.act, sizeof(unsigned short) * WS_MAX_BATCH * 3 *
                                   WS_HIDDEN));
  WS_HIP(ws, hipMalloc(&T.delta, sizeof(float) * WS_MAX_BATCH * 3 *
                                     WS_HIDDEN));
  WS_HIP(ws, hipMalloc(&T.dy, sizeof(float) * WS_MAX_BATCH));
  WS_HIP(ws, hipMalloc(&T.loss_part, sizeof(float) * (WS_MAX_BATCH / 256)));
  WS_HIP(ws, hipMalloc(&T.iota, sizeof(int) * WS_MAX_BATCH));
  WS_HIP(ws, hipMalloc(&T.wsplit, sizeof(unsigned short) * WS_SPLIT_ELEMS));
  std::vector<int> iota(WS_MAX_BATCH);
  std::iota(iota.begin(), iota.end(), 0);
  // He-uniform weights, zero biases, from the seed
  std::vector<float> P(WS_NPARAM, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_in, int count, float gain) {
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    const float a = gain * std::sqrt(6.0f / (float)fan_in);
    for (int k = 0; k < count; ++k) P[off + k] = a * U(rng);
  };
  fill(WS_OFF_W1, WS_FEAT, WS_FEAT * WS_HIDDEN, 1.0f);
  fill(WS_OFF_W2, WS_HIDDEN, WS_HIDDEN * WS_HIDDEN, 1.0f);
  fill(WS_OFF_W3, WS_HIDDEN, WS_HIDDEN * WS_HIDDEN, 1.0f);
  fill(WS_OFF_W4, WS_HIDDEN, WS_HIDDEN, 0.1f);
  std::vector<float> zero(WS_NPARAM, 0.0f);
  hipStream_t st = ctx->stream;
  WS_HIP(ws, hipMemcpyAsync(T.P, P.data(), sizeof(float) * WS_NPARAM,
                            hipMemcpyHostToDevice, st));
  WS_HIP(ws, hipMemcpyAsync(T.m, zero.data(), sizeof(float) * WS_NPARAM,
                            hipMemcpyHostToDevice, st));
  WS_HIP(ws, hipMemcpyAsync(T.v, zero.data(), sizeof(float) * WS_NPARAM,
                            hipMemcpyHostToDevice, st));
  WS_HIP(ws, hipMemcpyAsync(T.iota, iota.data(), sizeof(int) * WS_MAX_BATCH,
                            hipMemcpyHostToDevice, st));
  WS_HIP(ws, hipStreamSynchronize(st));
  ds::ctx_retain(ctx);
  *out = ws.release();
  return DS_OK;
}

int ds_ws_destroy(ds_warmstart* ws) {
  if (!ws) return DS_E_ARG;
  (void)hipStreamSynchronize(ws->ctx->stream);
  free_trainer(ws->T);
  for (auto& s : ws->samples) {
    if (s.feat) (void)hipFree(s.feat);
    if (s.target) (void)hipFree(s.target);
  }
  if (ws->fscratch) (void)hipFree(ws->fscratch);
  if (ws->tscratch) (void)hipFree(ws->tscratch);
  if (ws->iscratch) (void)hipFree(ws->iscratch);
  ds_ctx* ctx = ws->ctx;
  delete ws;
  ds::ctx_release(ctx);
  return DS_OK;
}

int ds_ws_set_params(ds_warmstart* ws, const float* P, int n) {
  if (!ws || !P || n != WS_NPARAM) return DS_E_ARG;
  WS_HIP(ws, hipMemcpyAsync(ws->T.P, P, sizeof(float) * n,
                            hipMemcpyHostToDevice, ws->ctx->stream));
  std::vector<float> zero(WS_NPARAM, 0.0f);
  WS_HIP(ws, hipMemcpyAsync(ws->T.m, zero.data(), sizeof(float) * n,
                            hipMemcpyHostToDevice, ws->ctx->stream));
  WS_HIP(ws, hipMemcpyAsync(ws->T.v, zero.data(), sizeof(float) * n,
                            hipMemcpyHostToDevice, ws->ctx->stream));
  WS_HIP(ws, hipStreamSynchronize(ws->ctx->stream));
  ws->T.step = 0;
  return DS_OK;
}

int ds_ws_get_params(ds_warmstart* ws, float* P, int n) {
  if (!ws || !P || n != WS_NPARAM) return DS_E_ARG;
  WS_HIP(ws, hipMemcpyAsync(P, ws->T.P, sizeof(float) * n,
                            hipMemcpyDeviceToHost, ws->ctx->stream));
  WS_HIP(ws, hipStreamSynchronize(ws->ctx->stream));
  return DS_OK;
}

int ds_ws_set_quantized(ds_warmstart* ws, int on) {
  if (!ws) return DS_E_ARG;
  ws->quant = on ? 1 : 0;
  return DS_OK;
}

int ds_ws_features(ds_warmstart* ws, ds_problem* p, float* feat_out) {
  if (!ws || !p || !feat_out) return DS_E_ARG;
  if (p->n_sub != 1)
    return ws_fail(ws->ctx, DS_E_STATE, "warm start needs a single volume");
  if (int rc = ensure_scratch(ws, p->npad)) return rc;
  hipStream_t st = ws->ctx->stream;
  WS_HIP(ws, ds::ws_features(p, ws->fscratch, st));
  std::vector<unsigned short> h((size_t)p->npad * WS_FEAT);
  WS_HIP(ws, hipMemcpyAsync(h.data(), ws->fscratch,
                            sizeof(unsigned short) * h.size(),
                            hipMemcpyDeviceToHost, st));
  WS_HIP(ws, hipStreamSynchronize(st));
  for (int k = 0; k < p->n; ++k) {
    float* dst = feat_out + (size_t)p->perm[k] * WS_FEAT;
    const unsigned short* src = h.data() + (size_t)k * WS_FEAT;
    for (int c = 0; c < WS_FEAT; ++c) {
      const unsigned u = ((unsigned)src[c]) << 16;
      std::memcpy(dst + c, &u, 4);
    }
  }
  return DS_OK;
}

// Work enqueued on `later` from now on runs after what `earlier` holds.
static hipError_t order_after(hipStream_t later, hipStream_t earlier) {
  hipEvent_t ev;
  hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e != hipSuccess) return e;
  e = hipEventRecord(ev, earlier);
  if (e == hipSuccess) e = hipStreamWaitEvent(later, ev, 0);
  const hipError_t d = hipEventDestroy(ev);
  return e != hipSuccess ? e : d;
}

int ds_ws_predict(ds_warmstart* ws, ds_problem* p) {
  if (!ws || !p) return DS_E_ARG;
  if (p->n_sub != 1)
    return ws_fail(ws->ctx, DS_E_STATE, "warm start needs a single volume");
  if (int rc = ensure_scratch(ws, p->npad)) return rc;
  hipStream_t st = ws->ctx->stream;
  // Stream-ordered, no host wait: the prediction is in place for every
  // later call on the problem's context.  A warm start on another context
  // (another stream) is ordered against the problem's stream both ways.
  hipStream_t ps = p->ctx->stream;
  if (ps != st) WS_HIP(ws, order_after(st, ps));
  WS_HIP(ws, ds::ws_predict(ws->T.P, ws->T.wsplit, ws->quant, p,
                            ws->fscratch, st));
  if (ps != st) WS_HIP(ws, order_after(ps, st));
  return DS_OK;
}

int ds_ws_add_sample(ds_warmstart* ws, ds_problem* p,
                     unsigned long long seed) {
  if (!ws || !p) return DS_E_ARG;
  if (p->n_sub != 1)
    return ws_fail(ws->ctx, DS_E_STATE, "warm start needs a single volume");
  if (int rc = ensure_scratch(ws, p->npad)) return rc;
  hipStream_t st = ws->ctx->stream;
  WsSample s;
  s.n = p->n;
  WS_HIP(ws, hipMalloc(&s.feat, sizeof(unsigned short) * (size_t)s.n *
                                    WS_FEAT));
  WS_HIP(ws, hipMalloc(&s.target, sizeof(float) * (size_t)s.n));
  ws->samples.push_back(s);
  std::vector<int> perm(s.n);
  std::iota(perm.begin(), perm.end(), 0);
  std::shuffle(perm.begin(), perm.end(), std::mt19937_64(seed));
  WS_HIP(ws, ds::ws_features(p, ws->fscratch, st));
  float* traw = ws->tscratch;
  WS_HIP(ws, ds::ws_targets(p, traw, st));
  WS_HIP(ws, hipMemcpyAsync(ws->iscratch, perm.data(), sizeof(int) * s.n,
                            hipMemcpyHostToDevice, st));
  WS_HIP(ws, ds::ws_gather(ws->fscratch, traw, ws->iscratch, s.n, s.feat,
                           s.target, st));
  WS_HIP(ws, hipStreamSynchronize(st));
  return DS_OK;
}

int ds_ws_num_samples(ds_warmstart* ws, long long* count) {
  if (!ws || !count) return DS_E_ARG;
  long long c = 0;
  for (auto& s : ws->samples) c += s.n;
  *count = c;
  return DS_OK;
}

int ds_ws_clear_samples(ds_warmstart* ws) {
  if (!ws) return DS_E_ARG;
  (void)hipStreamSynchronize(ws->ctx->stream);
  for (auto& s : ws->samples) {
    if (s.feat) (void)hipFree(s.feat);
    if (s.target) (void)hipFree(s.target);
  }
  ws->samples.clear();
  return DS_OK;
}

int ds_ws_train(ds_warmstart* ws, int epochs, int batch, float lr,
                float lr_decay, unsigned long long seed,
                double* loss_per_epoch) {
  if (!ws || epochs < 0 || batch < 1 || batch > WS_MAX_BATCH || !(lr > 0))
    return DS_E_ARG;
  if (ws->samples.empty())
    return ws_fail(ws->ctx, DS_E_STATE, "no training samples");
  hipStream_t st = ws->ctx->stream;
  std::vector<std::pair<int, int>> steps;   // (sample, start)
  for (int s = 0; s < (int)ws->samples.size(); ++s)
    for (int b = 0; b < ws->samples[s].n; b += batch) steps.push_back({s, b});
  const int per = (batch + 255) / 256;
  float* lossbuf = nullptr;
  WS_HIP(ws, hipMalloc(&lossbuf, sizeof(float) * steps.size() * per));
  std::unique_ptr<float, decltype(&hipFree)> guard(lossbuf, hipFree);
  std::mt19937_64 rng(seed);
  std::vector<float> hl(steps.size() * per);
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(steps.begin(), steps.end(), rng);
    const float lr_e = lr * std::pow(lr_decay, (float)e);
    WS_HIP(ws, hipMemsetAsync(lossbuf, 0, sizeof(float) * hl.size(), st));
    for (size_t k = 0; k < steps.size(); ++k) {
      const WsSample& s = ws->samples[steps[k].first];
      const int b0 = steps[k].second;
      const int B = std::min(batch, s.n - b0);
      WS_HIP(ws, ds::ws_batch(ws->T, s.feat + (size_t)b0 * WS_FEAT,
                              s.target + b0, ws->T.iota, B, ws->quant,
                              lossbuf + k * per, st));
      WS_HIP(ws, ds::ws_adam(ws->T, lr_e, st));
    }
    WS_HIP(ws, hipMemcpyAsync(hl.data(), lossbuf, sizeof(float) * hl.size(),
                              hipMemcpyDeviceToHost, st));
    WS_HIP(ws, hipStreamSynchronize(st));
    double sum = 0.0;
    long long cnt = 0;
    for (float v : hl) sum += v;
    for (auto& s : ws->samples) cnt += s.n;
    if (loss_per_epoch) loss_per_epoch[e] = sum / (double)cnt;
  }
  return DS_OK;
}

int ds_ws_eval_batch(ds_warmstart* ws, int n, const float* feat,
                     const float* target, float* pred, float* grad,
                     double* loss) {
  if (!ws || n < 1 || n > WS_MAX_BATCH || !feat || !target) return DS_E_ARG;
  if (int rc = ensure_scratch(ws, n)) return rc;
  hipStream_t st = ws->ctx->stream;
  std::vector<unsigned short> h((size_t)n * WS_FEAT);
  for (size_t k = 0; k < h.size(); ++k) h[k] = host_f2bf(feat[k]);
  WS_HIP(ws, hipMemcpyAsync(ws->fscratch, h.data(),
                            sizeof(unsigned short) * h.size(),
                            hipMemcpyHostToDevice, st));
  WS_HIP(ws, hipMemcpyAsync(ws->tscratch, target, sizeof(float) * n,
                            hipMemcpyHostToDevice, st));
  float* t_out = nullptr;
  WS_HIP(ws, hipMalloc(&t_out, sizeof(float) * n));
  std::unique_ptr<float, decltype(&hipFree)> g1(t_out, hipFree);
  WS_HIP(ws, ds::ws_forward(ws->T.P, ws->T.wsplit, ws->fscratch, n, n,
                            ws->quant, nullptr,
                            nullptr, t_out, st));
  WS_HIP(ws, ds::ws_batch(ws->T, ws->fscratch, ws->tscratch, ws->T.iota, n,
                          ws->quant, ws->T.loss_part, st));
  std::vector<float> lp((n + 255) / 256);
  WS_HIP(ws, hipMemcpyAsync(lp.data(), ws->T.loss_part,
                            sizeof(float) * lp.size(), hipMemcpyDeviceToHost,
                            st));
  if (pred)
    WS_HIP(ws, hipMemcpyAsync(pred, t_out, sizeof(float) * n,
                              hipMemcpyDeviceToHost, st));
  if (grad)
    WS_HIP(ws, hipMemcpyAsync(grad, ws->T.grad, sizeof(float) * WS_NPARAM,
                              hipMemcpyDeviceToHost, st));
  WS_HIP(ws, hipStreamSynchronize(st));
  if (loss) {
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = s / n;
  }
  return DS_OK;
}

}  // extern "C"
