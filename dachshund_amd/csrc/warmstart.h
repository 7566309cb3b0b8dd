// warmstart.h -- sizes and device state of the learned initial guess.
#pragma once
#include <hip/hip_runtime.h>

#define WS_K 32                       // nearest neighbours per pixel
#define WS_FEAT (5 * WS_K + 2)        // 162 inputs
#define WS_HIDDEN 64
// flat parameter vector (FP32); weights are stored [in][out]
#define WS_OFF_W1 0
#define WS_OFF_B1 (WS_OFF_W1 + WS_FEAT * WS_HIDDEN)
#define WS_OFF_W2 (WS_OFF_B1 + WS_HIDDEN)
#define WS_OFF_B2 (WS_OFF_W2 + WS_HIDDEN * WS_HIDDEN)
#define WS_OFF_W3 (WS_OFF_B2 + WS_HIDDEN)
#define WS_OFF_B3 (WS_OFF_W3 + WS_HIDDEN * WS_HIDDEN)
#define WS_OFF_W4 (WS_OFF_B3 + WS_HIDDEN)
#define WS_OFF_B4 (WS_OFF_W4 + WS_HIDDEN)
#define WS_NPARAM (WS_OFF_B4 + 1)     // 18817
#define WS_CHUNK 1024                 // samples per weight-gradient partial
#define WS_MAX_BATCH 65536

struct ds_problem;

namespace ds {

struct WsTrainer {
  float* P = nullptr;       // parameters
  float* m = nullptr;       // Adam moments
  float* v = nullptr;
  float* grad = nullptr;
  float* part = nullptr;    // per-chunk gradient partials
  unsigned short* act = nullptr;   // minibatch activations (BF16)
  float* delta = nullptr;   // minibatch deltas
  float* dy = nullptr;
  float* loss_part = nullptr;
  int* iota = nullptr;      // 0..WS_MAX_BATCH-1
  unsigned short* wsplit = nullptr;   // layers 1-3 as 3 BF16 parts (MFMA)
  long long step = 0;
};

// Padded input widths of layers 1-3 in the MFMA forward and the size of
// the split-weight buffer (3 BF16 parts x 64 units x padded width).
#define WS_K1P 176
#define WS_SPLIT_ELEMS (3 * WS_HIDDEN * (WS_K1P + 2 * WS_HIDDEN))

hipError_t ws_features(ds_problem* p, unsigned short* feat,
                       hipStream_t st);
// Inference; with quant (BF16 activations) and wsplit it runs on the
// matrix cores, else on the VALU.
// Prediction of the problem's x: with quant and wsplit one fused kernel
// (features in LDS, network on the matrix cores), else features into feat
// and the VALU forward.
hipError_t ws_predict(const float* P, unsigned short* wsplit, int quant,
                      ds_problem* p, unsigned short* feat, hipStream_t st);
hipError_t ws_forward(const float* P, unsigned short* wsplit,
                      const unsigned short* feat, int n, int npad, int quant,
                      const ds_problem* p, double* x_out, float* t_out,
                      hipStream_t st);
hipError_t ws_targets(const ds_problem* p, float* t, hipStream_t st);
hipError_t ws_batch(WsTrainer& T, const unsigned short* feat,
                    const float* target, const int* idx, int B, int quant,
                    float* loss_part, hipStream_t st);
hipError_t ws_adam(WsTrainer& T, float lr, hipStream_t st);

}  // namespace ds

namespace ds {
hipError_t ws_gather(const unsigned short* fsrc, const float* tsrc,
                     const int* perm, int n, unsigned short* fdst,
                     float* tdst, hipStream_t st);
}  // namespace ds
