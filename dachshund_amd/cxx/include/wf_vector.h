// wf_vector.h -- drop-in of lib/wf_vector.h: the Wiener-filter vector
// x = A^{-1} b in its two algebraic forms, backed by the gfx950 library.
//
// Construction takes the same non-owning pixel pointer as the reference;
// the pixels are copied to the GPU (ordered and culled) on first use, so
// they must not change after the first operator call.  A_product,
// A_residual, preconditioner_solve, norm_b and solve_pcg run on the GPU in
// FP64; A(i, j, exact) is the scalar host element used by dense callers.
#ifndef DS_DROPIN_WF_VECTOR_H
#define DS_DROPIN_WF_VECTOR_H

#include <functional>

#include "pcg.h"
#include "pixel.h"
#include "signal_covar.h"

struct ds_problem;

class WFVector {
 public:
  WFVector(const int num_pixels, const SignalCovarParams* const s_params)
      : n(num_pixels), s_params(s_params) {}

  const int n;
  const SignalCovarParams* const s_params;
};

// x = (S + N)^{-1} d  (lib/wf_vector.cc:16-155)
class WFVectorSN : public WFVector {
 public:
  const NPixel* const pixels;

  WFVectorSN(const int num_pixels, const NPixel* const pixels,
             const SignalCovarParams* const s_params);
  ~WFVectorSN();
  WFVectorSN(const WFVectorSN&) = delete;
  WFVectorSN& operator=(const WFVectorSN&) = delete;

  double A(const int i, const int j, const bool exact);
  double norm_b();

  void preconditioner_solve(const double* const b, double* const x);
  void A_product(const double* const a, double* const b);
  void A_residual(const double* const x, double* const r);

  std::function<void(const double* const, double* const)> this_pc_solve;
  std::function<void(const double* const, double* const)> this_A_product;
  std::function<void(const double* const, double* const)> this_A_residual;

  PCGResult solve_pcg(const PCGParams& pcg_params, double* const x);
  void solve_cf(double* const x, const bool verbose, const bool exact);

  ds_problem* gpu();   // the device problem (created on first use)

 private:
  ds_problem* gpu_ = nullptr;
};

// x = w (w S w + I)^{-1} w d  (lib/wf_vector.cc:157-319)
//   A_ij = w_i w_j S_ij + delta_ij,  b_i = w_i d_i
class WFVectorW : public WFVector {
 public:
  const WPixel* const pixels;

  WFVectorW(const int num_pixels, const WPixel* const pixels,
            const SignalCovarParams* const s_params);
  ~WFVectorW();
  WFVectorW(const WFVectorW&) = delete;
  WFVectorW& operator=(const WFVectorW&) = delete;

  double A(const int i, const int j, const bool exact);
  double norm_b();

  void preconditioner_solve(const double* const b, double* const x);
  void A_product(const double* const a, double* const b);
  void A_residual(const double* const x, double* const r);

  std::function<void(const double* const, double* const)> this_pc_solve;
  std::function<void(const double* const, double* const)> this_A_product;
  std::function<void(const double* const, double* const)> this_A_residual;

  // x: in = initial guess y0 (W space), out = x = w .* y.
  PCGResult solve_pcg(const PCGParams& pcg_params, double* const x);
  // Mixed FP32/FP64 solve to the same FP64 tolerance (no reference
  // counterpart; same x convention).
  PCGResult solve_pcg_mixed(const PCGParams& pcg_params, double* const x,
                            const double refine_drop = 1e-3);
  void solve_cf(double* const x, const bool verbose, const bool exact);

  ds_problem* gpu();

 private:
  ds_problem* gpu_ = nullptr;
};

#endif
