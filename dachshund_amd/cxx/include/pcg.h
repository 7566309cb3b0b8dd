// pcg.h -- drop-in of lib/pcg.h: PCGParams / PCGResult and the generic
// Jacobi-preconditioned CG over std::function callbacks, for any C++ caller
// with its own operator.  The Wiener-filter classes do not use it: their
// solve_pcg runs the whole iteration on the GPU (ds_solve_pcg) with the same
// numbering, refresh and stop rules as below.
#ifndef DS_DROPIN_PCG_H
#define DS_DROPIN_PCG_H

#include <cmath>
#include <cstdio>
#include <functional>
#include <vector>

#include "linalg.h"

struct PCGParams {
  int n;
  int max_iter;
  int step_r;
  double tol;
  bool verbose;
};

struct PCGResult {
  bool success;
  int num_iters;
  std::vector<double> residual_norms;
};

typedef std::function<void(const double*, double*)> ds_operator;

// One CG step (lib/pcg.h:27-65): q = A d, alpha, x update, residual (full
// or recurrence), preconditioning, new delta and direction.
static inline void pcg_step(const int n, ds_operator A_product, ds_operator A_residual,
                     ds_operator preconditioner_solve, double* const d,
                     double* const q, double* const r, double* const s,
                     double* const x, const bool full_residual,
                     double* const delta) {
  A_product(d, q);
  const double alpha = *delta / ds_vector_dot(n, d, q);
  ds_vector_add_cv(n, alpha, d, x);
  if (full_residual)
    A_residual(x, r);
  else
    ds_vector_sub_cv(n, alpha, q, r);
  preconditioner_solve(r, s);
  const double delta_old = *delta;
  *delta = ds_vector_dot(n, r, s);
  const double beta = *delta / delta_old;
  for (int i = 0; i < n; ++i) d[i] = s[i] + beta * d[i];
}

// lib/pcg.h:67-160.  x: in = initial guess, out = solution.  Steps are
// numbered from 1; every step_r-th recomputes r = b - A x; the solve stops
// when |r| <= tol |b| (success) or after max_iter - 1 steps
// (success = false, num_iters = max_iter).  |b| <= 0 returns at once with
// success = false.
static inline PCGResult pcg_solve(const PCGParams& params, const double norm_b,
                           ds_operator A_product, ds_operator A_residual,
                           ds_operator preconditioner_solve, double* const x) {
  PCGResult res;
  res.success = false;
  res.num_iters = 0;
  const int n = params.n;
  if (norm_b <= 0.0) {
    puts("[WARNING] Requested PCG solve where |b| <= 0.0. Returning early.\n");
    return res;
  }
  std::vector<double> d(n), r(n), s(n), q(n);
  const double stop = params.tol * norm_b;
  A_residual(x, r.data());
  preconditioner_solve(r.data(), d.data());
  double delta = ds_vector_dot(n, r.data(), d.data());
  double norm_r = ds_vector_norm(n, r.data());
  res.residual_norms.push_back(norm_r);
  if (params.verbose) {
    printf("[PCG] Solving %i x %i problem.\n", n, n);
    printf("    Goal |r| < (tol) |b| = %e,  delta = %e\n", stop, delta);
  }
  int iter = 1;
  for (; iter < params.max_iter; ++iter) {
    pcg_step(n, A_product, A_residual, preconditioner_solve, d.data(),
             q.data(), r.data(), s.data(), x, iter % params.step_r == 0,
             &delta);
    norm_r = ds_vector_norm(n, r.data());
    res.residual_norms.push_back(norm_r);
    if (norm_r <= stop) {
      if (params.verbose)
        printf("    iter %i reached |r| = %e < tol |b| = %e\n", iter, norm_r,
               stop);
      break;
    }
    if (params.verbose)
      printf("    iter %i, |r| %e, delta %e\n", iter, norm_r, delta);
  }
  res.success = iter != params.max_iter;
  if (!res.success)
    puts("[WARNING] PCG reached max iteration before stop condition.");
  res.num_iters = iter;
  return res;
}

#endif
