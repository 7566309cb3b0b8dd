// signal_covar.h -- drop-in of lib/signal_covar.h: the covariance
// parameters with their 200-point table, and the scalar host evaluators.
// The hot loops never call these: they run in the gfx950 kernels, which
// receive the same constants and table through ds_covar.
#ifndef DS_DROPIN_SIGNAL_COVAR_H
#define DS_DROPIN_SIGNAL_COVAR_H

#include <cmath>

const double default_gaussian_table_x_max = 4.0;
const int default_gaussian_table_n = 200;

class SignalCovarParams {
 public:
  double var_s;
  double gaussian_const_perp;   // 1 / (2 l_perp^2)
  double gaussian_const_para;   // 1 / (2 l_para^2)
  int func_table_n;
  double func_table_dx_inv;
  double* func_table;           // exp(-i dx), i < func_table_n

  SignalCovarParams(const double var_s, const double l_perp,
                    const double l_para);
  SignalCovarParams(const SignalCovarParams& o);
  SignalCovarParams& operator=(const SignalCovarParams& o);
  ~SignalCovarParams();
};

// lib/signal_covar.h:53-65: linear interpolation in a table f_i = f(i dx);
// exactly 0 from the last tabulated point on.
static inline double linterp(const double x, const int n, const double dx_inv,
                             const double* const f) {
  const double xx = x * dx_inv;
  const int ix = (int)xx;
  if (ix > n - 2) return 0.0;
  return f[ix] * (ix + 1 - xx) + f[ix + 1] * (xx - ix);
}

// lib/signal_covar.h:67-77
static inline double signal_covar(const double x_perp_2, const double x_para_2,
                                  const SignalCovarParams* const p) {
  const double g = p->gaussian_const_perp * x_perp_2 +
                   p->gaussian_const_para * x_para_2;
  return p->var_s * linterp(g, p->func_table_n, p->func_table_dx_inv,
                            p->func_table);
}

// lib/signal_covar.h:79-86
static inline double signal_covar_exact(const double x_perp_2,
                                        const double x_para_2,
                                        const SignalCovarParams* const p) {
  const double g = p->gaussian_const_perp * x_perp_2 +
                   p->gaussian_const_para * x_para_2;
  return p->var_s * std::exp(-g);
}

#endif
