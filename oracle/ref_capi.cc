// ref_capi.cc -- a C surface over the reference library, built from the
// reference's own sources into oracle/_ref/libdsref.so.  TEST AND CPU
// TIMING INFRASTRUCTURE ONLY: the benchmark times these calls as the
// reference OpenMP CPU path; the product never loads this library.
//
// Pixels are reference WPixel records {x,y,z,w,wd} (lib/pixel.h:39-45),
// points are {x,y,z} (lib/pixel.h:8-12).
#include <cstring>
#include <vector>
#if defined(_OPENMP)
#include <omp.h>
#endif

#include "dachshund.h"

extern "C" {

// lib/wf_vector.cc:196-220
void dsref_w_A_product(int n, const double* wpix, double var_s, double l_perp,
                       double l_para, const double* a, double* b) {
  SignalCovarParams p(var_s, l_perp, l_para);
  WFVectorW wf(n, (const WPixel*)wpix, &p);
  wf.A_product(a, b);
}

// lib/wf_vector.cc:222-251
void dsref_w_A_residual(int n, const double* wpix, double var_s,
                        double l_perp, double l_para, const double* x,
                        double* r) {
  SignalCovarParams p(var_s, l_perp, l_para);
  WFVectorW wf(n, (const WPixel*)wpix, &p);
  wf.A_residual(x, r);
}

// lib/map.cc:20-43
void dsref_smp_product(int n_pix, const double* pts, int n_map,
                       const double* map_pts, double var_s, double l_perp,
                       double l_para, const double* x, double* m) {
  SignalCovarParams p(var_s, l_perp, l_para);
  smp_product(n_pix, (const Point*)pts, n_map, (const Point*)map_pts, &p, x,
              m);
}

// lib/wf_vector.cc:263-274 over lib/pcg.h:67-160.  Returns num_iters;
// *success set; x in/out.
int dsref_w_solve_pcg(int n, const double* wpix, double var_s, double l_perp,
                      double l_para, int max_iter, int step_r, double tol,
                      double* x, int* success) {
  SignalCovarParams p(var_s, l_perp, l_para);
  WFVectorW wf(n, (const WPixel*)wpix, &p);
  PCGParams pp;
  pp.n = n;
  pp.max_iter = max_iter;
  pp.step_r = step_r;
  pp.tol = tol;
  pp.verbose = false;
  PCGResult r = wf.solve_pcg(pp, x);
  if (success) *success = r.success ? 1 : 0;
  return r.num_iters;
}

int dsref_num_threads(void) {
#if defined(_OPENMP)
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
