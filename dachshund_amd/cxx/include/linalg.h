// linalg.h -- drop-in of lib/linalg.h: serial FP64 vector helpers for host
// vectors (lib/linalg.cc:9-49; add_cv/sub_cv compute c += a b / c -= a b).
#ifndef DS_DROPIN_LINALG_H
#define DS_DROPIN_LINALG_H

double ds_vector_dot(const int n, const double* const a,
                     const double* const b);
double ds_vector_norm(const int n, const double* const x);
void ds_vector_add_cv(const int n, const double a, const double* const b,
                      double* const c);
void ds_vector_sub_cv(const int n, const double a, const double* const b,
                      double* const c);

#endif
