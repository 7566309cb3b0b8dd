// ds_internal.h -- shared types of the gfx950 Wiener-filter library.
//
// Layout in HBM (all per problem, in *sorted* pixel order):
//   pos64   double4[npad]  {x, y, z, w}   w = 1/n (W form) or N = n^2 (SN form)
//   rhs64   double[npad]   wd (W form) or d (SN form)
//   loc32   float4[npad]   pre-scaled coordinates relative to the block centre
//   blk_c   double4[nblk]  block centre in pre-scaled coordinates
//   pairs   CSR (pair_ptr[nblk+1], pair_idx[npairs]) of source blocks per
//           target block that survive the exact covariance cutoff
// Pixels are sorted by a k-d tree into blocks of DS_BS = 64 pixels (one
// wavefront of target rows); the tail block is padded with copies of the
// last pixel whose vector entries are kept at zero.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dachshund_gpu.h"

#define DS_BS 64
#define DS_MAX_TABLE 1024

namespace ds {

// Covariance parameters as the kernels see them.
struct CovarDev {
  double var_s;
  double c_perp;
  double c_para;
  double dx_inv;
  int table_n;
  float s_perp;   // sqrt(c_perp * dx_inv): scaled perpendicular coordinate
  float s_para;   // sqrt(c_para * dx_inv)
};

// Host-side result of spatial blocking.
struct Blocking {
  int n = 0;
  int nblk = 0;
  int npad = 0;
  std::vector<int> perm;        // sorted slot -> original index (npad, tail = last)
  std::vector<double> lo, hi;   // nblk * 3, physical AABB of the real pixels
};

// Sort points (stride doubles apart, x,y,z first) into k-d blocks of DS_BS.
// scale[3] is the metric used to choose split axes (isotropic in it).
void build_blocking(const double* pts, int stride, int n, const double scale[3],
                    Blocking& out);

// Exact culling: keep (I, J) unless every pair between the two AABBs has
// table index >= table_n - 1 under the reference's own arithmetic.
void build_pair_lists(const Blocking& tgt, const Blocking& src,
                      const CovarDev& cv, std::vector<int>& ptr,
                      std::vector<int>& idx);
// Same with targets given directly as AABBs (map tiles).
void build_pair_lists_boxes(int ntgt, const double* tlo, const double* thi,
                            const Blocking& src, const CovarDev& cv,
                            std::vector<int>& ptr, std::vector<int>& idx);

// Lower bound of the table index for any pair drawn from two boxes, with the
// operation order of lib/signal_covar.h:67-77 and lib/wf_vector.cc:62-66.
double box_gap_xx(const double* alo, const double* ahi, const double* blo,
                  const double* bhi, const CovarDev& cv);

}  // namespace ds

// Device-resident scalars of one CG solve.
struct ds_cg_scalars {
  double delta;        // r . M^-1 r
  double alpha;
  double beta;
  double norm_r;       // |r|_2 of the current residual
  double tol_norm_b;
  double max_r;        // mixed solver: largest |r| since the last true residual
  int iter;            // current iteration number (reference numbering)
  int done;            // 1 once converged or out of iterations
  int success;
  int num_iters;
  int n_norms;
  int full_residual;   // this iteration recomputes r = b - A x
  int n_true;          // mixed solver: number of FP64 true residuals
  int pad;
};

struct ds_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int num_cu = 0;
};

struct ds_problem {
  ds_ctx* ctx = nullptr;
  int form = DS_FORM_W;
  int n = 0, nblk = 0, npad = 0;
  ds::CovarDev cv{};
  std::vector<double> table;           // host copy, table_n entries
  ds::Blocking blk;
  std::vector<int> pair_ptr_h;
  long long npairs = 0;                // block pairs after culling
  // device buffers
  double4* pos64 = nullptr;
  double* rhs64 = nullptr;
  float4* loc32 = nullptr;
  double4* blk_c = nullptr;
  int* pair_ptr = nullptr;
  int* pair_idx = nullptr;
  double2* table64 = nullptr;          // {T[i], T[i+1]}, last = {0, 0}
  float2* table32 = nullptr;           // {var_s T[i], var_s (T[i+1]-T[i])}
  int* perm = nullptr;                 // sorted -> original
  // CG work vectors (npad each)
  double *x = nullptr, *r = nullptr, *d = nullptr, *s = nullptr, *q = nullptr,
         *v = nullptr, *tmp = nullptr;
  float* v32 = nullptr;
  double* partials = nullptr;          // 4 * nblk
  ds_cg_scalars* sc = nullptr;
  double* norms = nullptr;             // device residual history
  int norms_cap = 0;
  ds_cg_scalars* sc_host = nullptr;    // pinned mirror
};
