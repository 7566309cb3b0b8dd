// dense.hip -- explicit covariance operators on the GPU.
//
//  * k_dense_A: the dense n x n matrix A of lib/wf_vector.cc:16-40 (SN) and
//    :157-183 (W), element by element with the table interpolant or the
//    exact exponential, for the closed-form solve (solve_cf,
//    lib/wf_vector.cc:123-155, :276-319; the reference factors it with
//    Eigen's LDLT, here rocSOLVER's Cholesky on the device).
//  * k_column: one column of S_pm for a map point (lib/map_covar.cc:47-61).
//  * k_dot1: a single-block FP64 dot product in a fixed order
//    (lib/map_covar.cc:95-109).
//  * k_fill: vector fill.
// COMPILE WITH -ffp-contract=off (the element arithmetic is the
// reference's).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "cg_device.h"
#include "ds_internal.h"

namespace ds {

namespace {

struct ElemArgs {
  double var_s, c_perp, c_para, dx_inv;
  const double* table;
  int table_n;
};

// lib/signal_covar.h:53-86 on one pair of coordinates (FP64).
__device__ __forceinline__ double pair_covar(double dx, double dy, double dz,
                                             const ElemArgs& e, int exact) {
  const double x_perp_2 = (dx * dx + dy * dy);
  const double x_para_2 = dz * dz;
  const double ga_perp = e.c_perp * x_perp_2;
  const double ga_para = e.c_para * x_para_2;
  if (exact) return e.var_s * exp(-(ga_perp + ga_para));
  const double xx = (ga_perp + ga_para) * e.dx_inv;
  if (!(xx < (double)(e.table_n - 1))) return 0.0;
  const int ix = (int)xx;
  const double gp = e.table[ix] * ((double)(ix + 1) - xx) +
                    e.table[ix + 1] * (xx - (double)ix);
  return e.var_s * gp;
}

__global__ __launch_bounds__(256) void k_dense_A(const double4* pos, int n,
                                                 int form, int exact,
                                                 ElemArgs e, double* A,
                                                 size_t ld) {
  const int i = blockIdx.x;   // rows on x: grid y is limited to 65536
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double4 pi = pos[i], pj = pos[j];
  const double S = pair_covar(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, e, exact);
  const double d = (i == j) ? 1.0 : 0.0;
  A[(size_t)i * ld + j] = (form == DS_FORM_W) ? pi.w * S * pj.w + d
                                              : S + pi.w * d;
}

__global__ __launch_bounds__(256) void k_column(const double4* pos, int n,
                                                int npad, double mx, double my,
                                                double mz, ElemArgs e,
                                                double* col) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npad) return;
  if (k >= n) {
    col[k] = 0.0;
    return;
  }
  const double4 p = pos[k];
  // map point minus pixel, lib/map_covar.cc:50-55
  col[k] = pair_covar(mx - p.x, my - p.y, mz - p.z, e, 0);
}

__global__ __launch_bounds__(1024) void k_dot1(const double* a,
                                               const double* b, int n,
                                               double* out) {
  __shared__ double red[16];
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[0] += a[i] * b[i];
  block_sum<1>(v, red);
  if (threadIdx.x == 0) *out = v[0];
}

__global__ __launch_bounds__(256) void k_fill(double* v, double value, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = value;
}

ElemArgs elem_args(const ds_problem* p, const double* table_dev) {
  ElemArgs e;
  e.var_s = p->cv.var_s;
  e.c_perp = p->cv.c_perp;
  e.c_para = p->cv.c_para;
  e.dx_inv = p->cv.dx_inv;
  e.table = table_dev;
  e.table_n = p->cv.table_n;
  return e;
}

}  // namespace

hipError_t launch_fill(double* v, double value, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, st, v, value,
                     n);
  return hipGetLastError();
}

hipError_t launch_column(const ds_problem* p, const double* table_dev,
                         const double* map_pt, double* col, hipStream_t st) {
  hipLaunchKernelGGL(k_column, dim3((p->npad + 255) / 256), dim3(256), 0, st,
                     p->pos64, p->n, p->npad, map_pt[0], map_pt[1], map_pt[2],
                     elem_args(p, table_dev), col);
  return hipGetLastError();
}

hipError_t launch_column_sub(const ds_problem* p, int k,
                             const double* table_dev, const double* map_pt,
                             double* col, hipStream_t st) {
  const int b0 = p->sub_blk0_h[k], b1 = p->sub_blk0_h[k + 1];
  const int npad1 = (b1 - b0) * DS_BS;
  hipLaunchKernelGGL(k_column, dim3((npad1 + 255) / 256), dim3(256), 0, st,
                     p->pos64 + (size_t)b0 * DS_BS, p->sub_n[k], npad1,
                     map_pt[0], map_pt[1], map_pt[2], elem_args(p, table_dev),
                     col);
  return hipGetLastError();
}

hipError_t launch_dot1(const double* a, const double* b, int n, double* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_dot1, dim3(1), dim3(1024), 0, st, a, b, n, out);
  return hipGetLastError();
}

// Dense solve of A y = rhs in sorted order (first n slots), y into p->x.
// Returns a HIP error, or hipErrorInvalidValue with *info > 0 when A is not
// positive definite.
hipError_t dense_solve(ds_problem* p, int exact, const double* table_dev,
                       int* info_out) {
  hipStream_t st = p->ctx->stream;
  const int n = p->n;
  const size_t ld = (size_t)n;
  double* A = nullptr;
  int* info = nullptr;
  hipError_t e = hipMalloc(&A, sizeof(double) * ld * n);
  if (e != hipSuccess) return e;
  e = hipMalloc(&info, sizeof(int));
  if (e != hipSuccess) {
    (void)hipFree(A);
    return e;
  }
  rocblas_handle h = nullptr;
  rocblas_status rs = rocblas_create_handle(&h);
  if (rs == rocblas_status_success) rs = rocblas_set_stream(h, st);
  hipLaunchKernelGGL(k_dense_A, dim3(n, (n + 255) / 256), dim3(256), 0, st,
                     p->pos64, n, p->form, exact, elem_args(p, table_dev), A,
                     ld);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->x, p->rhs64, sizeof(double) * n,
                       hipMemcpyDeviceToDevice, st);
  // Row-major A is its own transpose up to the rounding of (w_i S) w_j; the
  // factorisation reads one triangle.
  if (e == hipSuccess && rs == rocblas_status_success)
    rs = rocsolver_dpotrf(h, rocblas_fill_lower, n, A, n, info);
  int hinfo = 0;
  if (e == hipSuccess && rs == rocblas_status_success)
    e = hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess && rs == rocblas_status_success && hinfo == 0)
    rs = rocsolver_dpotrs(h, rocblas_fill_lower, n, 1, A, n, p->x, n);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (h) rocblas_destroy_handle(h);
  (void)hipFree(A);
  (void)hipFree(info);
  *info_out = hinfo;
  if (e != hipSuccess) return e;
  if (rs != rocblas_status_success) return hipErrorUnknown;
  if (hinfo != 0) return hipErrorInvalidValue;
  return hipSuccess;
}

}  // namespace ds
