// FP64 kernels for the Krylov drivers: CSR SpMV, Gram-Schmidt GEMVs and
// vector updates.  Reductions are deterministic (fixed workgroup order, no
// atomics): one workgroup per output dot product.
#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

// y = alpha * A x + beta * y  (CSR, one wavefront per row)
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                              const double* __restrict__ v, const double* __restrict__ x,
                                              double* __restrict__ y, double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  double s = 0;
  for (int p = rp[row] + lane; p < rp[row + 1]; p += 64) s += v[p] * x[ci[p]];
  s = wave_sum(s);
  if (lane == 0) y[row] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[row]);
}

// out[j] = V(:, j) . w  for j < ncol (V is n x ncol, ld)
// out[j] = V(:, j) . w for j < ncol, in two passes: blocks (b, j) sum rows
// [b*chunk, (b+1)*chunk) into part[j*nb + b], then one wave per column adds
// the nb partials in a fixed butterfly (the result does not depend on
// scheduling).
constexpr int kDotParts = 64;
__global__ __launch_bounds__(256) void k_dots_part(int n, int chunk, int nb, const double* __restrict__ V, int64_t ld,
                                                   const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double red[4];
  const int b = blockIdx.x, j = blockIdx.y;
  const double* __restrict__ vj = V + (int64_t)j * ld;
  const int lo = b * chunk, hi = min(n, lo + chunk);
  double s = 0;
  for (int i = lo + threadIdx.x; i < hi; i += 256) s += vj[i] * w[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[(int64_t)j * nb + b] = s;
}
__global__ __launch_bounds__(64) void k_dots_fin(int nb, const double* __restrict__ part, double* __restrict__ out) {
  const int j = blockIdx.x, l = threadIdx.x;
  double s = l < nb ? part[(int64_t)j * nb + l] : 0.0;
  s = wave_sum(s);
  if (l == 0) out[j] = s;
}

// w -= V(:, 0:ncol) h
__global__ __launch_bounds__(256) void k_gemv_sub(int n, int ncol, const double* __restrict__ V, int64_t ld,
                                                  const double* __restrict__ h, double* __restrict__ w) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int j = 0; j < ncol; ++j) s += V[i + (int64_t)j * ld] * h[j];
  w[i] -= s;
}

// y = a*x + b*y
__global__ void k_axpby(int n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

// y = V(:, 0:ncol) h   (overwrite)
__global__ __launch_bounds__(256) void k_gemv(int n, int ncol, const double* __restrict__ V, int64_t ld,
                                              const double* __restrict__ h, double* __restrict__ y) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int j = 0; j < ncol; ++j) s += V[i + (int64_t)j * ld] * h[j];
  y[i] = s;
}

// out[k] = alpha * u[rowind[k]] * v[j] for the nonzeros k of column j of a
// CSC pattern (one thread per nonzero; j by binary search in colptr).  With
// u = lambda, v = x and alpha = -1 this is dL/dA on A's pattern for
// x = A^{-1} b and A^T lambda = dL/dx.
__global__ __launch_bounds__(256) void k_csc_outer(int n, int nnz, const int* __restrict__ colptr,
                                                   const int* __restrict__ rowind, double alpha,
                                                   const double* __restrict__ u, const double* __restrict__ v,
                                                   double* __restrict__ out) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nnz) return;
  int lo = 0, hi = n - 1;  // last column with colptr[j] <= k
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (colptr[mid] <= k) lo = mid; else hi = mid - 1;
  }
  out[k] = alpha * u[rowind[k]] * v[lo];
}
void launch_csc_outer(hipStream_t s, int n, int nnz, const int* colptr, const int* rowind, double alpha,
                      const double* u, const double* v, double* out) {
  if (n <= 0 || nnz <= 0) return;
  hipLaunchKernelGGL(k_csc_outer, dim3((nnz + 255) / 256), dim3(256), 0, s, n, nnz, colptr, rowind, alpha, u, v, out);
}

void launch_spmv(hipStream_t s, int n, const int* rp, const int* ci, const double* v, const double* x, double* y,
                 double alpha, double beta) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_spmv, dim3((n + 3) / 4), dim3(256), 0, s, n, rp, ci, v, x, y, alpha, beta);
}
size_t dots_part_elems(int ncol) { return (size_t)std::max(ncol, 1) * kDotParts; }
void launch_dots(hipStream_t s, int n, int ncol, const double* V, int64_t ld, const double* w, double* out,
                 double* part) {
  if (ncol <= 0) return;
  const int nb = std::max(1, std::min(kDotParts, (n + 4095) / 4096));
  const int chunk = std::max(1, (n + nb - 1) / nb);
  hipLaunchKernelGGL(k_dots_part, dim3(nb, ncol), dim3(256), 0, s, n, chunk, nb, V, ld, w, part);
  hipLaunchKernelGGL(k_dots_fin, dim3(ncol), dim3(64), 0, s, nb, part, out);
}
void launch_gemv_sub(hipStream_t s, int n, int ncol, const double* V, int64_t ld, const double* h, double* w) {
  if (n <= 0 || ncol <= 0) return;
  hipLaunchKernelGGL(k_gemv_sub, dim3((n + 255) / 256), dim3(256), 0, s, n, ncol, V, ld, h, w);
}
void launch_gemv(hipStream_t s, int n, int ncol, const double* V, int64_t ld, const double* h, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gemv, dim3((n + 255) / 256), dim3(256), 0, s, n, ncol, V, ld, h, y);
}
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3((n + 255) / 256), dim3(256), 0, s, n, a, x, b, y);
}

}  // namespace spq
