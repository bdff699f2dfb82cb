// Smallest singular value of upper-triangular factors, one workgroup per
// problem (the sigma_min normalisation of the --scale 0 sparsification: the
// QR of Ap = [App; Ampc] leaves R with the singular values of Ap).
//
//   sigma_min(R)^2 = 1 / lambda_max((R^T R)^{-1})
//
// lambda_max comes from Lanczos with full reorthogonalisation on the operator
// v -> R^{-1} R^{-T} v (a forward and a backward blocked triangular solve),
// and the largest Ritz value of the Lanczos tridiagonal by multisection with
// Sturm counts.  Lanczos runs until the Ritz value is stationary to 4 ulps
// over three steps, the basis spans the space (k steps) or kSminSteps steps.
// Everything is computed in FP64 (R of either precision is read as double).
#include <cfloat>

#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

constexpr int kSminThreads = 256;
constexpr int kSminNB = 32;  // triangular block

// Sturm count: number of eigenvalues of the tridiagonal (a[0..n), b[0..n-1))
// below x.
__device__ __forceinline__ int sturm_below(const double* a, const double* b, int n, double x) {
  int cnt = 0;
  double d = a[0] - x;
  if (d < 0) ++cnt;
  for (int i = 1; i < n; ++i) {
    if (d == 0.0) d = -DBL_MIN * 1e3;
    d = (a[i] - x) - b[i - 1] * b[i - 1] / d;
    if (d < 0) ++cnt;
  }
  return cnt;
}

// Largest eigenvalue of the tridiagonal by multisection on one wavefront
// (64 Sturm counts per round); the result is broadcast to every lane.
__device__ double tridiag_max_eig(const double* a, const double* b, int n) {
  const int lane = threadIdx.x & (kWave - 1);
  double lo = a[0], hi = a[0];
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i + 1 < n ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  // invariant: count(lo) < n, count(hi) == n (the largest eigenvalue is in [lo, hi])
  hi = hi + fabs(hi) * 4 * DBL_EPSILON + DBL_MIN;
  for (int it = 0; it < 40; ++it) {
    if (hi - lo <= 2 * DBL_EPSILON * fmax(fabs(lo), fabs(hi))) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) / (double)(kWave + 1);
    const bool all_below = sturm_below(a, b, n, x) >= n;
    const unsigned long long m = __ballot(all_below);
    // lowest lane whose point is above the largest eigenvalue
    const int first = m ? __ffsll((long long)m) - 1 : kWave;
    const double nlo = first == 0 ? lo : lo + (hi - lo) * (double)first / (double)(kWave + 1);
    const double nhi = first == kWave ? hi : lo + (hi - lo) * (double)(first + 1) / (double)(kWave + 1);
    lo = nlo;
    hi = nhi;
  }
  return 0.5 * (lo + hi);
}

// x <- R^{-T} x (trans) or R^{-1} x, x in LDS (double), R k x k upper.
template <typename T>
__device__ void tri_solve(const T* __restrict__ R, int64_t ld, int k, double* x, int trans) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int nblk = (k + kSminNB - 1) / kSminNB;
  for (int t = 0; t < nblk; ++t) {
    const int bi = trans ? t : nblk - 1 - t;
    const int jb = bi * kSminNB, w = min(kSminNB, k - jb);
    if (wv == 0) {
      // diagonal block by one wavefront: lane l holds x[jb + l]
      double v = lane < w ? x[jb + lane] : 0.0;
      if (trans) {
        for (int c = 0; c < w; ++c) {
          if (lane == c) v /= (double)R[(jb + c) + (jb + c) * ld];
          const double xc = __shfl(v, c, kWave);
          if (lane > c && lane < w) v -= (double)R[(jb + c) + (jb + lane) * ld] * xc;
        }
      } else {
        for (int c = w - 1; c >= 0; --c) {
          if (lane == c) v /= (double)R[(jb + c) + (jb + c) * ld];
          const double xc = __shfl(v, c, kWave);
          if (lane < c) v -= (double)R[(jb + lane) + (jb + c) * ld] * xc;
        }
      }
      if (lane < w) x[jb + lane] = v;
    }
    __syncthreads();
    if (trans) {
      // x[r] -= sum_c R[jb + c, r] x[jb + c] for r past the block: a half-wave
      // per r reads the contiguous 32-entry segment of column r
      const int hl = tid & (kSminNB - 1), hw = tid / kSminNB;
      for (int rb = jb + w; rb < k; rb += kSminThreads / kSminNB) {
        const int r = rb + hw;
        double v = (r < k && hl < w) ? (double)R[(jb + hl) + r * ld] * x[jb + hl] : 0.0;
#pragma unroll
        for (int o = kSminNB / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        if (hl == 0 && r < k) x[r] -= v;
      }
    } else {
      for (int r = tid; r < jb; r += kSminThreads) {
        double s = 0;
        for (int c = 0; c < w; ++c) s += (double)R[r + (jb + c) * ld] * x[jb + c];
        x[r] -= s;
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ double block_dot(const double* u, const double* v, int k, double* red) {
  double s = 0;
  for (int i = threadIdx.x; i < k; i += kSminThreads) s += u[i] * v[i];
  return block_sum(s, red);
}

template <typename T>
__global__ __launch_bounds__(kSminThreads) void k_smin(const SminDesc<T>* __restrict__ descs) {
  const SminDesc<T> d = descs[blockIdx.x];
  const int k = d.k, tid = threadIdx.x;
  extern __shared__ double sm[];
  double* x = sm;                  // k: operator input / output
  __shared__ double red[kSminThreads / kWave];
  __shared__ double al[kSminSteps], be[kSminSteps], hv[kSminSteps];
  __shared__ double s_theta;
  __shared__ int s_stop;
  double* Q = d.work;  // k x (kSminSteps + 1) Lanczos basis, ld k
  const int64_t ld = d.ld;
  if (k <= 0) {
    if (tid == 0) *d.out = 0.0;
    return;
  }
  // start vector: fixed, dense, without symmetry
  for (int i = tid; i < k; i += kSminThreads) {
    const unsigned int h = (unsigned int)(i + 1) * 2654435761u;
    x[i] = 0.5 + (double)(h >> 8) / (double)(1u << 24);
  }
  __syncthreads();
  {
    const double nn = sqrt(block_dot(x, x, k, red));
    for (int i = tid; i < k; i += kSminThreads) Q[i] = x[i] / nn;
  }
  const int mmax = min(k, kSminSteps);
  double theta = 0, prev = -1, prev2 = -2;
  int m = 0;
  for (int j = 0; j < mmax; ++j) {
    const double* qj = Q + (int64_t)j * k;
    for (int i = tid; i < k; i += kSminThreads) x[i] = qj[i];
    __syncthreads();
    tri_solve(d.R, ld, k, x, 1);
    tri_solve(d.R, ld, k, x, 0);
    const double a = block_dot(qj, x, k, red);
    // full reorthogonalisation against q_0 .. q_j, classical Gram-Schmidt
    // twice: one wavefront per dot product, then one update sweep
    for (int pass = 0; pass < 2; ++pass) {
      for (int i = tid / kWave; i <= j; i += kSminThreads / kWave) {
        const double* qi = Q + (int64_t)i * k;
        double sdot = 0;
        for (int r = tid & (kWave - 1); r < k; r += kWave) sdot += qi[r] * x[r];
        sdot = wave_sum(sdot);
        if ((tid & (kWave - 1)) == 0) hv[i] = sdot;
      }
      __syncthreads();
      for (int r = tid; r < k; r += kSminThreads) {
        double acc = 0;
        for (int i = 0; i <= j; ++i) acc += hv[i] * Q[r + (int64_t)i * k];
        x[r] -= acc;
      }
      __syncthreads();
    }
    const double b = sqrt(block_dot(x, x, k, red));
    if (tid == 0) { al[j] = a; be[j] = b; }
    __syncthreads();
    m = j + 1;
    if (tid < kWave) {
      const double th = tridiag_max_eig(al, be, m);
      if (tid == 0) s_theta = th;
    }
    __syncthreads();
    theta = s_theta;
    const bool flat = fabs(theta - prev) <= 4 * DBL_EPSILON * theta && fabs(prev - prev2) <= 4 * DBL_EPSILON * theta;
    if (tid == 0) s_stop = (flat || !(b > 1e-14 * theta) || m == mmax) ? 1 : 0;
    __syncthreads();
    if (s_stop) break;
    prev2 = prev;
    prev = theta;
    double* qn = Q + (int64_t)(j + 1) * k;
    for (int i = tid; i < k; i += kSminThreads) qn[i] = x[i] / b;
    __syncthreads();
  }
  if (tid == 0) *d.out = (theta > 0 && theta == theta) ? 1.0 / sqrt(theta) : 0.0;
}

int smin_max_k() { return kSminMaxK; }
size_t smin_work_elems(int k) { return (size_t)k * (kSminSteps + 1); }

template <typename T>
void launch_smin(hipStream_t s, const SminDesc<T>* d, int nd, int kmax) {
  if (nd <= 0) return;
  hipLaunchKernelGGL((k_smin<T>), dim3(nd), dim3(kSminThreads), sizeof(double) * (size_t)std::max(kmax, 1), s, d);
}
template void launch_smin<double>(hipStream_t, const SminDesc<double>*, int, int);
template void launch_smin<float>(hipStream_t, const SminDesc<float>*, int, int);

// C[r0:r1, 0:cols] *= 1 / s[p]  (the (1/s) App^T Apmc block of --scale 0)
template <typename T>
__global__ __launch_bounds__(256) void k_scale_inv(const ScaleInvDesc<T>* __restrict__ descs,
                                                   const int* __restrict__ tstart, int nd) {
  const int p = find_owner(tstart, nd, blockIdx.x);
  const ScaleInvDesc<T> d = descs[p];
  const int64_t e = (int64_t)(blockIdx.x - tstart[p]) * 256 + threadIdx.x;
  const int rows = d.r1 - d.r0;
  if (e >= (int64_t)rows * d.cols) return;
  const int i = d.r0 + (int)(e % rows), j = (int)(e / rows);
  const T inv = (T)(1.0 / *d.s);
  T* c = d.C + i + (int64_t)j * d.ld;
  *c = inv * *c;
}
template <typename T>
void launch_scale_inv(hipStream_t s, const ScaleInvDesc<T>* d, const int* tstart, int nd, int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL((k_scale_inv<T>), dim3(ntiles), dim3(256), 0, s, d, tstart, nd);
}
template void launch_scale_inv<double>(hipStream_t, const ScaleInvDesc<double>*, const int*, int, int);
template void launch_scale_inv<float>(hipStream_t, const ScaleInvDesc<float>*, const int*, int, int);

}  // namespace spq
