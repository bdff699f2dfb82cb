// Householder panel kernels (gfx950).
//
// k_panel: factors one NB-wide column block [jb, jb+w) of every panel in the
// batch (one workgroup per panel): dlarfg-convention reflectors
// (beta = -sign(alpha) * ||x||, tau = (beta - alpha) / beta, v(0) = 1), the
// reflector applied to the remaining columns of the block, and the block's
// compact-WY factor T (forward, columnwise: T(0:j,j) = -tau_j T(0:j,0:j) V^T v_j).
// Each column step needs two workgroup reductions: the norm of the column
// below the diagonal and the NB dot products of v_j with every column of the
// block (those against earlier reflectors feed T, those against later columns
// feed the update).  The caller applies the block to the trailing columns with
// the batched GEMM.
#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

constexpr int kNB = 32;
constexpr int kThreads = 256;

template <typename T>
__device__ __forceinline__ void tcol_from_dots(T* Ts, int jj, T tau, const T* g) {
  // Ts is NB x NB (ld NB) in LDS; column jj from dots g[0..jj)
  if ((int)threadIdx.x < jj) {
    const int r = threadIdx.x;
    T s = 0;
    for (int l = r; l < jj; ++l) s += Ts[r + l * kNB] * g[l];
    Ts[r + jj * kNB] = -tau * s;
  }
  if ((int)threadIdx.x == jj) Ts[jj + jj * kNB] = tau;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_panel(const PanelDesc<T>* __restrict__ descs, int jb) {
  const PanelDesc<T> d = descs[blockIdx.x];
  const int m = d.m;
  const int kk = min(d.m, d.k);
  if (jb >= kk) return;
  const int w = min(kNB, kk - jb);
  const int ncb = min(kNB, d.k - jb);
  T* __restrict__ P = d.P;
  const int64_t ld = d.ld;
  const int tid = threadIdx.x;

  __shared__ T red[kNB * (kThreads / kWave)];
  __shared__ T Ts[kNB * kNB];
  __shared__ T gsh[kNB];

  for (int i = tid; i < kNB * kNB; i += kThreads) Ts[i] = T(0);

  for (int j = jb; j < jb + w; ++j) {
    const int jj = j - jb;
    T* __restrict__ pj = P + (int64_t)j * ld;
    // (1) squared norm of the column below the diagonal
    T s = 0;
    for (int i = j + 1 + tid; i < m; i += kThreads) { const T v = pj[i]; s += v * v; }
    s = block_sum(s, red);
    const T alpha = pj[j];
    T tau = 0, beta = alpha, scal = 0;
    if (s > T(0)) {
      const T xn = sqrt(s);
      beta = -copysign(hypot(alpha, xn), alpha);
      tau = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
      for (int i = j + 1 + tid; i < m; i += kThreads) pj[i] *= scal;
    }
    // row j of the block, read before anyone updates it
    T rowj[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) rowj[c] = (c < ncb && c != jj) ? P[j + (jb + c) * ld] : T(0);
    __syncthreads();
    // (2) dots of v_j with every column of the block (rows > j)
    T acc[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) acc[c] = T(0);
    for (int i = j + 1 + tid; i < m; i += kThreads) {
      const T vi = pj[i];
#pragma unroll
      for (int c = 0; c < kNB; ++c)
        if (c < ncb && c != jj) acc[c] += vi * P[i + (jb + c) * ld];
    }
    block_sum_n<T, kNB>(acc, red, ncb);
    if (tid == 0) { pj[j] = beta; d.tau[j] = tau; }
    // (3) apply H_j to the later columns of the block
    for (int c = jj + 1; c < ncb; ++c) {
      const T wc = tau * (rowj[c] + acc[c]);
      if (wc == T(0)) continue;
      T* __restrict__ pc = P + (int64_t)(jb + c) * ld;
      if (tid == 0) pc[j] -= wc;
      for (int i = j + 1 + tid; i < m; i += kThreads) pc[i] -= wc * pj[i];
    }
    // (4) T column from the dots against earlier reflectors
    if (tid < kNB) gsh[tid] = (tid < jj) ? (rowj[tid] + acc[tid]) : T(0);
    __syncthreads();
    tcol_from_dots(Ts, jj, tau, gsh);
    __syncthreads();
  }
  T* __restrict__ Tb = d.Tb + (int64_t)jb * kNB;
  for (int i = tid; i < kNB * kNB; i += kThreads) Tb[i] = Ts[i];
}

// Block T factors for an existing reflector set (QP3 output): one workgroup
// per (descriptor, NB block).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_larft(const LarftDesc<T>* __restrict__ descs, const int* __restrict__ tstart,
                                                    int nd) {
  const int t = blockIdx.x;
  const int p = find_owner(tstart, nd, t);
  const LarftDesc<T> d = descs[p];
  const int jb = (t - tstart[p]) * kNB;
  const int w = min(kNB, d.r - jb);
  const int m = d.m, tid = threadIdx.x;
  const int64_t ld = d.ldv;
  __shared__ T red[kNB * (kThreads / kWave)];
  __shared__ T Ts[kNB * kNB];
  __shared__ T gsh[kNB];
  for (int i = tid; i < kNB * kNB; i += kThreads) Ts[i] = T(0);
  __syncthreads();
  for (int jj = 0; jj < w; ++jj) {
    const int j = jb + jj;
    const T* __restrict__ vj = d.V + (int64_t)j * ld;
    T acc[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) acc[c] = T(0);
    // v_c . v_j for c < jj: v_c(j) * 1 + sum_{i>j} v_c(i) v_j(i)
    for (int i = j + 1 + tid; i < m; i += kThreads) {
      const T vi = vj[i];
#pragma unroll
      for (int c = 0; c < kNB; ++c)
        if (c < jj) acc[c] += vi * d.V[i + (int64_t)(jb + c) * ld];
    }
    block_sum_n<T, kNB>(acc, red, jj);
    if (tid < kNB) gsh[tid] = (tid < jj) ? (d.V[j + (int64_t)(jb + tid) * ld] + acc[tid]) : T(0);
    __syncthreads();
    tcol_from_dots(Ts, jj, d.tau[j], gsh);
    __syncthreads();
  }
  T* __restrict__ Tb = d.Tb + (int64_t)jb * kNB;
  for (int i = tid; i < kNB * kNB; i += kThreads) Tb[i] = Ts[i];
}

// B[:, jb:jb+w) <- B[:, jb:jb+w) * inv(R[jb:jb+w, jb:jb+w]) ; one row per thread.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_trsm_block(const TrsmDesc<T>* __restrict__ descs,
                                                         const int* __restrict__ tstart, int nd, int jb) {
  const int t = blockIdx.x;
  const int p = find_owner(tstart, nd, t);
  const TrsmDesc<T> d = descs[p];
  if (jb >= d.k) return;
  const int w = min(kNB, d.k - jb);
  __shared__ T Rs[kNB * kNB];
  for (int i = threadIdx.x; i < kNB * kNB; i += kThreads) {
    const int r = i % kNB, c = i / kNB;
    Rs[i] = (r < w && c < w) ? d.R[(jb + r) + (int64_t)(jb + c) * d.ldr] : T(0);
  }
  __syncthreads();
  const int row = (t - tstart[p]) * kThreads + threadIdx.x;
  if (row >= d.rows) return;
  T x[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) x[c] = (c < w) ? d.B[row + (int64_t)(jb + c) * d.ldb] : T(0);
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (c < w) {
      T s = x[c];
#pragma unroll
      for (int l = 0; l < kNB; ++l)
        if (l < c) s -= x[l] * Rs[l + c * kNB];
      x[c] = s / Rs[c + c * kNB];
    }
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c)
    if (c < w) d.B[row + (int64_t)(jb + c) * d.ldb] = x[c];
}

int panel_nb() { return kNB; }

template <typename T>
void launch_panel(hipStream_t s, const PanelDesc<T>* d, int nd, int jb) {
  if (nd <= 0) return;
  hipLaunchKernelGGL((k_panel<T>), dim3(nd), dim3(kThreads), 0, s, d, jb);
}
template <typename T>
void launch_larft(hipStream_t s, const LarftDesc<T>* d, const int* tstart, int nd, int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL((k_larft<T>), dim3(ntiles), dim3(kThreads), 0, s, d, tstart, nd);
}
template <typename T>
void launch_trsm_block(hipStream_t s, const TrsmDesc<T>* d, const int* tstart, int nd, int ntiles, int jb) {
  if (nd <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL((k_trsm_block<T>), dim3(ntiles), dim3(kThreads), 0, s, d, tstart, nd, jb);
}
int trsm_rows_per_tile() { return kThreads; }

template void launch_panel<double>(hipStream_t, const PanelDesc<double>*, int, int);
template void launch_panel<float>(hipStream_t, const PanelDesc<float>*, int, int);
template void launch_larft<double>(hipStream_t, const LarftDesc<double>*, const int*, int, int);
template void launch_larft<float>(hipStream_t, const LarftDesc<float>*, const int*, int, int);
template void launch_trsm_block<double>(hipStream_t, const TrsmDesc<double>*, const int*, int, int, int);
template void launch_trsm_block<float>(hipStream_t, const TrsmDesc<float>*, const int*, int, int, int);

}  // namespace spq
