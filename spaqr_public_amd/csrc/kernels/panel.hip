// Householder panel kernels (gfx950).
//
// k_panel_r: factors one NB-wide column block [jb, jb+w) of every panel in
// the batch: dlarfg-convention reflectors (beta = -sign(alpha) * ||x||,
// tau = (beta - alpha) / beta, v(0) = 1), the reflector applied to the
// remaining columns of the block, and the block's compact-WY factor T
// (forward, columnwise: T(0:j,j) = -tau_j T(0:j,0:j) V^T v_j).  The caller
// applies the block to the trailing columns with the batched GEMM.
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

// Register form of the panel step.  Rows [jb, m) of the block are split over
// nwg workgroups, one row per thread with the block's 32 columns in
// registers, so the panel is read once and written once.  Column step j needs
// one combined reduction of d_c = sum_{r>j} x_r A(r, c) for all 32 columns
// (x = column j below the diagonal): d_j is ||x||^2, and with the pivot row
// A(j, :) (held by workgroup 0, published with the partial sums)
//   v_j^T A(:, c) = A(j, c) + scal * d_c
// gives both the update of the later columns and the T column from the
// earlier reflectors.  One inter-workgroup barrier per column; a single
// workgroup needs none.  Same dlarfg convention as k_panel.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_panel_r(const PanelDesc<T>* __restrict__ descs,
                                                      const int* __restrict__ wstart, int nd, int jb) {
  constexpr int NW = kThreads / kWave;
  __shared__ T red[kNB * NW];
  __shared__ T Dsh[kNB], rowsh[kNB], wsh[kNB], gsh[kNB];
  __shared__ T Ts[kNB * kNB];
  const int p = find_owner(wstart, nd, blockIdx.x);
  const PanelDesc<T> d = descs[p];
  const int g = blockIdx.x - wstart[p], nwg = d.nwg;
  const int m = d.m, kk = min(d.m, d.k);
  if (jb >= kk) return;  // uniform over the problem's workgroups
  const int w = min(kNB, kk - jb), ncb = min(kNB, d.k - jb);
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int r = jb + g * kThreads + tid;  // this thread's row
  const bool has = r < m;
  const int64_t ld = d.ld;
  T a[kNB];
#pragma unroll
  for (int c = 0; c < kNB; ++c) a[c] = (has && c < ncb) ? d.P[r + (int64_t)(jb + c) * ld] : T(0);
  if (g == 0)
    for (int i = tid; i < kNB * kNB; i += kThreads) Ts[i] = T(0);
  unsigned int target = 0;
  for (int jj = 0; jj < w; ++jj) {
    const int j = jb + jj;
    T x = T(0);
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      if (c == jj) x = a[c];
    const bool below = has && r > j;
    T dd[kNB];
#pragma unroll
    for (int c = 0; c < kNB; ++c) dd[c] = below ? x * a[c] : T(0);
    const T sw = wave_reduce_scatter32_nolds(dd);
    if ((lane & 1) == 0) red[wv * kNB + (lane >> 1)] = sw;
    if (g == 0 && tid == jj) {
#pragma unroll
      for (int c = 0; c < kNB; ++c) rowsh[c] = a[c];
    }
    __syncthreads();
    T sum = T(0);
    if (tid < kNB) {
#pragma unroll
      for (int q = 0; q < NW; ++q) sum += red[q * kNB + tid];
    }
    if (nwg > 1) {
      const int buf = jj & 1;
      // handed over through agent-scope atomics (no cache maintenance at the barrier)
      if (tid < kNB) {
        T* __restrict__ pp = d.part + ((int64_t)buf * nwg + g) * (2 * kNB);
        agent_store(pp + tid, sum);
        if (g == 0) agent_store(pp + kNB + tid, rowsh[tid]);
      }
      if (nwg > kFanIn) {
        group_barrier2_rx(d.sync, g, nwg, (unsigned int)(jj + 1), d.err);
      } else {
        target += nwg;
        group_barrier_rx(d.sync, target, d.err);
      }
      if (tid < kNB) {
        // the partials of all workgroups, 8 loads in flight at a time, summed
        // in workgroup order
        T s = T(0);
        const T* pb = d.part + (int64_t)buf * nwg * (2 * kNB) + tid;
        for (int q0 = 0; q0 < nwg; q0 += 8) {
          T pv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) pv[u] = agent_load(pb + (int64_t)min(q0 + u, nwg - 1) * (2 * kNB));
#pragma unroll
          for (int u = 0; u < 8; ++u) s += (q0 + u < nwg) ? pv[u] : T(0);
        }
        Dsh[tid] = s;
        if (g != 0) rowsh[tid] = agent_load(d.part + (int64_t)buf * nwg * (2 * kNB) + kNB + tid);
      }
    } else if (tid < kNB) {
      Dsh[tid] = sum;
    }
    __syncthreads();
    // reflector scalars (identical in every thread and workgroup)
    const T ssq = Dsh[jj], alpha = rowsh[jj];
    T tau = T(0), beta = alpha, scal = T(0);
    if (ssq > T(0)) {
      beta = -copysign(hypot(alpha, sqrt(ssq)), alpha);
      tau = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
    }
    if (tid < kNB) {
      const T vdot = rowsh[tid] + scal * Dsh[tid];
      wsh[tid] = (tid > jj && tid < ncb) ? tau * vdot : T(0);
      gsh[tid] = (tid < jj) ? vdot : T(0);
    }
    __syncthreads();
    if (g == 0) {
      tcol_from_dots(Ts, jj, tau, gsh);
      if (tid == 0) d.tau[j] = tau;
    }
    // apply H_j to the later columns of this row; column j becomes v_j / beta.
    // No barrier closes the step: the next step rewrites red after this
    // step's readers passed the barrier above, and rewrites rowsh, Dsh, wsh
    // and gsh only after its own first barrier, which every thread reaches
    // after its reads here.
    if (has && r >= j) {
      const T v = (r == j) ? T(1) : ((ssq > T(0)) ? x * scal : x);
#pragma unroll
      for (int c = 0; c < kNB; ++c) {
        if (c > jj) a[c] -= wsh[c] * v;
        else if (c == jj) a[c] = (r == j) ? beta : v;
      }
    }
  }
  __syncthreads();  // the last step's T column is in Ts
  if (has) {
#pragma unroll
    for (int c = 0; c < kNB; ++c)
      if (c < ncb) d.P[r + (int64_t)(jb + c) * ld] = a[c];
  }
  if (g == 0) {
    T* __restrict__ Tb = d.Tb + (int64_t)jb * kNB;
    for (int i = tid; i < kNB * kNB; i += kThreads) Tb[i] = Ts[i];
  }
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
    // one reduce-scatter for all jj dots (entries c >= jj are zero), then
    // the waves' partials
    const int lane = tid & (kWave - 1), wv = tid / kWave;
    const T sw = wave_reduce_scatter32_nolds(acc);
    if ((lane & 1) == 0) red[wv * kNB + (lane >> 1)] = sw;
    __syncthreads();
    if (tid < kNB) {
      T s = T(0);
#pragma unroll
      for (int q = 0; q < kThreads / kWave; ++q) s += red[q * kNB + tid];
      gsh[tid] = (tid < jj) ? (d.V[j + (int64_t)(jb + tid) * ld] + s) : T(0);
    }
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
int panel_rows_per_wg() { return kThreads; }

template <typename T>
int panel_r_capacity() {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_panel_r<T>, kThreads, 0) != hipSuccess)
    return 0;
  return per_cu * cus;
}
template <typename T>
int launch_panel_r(hipStream_t s, const PanelDesc<T>* d, const int* wstart, int nd, int nwg_total, int jb, int coop) {
  if (nd <= 0) return 0;
  if (!coop) {
    hipLaunchKernelGGL((k_panel_r<T>), dim3(nwg_total), dim3(kThreads), 0, s, d, wstart, nd, jb);
    return (int)hipGetLastError();
  }
  void* args[] = {(void*)&d, (void*)&wstart, (void*)&nd, (void*)&jb};
  return (int)launch_coresident((const void*)k_panel_r<T>, dim3(nwg_total), dim3(kThreads), args, 0, s);
}
template int panel_r_capacity<double>();
template int panel_r_capacity<float>();
template int launch_panel_r<double>(hipStream_t, const PanelDesc<double>*, const int*, int, int, int, int);
template int launch_panel_r<float>(hipStream_t, const PanelDesc<float>*, const int*, int, int, int, int);

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

template void launch_larft<double>(hipStream_t, const LarftDesc<double>*, const int*, int, int);
template void launch_larft<float>(hipStream_t, const LarftDesc<float>*, const int*, int, int);
template void launch_trsm_block<double>(hipStream_t, const TrsmDesc<double>*, const int*, int, int, int);
template void launch_trsm_block<float>(hipStream_t, const TrsmDesc<float>*, const int*, int, int, int);

}  // namespace spq
