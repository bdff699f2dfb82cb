// Solve-side kernels: the stored spaQR factors applied to the cluster vector
// buffer, one batched launch per (level, op type) group.
//   k_refl   : x <- H^T x or H x for blocked reflectors (QR fwd over gathered
//              neighbour segments, ScaleD fwd, OrthogonalD fwd/bwd)
//   k_bwd    : x_c -= sum R12_n x_n ; x_c <- R^{-1} x_c   (QR bwd, ScaleD bwd)
//              or x_c <- R^{-T} x_c ; x_n -= R12_n^T x_c  (transposed apply)
//   k_move   : segment copies (Merge, SplitD)
//   k_gather_in / k_scatter_out : permuted entry/exit of the apply
#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

constexpr int kSThreads = 256;
constexpr int kRNB = 32;  // reflector block width (matches the panel kernel)

template <typename T>
__global__ __launch_bounds__(kSThreads) void k_refl(const ReflDesc<T>* __restrict__ descs, const Seg* __restrict__ segs,
                                                    T* __restrict__ xbuf, T* __restrict__ work, int lds_cap,
                                                    int mmax) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ T red[kRNB * (kSThreads / kWave)];
  __shared__ T zsh[kRNB];
  const ReflDesc<T> d = descs[blockIdx.x];
  const int m = d.m, tid = threadIdx.x;
  // LDS when every problem of the launch fits, else a per-workgroup slice of
  // the global work buffer (nd * mmax elements).
  T* x = (mmax <= lds_cap) ? reinterpret_cast<T*>(smem_raw) : (work + (int64_t)blockIdx.x * mmax);
  // gather
  {
    int base = 0;
    for (int s = 0; s < d.nseg; ++s) {
      const Seg g = segs[d.seg0 + s];
      for (int i = tid; i < g.len; i += kSThreads) x[base + i] = xbuf[g.off + i];
      base += g.len;
    }
  }
  __syncthreads();
  const int nblk = (d.k + kRNB - 1) / kRNB;
  for (int t = 0; t < nblk; ++t) {
    const int bi = d.trans ? t : nblk - 1 - t;
    const int jb = bi * kRNB, w = min(kRNB, d.k - jb);
    T acc[kRNB];
#pragma unroll
    for (int c = 0; c < kRNB; ++c) acc[c] = T(0);
    for (int i = jb + tid; i < m; i += kSThreads) {
      const T xi = x[i];
#pragma unroll
      for (int c = 0; c < kRNB; ++c) {
        if (c < w) {
          const int col = jb + c;
          if (i > col) acc[c] += d.V[i + (int64_t)col * d.ldv] * xi;
          else if (i == col) acc[c] += xi;
        }
      }
    }
    block_sum_n<T, kRNB>(acc, red, w);
    // z = T^T y (trans) or T y ; T block upper w x w, ld kRNB
    const T* __restrict__ Tb = d.Tb + (int64_t)jb * kRNB;
    if (tid < w) {
      T z = 0;
      if (d.trans) { for (int l = 0; l <= tid; ++l) z += Tb[l + tid * kRNB] * acc[l]; }
      else { for (int l = tid; l < w; ++l) z += Tb[tid + l * kRNB] * acc[l]; }
      zsh[tid] = z;
    }
    __syncthreads();
    for (int i = jb + tid; i < m; i += kSThreads) {
      T s = 0;
#pragma unroll
      for (int c = 0; c < kRNB; ++c) {
        if (c < w) {
          const int col = jb + c;
          if (i > col) s += d.V[i + (int64_t)col * d.ldv] * zsh[c];
          else if (i == col) s += zsh[c];
        }
      }
      x[i] -= s;
    }
    __syncthreads();
  }
  // scatter
  {
    int base = 0;
    for (int s = 0; s < d.nseg; ++s) {
      const Seg g = segs[d.seg0 + s];
      for (int i = tid; i < g.len; i += kSThreads) xbuf[g.off + i] = x[base + i];
      base += g.len;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kSThreads) void k_bwd(const BwdDesc<T>* __restrict__ descs, const R12Desc<T>* __restrict__ r12,
                                                   T* __restrict__ xbuf) {
  __shared__ T xk;
  const BwdDesc<T> d = descs[blockIdx.x];
  const int k = d.k, tid = threadIdx.x;
  T* __restrict__ xs = xbuf + d.xoff;
  if (!d.trans) {
    // x_c[0:k] -= sum_n A_n[0:k, 0:s] x_n[0:s]
    for (int r = tid; r < k; r += kSThreads) {
      T acc = 0;
      for (int b = 0; b < d.nblk; ++b) {
        const R12Desc<T> q = r12[d.blk0 + b];
        const T* __restrict__ xn = xbuf + q.xoff;
        for (int j = 0; j < q.s; ++j) acc += q.A[r + (int64_t)j * q.ld] * xn[j];
      }
      xs[r] -= acc;
    }
    __syncthreads();
    // back substitution with R (upper, ld)
    for (int i = k - 1; i >= 0; --i) {
      if (tid == 0) { xs[i] = xs[i] / d.R[i + (int64_t)i * d.ldr]; xk = xs[i]; }
      __syncthreads();
      const T v = xk;
      for (int r = tid; r < i; r += kSThreads) xs[r] -= d.R[r + (int64_t)i * d.ldr] * v;
      __syncthreads();
    }
  } else {
    // forward substitution with R^T: x_i = (x_i - sum_{r<i} R(r,i) x_r) / R(i,i)
    for (int i = 0; i < k; ++i) {
      if (tid == 0) { xs[i] = xs[i] / d.R[i + (int64_t)i * d.ldr]; xk = xs[i]; }
      __syncthreads();
      const T v = xk;
      for (int r = i + 1 + tid; r < k; r += kSThreads) xs[r] -= d.R[i + (int64_t)r * d.ldr] * v;
      __syncthreads();
    }
    // x_n[0:s] -= A_n[0:k, 0:s]^T x_c[0:k]   (atomic: several clusters share x_n)
    for (int b = 0; b < d.nblk; ++b) {
      const R12Desc<T> q = r12[d.blk0 + b];
      for (int j = tid; j < q.s; j += kSThreads) {
        T acc = 0;
        for (int r = 0; r < k; ++r) acc += q.A[r + (int64_t)j * q.ld] * xs[r];
        atomicAdd(xbuf + q.xoff + j, -acc);
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kSThreads) void k_move(const MoveDesc* __restrict__ d, T* __restrict__ dst_base,
                                                    const T* __restrict__ src_base) {
  const MoveDesc m = d[blockIdx.x];
  for (int i = threadIdx.x; i < m.len; i += kSThreads) dst_base[m.dst + i] = src_base[m.src + i];
}

// xbuf[dst[i]] = (T) b[src[i]]
template <typename T>
__global__ void k_gather_in(const int64_t* __restrict__ dst, const int* __restrict__ src, int n, const double* __restrict__ b,
                            T* __restrict__ xbuf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xbuf[dst[i]] = (T)b[src[i]];
}
// x[dst[i]] = (double) xbuf[src[i]]
template <typename T>
__global__ void k_scatter_out(const int64_t* __restrict__ src, const int* __restrict__ dst, int n, const T* __restrict__ xbuf,
                              double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[dst[i]] = (double)xbuf[src[i]];
}

int refl_lds_cap(int elem) { return (48 * 1024) / elem; }

template <typename T>
void launch_refl(hipStream_t s, const ReflDesc<T>* d, int nd, const Seg* segs, T* xbuf, T* work, int mmax) {
  if (nd <= 0) return;
  const int cap = refl_lds_cap(sizeof(T));
  size_t shm = (mmax <= cap) ? sizeof(T) * (size_t)mmax : 0;
  hipLaunchKernelGGL((k_refl<T>), dim3(nd), dim3(kSThreads), shm, s, d, segs, xbuf, work, cap, mmax);
}
template <typename T>
void launch_bwd(hipStream_t s, const BwdDesc<T>* d, int nd, const R12Desc<T>* r12, T* xbuf) {
  if (nd <= 0) return;
  hipLaunchKernelGGL((k_bwd<T>), dim3(nd), dim3(kSThreads), 0, s, d, r12, xbuf);
}
template <typename T>
void launch_move(hipStream_t s, const MoveDesc* d, int nd, T* dst, const T* src) {
  if (nd <= 0) return;
  hipLaunchKernelGGL((k_move<T>), dim3(nd), dim3(kSThreads), 0, s, d, dst, src);
}
template <typename T>
void launch_gather_in(hipStream_t s, const int64_t* dst, const int* src, int n, const double* b, T* xbuf) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_gather_in<T>), dim3((n + 255) / 256), dim3(256), 0, s, dst, src, n, b, xbuf);
}
template <typename T>
void launch_scatter_out(hipStream_t s, const int64_t* src, const int* dst, int n, const T* xbuf, double* x) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_scatter_out<T>), dim3((n + 255) / 256), dim3(256), 0, s, src, dst, n, xbuf, x);
}

#define SPQ_INST(T)                                                                                   \
  template void launch_refl<T>(hipStream_t, const ReflDesc<T>*, int, const Seg*, T*, T*, int);        \
  template void launch_bwd<T>(hipStream_t, const BwdDesc<T>*, int, const R12Desc<T>*, T*);            \
  template void launch_move<T>(hipStream_t, const MoveDesc*, int, T*, const T*);                      \
  template void launch_gather_in<T>(hipStream_t, const int64_t*, const int*, int, const double*, T*); \
  template void launch_scatter_out<T>(hipStream_t, const int64_t*, const int*, int, const T*, double*);
SPQ_INST(double)
SPQ_INST(float)

}  // namespace spq
