// Batched uneven GEMM with structured operands for the block-reflector
// updates (larfb), T assembly and blocked TRSM of the spaQR phases.
//   C = alpha * op(A) * op(B) + beta * C
// One 64x64 C tile per workgroup (256 threads, 4x4 register micro-tile),
// K staged through LDS in chunks of 16.  Structure masks (unit-lower V,
// triangular T/R) are applied while staging, so the stored upper triangle of a
// Householder panel (which holds R) is never read as part of V.
#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

template <typename T>
__device__ __forceinline__ T structured(const T* __restrict__ M, int ld, int r, int c, int s) {
  if (s == kStructUnitLower) {
    if (r < c) return T(0);
    if (r == c) return T(1);
  } else if (s == kStructUpper) {
    if (r > c) return T(0);
  } else if (s == kStructLower) {
    if (r < c) return T(0);
  }
  return M[r + (int64_t)c * ld];
}

constexpr int TM = 64, TN = 64, TK = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc<T>* __restrict__ descs, const int* __restrict__ tstart,
                                              int nd) {
  const int t = blockIdx.x;
  const int p = find_owner(tstart, nd, t);
  const GemmDesc<T> D = descs[p];
  int local = t - tstart[p];
  const int ntm = (D.m + TM - 1) / TM, ntn = (D.n + TN - 1) / TN;
  const int split = local / (ntm * ntn);  // K chunk (split-K)
  local -= split * ntm * ntn;
  const int i0 = (local % ntm) * TM, j0 = (local / ntm) * TN;
  const int kc = ((D.k + D.ksplit - 1) / D.ksplit + TK - 1) / TK * TK;
  const int kbeg = split * kc, kend = min(D.k, kbeg + kc);
  const bool ta = D.flags & kTransA, tb = D.flags & kTransB;
  const int sa = (D.flags >> 4) & 3, sb = (D.flags >> 6) & 3;

  __shared__ T As[TK][TM + 1];
  __shared__ T Bs[TK][TN + 1];

  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = T(0);

  for (int l0 = kbeg; l0 < kend; l0 += TK) {
    // stage op(A)[i0:i0+64, l0:l0+16] into As[l][i]
    if (!ta) {
      const int i = tid & 63;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = (tid >> 6) + 4 * q;
        const int gi = i0 + i, gl = l0 + l;
        As[l][i] = (gi < D.m && gl < kend) ? structured(D.A, D.lda, gi, gl, sa) : T(0);
      }
    } else {
      const int l = tid & 15;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = (tid >> 4) + 16 * q;
        const int gi = i0 + i, gl = l0 + l;
        As[l][i] = (gi < D.m && gl < kend) ? structured(D.A, D.lda, gl, gi, sa) : T(0);
      }
    }
    // stage op(B)[l0:l0+16, j0:j0+64] into Bs[l][j]
    if (!tb) {
      const int l = tid & 15;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = (tid >> 4) + 16 * q;
        const int gl = l0 + l, gj = j0 + j;
        Bs[l][j] = (gl < kend && gj < D.n) ? structured(D.B, D.ldb, gl, gj, sb) : T(0);
      }
    } else {
      const int j = tid & 63;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int l = (tid >> 6) + 4 * q;
        const int gl = l0 + l, gj = j0 + j;
        Bs[l][j] = (gl < kend && gj < D.n) ? structured(D.B, D.ldb, gj, gl, sb) : T(0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < TK; ++l) {
      T av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[l][tx + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[l][ty + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
    }
    __syncthreads();
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int gj = j0 + ty + 16 * b;
    if (gj >= D.n) continue;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int gi = i0 + tx + 16 * a;
      if (gi >= D.m) continue;
      const T v = D.alpha * acc[a][b];
      if (D.ksplit > 1) {
        D.work[(int64_t)split * D.m * D.n + gi + (int64_t)gj * D.m] = v;
      } else {
        T* c = D.C + gi + (int64_t)gj * D.ldc;
        *c = (D.beta == T(0)) ? v : v + D.beta * (*c);
      }
    }
  }
}

// C = sum_s work_s + beta * C for split-K descriptors: 256 elements of C per
// workgroup (tiles: prefix sums of ceil(m*n / 256)).
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_reduce(const GemmDesc<T>* __restrict__ descs,
                                                     const int* __restrict__ tstart, int nd) {
  const int p = find_owner(tstart, nd, blockIdx.x);
  const GemmDesc<T> D = descs[p];
  const int64_t e = (int64_t)(blockIdx.x - tstart[p]) * 256 + threadIdx.x;
  const int64_t mn = (int64_t)D.m * D.n;
  if (e >= mn) return;
  const int i = (int)(e % D.m), j = (int)(e / D.m);
  T s = 0;
  for (int q = 0; q < D.ksplit; ++q) s += D.work[q * mn + e];
  T* c = D.C + i + (int64_t)j * D.ldc;
  *c = (D.beta == T(0)) ? s : s + D.beta * (*c);
}

int gemm_tiles(int m, int n) { return ((m + TM - 1) / TM) * ((n + TN - 1) / TN); }
int gemm_tile_k() { return TK; }

template <typename T>
void launch_gemm_reduce(hipStream_t s, const GemmDesc<T>* d, const int* tstart, int nd, int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL((k_gemm_reduce<T>), dim3(ntiles), dim3(256), 0, s, d, tstart, nd);
}
template void launch_gemm_reduce<double>(hipStream_t, const GemmDesc<double>*, const int*, int, int);
template void launch_gemm_reduce<float>(hipStream_t, const GemmDesc<float>*, const int*, int, int);

template <typename T>
void launch_gemm(hipStream_t s, const GemmDesc<T>* d, const int* tstart, int nd, int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL((k_gemm<T>), dim3(ntiles), dim3(256), 0, s, d, tstart, nd);
}
template void launch_gemm<double>(hipStream_t, const GemmDesc<double>*, const int*, int, int);
template void launch_gemm<float>(hipStream_t, const GemmDesc<float>*, const int*, int, int);

}  // namespace spq
