// Batched uneven GEMM with structured operands for the block-reflector
// updates (larfb), T assembly and blocked TRSM of the spaQR phases.
//   C = alpha * op(A) * op(B) + beta * C
// Structure masks (unit-lower V, triangular T/R) are applied while staging,
// so the stored upper triangle of a Householder panel (which holds R) is
// never read as part of V.
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

// Matrix cores: v_mfma_f64_16x16x4f64 (FP64) / v_mfma_f32_16x16x4f32 (FP32).
// A workgroup of 4 waves computes one C tile; each wave owns a 32 x 32 block
// of it as 2 x 2 MFMA tiles of 16 x 16, accumulating over K in steps of 4.
// The product is formed transposed (op(B)^T op(A)^T on the matrix core), so
// a lane's four accumulators lie in one row of C and the 16 lanes storing
// one column cover 16 contiguous rows: the col-major stores of C are 128-byte
// segments.
// The tile shape is chosen per descriptor from (m, n): 64 x 64, 32 x 128
// (m <= 32: the V_b^T C products of the larfb) or 128 x 32 (n <= 32: the
// V_b T_b^T products).  K is staged through LDS in chunks of 16, the next
// chunk's global loads issued before the current chunk's MFMAs.
constexpr int TK = 16;
constexpr int kTileMaxM = 128, kTileMaxN = 128;


// Operands: lane l holds A(l % 16, l / 16) of the 16 x 4 A block and
// B(l / 16, l % 16) of the 4 x 16 B block.  Accumulator c of lane l holds
// D(row(l / 16, c), l % 16): row = l / 16 + 4 c for FP64, 4 (l / 16) + c for
// FP32 (the two instructions interleave their result rows differently).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int k4, int c) { return k4 + 4 * c; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int k4, int c) { return 4 * k4 + c; }
};

// One TM x TN tile (TM = 32 WM, TN = 32 WN, WM * WN = 4 waves).
template <typename T, int WM, int WN>
__device__ __forceinline__ void gemm_tile(const GemmDesc<T>& D, int local, T (*As)[kTileMaxM + 2],
                                          T (*Bs)[kTileMaxN + 2]) {
  constexpr int TM = 32 * WM, TN = 32 * WN;
  constexpr int NA = TM * TK / 256, NB = TN * TK / 256;  // staged elements per thread
  using V4 = typename Mfma<T>::v4;
  const int ntm = (D.m + TM - 1) / TM, ntn = (D.n + TN - 1) / TN;
  const int split = local / (ntm * ntn);  // K chunk (split-K)
  local -= split * ntm * ntn;
  const int i0 = (local % ntm) * TM, j0 = (local / ntm) * TN;
  const int kc = ((D.k + D.ksplit - 1) / D.ksplit + TK - 1) / TK * TK;
  const int kbeg = split * kc, kend = min(D.k, kbeg + kc);
  const bool ta = D.flags & kTransA, tb = D.flags & kTransB;
  const int sa = (D.flags >> 4) & 3, sb = (D.flags >> 6) & 3;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w % WM, wn = w / WM;

  // staging maps: element q of this thread is (row ia[q] of the tile, k la[q])
  // for A and (k lb[q], column jb[q]) for B; consecutive threads walk the
  // contiguous direction of the stored operand
  T ra[NA], rb[NB];
  auto load = [&](int l0) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      const int i = ta ? e / TK : e % TM, l = ta ? e % TK : e / TM;
      const int gi = i0 + i, gl = l0 + l;
      ra[q] = (gi < D.m && gl < kend) ? (ta ? structured(D.A, D.lda, gl, gi, sa) : structured(D.A, D.lda, gi, gl, sa))
                                      : T(0);
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      const int j = tb ? e % TN : e / TK, l = tb ? e / TN : e % TK;
      const int gj = j0 + j, gl = l0 + l;
      rb[q] = (gl < kend && gj < D.n) ? (tb ? structured(D.B, D.ldb, gj, gl, sb) : structured(D.B, D.ldb, gl, gj, sb))
                                      : T(0);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      const int i = ta ? e / TK : e % TM, l = ta ? e % TK : e / TM;
      As[l][i] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      const int j = tb ? e % TN : e / TK, l = tb ? e / TN : e % TK;
      Bs[l][j] = rb[q];
    }
  };

  V4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V4{T(0), T(0), T(0), T(0)};

  const int r16 = lane & 15, k4 = lane >> 4;
  if (kbeg < kend) load(kbeg);
  for (int l0 = kbeg; l0 < kend; l0 += TK) {
    store();
    __syncthreads();
    if (l0 + TK < kend) load(l0 + TK);  // next chunk in flight during the MFMAs
#pragma unroll
    for (int kk = 0; kk < TK / 4; ++kk) {
      const int l = kk * 4 + k4;
      T av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[l][wm * 32 + a * 16 + r16];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = Bs[l][wn * 32 + b * 16 + r16];
      // transposed product: the op(B) values take the A role
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Mfma<T>::mac(bv[b], av[a], acc[a][b]);
    }
    __syncthreads();
  }
  // the accumulators hold the transposed tile: lane l has row r16 of each
  // 16-row block of C and columns Mfma::row(k4, c) of each 16-column block
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int gi = i0 + wm * 32 + a * 16 + r16;
    if (gi >= D.m) continue;
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int gj = j0 + wn * 32 + b * 16 + Mfma<T>::row(k4, c);
        if (gj >= D.n) continue;
        const T v = D.alpha * acc[a][b][c];
        if (D.ksplit > 1) {
          D.work[(int64_t)split * D.m * D.n + gi + (int64_t)gj * D.m] = v;
        } else {
          T* cp = D.C + gi + (int64_t)gj * D.ldc;
          *cp = (D.beta == T(0)) ? v : v + D.beta * (*cp);
        }
      }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc<T>* __restrict__ descs, const int* __restrict__ tstart,
                                              int nd) {
  __shared__ T As[TK][kTileMaxM + 2];
  __shared__ T Bs[TK][kTileMaxN + 2];
  const int t = blockIdx.x;
  const int p = find_owner(tstart, nd, t);
  const GemmDesc<T> D = descs[p];
  const int local = t - tstart[p];
  switch (gemm_shape(D.m, D.n)) {
    case 1: gemm_tile<T, 1, 4>(D, local, As, Bs); break;
    case 2: gemm_tile<T, 4, 1>(D, local, As, Bs); break;
    default: gemm_tile<T, 2, 2>(D, local, As, Bs); break;
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
