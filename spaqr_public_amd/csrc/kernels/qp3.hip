// Truncated column-pivoted Householder QR for the sparsify phase (gfx950).
//
// C (k x N, usually wide) is stored transposed: Ct is N x k col-major, so row
// r of C is a contiguous column of Ct and a wavefront reading C(r, j..j+63)
// is coalesced.  Threads own columns of C.  Column data never moves: the
// LAPACK column swap is kept as a position permutation (perm/ipos), so the
// pivot rule is exactly idamax over positions (largest partial norm, first
// position on ties) and the partial-norm downdate is the dlaqp2 rule with
// tol3z = sqrt(eps) recomputation.  Norms, the pivot choice and the rank test
// run in FP64 whatever the storage type.
//
// The factorization stops at the first step i >= 1 whose |r_ii| / |r_00| is
// not >= tol (the reference's choose_rank on the full geqp3 diagonal), so
// rank reflectors are generated.  On exit, for every original column j,
// C(0:rank, j) holds triu(R(0:rank,:)) P^T (pivoted columns have R entries
// above and zero below their pivot row), V(:, i) the unit-lower reflector i
// and tau(i) its scalar.
#include <limits.h>

#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

constexpr int kQThreads = 256;

__device__ __forceinline__ void argmax_combine(double& v, int& p, double ov, int op) {
  if (ov > v || (ov == v && op < p)) { v = ov; p = op; }
}

template <typename T>
__global__ __launch_bounds__(kQThreads) void k_qp3(const Qp3Desc<T>* __restrict__ descs, int use_lds) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ double redv[kQThreads / kWave];
  __shared__ int redp[kQThreads / kWave];
  __shared__ double sred[kQThreads / kWave];
  __shared__ int s_piv, s_pos, s_stop;

  const Qp3Desc<T> d = descs[blockIdx.x];
  const int k = d.k, N = d.N, tid = threadIdx.x;
  const int64_t ldc = d.ldc;
  T* __restrict__ Ct = d.Ct;
  double* __restrict__ vn1 = d.vn;
  double* __restrict__ vn2 = d.vn + N;
  int* __restrict__ ipos = d.pos;
  int* __restrict__ perm = d.pos + N;
  const int kmin = min(k, N);
  const double tol3z = sqrt(ldexp(1.0, -53));

  for (int j = tid; j < N; j += kQThreads) {
    double s = 0;
    for (int r = 0; r < k; ++r) { const double v = (double)Ct[j + r * ldc]; s += v * v; }
    vn1[j] = sqrt(s);
    vn2[j] = vn1[j];
    ipos[j] = j;
    perm[j] = j;
  }
  __syncthreads();

  double r00 = 0;
  int rank = kmin;
  for (int i = 0; i < kmin; ++i) {
    // ---- pivot: max vn1 over positions >= i, first position on ties
    // NaN norms rank below every number, so a position is always selected
    double bv = -2.0;
    int bp = INT_MAX;
    for (int j = tid; j < N; j += kQThreads) {
      const int pj = ipos[j];
      const double v = vn1[j];
      if (pj >= i) argmax_combine(bv, bp, (v == v) ? v : -1.0, pj);
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o, kWave);
      const int op = __shfl_xor(bp, o, kWave);
      argmax_combine(bv, bp, ov, op);
    }
    if ((tid & (kWave - 1)) == 0) { redv[tid / kWave] = bv; redp[tid / kWave] = bp; }
    __syncthreads();
    if (tid == 0) {
      double v = redv[0];
      int p = redp[0];
      for (int w = 1; w < kQThreads / kWave; ++w) argmax_combine(v, p, redv[w], redp[w]);
      s_pos = p;
      s_piv = perm[p];
    }
    __syncthreads();
    const int piv = s_piv, ppos = s_pos;
    // ---- Householder vector of column piv, rows i..k-1
    double s = 0;
    for (int r = i + 1 + tid; r < k; r += kQThreads) { const double v = (double)Ct[piv + r * ldc]; s += v * v; }
    s = block_sum(s, sred);
    const double alpha = (double)Ct[piv + (int64_t)i * ldc];
    double beta = alpha, tau = 0, scal = 0;
    if (s > 0) {
      beta = -copysign(hypot(alpha, sqrt(s)), alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (i == 0) r00 = fabs(beta);
    if (tid == 0) s_stop = (i >= 1 && !(fabs(beta) / r00 >= d.tol)) ? 1 : 0;
    __syncthreads();
    if (s_stop) { rank = i; break; }
    // reflector i: in LDS when the launch's largest k fits, else read back
    // from its column of V (written here, visible after the barrier)
    T* vsh = use_lds ? reinterpret_cast<T*>(smem_raw) : (d.V + (int64_t)i * d.ldv);
    for (int r = tid; r < k; r += kQThreads) {
      T v = (r < i) ? T(0) : (r == i ? T(1) : (T)((double)Ct[piv + r * ldc] * scal));
      if (use_lds) vsh[r] = v;
      d.V[r + (int64_t)i * d.ldv] = v;
    }
    if (tid == 0) {
      d.tau[i] = (T)tau;
      // LAPACK swap of positions i and ppos, expressed on the permutation
      const int q = perm[i];
      perm[ppos] = q;
      ipos[q] = ppos;
      perm[i] = piv;
      ipos[piv] = i;
    }
    __syncthreads();
    // ---- pivot column becomes R(0:i, i) = (already final), R(i,i) = beta, zeros below
    for (int r = i + tid; r < k; r += kQThreads) Ct[piv + r * ldc] = (r == i) ? (T)beta : T(0);
    // ---- apply H_i to the remaining columns and downdate their norms
    for (int j = tid; j < N; j += kQThreads) {
      if (ipos[j] <= i) continue;
      T* __restrict__ cj = Ct + j;
      double w = (double)cj[(int64_t)i * ldc];
      for (int r = i + 1; r < k; ++r) w += (double)vsh[r] * (double)cj[r * ldc];
      w *= tau;
      if (w != 0.0) {
        cj[(int64_t)i * ldc] = (T)((double)cj[(int64_t)i * ldc] - w);
        for (int r = i + 1; r < k; ++r) cj[r * ldc] = (T)((double)cj[r * ldc] - w * (double)vsh[r]);
      }
      double a = vn1[j];
      if (a != 0.0) {
        double tt = fabs((double)cj[(int64_t)i * ldc]) / a;
        tt = 1.0 - tt * tt;
        tt = tt < 0.0 ? 0.0 : tt;
        const double q = a / vn2[j];
        const double tt2 = tt * q * q;
        if (tt2 <= tol3z) {
          if (i < k - 1) {
            double ss = 0;
            for (int r = i + 1; r < k; ++r) { const double v = (double)cj[r * ldc]; ss += v * v; }
            vn1[j] = sqrt(ss);
            vn2[j] = vn1[j];
          } else {
            vn1[j] = 0.0;
            vn2[j] = 0.0;
          }
        } else {
          vn1[j] = a * sqrt(tt);
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) *d.rank_out = rank;
}

template <typename T>
void launch_qp3(hipStream_t s, const Qp3Desc<T>* d, int nd, int kmax) {
  if (nd <= 0) return;
  const size_t need = sizeof(T) * (size_t)(kmax > 0 ? kmax : 1);
  const int use_lds = need <= (size_t)48 * 1024;
  hipLaunchKernelGGL((k_qp3<T>), dim3(nd), dim3(kQThreads), use_lds ? need : 0, s, d, use_lds);
}
template void launch_qp3<double>(hipStream_t, const Qp3Desc<double>*, int, int);
template void launch_qp3<float>(hipStream_t, const Qp3Desc<float>*, int, int);

}  // namespace spq
