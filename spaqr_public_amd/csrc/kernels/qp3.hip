// Truncated column-pivoted Householder QR for the sparsify phase (gfx950).
//
// C (k x N, usually wide) is stored transposed: Ct is N x k col-major, so row
// r of C is a contiguous column of Ct and a wavefront reading C(r, j..j+63)
// is coalesced.  Threads own columns of C; a problem is split over `nwg`
// workgroups owning contiguous column slices, which meet at one
// inter-workgroup barrier per pivot step (a single workgroup needs none).
// Column data never moves: the LAPACK column swap is kept as a position
// permutation (each workgroup applies the same swaps to its own copy), so the
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
//
// Every workgroup of a multi-workgroup problem must be resident at once: the
// host launches those problems cooperatively within the device's capacity.
// Barrier waits are bounded, so a kernel always terminates; a wait that gives
// up sets `err` and the host reports it.
#include <limits.h>

#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

constexpr int kQThreads = 256;

__device__ __forceinline__ void argmax_combine(double& v, int& p, int& c, double ov, int op, int oc) {
  if (ov > v || (ov == v && op < p)) { v = ov; p = op; c = oc; }
}

template <typename T>
__global__ __launch_bounds__(kQThreads) void k_qp3(const Qp3Desc<T>* __restrict__ descs, const int* __restrict__ wstart,
                                                   int nd, int use_lds) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ double redv[kQThreads / kWave];
  __shared__ int redp[kQThreads / kWave], redc[kQThreads / kWave];
  __shared__ double sred[kQThreads / kWave];
  __shared__ int s_piv, s_pos, s_stop;

  const int p = find_owner(wstart, nd, blockIdx.x);
  const Qp3Desc<T> d = descs[p];
  const int g = blockIdx.x - wstart[p], nwg = d.nwg;
  const int k = d.k, N = d.N, tid = threadIdx.x;
  const int c0 = (int)((int64_t)g * N / nwg), c1 = (int)((int64_t)(g + 1) * N / nwg);
  const int64_t ldc = d.ldc;
  T* __restrict__ Ct = d.Ct;
  double* __restrict__ vn1 = d.vn;
  double* __restrict__ vn2 = d.vn + N;
  int* __restrict__ ipos = d.pos + (int64_t)g * 2 * N;
  int* __restrict__ perm = ipos + N;
  const int kmin = min(k, N);
  const double tol3z = sqrt(ldexp(1.0, -53));
  unsigned int target = 0;

  for (int j = tid; j < N; j += kQThreads) { ipos[j] = j; perm[j] = j; }
  // initial norms and the local argmax of step 0
  double bv = -2.0;
  int bp = INT_MAX, bc = -1;
  for (int j = c0 + tid; j < c1; j += kQThreads) {
    double s = 0;
    for (int r = 0; r < k; ++r) { const double v = (double)Ct[j + r * ldc]; s += v * v; }
    const double nv = sqrt(s);
    vn1[j] = nv;
    vn2[j] = nv;
    argmax_combine(bv, bp, bc, (nv == nv) ? nv : -1.0, j, j);
  }

  double r00 = 0;
  int rank = kmin;
  for (int i = 0; i < kmin; ++i) {
    const int buf = i & 1;
    // ---- publish the local argmax, meet, reduce over the problem's workgroups
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o, kWave);
      const int op = __shfl_xor(bp, o, kWave), oc = __shfl_xor(bc, o, kWave);
      argmax_combine(bv, bp, bc, ov, op, oc);
    }
    if ((tid & (kWave - 1)) == 0) { redv[tid / kWave] = bv; redp[tid / kWave] = bp; redc[tid / kWave] = bc; }
    __syncthreads();
    if (tid == 0) {
      double v = redv[0];
      int pp = redp[0], cc = redc[0];
      for (int w = 1; w < kQThreads / kWave; ++w) argmax_combine(v, pp, cc, redv[w], redp[w], redc[w]);
      d.part_v[buf * nwg + g] = v;
      d.part_p[2 * (buf * nwg + g)] = pp;
      d.part_p[2 * (buf * nwg + g) + 1] = cc;
    }
    if (nwg > 1) {
      target += nwg;
      group_barrier(d.sync, target, d.err);
    } else {
      __syncthreads();
    }
    if (tid == 0) {
      double v = -3.0;
      int pp = INT_MAX, cc = -1;
      for (int w = 0; w < nwg; ++w)
        argmax_combine(v, pp, cc, d.part_v[buf * nwg + w], d.part_p[2 * (buf * nwg + w)],
                       d.part_p[2 * (buf * nwg + w) + 1]);
      s_pos = pp;
      s_piv = cc;
    }
    __syncthreads();
    const int piv = s_piv, ppos = s_pos;
    // ---- Householder vector of column piv, rows i..k-1 (identical in every workgroup)
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
    // reflector i: LDS, or (single-workgroup problems only) its column of V
    T* vsh = use_lds ? reinterpret_cast<T*>(smem_raw) : (d.V + (int64_t)i * d.ldv);
    for (int r = tid; r < k; r += kQThreads) {
      const T v = (r < i) ? T(0) : (r == i ? T(1) : (T)((double)Ct[piv + r * ldc] * scal));
      if (use_lds) vsh[r] = v;
      if (g == 0) d.V[r + (int64_t)i * d.ldv] = v;
    }
    if (tid == 0) {
      if (g == 0) { d.tau[i] = (T)tau; d.beta[i] = beta; }
      // LAPACK swap of positions i and ppos, on this workgroup's permutation copy
      const int q = perm[i];
      perm[ppos] = q;
      ipos[q] = ppos;
      perm[i] = piv;
      ipos[piv] = i;
    }
    __syncthreads();
    // ---- apply H_i to the owned remaining columns, downdate, next local argmax
    bv = -2.0;
    bp = INT_MAX;
    bc = -1;
    for (int j = c0 + tid; j < c1; j += kQThreads) {
      const int pj = ipos[j];
      if (pj <= i) continue;
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
            a = sqrt(ss);
            vn2[j] = a;
          } else {
            a = 0.0;
            vn2[j] = 0.0;
          }
        } else {
          a = a * sqrt(tt);
        }
        vn1[j] = a;
      }
      argmax_combine(bv, bp, bc, (a == a) ? a : -1.0, pj, j);
    }
    __syncthreads();
  }
  // every workgroup is done reading pivot columns before any is overwritten
  if (nwg > 1) {
    target += nwg;
    group_barrier(d.sync, target, d.err);
  }
  // pivoted columns: R(i,i) = beta_i on the pivot row, zeros below
  for (int j = c0 + tid; j < c1; j += kQThreads) {
    const int pj = ipos[j];
    if (pj >= rank) continue;
    for (int r = pj; r < k; ++r) Ct[j + r * ldc] = (r == pj) ? (T)d.beta[pj] : T(0);
  }
  if (g == 0 && tid == 0) *d.rank_out = rank;
}

int qp3_threads() { return kQThreads; }

// Workgroups of one launch that can be resident at once (cooperative bound).
template <typename T>
int qp3_capacity(int kmax) {
  const size_t need = sizeof(T) * (size_t)(kmax > 0 ? kmax : 1);
  const int use_lds = need <= (size_t)48 * 1024;
  int per_cu = 0, dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_qp3<T>, kQThreads, use_lds ? need : 0) !=
      hipSuccess)
    return 0;
  return per_cu * cus;
}

template <typename T>
int launch_qp3(hipStream_t s, const Qp3Desc<T>* d, const int* wstart, int nd, int nwg_total, int kmax, int coop) {
  if (nd <= 0) return 0;
  const size_t need = sizeof(T) * (size_t)(kmax > 0 ? kmax : 1);
  int use_lds = need <= (size_t)48 * 1024;
  const size_t shm = use_lds ? need : 0;
  if (!coop) {
    hipLaunchKernelGGL((k_qp3<T>), dim3(nwg_total), dim3(kQThreads), shm, s, d, wstart, nd, use_lds);
    return (int)hipGetLastError();
  }
  void* args[] = {(void*)&d, (void*)&wstart, (void*)&nd, (void*)&use_lds};
  return (int)hipLaunchCooperativeKernel((const void*)k_qp3<T>, dim3(nwg_total), dim3(kQThreads), args, shm, s);
}
template int qp3_capacity<double>(int);
template int qp3_capacity<float>(int);
template int launch_qp3<double>(hipStream_t, const Qp3Desc<double>*, const int*, int, int, int, int);
template int launch_qp3<float>(hipStream_t, const Qp3Desc<float>*, const int*, int, int, int, int);

}  // namespace spq
