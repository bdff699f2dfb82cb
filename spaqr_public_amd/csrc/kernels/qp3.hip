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

#include <algorithm>

#include <type_traits>

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

// ---------------------------------------------------------------------------
// Register form.  Column j of C lives in the registers of an aligned group of
// G lanes (G a power of two): lane l of the group holds the RC contiguous
// rows [l*RC, (l+1)*RC) (rows >= k zero; RC from 16 to 100, chosen per
// batch), so C is read once and written
// once and a pivot step moves only O(k) values per workgroup.  Per step every
// workgroup publishes its local-best column next to its (norm, position,
// column) partial; after one inter-workgroup barrier all workgroups pick the
// same winner and read its column from the winner's slot.  Positions are
// tracked per column in registers; the position -> column table `pos` is
// written by workgroup 0 only (entry p is read at step p, after the barrier
// that follows its last write).  Arithmetic, pivot rule, downdate and stop
// test are those of the streaming form.
constexpr int kPosLds = 4096; // largest N whose position table lives in LDS
// LDS bytes of the reflector column: G lanes x (RC + 1) (one pad element per lane)
template <typename T>
__host__ __device__ constexpr int vsh_bytes(int G, int RC) { return (int)(((sizeof(T) * G * (RC + 1)) + 15) / 16 * 16); }
template <typename T>
constexpr size_t qp3r_lds(int gmax, int rc, int nmax) {
  return (size_t)vsh_bytes<T>(gmax, rc) + (nmax <= kPosLds ? sizeof(int) * (size_t)nmax : 0);
}

__device__ __forceinline__ void argmax4(double& v, int& p, int& c, int& w, double ov, int op, int oc, int ow) {
  if (ov > v || (ov == v && op < p)) { v = ov; p = op; c = oc; w = ow; }
}
// Butterfly sum over aligned groups of G lanes; every lane of a group ends
// with the same bits (IEEE addition is commutative).
__device__ __forceinline__ double group_sum(double v, int G) {
  if (G > 1) v += lane_xor<1>(v);
  if (G > 2) v += lane_xor<2>(v);
  if (G > 4) v += lane_xor<4>(v);
  if (G > 8) v += lane_xor<8>(v);
  if (G > 16) v += lane_xor<16>(v);
  if (G > 32) v += lane_xor<32>(v);
  return v;
}
// LDS slot of row r in the padded reflector layout
template <int RC>
__device__ __forceinline__ int vslot(int r) { return r + r / RC; }

// resident workgroups per CU the register budget is sized for: 4 waves per
// SIMD up to 32 rows per lane, 3 up to 48, 2 beyond (256-thread workgroups
// only)
constexpr int qp3r_min_blocks(int nt, int rc) { return rc <= 32 ? 1024 / nt : (rc <= 48 ? 768 / nt : 512 / nt); }
template <typename T, int NT, int RC>
__global__ __launch_bounds__(NT, qp3r_min_blocks(NT, RC)) void k_qp3r(const Qp3Desc<T>* __restrict__ descs,
                                                     const int* __restrict__ wstart, int nd) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int NW = NT / kWave;
  __shared__ double redv[NW];
  __shared__ int redp[NW], redc[NW], redw[NW];
  __shared__ double sred[NW];
  __shared__ double s_bv;
  __shared__ int s_bp, s_bc, s_piv, s_pos, s_gw, s_q;

  const int p = find_owner(wstart, nd, blockIdx.x);
  const Qp3Desc<T> d = descs[p];
  const int g = blockIdx.x - wstart[p], nwg = d.nwg;
  const int k = d.k, N = d.N, tid = threadIdx.x, lane = tid & (kWave - 1);
  const int G = d.G, gl = tid & (G - 1), rbase = gl * RC;
  const int j = g * (NT / G) + tid / G;  // this group's column
  const bool has = j < N;
  T* vsh = reinterpret_cast<T*>(smem_raw);  // G*(RC+1) entries: pivot column, then reflector
  const T* vmy = vsh + gl * (RC + 1);           // this lane's rows
  // position -> column table: in LDS (every workgroup applies the same swaps)
  // when N <= kPosLds, else in memory (written by workgroup 0)
  const bool lpos = N <= kPosLds;
  int* ptab = reinterpret_cast<int*>(smem_raw + vsh_bytes<T>(G, RC));
  const int kmin = min(k, N);
  const double tol3z = sqrt(ldexp(1.0, -53));

  T c[RC];
  {
    const T* src = d.Ct + j + (int64_t)rbase * d.ldc;
#pragma unroll
    for (int q = 0; q < RC; ++q) c[q] = (has && rbase + q < k) ? src[(int64_t)q * d.ldc] : T(0);
  }
  for (int r = k + tid; r < G * RC; r += NT) vsh[vslot<RC>(r)] = T(0);
  double s = 0;
#pragma unroll
  for (int q = 0; q < RC; ++q) s += (double)c[q] * (double)c[q];
  s = group_sum(s, G);
  double vn1 = sqrt(s), vn2 = vn1, beta_own = 0;
  int pos = j;
  if (lpos) {
    for (int q = tid; q < N; q += NT) ptab[q] = q;
  } else if (has && gl == 0) {
    agent_store(d.pos + j, j);
  }
  double bv = -2.0;
  int bp = INT_MAX, bc = -1;
  if (has && gl == 0) { bv = (vn1 == vn1) ? vn1 : -1.0; bp = j; bc = j; }

  double r00 = 0;
  int rank = kmin;
  unsigned int target = 0;
  for (int i = 0; i < kmin; ++i) {
    const int buf = i & 1;
    // ---- local argmax of the workgroup
    auto am_round = [&](auto xo) {
      constexpr int O = decltype(xo)::value;
      const double ov = lane_xor<O>(bv);
      const int op = lane_xor<O>(bp), oc = lane_xor<O>(bc);
      argmax_combine(bv, bp, bc, ov, op, oc);
    };
    am_round(std::integral_constant<int, 32>{});
    am_round(std::integral_constant<int, 16>{});
    am_round(std::integral_constant<int, 8>{});
    am_round(std::integral_constant<int, 4>{});
    am_round(std::integral_constant<int, 2>{});
    am_round(std::integral_constant<int, 1>{});
    if (lane == 0) { redv[tid / kWave] = bv; redp[tid / kWave] = bp; redc[tid / kWave] = bc; }
    __syncthreads();
    if (tid == 0) {
      double v = redv[0];
      int pp = redp[0], cc = redc[0];
      for (int w = 1; w < NW; ++w) argmax_combine(v, pp, cc, redv[w], redp[w], redc[w]);
      s_bv = v; s_bp = pp; s_bc = cc;
    }
    __syncthreads();
    // ---- publish the local-best column and its partial; find the winner
    const int wbest = s_bc;
    if (nwg == 1) {
      // one workgroup: the winner is local and its column goes straight to LDS
      if (has && j == wbest) {
#pragma unroll
        for (int q = 0; q < RC; ++q)
          if (rbase + q < k) vsh[vslot<RC>(rbase + q)] = c[q];
      }
      if (tid == 0) {
        s_piv = wbest; s_pos = s_bp; s_gw = 0;
        s_q = lpos ? ptab[i] : agent_load(d.pos + i);
      }
      __syncthreads();
    } else {
      // everything handed to the other workgroups goes through agent-scope
      // atomics, so the barrier needs no cache maintenance
      // rows [i, k) of the candidate (the rows above i are never read)
      if (has && j == wbest) {
        T* __restrict__ pc = d.part_col + ((int64_t)buf * nwg + g) * k + rbase;
#pragma unroll
        for (int q = 0; q < RC; ++q)
          if (rbase + q >= i && rbase + q < k) agent_store(pc + q, c[q]);
      }
      if (tid == 0) {
        agent_store(d.part_v + buf * nwg + g, s_bv);
        agent_store(d.part_p + 2 * (buf * nwg + g), s_bp);
        agent_store(d.part_p + 2 * (buf * nwg + g) + 1, s_bc);
      }
      if (nwg > kFanIn) {
        group_barrier2_rx(d.sync, g, nwg, (unsigned int)(i + 1), d.err);
      } else {
        target += nwg;
        group_barrier_rx(d.sync, target, d.err);
      }
      {
        double v = -3.0;
        int pp = INT_MAX, cc = -1, ww = -1;
        for (int w = tid; w < nwg; w += NT)
          argmax4(v, pp, cc, ww, agent_load(d.part_v + buf * nwg + w), agent_load(d.part_p + 2 * (buf * nwg + w)),
                  agent_load(d.part_p + 2 * (buf * nwg + w) + 1), w);
        auto am4_round = [&](auto xo) {
          constexpr int O = decltype(xo)::value;
          const double ov = lane_xor<O>(v);
          const int op = lane_xor<O>(pp), oc = lane_xor<O>(cc), ow = lane_xor<O>(ww);
          argmax4(v, pp, cc, ww, ov, op, oc, ow);
        };
        am4_round(std::integral_constant<int, 32>{});
        am4_round(std::integral_constant<int, 16>{});
        am4_round(std::integral_constant<int, 8>{});
        am4_round(std::integral_constant<int, 4>{});
        am4_round(std::integral_constant<int, 2>{});
        am4_round(std::integral_constant<int, 1>{});
        if (lane == 0) { redv[tid / kWave] = v; redp[tid / kWave] = pp; redc[tid / kWave] = cc; redw[tid / kWave] = ww; }
        __syncthreads();
        if (tid == 0) {
          for (int w = 1; w < NW; ++w) argmax4(v, pp, cc, ww, redv[w], redp[w], redc[w], redw[w]);
          s_piv = cc; s_pos = pp; s_gw = ww;
          s_q = lpos ? ptab[i] : agent_load(d.pos + i);
        }
        __syncthreads();
      }
      const T* __restrict__ pcol = d.part_col + ((int64_t)buf * nwg + s_gw) * k;
      for (int r = i + tid; r < k; r += NT) vsh[vslot<RC>(r)] = agent_load(pcol + r);
      __syncthreads();
    }
    const int piv = s_piv, ppos = s_pos, qcol = s_q;
    // ---- Householder vector of the pivot column (identical in every workgroup)
    double ss = 0;
    for (int r = i + 1 + tid; r < k; r += NT) { const double v = (double)vsh[vslot<RC>(r)]; ss += v * v; }
    ss = block_sum_x(ss, sred);
    const double alpha = (double)vsh[vslot<RC>(i)];
    double beta = alpha, tau = 0, scal = 0;
    if (ss > 0) {
      beta = -copysign(hypot(alpha, sqrt(ss)), alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (i == 0) r00 = fabs(beta);
    if (i >= 1 && !(fabs(beta) / r00 >= d.tol)) { rank = i; break; }
    __syncthreads();  // alpha read everywhere before vsh is overwritten
    for (int r = tid; r < k; r += NT) {
      const T v = (r < i) ? T(0) : (r == i ? T(1) : (T)((double)vsh[vslot<RC>(r)] * scal));
      vsh[vslot<RC>(r)] = v;
      if (g == 0) d.V[r + (int64_t)i * d.ldv] = v;
    }
    if (tid == 0) {
      // LAPACK swap of positions i and ppos (entry i is never read again)
      if (lpos) ptab[ppos] = qcol;
      else if (g == 0) agent_store(d.pos + ppos, qcol);
      if (g == 0) { d.tau[i] = (T)tau; d.beta[i] = beta; }
    }
    if (has) {
      if (j == qcol) pos = ppos;
      if (j == piv) { pos = i; beta_own = beta; }
    }
    __syncthreads();
    // ---- apply H_i to the live columns, downdate, next local candidate
    bv = -2.0;
    bp = INT_MAX;
    bc = -1;
    if (has && pos > i) {  // uniform over the group
      const int li = i - rbase;  // row i relative to this lane's first row
      // The reflector is read from LDS in chunks of CH rows, the next chunk's
      // reads issued before the current chunk is used (two chunks in flight
      // bound the live registers; the compiler fence keeps later chunks from
      // being hoisted).  Rows above i hold zeros in the reflector, so the dot
      // product and the update run over every row unmasked: those rows add
      // nothing to w and are left unchanged (c - w * 0 = c).
      constexpr int CH = sizeof(T) == 4 && RC % 8 == 0 ? 8 : 4;  // 16 VGPRs of buffers
      T vb[2][CH];
      auto fetch = [&](int q0, int buf) {
#pragma unroll
        for (int q = 0; q < CH; ++q) vb[buf][q] = vmy[q0 + q];
      };
      double w = 0;
      fetch(0, 0);
#pragma unroll
      for (int q0 = 0; q0 < RC; q0 += CH) {
        const int cur = (q0 / CH) & 1;
        if (q0 + CH < RC) fetch(q0 + CH, cur ^ 1);
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < CH; ++q) w += (double)vb[cur][q] * (double)c[q0 + q];
      }
      w = group_sum(w, G) * tau;
      if (w != 0.0) {
        // FP64 storage: the update in FP64.  FP32 storage: one FP32 FMA with
        // w rounded once (the stored result is FP32 either way); forming it
        // in FP64 would keep an FP64 copy of every value of the column live
        // from the dot product above, doubling the register footprint.
        const T wt = (T)w;
        fetch(0, 0);
#pragma unroll
        for (int q0 = 0; q0 < RC; q0 += CH) {
          const int cur = (q0 / CH) & 1;
          if (q0 + CH < RC) fetch(q0 + CH, cur ^ 1);
          asm volatile("" ::: "memory");
#pragma unroll
          for (int q = 0; q < CH; ++q) {
            if constexpr (sizeof(T) == 8) c[q0 + q] = (T)((double)c[q0 + q] - w * (double)vb[cur][q]);
            else c[q0 + q] = fmaf(-wt, vb[cur][q], c[q0 + q]);
          }
        }
      }
      T mine = T(0);
#pragma unroll
      for (int q = 0; q < RC; ++q)
        if (q == li) mine = c[q];
      const double cii = (double)__shfl(mine, (lane & ~(G - 1)) | (i / RC), kWave);
      double a = vn1;
      if (a != 0.0) {
        double tt = fabs(cii) / a;
        tt = 1.0 - tt * tt;
        tt = tt < 0.0 ? 0.0 : tt;
        const double qq = a / vn2;
        const double tt2 = tt * qq * qq;
        if (tt2 <= tol3z) {
          if (i < k - 1) {
            double s2 = 0;
#pragma unroll
            for (int q = 0; q < RC; ++q)
              if (q > li) s2 += (double)c[q] * (double)c[q];
            a = sqrt(group_sum(s2, G));
            vn2 = a;
          } else {
            a = 0.0;
            vn2 = 0.0;
          }
        } else {
          a = a * sqrt(tt);
        }
        vn1 = a;
      }
      if (gl == 0) { bv = (a == a) ? a : -1.0; bp = pos; bc = j; }
    }
  }
  // pivoted columns: R(0:pos, j), beta on the pivot row, zeros below; the
  // others keep their updated values
  if (has) {
    int64_t ldc = d.ldc;
    asm volatile("" : "+s"(ldc));  // recompute the addresses here instead of keeping them live
    T* dst = d.Ct + j + (int64_t)rbase * ldc;
    const int lp = (pos < rank) ? pos - rbase   : RC;  // first row (relative) replaced by beta / zero
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      if (rbase + q < k) {
        T v = c[q];
        if (q >= lp) v = (q == lp) ? (T)beta_own : T(0);
        dst[(int64_t)q * ldc] = v;
      }
    }
  }
  if (g == 0 && tid == 0) *d.rank_out = rank;
}

// Register-form lanes per column for k rows at rc rows per lane: the least
// power of two G with G * rc >= k, or 0 beyond 64 lanes.
int qp3r_lanes(int k, int rc) {
  int g = 1;
  while (g * rc < k) g <<= 1;
  return g <= 64 ? g : 0;
}

namespace {
template <typename T, int NT>
const void* qp3r_fn(int rc) {
  switch (rc) {
    case 16: return (const void*)k_qp3r<T, NT, 16>;
    case 20: return (const void*)k_qp3r<T, NT, 20>;
    case 24: return (const void*)k_qp3r<T, NT, 24>;
    case 28: return (const void*)k_qp3r<T, NT, 28>;
    case 32: return (const void*)k_qp3r<T, NT, 32>;
  }
  return nullptr;
}
template <typename T>
const void* qp3r_fn(int nt, int rc) {
  if (nt == 256 && rc == 40) return (const void*)k_qp3r<T, 256, 40>;
  if (nt == 256 && rc == 48) return (const void*)k_qp3r<T, 256, 48>;
  if (nt == 256 && rc == 64) return (const void*)k_qp3r<T, 256, 64>;
  if (nt == 256 && rc == 80) return (const void*)k_qp3r<T, 256, 80>;
  if (nt == 256 && rc == 100) return (const void*)k_qp3r<T, 256, 100>;
  return nt == 1024 ? qp3r_fn<T, 1024>(rc) : (nt == 256 ? qp3r_fn<T, 256>(rc) : nullptr);
}
}  // namespace

template <typename T>
int qp3r_capacity(int G, int nt, int rc) {
  int per_cu = 0, dev = 0, cus = 0;
  const void* f = qp3r_fn<T>(nt, rc);
  if (!f) return 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, nt, qp3r_lds<T>(G, rc, kPosLds)) != hipSuccess) return 0;
  return per_cu * cus;
}

template <typename T>
int launch_qp3r(hipStream_t s, const Qp3Desc<T>* d, const int* wstart, int nd, int nwg_total, int gmax, int nmax,
                int nt, int rc, int coop) {
  if (nd <= 0) return 0;
  const void* f = qp3r_fn<T>(nt, rc);
  if (!f) return (int)hipErrorInvalidValue;
  // the LDS table is sized for the largest N that uses it
  const size_t shm = qp3r_lds<T>(gmax, rc, std::min(nmax, kPosLds));
  void* args[] = {(void*)&d, (void*)&wstart, (void*)&nd};
  if (!coop) return (int)hipLaunchKernel(f, dim3(nwg_total), dim3(nt), args, shm, s);
  return (int)launch_coresident(f, dim3(nwg_total), dim3(nt), args, shm, s);
}
template int qp3r_capacity<double>(int, int, int);
template int qp3r_capacity<float>(int, int, int);
template int launch_qp3r<double>(hipStream_t, const Qp3Desc<double>*, const int*, int, int, int, int, int, int, int);
template int launch_qp3r<float>(hipStream_t, const Qp3Desc<float>*, const int*, int, int, int, int, int, int, int);

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
  return (int)launch_coresident((const void*)k_qp3<T>, dim3(nwg_total), dim3(kQThreads), args, shm, s);
}
template int qp3_capacity<double>(int);
template int qp3_capacity<float>(int);
template int launch_qp3<double>(hipStream_t, const Qp3Desc<double>*, const int*, int, int, int, int);
template int launch_qp3<float>(hipStream_t, const Qp3Desc<float>*, const int*, int, int, int, int);

}  // namespace spq
