// Solve-side kernels: the stored spaQR factors applied to the cluster vector
// buffer, one batched launch per (level, op type) group.
//   k_refl   : x <- H^T x or H x for blocked reflectors (QR fwd over gathered
//              neighbour segments, ScaleD fwd, OrthogonalD fwd/bwd)
//   k_bwd    : x_c -= sum R12_n x_n ; x_c <- R^{-1} x_c   (QR bwd, ScaleD bwd)
//              or x_c <- R^{-T} x_c ; x_n -= R12_n^T x_c  (transposed apply)
//   k_move   : segment copies (Merge, SplitD)
//   k_gather_in / k_scatter_out : permuted entry/exit of the apply
//
// A launch mixes problem sizes: each descriptor gets `nwg` workgroups
// (wstart = exclusive prefix sums).  Small problems use one workgroup with the
// vector in LDS; large ones are split into row slices that meet at one
// inter-workgroup barrier per 32-column block.  The host places the
// multi-workgroup descriptors first and keeps their total within one
// workgroup per CU, so they are resident together.
#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

constexpr int kSThreads = 256;
constexpr int kSWaves = kSThreads / kWave;
constexpr int kRNB = 32;      // reflector / triangular block width (matches the panel kernel)
constexpr int kTLd = kRNB + 1;  // padded LDS leading dimension of a staged 32x32 block

// Entry (i, col) of a unit-lower reflector block; the load is in bounds for
// every i < m, col < k (the strict upper part holds R or is unused).
template <typename T>
__device__ __forceinline__ T vval(const T* __restrict__ V, int64_t ldv, int i, int col) {
  const T v = V[i + (int64_t)col * ldv];
  return (i > col) ? v : (i == col ? T(1) : T(0));
}

// acc[c] += V(i, jb+c) * x[i] over rows i in [lo, hi)
template <typename T>
__device__ __forceinline__ void refl_partial(const T* __restrict__ V, int64_t ldv, const T* __restrict__ x, int jb, int w,
                                             int lo, int hi, T (&acc)[kRNB]) {
  if (w == kRNB) {
    for (int i = lo + threadIdx.x; i < hi; i += kSThreads) {
      const T xi = x[i];
#pragma unroll
      for (int c = 0; c < kRNB; ++c) acc[c] += vval(V, ldv, i, jb + c) * xi;
    }
  } else {
    for (int i = lo + threadIdx.x; i < hi; i += kSThreads) {
      const T xi = x[i];
#pragma unroll
      for (int c = 0; c < kRNB; ++c)
        if (c < w) acc[c] += vval(V, ldv, i, jb + c) * xi;
    }
  }
}
// x[i] -= sum_c V(i, jb+c) z[c] over rows i in [lo, hi)
template <typename T>
__device__ __forceinline__ void refl_update(const T* __restrict__ V, int64_t ldv, T* __restrict__ x, int jb, int w, int lo,
                                            int hi, const T* zsh) {
  if (w == kRNB) {
    for (int i = lo + threadIdx.x; i < hi; i += kSThreads) {
      T s = 0;
#pragma unroll
      for (int c = 0; c < kRNB; ++c) s += vval(V, ldv, i, jb + c) * zsh[c];
      x[i] -= s;
    }
  } else {
    for (int i = lo + threadIdx.x; i < hi; i += kSThreads) {
      T s = 0;
#pragma unroll
      for (int c = 0; c < kRNB; ++c)
        if (c < w) s += vval(V, ldv, i, jb + c) * zsh[c];
      x[i] -= s;
    }
  }
}
// Stage the block T factor of columns [jb, jb+32) (ld 32) into LDS (ld kTLd).
template <typename T>
__device__ __forceinline__ void stage_t(const T* __restrict__ Tb, int jb, T* tsh) {
  for (int e = threadIdx.x; e < kRNB * kRNB; e += kSThreads) tsh[(e & (kRNB - 1)) + (e / kRNB) * kTLd] = Tb[(int64_t)jb * kRNB + e];
}
// z = T_b^T y (trans) or T_b y, threads c < w
template <typename T>
__device__ __forceinline__ void tmul(const T* tsh, int w, int trans, const T* y, T* zsh) {
  const int tid = threadIdx.x;
  if (tid < w) {
    T z = 0;
    if (trans) { for (int l = 0; l <= tid; ++l) z += tsh[l + tid * kTLd] * y[l]; }
    else { for (int l = tid; l < w; ++l) z += tsh[tid + l * kTLd] * y[l]; }
    zsh[tid] = z;
  }
}

template <typename T>
__global__ __launch_bounds__(kSThreads) void k_refl(const ReflDesc<T>* __restrict__ descs, const int* __restrict__ wstart,
                                                    int nd, const Seg* __restrict__ segs, T* __restrict__ xbuf,
                                                    T* __restrict__ work, int lds_cap, int mmax) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ T red[kRNB * kSWaves];
  __shared__ T tsh[kRNB * kTLd];
  __shared__ T ysh[kRNB];
  __shared__ T zsh[kRNB];
  const int p = ((int)gridDim.x == nd) ? (int)blockIdx.x : find_owner(wstart, nd, blockIdx.x);
  const ReflDesc<T> d = descs[p];
  const int g = blockIdx.x - wstart[p];
  const int m = d.m, tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int nblk = (d.k + kRNB - 1) / kRNB;
  const bool multi = d.nwg > 1;
  // rows [r0, r1) of the gathered vector belong to this workgroup
  const int r0 = multi ? (int)((int64_t)g * m / d.nwg) : 0;
  const int r1 = multi ? (int)((int64_t)(g + 1) * m / d.nwg) : m;
  T* __restrict__ x = multi ? d.work
                            : ((mmax <= lds_cap) ? reinterpret_cast<T*>(smem_raw) : (work + (int64_t)blockIdx.x * mmax));
  {
    int base = 0;
    for (int s = 0; s < d.nseg; ++s) {
      const Seg sg = segs[d.seg0 + s];
      const int lo = max(base, r0), hi = min(base + sg.len, r1);
      for (int i = lo + tid; i < hi; i += kSThreads) x[i] = xbuf[sg.off + (i - base)];
      base += sg.len;
    }
  }
  __syncthreads();
  unsigned int target = 0;
  for (int t = 0; t < nblk; ++t) {
    const int bi = d.trans ? t : nblk - 1 - t;
    const int jb = bi * kRNB, w = min(kRNB, d.k - jb);
    stage_t(d.Tb, jb, tsh);
    T acc[kRNB];
#pragma unroll
    for (int c = 0; c < kRNB; ++c) acc[c] = T(0);
    refl_partial(d.V, d.ldv, x, jb, w, max(r0, jb), r1, acc);
    const T sw = wave_reduce_scatter32(acc);
    if (lane < kRNB) red[wv * kRNB + lane] = sw;
    __syncthreads();
    if (tid < kRNB) {
      T y = 0;
#pragma unroll
      for (int q = 0; q < kSWaves; ++q) y += red[q * kRNB + tid];
      if (multi) agent_store(d.part + ((int64_t)(t & 1) * d.nwg + g) * kRNB + tid, y);
      else ysh[tid] = y;
    }
    if (multi) {
      // partial sums go through agent-scope atomics (no cache maintenance)
      target += d.nwg;
      group_barrier_rx(d.sync, target, d.err);
      if (tid < kRNB) {
        T y = 0;
        for (int q = 0; q < d.nwg; ++q) y += agent_load(d.part + ((int64_t)(t & 1) * d.nwg + q) * kRNB + tid);
        ysh[tid] = y;
      }
    }
    __syncthreads();
    tmul(tsh, w, d.trans, ysh, zsh);
    __syncthreads();
    refl_update(d.V, d.ldv, x, jb, w, max(r0, jb), r1, zsh);
    __syncthreads();
  }
  {
    int base = 0;
    for (int s = 0; s < d.nseg; ++s) {
      const Seg sg = segs[d.seg0 + s];
      const int lo = max(base, r0), hi = min(base + sg.len, r1);
      for (int i = lo + tid; i < hi; i += kSThreads) xbuf[sg.off + (i - base)] = x[i];
      base += sg.len;
    }
  }
}

// ---------------------------------------------------------------------------
// Register form of the reflector apply, for stages whose ops are few (the
// dependency waves of the QR forward sweep and the large ops of the upper
// levels), where one op's time, not the stage's throughput, sets the pace.
// Each row thread owns one row of the op's slice (<= 64 NRW rows per
// workgroup): its x entry and its 32 entries of the current reflector block
// stay in registers, so V is read once, and the next block's row is loaded
// while the block's sums are combined.  A leader wave owning no rows sums
// the row waves' partials, exchanges them between the workgroups of a
// split op (agent-scope stores and one relaxed barrier per block) and forms
// z = T_b^T y (or T_b y) from T_b, which it stages into LDS one block ahead.
// NRW row waves + the leader per workgroup: 7 (448 rows; 8 waves, 2 per
// SIMD, 256 VGPRs each) for stages of few ops, 1 or 2 (64 / 128 rows,
// several workgroups per CU) for stages of many small ops.
int reflr_rows() { return 7 * kWave; }
int reflr_row_waves(int mmax) { return mmax <= kWave ? 1 : (mmax <= 2 * kWave ? 2 : 7); }
size_t reflr_lds(int nseg_max) {
  return (((size_t)nseg_max + 1) * sizeof(int) + 7) / 8 * 8 + (size_t)std::max(nseg_max, 1) * sizeof(int64_t);
}

// Global-address-space view of a pointer read from a descriptor (global
// loads instead of flat ones, which would also count against the LDS wait
// counter).
template <typename T>
using gptr = const __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(const T* p) {
  return (gptr<T>)p;
}
template <typename T>
using gmut = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gmut<T> to_global_mut(T* p) {
  return (gmut<T>)p;
}
// The relaxed-barrier protocol of kcommon.hpp (agent-scope stores and loads
// of the exchanged values, relaxed counter) on global pointers: flat
// accesses would make every later LDS wait also wait for vector memory.
__device__ __forceinline__ void barrier_arrive_wait_g(unsigned int* ctr, unsigned int target, int* err) {
  const gmut<unsigned int> c = to_global_mut(ctr);
  const gmut<int> e = to_global_mut(err);
  __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned int spins = 0;
  while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(2);
    ++spins;
    if ((spins & 255u) == 0 && __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    if (spins > (1u << 20)) {
      __hip_atomic_fetch_or(e, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
  }
}
// base[byte_off / sizeof(T)] for a uniform base and a 32-bit byte offset
// (one scalar base register pair plus one vector register per load)
template <typename T>
__device__ __forceinline__ T load_at(const T* base, unsigned byte_off) {
  return *(gptr<T>)((const __attribute__((address_space(1))) char*)to_global(base) + byte_off);
}

// Value of v in lane l (l uniform) for every lane of the wave.
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Orders one wavefront's LDS and memory accesses around an exchange among its
// own lanes (no other wave takes part).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Descriptor owning workgroup `t` (start[] exclusive prefix sums, strictly
// increasing): every lane of the wave tests 64 entries per load round, so
// a launch of up to 256 descriptors costs one round of loads.
__device__ __forceinline__ int owner_by_ballot(const int* __restrict__ start, int n, int t) {
  const int lane = threadIdx.x & (kWave - 1);
  int cnt = 0;
  for (int s0 = 0; s0 < n; s0 += 4 * kWave) {
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int s = s0 + q * kWave + lane;
      v[q] = s < n ? start[s] : INT_MAX;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) cnt += __popcll(__ballot(v[q] <= t));
  }
  return cnt - 1;
}

template <typename T, int NRW>
__global__ __launch_bounds__((NRW + 1) * kWave) void k_reflr(const ReflDesc<T>* __restrict__ descs,
                                                             const int* __restrict__ wstart, int nd,
                                                             const Seg* __restrict__ segs, T* __restrict__ xbuf) {
  constexpr int kRRowWaves = NRW;
  extern __shared__ unsigned char smem_raw[];
  __shared__ T red[kRNB * kRRowWaves];
  __shared__ T tsh[kRNB * kTLd];
  __shared__ T ysh[kRNB];
  __shared__ T zsh[kRNB];
  // one workgroup per op (no op split): the descriptor is the block's own
  const int p = ((int)gridDim.x == nd) ? (int)blockIdx.x : owner_by_ballot(wstart, nd, blockIdx.x);
  const ReflDesc<T> d = descs[p];
  const int g = blockIdx.x - wstart[p];
  const int m = d.m, tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const bool leader = wv == kRRowWaves;
  const int nblk = (d.k + kRNB - 1) / kRNB;
  const bool multi = d.nwg > 1;
  const int r0 = multi ? (int)((int64_t)g * m / d.nwg) : 0;
  const int r1 = multi ? (int)((int64_t)(g + 1) * m / d.nwg) : m;
  const int i = r0 + tid;
  const bool own = !leader && i < r1;
  const int nseg = d.nseg;
  const T* __restrict__ V = d.V;
  const int64_t ldv = d.ldv;
  // V(i, jb + c) of block t: load_v issues 32 unconditional global loads,
  // each from a uniform column base plus the thread's row (rows above the
  // block, and idle threads, read row jb, one line per wave), and mask_v,
  // run when they are needed, applies the unit-lower structure (1 on the
  // diagonal, 0 above, 0 past column k).
  auto load_v = [&](int t, T (&raw)[kRNB]) {
    const int bi = d.trans ? t : nblk - 1 - t;
    const int jb = bi * kRNB, w = min(kRNB, d.k - jb);
    const unsigned off = (unsigned)((own && i > jb) ? i : min(jb, m - 1)) * (unsigned)sizeof(T);
    const T* col = V + (int64_t)jb * ldv;  // uniform; advanced column by column
    if (w == kRNB) {
#pragma unroll
      for (int c = 0; c < kRNB; ++c) {
        raw[c] = load_at(col, off);
        col += ldv;
      }
    } else {
#pragma unroll
      for (int c = 0; c < kRNB; ++c) {
        raw[c] = load_at(col, off);
        if (c + 1 < w) col += ldv;
      }
    }
  };
  // which rows of this wave the block touches (uniform per wave): none (all
  // rows above the block), all strictly below a full block with every row
  // owned (no masking), or some (per-entry masking)
  const int wrow0 = r0 + wv * kWave;
  auto wave_class = [&](int t) {
    const int bi = d.trans ? t : nblk - 1 - t;
    const int jb = bi * kRNB, w = min(kRNB, d.k - jb);
    if (leader || wrow0 >= r1 || wrow0 + kWave - 1 < jb) return 0;
    if (w == kRNB && wrow0 >= jb + kRNB && wrow0 + kWave <= r1) return 1;
    return 2;
  };
  auto mask_v = [&](int t, T (&v)[kRNB]) {
    const int bi = d.trans ? t : nblk - 1 - t;
    const int jb = bi * kRNB, w = min(kRNB, d.k - jb);
    const int dc = own ? i - jb : -1;  // the diagonal's column in this block, if any
#pragma unroll
    for (int c = 0; c < kRNB; ++c) v[c] = (c < w && c < dc) ? v[c] : ((c < w && c == dc) ? T(1) : T(0));
  };
  // the leader loads T_b of block t into registers (entry e = lane + 64 q is
  // (e % 32, e / 32) = (lane % 32, lane / 32 + 2 q)) and later stages it
  // into LDS (ld kTLd)
  constexpr int kTQ = kRNB * kRNB / kWave;
  auto load_t = [&](int t, T (&tv)[kTQ]) {
    const int bi = d.trans ? t : nblk - 1 - t;
    const gptr<T> tb = to_global(d.Tb + (int64_t)bi * kRNB * kRNB + lane);
#pragma unroll
    for (int q = 0; q < kTQ; ++q) tv[q] = tb[q * kWave];
  };
  auto store_t = [&](const T (&tv)[kTQ]) {
    T* dst = tsh + (lane & (kRNB - 1)) + (lane / kRNB) * kTLd;
#pragma unroll
    for (int q = 0; q < kTQ; ++q) dst[q * 2 * kTLd] = tv[q];
  };
  // the first block's loads go out before the gather; every wave, the
  // leader included (its loads all read row jb), keeps the register sets
  // defined alike on all paths, so the compiler's wait counters see them
  // as complete wherever they are read
  T va[kRNB], vb[kRNB];
  T tv[kTQ];
  if (nblk > 0) {
    load_v(0, va);
    if (leader) load_t(0, tv);
  }
  // segment table in LDS: row base (exclusive prefix of the lengths) and
  // x buffer offset of every segment, built by the first wave
  int* sbase = reinterpret_cast<int*>(smem_raw);
  int64_t* soff = reinterpret_cast<int64_t*>(smem_raw + (((size_t)nseg + 1) * sizeof(int) + 7) / 8 * 8);
  if (wv == 0) {
    int carry = 0;
    for (int s0 = 0; s0 < nseg; s0 += kWave) {
      const int s = s0 + lane;
      int len = 0;
      if (s < nseg) {
        const Seg sg = segs[d.seg0 + s];
        len = sg.len;
        soff[s] = sg.off;
      }
      int v = len;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const int u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
      }
      if (s < nseg) sbase[s] = carry + v - len;
      carry += __shfl(v, kWave - 1, kWave);
    }
  }
  __syncthreads();
  int64_t xo = 0;
  T xr = T(0);
  if (own) {
    int lo = 0, hi = nseg - 1;  // the last segment whose base is <= i holds row i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sbase[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    xo = soff[lo] + (i - sbase[lo]);
    xr = xbuf[xo];
  }
  if (leader && nblk > 0) store_t(tv);
  unsigned int target = 0;
  T* __restrict__ part = d.part;
  // one block: partial sums from `cur` (block t), prefetch of block t + 1
  // into `nxt` (and of T_{t+1} by the leader, staged after the second
  // barrier), the leader's reduction / exchange / T_b product, update
  auto step = [&](int t, T (&cur)[kRNB], T (&nxt)[kRNB]) {
    const int bi = d.trans ? t : nblk - 1 - t;
    const int w = min(kRNB, d.k - bi * kRNB);
    const int cls = wave_class(t);
    if (cls == 2) mask_v(t, cur);
    if (t + 1 < nblk) load_v(t + 1, nxt);  // in flight across this block's work
    if (!leader) {
      T sw = T(0);
      if (cls != 0) {
        T acc[kRNB];
#pragma unroll
        for (int c = 0; c < kRNB; ++c) acc[c] = cur[c] * xr;
        sw = wave_reduce_scatter32_nolds(acc);
      }
      if ((lane & 1) == 0) red[wv * kRNB + (lane >> 1)] = sw;
    }
    __syncthreads();
    if (leader) {
      const int c = lane;
      T y = T(0);
      if (c < kRNB) {
#pragma unroll
        for (int q = 0; q < kRRowWaves; ++q) y += red[q * kRNB + c];
      }
      if (multi) {
        const gmut<T> pb = to_global_mut(part) + (int64_t)(t & 1) * d.nwg * kRNB + c;
        if (c < kRNB) __hip_atomic_store(pb + g * kRNB, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);  // this wave's partials have landed
        target += d.nwg;
        if (lane == 0) barrier_arrive_wait_g(d.sync, target, d.err);
        wave_sync();
        // the partials of all workgroups, 8 loads in flight at a time,
        // summed in workgroup order (every workgroup forms the same y)
        y = T(0);
        for (int q0 = 0; q0 < d.nwg; q0 += 8) {
          T pv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u)
            pv[u] = __hip_atomic_load(pb + min(q0 + u, d.nwg - 1) * kRNB, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
          for (int u = 0; u < 8; ++u) y += (q0 + u < d.nwg) ? pv[u] : T(0);
        }
      }
      // z = T_b^T y (trans) or T_b y.  y goes through LDS; lane c (and its
      // partner c + 32) combines entries l of column c of T_b (trans) or of
      // row c (no trans), the lower lanes l < 16 and the upper ones l >= 16,
      // with all 16 LDS reads of T_b (conflict-free through the padded
      // leading dimension) and of y issued before the FMAs; entries outside
      // the triangle [lo, hi] are never combined
      if (c < kRNB) ysh[c] = y;
      wave_sync();
      const int cc = c & (kRNB - 1), half = (c >> 5) * (kRNB / 2);
      const int tb0 = d.trans ? cc * kTLd : cc, tstr = d.trans ? 1 : kTLd;
      const int lo = d.trans ? 0 : cc, hi = d.trans ? cc : w - 1;
      T tl[kRNB / 2], yl[kRNB / 2];
#pragma unroll
      for (int u = 0; u < kRNB / 2; ++u) {
        tl[u] = tsh[tb0 + (half + u) * tstr];
        yl[u] = ysh[half + u];
      }
      T z0 = T(0), z1 = T(0);
#pragma unroll
      for (int u = 0; u < kRNB / 2; u += 2) {
        const int l = half + u;
        z0 = (l >= lo && l <= hi) ? fma(tl[u], yl[u], z0) : z0;
        z1 = (l + 1 >= lo && l + 1 <= hi) ? fma(tl[u + 1], yl[u + 1], z1) : z1;
      }
      T z = z0 + z1;
      z += __shfl_xor(z, kRNB, kWave);
      if (c < kRNB) zsh[c] = (c < w) ? z : T(0);
    }
    __syncthreads();
    if (cls != 0) {
      T s4[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int c = 0; c < kRNB; ++c) s4[c & 3] += cur[c] * zsh[c];
      xr -= (s4[0] + s4[1]) + (s4[2] + s4[3]);
    }
    // T_{t+1}, loaded while the row waves update and form the next sums;
    // read again only after the next block's first barrier
    if (leader && t + 1 < nblk) {
      T tn[kTQ];
      load_t(t + 1, tn);
      store_t(tn);
    }
  };
  // two register sets that swap roles block by block (no copies)
  for (int t = 0; t < nblk; t += 2) {
    step(t, va, vb);
    if (t + 1 < nblk) step(t + 1, vb, va);
  }
  if (own) xbuf[xo] = xr;
}

template <typename T>
void launch_reflr(hipStream_t s, const ReflDesc<T>* d, const int* wstart, int nd, int nwg_total, const Seg* segs, T* xbuf,
                  int nseg_max, int nrw) {
  if (nd <= 0) return;
  const size_t lds = reflr_lds(nseg_max);
  if (nrw == 1)
    hipLaunchKernelGGL((k_reflr<T, 1>), dim3(nwg_total), dim3(2 * kWave), lds, s, d, wstart, nd, segs, xbuf);
  else if (nrw == 2)
    hipLaunchKernelGGL((k_reflr<T, 2>), dim3(nwg_total), dim3(3 * kWave), lds, s, d, wstart, nd, segs, xbuf);
  else
    hipLaunchKernelGGL((k_reflr<T, 7>), dim3(nwg_total), dim3(8 * kWave), lds, s, d, wstart, nd, segs, xbuf);
}

// Solve the w x w diagonal block (staged in LDS, ld kTLd) in place on
// xs[jb:jb+w] with the first wavefront (lane l holds x_{jb+l});
// trans: R^T (forward substitution).
template <typename T>
__device__ __forceinline__ void diag_solve(const T* rsh, T* __restrict__ xs, int jb, int w, int trans, bool shared) {
  if (threadIdx.x >= kWave) return;
  const int l = threadIdx.x, lc = l & (kRNB - 1);
  T xl = (l < w) ? xs[jb + l] : T(0);
  // lane l's row of the block (its column for R^T) and 1 / R[l, l], read
  // together; x_i is broadcast by readlane (all 64 lanes of the wave run)
  T rr[kRNB];
#pragma unroll
  for (int i = 0; i < kRNB; ++i) rr[i] = trans ? rsh[i + lc * kTLd] : rsh[lc + i * kTLd];
  const T dinv = T(1) / rsh[lc + lc * kTLd];
  if (!trans) {
#pragma unroll
    for (int i = kRNB - 1; i >= 0; --i) {
      if (i < w) {
        if (l == i) xl *= dinv;
        const T xi = lane_bcast(xl, i);
        if (l < i) xl -= rr[i] * xi;
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < kRNB; ++i) {
      if (i < w) {
        if (l == i) xl *= dinv;
        const T xi = lane_bcast(xl, i);
        if (l > i && l < w) xl -= rr[i] * xi;
      }
    }
  }
  if (l < w) {
    if (shared) agent_store(xs + jb + l, xl);  // read by the other workgroups after the barrier
    else xs[jb + l] = xl;
  }
}

template <typename T>
__global__ __launch_bounds__(kSThreads) void k_bwd(const BwdDesc<T>* __restrict__ descs, const int* __restrict__ wstart,
                                                   int nd, const R12Desc<T>* __restrict__ r12, T* __restrict__ xbuf) {
  __shared__ T rsh[kRNB * kTLd];
  __shared__ T xsh[kRNB];
  __shared__ T red[kSThreads];
  const int p = ((int)gridDim.x == nd) ? (int)blockIdx.x : find_owner(wstart, nd, blockIdx.x);
  const BwdDesc<T> d = descs[p];
  const int g = blockIdx.x - wstart[p], nwg = d.nwg;
  const int k = d.k, tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int nblk = (k + kRNB - 1) / kRNB;
  const int64_t ldr = d.ldr;
  T* __restrict__ xs = xbuf + d.xoff;
  // row slices in whole 32-row blocks: blocks [b0, b1) belong to this workgroup
  const int b0 = (int)((int64_t)g * nblk / nwg), b1 = (int)((int64_t)(g + 1) * nblk / nwg);
  const int r0 = b0 * kRNB, r1 = min(k, b1 * kRNB);
  unsigned int target = 0;
  const bool shared = nwg > 1;
  auto meet = [&]() {
    if (shared) {
      target += nwg;
      group_barrier_rx(d.sync, target, d.err);
    } else {
      __syncthreads();
    }
  };
  auto xget = [&](int r) { return shared ? agent_load(xs + r) : xs[r]; };
  auto stage_diag = [&](int jb, int w) {
    for (int e = tid; e < kRNB * kRNB; e += kSThreads) {
      const int r = e & (kRNB - 1), c = e / kRNB;
      if (r < w && c < w) rsh[r + c * kTLd] = d.R[(jb + r) + (int64_t)(jb + c) * ldr];
    }
    __syncthreads();
  };
  if (!d.trans) {
    // x_c[r] -= sum_n A_n[r, 0:s] x_n[0:s] for the owned rows: lanes take
    // rows, waves take interleaved columns, LDS sums the waves
    for (int rb = r0; rb < r1; rb += kWave) {
      const int r = rb + lane;
      T acc = 0;
      if (r < r1) {
        for (int b = 0; b < d.nblk; ++b) {
          const R12Desc<T> q = r12[d.blk0 + b];
          const T* __restrict__ xn = xbuf + q.xoff;
          for (int j = wv; j < q.s; j += kSWaves) acc += q.A[r + (int64_t)j * q.ld] * xn[j];
        }
      }
      red[tid] = acc;
      __syncthreads();
      if (wv == 0 && r < r1) {
        T s = 0;
#pragma unroll
        for (int q = 0; q < kSWaves; ++q) s += red[q * kWave + lane];
        xs[r] -= s;
      }
      __syncthreads();
    }
    // blocked back substitution, last block first
    for (int bi = nblk - 1; bi >= 0; --bi) {
      const int jb = bi * kRNB, w = min(kRNB, k - jb);
      if (bi >= b0 && bi < b1) {
        stage_diag(jb, w);
        diag_solve(rsh, xs, jb, w, 0, shared);
      }
      meet();
      if (tid < w) xsh[tid] = xget(jb + tid);
      __syncthreads();
      for (int r = r0 + tid; r < min(r1, jb); r += kSThreads) {
        T s = 0;
        for (int c = 0; c < w; ++c) s += d.R[r + (int64_t)(jb + c) * ldr] * xsh[c];
        xs[r] -= s;
      }
      __syncthreads();
    }
  } else {
    // blocked forward substitution with R^T, first block first
    const int hl = tid & (kRNB - 1), hw = tid / kRNB;  // half-wave lane / index
    for (int bi = 0; bi < nblk; ++bi) {
      const int jb = bi * kRNB, w = min(kRNB, k - jb);
      if (bi >= b0 && bi < b1) {
        stage_diag(jb, w);
        diag_solve(rsh, xs, jb, w, 1, shared);
      }
      meet();
      if (tid < w) xsh[tid] = xget(jb + tid);
      __syncthreads();
      // x[r] -= sum_c R[jb+c, r] x[jb+c]: a half-wave reads the contiguous
      // 32-entry column segment of R for one row r
      for (int rb = max(r0, jb + w); rb < r1; rb += kSThreads / kRNB) {
        const int r = rb + hw;
        T v = (r < r1 && hl < w) ? d.R[(jb + hl) + (int64_t)r * ldr] * xsh[hl] : T(0);
#pragma unroll
        for (int o = kRNB / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        if (hl == 0 && r < r1) xs[r] -= v;
      }
      __syncthreads();
    }
    if (shared) {
      target += nwg;
      group_barrier_rx(d.sync, target, d.err);
    }
    // x_n[j] -= A_n[0:k, j]^T x_c[0:k]: one wavefront per column, columns
    // interleaved over all waves of the problem (atomic: several clusters of
    // one stage share x_n)
    for (int b = 0; b < d.nblk; ++b) {
      const R12Desc<T> q = r12[d.blk0 + b];
      for (int j = g * kSWaves + wv; j < q.s; j += nwg * kSWaves) {
        T acc = 0;
        for (int r = lane; r < k; r += kWave) acc += q.A[r + (int64_t)j * q.ld] * xget(r);
        acc = wave_sum(acc);
        if (lane == 0) atomicAdd(xbuf + q.xoff + j, -acc);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Latency form of the triangular sweep (ops without in-kernel R12 products,
// one workgroup each, k <= bwdr_max_k()): x_c stays in LDS; per 32-column
// block the first wave solves the diagonal block from registers (its row of
// R_bb, or column for R^T, loaded one block ahead; reciprocals of the
// diagonal formed in parallel; x_i broadcast by readlane), while the other
// waves load their R entries of the block's off-diagonal update, which they
// apply once x_b is known.
constexpr int kBThreads = 512;
constexpr int kBRows = kBThreads - kWave;  // update rows per pass (waves 1..7)
constexpr int kBMaxK = 8192;

int bwdr_max_k() { return kBMaxK; }

template <typename T>
__global__ __launch_bounds__(kBThreads) void k_bwdr(const BwdDesc<T>* __restrict__ descs, T* __restrict__ xbuf) {
  extern __shared__ unsigned char smem_raw[];
  T* xs = reinterpret_cast<T*>(smem_raw);  // k
  const BwdDesc<T> d = descs[blockIdx.x];
  const int k = d.k, tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int nblk = (k + kRNB - 1) / kRNB;
  const int64_t ldr = d.ldr;
  const T* __restrict__ R = d.R;
  T* __restrict__ xg = xbuf + d.xoff;
  for (int r = tid; r < k; r += kBThreads) xs[r] = xg[r];
  const bool tr = d.trans != 0;
  // wave 0, lane l < 32: entries (l, c) of the block (no trans: row l of
  // R_bb; trans: column l, i.e. R[jb + c, jb + l]); entries outside the
  // triangle or past k read an in-bounds element and are never used
  auto load_diag = [&](int bi, T (&rr)[kRNB]) {
    const int jb = bi * kRNB, w = min(kRNB, k - jb);
    const int l = min(lane & (kRNB - 1), w - 1);
#pragma unroll
    for (int c = 0; c < kRNB; ++c) {
      const int cc = min(c, w - 1);
      rr[c] = tr ? R[(jb + cc) + (int64_t)(jb + l) * ldr] : R[(jb + l) + (int64_t)(jb + cc) * ldr];
    }
  };
  // update thread (waves 1..7) q-th row of the pass starting at row0:
  // entries R[r, jb + c] (no trans) or R[jb + c, r] (trans)
  auto load_upd = [&](int bi, int r, T (&ru)[kRNB]) {
    const int jb = bi * kRNB, w = min(kRNB, k - jb);
    const int rc = min(max(r, 0), k - 1);
#pragma unroll
    for (int c = 0; c < kRNB; ++c) {
      const int cc = min(c, w - 1);
      ru[c] = tr ? R[(jb + cc) + (int64_t)rc * ldr] : R[rc + (int64_t)(jb + cc) * ldr];
    }
  };
  T rr[kRNB];
  if (wv == 0 && nblk > 0) load_diag(tr ? 0 : nblk - 1, rr);
  __syncthreads();
  for (int t = 0; t < nblk; ++t) {
    const int bi = tr ? t : nblk - 1 - t;
    const int jb = bi * kRNB, w = min(kRNB, k - jb);
    // rows updated by this block: [0, jb) (no trans) or [jb + w, k) (trans)
    const int ulo = tr ? jb + w : 0, uhi = tr ? k : jb;
    T ru[kRNB];
    const int r1st = ulo + (tid - kWave);  // first pass row of an update thread
    if (wv > 0 && r1st < uhi) load_upd(bi, r1st, ru);
    if (wv == 0) {
      // diagonal solve: lane l (< w) holds x_{jb+l}
      T xl = lane < w ? xs[jb + lane] : T(0);
      // 1 / R[jb+l, jb+l]: entry c = l of the lane's row (column)
      T diag = rr[0];
#pragma unroll
      for (int c = 1; c < kRNB; ++c) diag = (c == (lane & (kRNB - 1))) ? rr[c] : diag;
      const T dinv = T(1) / diag;
      if (!tr && w == kRNB) {
#pragma unroll
        for (int i = kRNB - 1; i >= 0; --i) {
          if (lane == i) xl *= dinv;
          const T xi = lane_bcast(xl, i);
          if (lane < i) xl -= rr[i] * xi;
        }
      } else if (!tr) {
#pragma unroll
        for (int i = kRNB - 1; i >= 0; --i) {
          if (i < w) {
            if (lane == i) xl *= dinv;
            const T xi = lane_bcast(xl, i);
            if (lane < i) xl -= rr[i] * xi;
          }
        }
      } else if (w == kRNB) {
#pragma unroll
        for (int i = 0; i < kRNB; ++i) {
          if (lane == i) xl *= dinv;
          const T xi = lane_bcast(xl, i);
          if (lane > i && lane < kRNB) xl -= rr[i] * xi;
        }
      } else {
#pragma unroll
        for (int i = 0; i < kRNB; ++i) {
          if (i < w) {
            if (lane == i) xl *= dinv;
            const T xi = lane_bcast(xl, i);
            if (lane > i && lane < w) xl -= rr[i] * xi;
          }
        }
      }
      if (lane < w) xs[jb + lane] = xl;
      // the next block's diagonal, in flight across the update
      if (t + 1 < nblk) load_diag(tr ? t + 1 : nblk - 2 - t, rr);
    }
    __syncthreads();
    if (wv > 0) {
      for (int r0 = ulo; r0 < uhi; r0 += kBRows) {
        const int r = r0 + (tid - kWave);
        if (r0 > ulo && r < uhi) load_upd(bi, r, ru);
        if (r < uhi) {
          // four chains; a full block needs no column test
          T s4[4] = {T(0), T(0), T(0), T(0)};
          if (w == kRNB) {
#pragma unroll
            for (int c = 0; c < kRNB; ++c) s4[c & 3] += ru[c] * xs[jb + c];
          } else {
#pragma unroll
            for (int c = 0; c < kRNB; ++c)
              if (c < w) s4[c & 3] += ru[c] * xs[jb + c];
          }
          xs[r] -= (s4[0] + s4[1]) + (s4[2] + s4[3]);
        }
      }
    }
    __syncthreads();
  }
  for (int r = tid; r < k; r += kBThreads) xg[r] = xs[r];
}

template <typename T>
void launch_bwdr(hipStream_t s, const BwdDesc<T>* d, int nd, int kmax, T* xbuf) {
  if (nd <= 0) return;
  hipLaunchKernelGGL((k_bwdr<T>), dim3(nd), dim3(kBThreads), sizeof(T) * (size_t)std::max(kmax, 1), s, d, xbuf);
}

// R12 products of a solve stage as one batched GEMV over column chunks
// (one workgroup per chunk), accumulated with atomics: the products of a
// stage are independent of its triangular sweeps and stream far more data.
template <typename T>
__global__ __launch_bounds__(kSThreads) void k_r12(const R12Chunk<T>* __restrict__ d, T* __restrict__ xbuf) {
  const R12Chunk<T> q = d[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  if (!q.trans) {
    const T* __restrict__ xn = xbuf + q.xn;
    for (int r = tid; r < q.k; r += kSThreads) {
      T acc = 0;
      for (int j = 0; j < q.nc; ++j) acc += q.A[r + (int64_t)j * q.ld] * xn[j];
      atomicAdd(xbuf + q.xc + r, -acc);
    }
  } else {
    const T* __restrict__ xc = xbuf + q.xc;
    for (int j = wv; j < q.nc; j += kSWaves) {
      T acc = 0;
      for (int r = lane; r < q.k; r += kWave) acc += q.A[r + (int64_t)j * q.ld] * xc[r];
      acc = wave_sum(acc);
      if (lane == 0) atomicAdd(xbuf + q.xn + j, -acc);
    }
  }
}
int r12_chunk_cols() { return 64; }
template <typename T>
void launch_r12(hipStream_t s, const R12Chunk<T>* d, int nd, T* xbuf) {
  if (nd <= 0) return;
  hipLaunchKernelGGL((k_r12<T>), dim3(nd), dim3(kSThreads), 0, s, d, xbuf);
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
void launch_refl(hipStream_t s, const ReflDesc<T>* d, const int* wstart, int nd, int nwg_total, const Seg* segs,
                 T* xbuf, T* work, int mmax) {
  if (nd <= 0) return;
  const int cap = refl_lds_cap(sizeof(T));
  size_t shm = (mmax <= cap) ? sizeof(T) * (size_t)mmax : 0;
  hipLaunchKernelGGL((k_refl<T>), dim3(nwg_total), dim3(kSThreads), shm, s, d, wstart, nd, segs, xbuf, work, cap, mmax);
}
template <typename T>
void launch_bwd(hipStream_t s, const BwdDesc<T>* d, const int* wstart, int nd, int nwg_total, const R12Desc<T>* r12,
                T* xbuf) {
  if (nd <= 0) return;
  hipLaunchKernelGGL((k_bwd<T>), dim3(nwg_total), dim3(kSThreads), 0, s, d, wstart, nd, r12, xbuf);
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

#define SPQ_INST(T)                                                                                            \
  template void launch_refl<T>(hipStream_t, const ReflDesc<T>*, const int*, int, int, const Seg*, T*, T*, int); \
  template void launch_reflr<T>(hipStream_t, const ReflDesc<T>*, const int*, int, int, const Seg*, T*, int, int); \
  template void launch_bwd<T>(hipStream_t, const BwdDesc<T>*, const int*, int, int, const R12Desc<T>*, T*);     \
  template void launch_bwdr<T>(hipStream_t, const BwdDesc<T>*, int, int, T*);                                 \
  template void launch_move<T>(hipStream_t, const MoveDesc*, int, T*, const T*);                               \
  template void launch_r12<T>(hipStream_t, const R12Chunk<T>*, int, T*);                                        \
  template void launch_gather_in<T>(hipStream_t, const int64_t*, const int*, int, const double*, T*);          \
  template void launch_scatter_out<T>(hipStream_t, const int64_t*, const int*, int, const T*, double*);
SPQ_INST(double)
SPQ_INST(float)

}  // namespace spq
