// Shared device helpers for the spaQR gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "descs.hpp"

namespace spq {

constexpr int kWave = 64;

// Launch of a grid whose workgroups wait for one another (sized by the
// caller within the device's resident capacity).  The cooperative launch
// guarantees co-residency but rests on a device-wide resource that another
// process using the GPU can hold, and the runtime then refuses it
// (hipErrorIllegalState).  The same grid is then launched plainly: on a
// device with no other work its workgroups are all resident (the solve's
// multi-workgroup stages rely on the same), and every barrier wait in these
// kernels is bounded, so a grid that is not co-resident sets the error flag
// and ends instead of hanging.
inline hipError_t launch_coresident(const void* f, dim3 grid, dim3 block, void** args, size_t shm, hipStream_t s) {
  hipError_t e = hipLaunchCooperativeKernel(f, grid, block, args, shm, s);
  if (e == hipErrorIllegalState) {
    (void)hipGetLastError();  // clear the refusal
    e = hipLaunchKernel(f, grid, block, args, shm, s);
  }
  return e;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// ---------------------------------------------------------------------------
// Cross-lane exchange without LDS (DPP and the gfx950 permlane swaps).
template <int N>
struct Dw {  // a value as N 32-bit words
  unsigned w[N];
};
template <typename T>
__device__ __forceinline__ Dw<sizeof(T) / 4> to_dw(T v) {
  Dw<sizeof(T) / 4> r;
  __builtin_memcpy(r.w, &v, sizeof(T));
  return r;
}
template <typename T>
__device__ __forceinline__ T from_dw(const Dw<sizeof(T) / 4>& r) {
  T v;
  __builtin_memcpy(&v, r.w, sizeof(T));
  return v;
}
template <int CTRL>
__device__ __forceinline__ unsigned dpp_w(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
// Value of lane l ^ X (X a power of two below 64): DPP within 16-lane rows
// (quad_perm for 1 and 2, row_half_mirror then quad_perm 3210 for 4,
// row_ror:8 for 8), a permlane16 / permlane32 swap of the value with itself
// for 16 / 32.  The partner lane must be active (callers run it converged,
// or within groups of lanes that are active together).
template <int X, typename T>
__device__ __forceinline__ T lane_xor(T v) {
  auto a = to_dw(v);
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int q = 0; q < (int)(sizeof(T) / 4); ++q) {
    if (X == 1) a.w[q] = dpp_w<0xB1>(a.w[q]);
    else if (X == 2) a.w[q] = dpp_w<0x4E>(a.w[q]);
    else if (X == 4) a.w[q] = dpp_w<0x1B>(dpp_w<0x141>(a.w[q]));
    else if (X == 8) a.w[q] = dpp_w<0x128>(a.w[q]);
    else if (X == 16) {
      const auto r = __builtin_amdgcn_permlane16_swap(a.w[q], a.w[q], false, false);
      a.w[q] = (lane & 16) ? r[0] : r[1];
    } else {
      const auto r = __builtin_amdgcn_permlane32_swap(a.w[q], a.w[q], false, false);
      a.w[q] = (lane & 32) ? r[0] : r[1];
    }
  }
  return from_dw<T>(a);
}
// wave_sum with the same pairing order through lane_xor (converged waves)
template <typename T>
__device__ __forceinline__ T wave_sum_x(T v) {
  v += lane_xor<32>(v);
  v += lane_xor<16>(v);
  v += lane_xor<8>(v);
  v += lane_xor<4>(v);
  v += lane_xor<2>(v);
  v += lane_xor<1>(v);
  return v;
}

// Reduce-scatter of 32 per-lane values over the wave without LDS: the
// distance-32 and distance-16 steps swap half-waves / rows with
// v_permlane32_swap / v_permlane16_swap (no selects), the steps within a
// 16-lane row exchange through DPP (lane_xor).  On return lanes 2c and
// 2c + 1 hold the sum over the wave of a[c].
// x <- x + y after swapping x's upper half-wave with y's lower one (32) or
// x's odd rows with y's even rows (16): lanes then hold the pair sums of x
// (lower half / even rows) and of y (upper half / odd rows)
template <typename T, int D>
__device__ __forceinline__ T swap_add(T x, T y) {
  auto a = to_dw(x), b = to_dw(y);
#pragma unroll
  for (int q = 0; q < (int)(sizeof(T) / 4); ++q) {
    const auto r = D == 32 ? __builtin_amdgcn_permlane32_swap(a.w[q], b.w[q], false, false)
                           : __builtin_amdgcn_permlane16_swap(a.w[q], b.w[q], false, false);
    a.w[q] = r[0];
    b.w[q] = r[1];
  }
  return from_dw<T>(a) + from_dw<T>(b);
}
template <typename T, int H, int X>
__device__ __forceinline__ void rs_row_step(T (&a)[32], int lane) {
  const bool up = (lane & X) != 0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const T keep = up ? a[j + H] : a[j];
    const T send = up ? a[j] : a[j + H];
    a[j] = keep + lane_xor<X>(send);
  }
}
template <typename T>
__device__ __forceinline__ T wave_reduce_scatter32_nolds(T (&a)[32]) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int j = 0; j < 16; ++j) a[j] = swap_add<T, 32>(a[j], a[j + 16]);
#pragma unroll
  for (int j = 0; j < 8; ++j) a[j] = swap_add<T, 16>(a[j], a[j + 8]);
  rs_row_step<T, 4, 8>(a, lane);
  rs_row_step<T, 2, 4>(a, lane);
  rs_row_step<T, 1, 2>(a, lane);
  return a[0] + lane_xor<1>(a[0]);
}

// Block-wide sum of one value; `red` must hold >= blockDim.x/64 entries.
template <typename T>
__device__ __forceinline__ T block_sum_x(T v, T* red) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  v = wave_sum_x(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// Block-wide sums of N values per thread (N small).  `red` holds N*nwaves.
template <typename T, int N>
__device__ __forceinline__ void block_sum_n(T (&v)[N], T* red, int n) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < N; ++c) {
    if (c < n) {
      T s = wave_sum(v[c]);
      if (lane == 0) red[c * nw + w] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < N; ++c) {
    if (c < n) {
      T s = 0;
      for (int i = 0; i < nw; ++i) s += red[c * nw + i];
      v[c] = s;
    }
  }
}

// Butterfly reduce-scatter of 32 per-lane partial sums over a wavefront: each
// step exchanges half of the live values with lane ^ H, so 31 shuffles leave
// lane l with the sum over its 32-lane half of a[l & 31]; one more folds the
// two halves.  All indices are compile-time, so a[] stays in registers.
template <typename T, int H>
__device__ __forceinline__ void rs_step(T (&a)[32], int lane) {
  const bool up = (lane & H) != 0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const T keep = up ? a[j + H] : a[j];
    const T send = up ? a[j] : a[j + H];
    a[j] = keep + __shfl_xor(send, H, kWave);
  }
}
template <typename T>
__device__ __forceinline__ T wave_reduce_scatter32(T (&a)[32]) {
  const int lane = threadIdx.x & (kWave - 1);
  rs_step<T, 16>(a, lane);
  rs_step<T, 8>(a, lane);
  rs_step<T, 4>(a, lane);
  rs_step<T, 2>(a, lane);
  rs_step<T, 1>(a, lane);
  return a[0] + __shfl_xor(a[0], 32, kWave);
}

// Inter-workgroup barrier on a monotonically increasing counter (target =
// arrivals expected so far).  All workgroups of the group must be resident;
// the wait is bounded so the kernel always terminates: a wait that gives up
// sets err, and a workgroup that sees err set stops waiting as well (the
// results are then invalid and the host reports the error).
__device__ __forceinline__ void group_barrier(unsigned int* ctr, unsigned int target, int* err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(ctr, 1u);
    unsigned int spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      ++spins;
      if ((spins & 255u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (spins > (1u << 20)) { atomicOr(err, 1); break; }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Two-level form for large groups.  Workgroup g arrives on fan-in counter
// 1 + g/16 (counters 64 bytes apart in `sync`); the last arrival of a fan-in
// group in a phase forwards one arrival to the root counter sync[0], which
// every workgroup polls.  All counters start at zero; `phase` numbers the
// barriers 1, 2, ...  Waits are bounded as in group_barrier.
__device__ __forceinline__ void group_barrier2(unsigned int* sync, int g, int nwg, unsigned int phase, int* err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const int grp = g / kFanIn, ngrp = (nwg + kFanIn - 1) / kFanIn;
    const unsigned int sz = (unsigned int)min(kFanIn, nwg - grp * kFanIn);
    __threadfence();
    const unsigned int old = atomicAdd(sync + 16 * (1 + grp), 1u);
    if (old + 1 == phase * sz) {
      // last of the fan-in group: acquire its members' writes, release them on
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __threadfence();
      atomicAdd(sync, 1u);
    }
    const unsigned int target = phase * (unsigned int)ngrp;
    unsigned int spins = 0;
    while (__hip_atomic_load(sync, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      ++spins;
      if ((spins & 255u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (spins > (1u << 20)) { atomicOr(err, 1); break; }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
// ---------------------------------------------------------------------------
// Barriers without cache maintenance.  Values a workgroup hands to the
// others go through agent-scope atomic stores and loads (coherent across
// XCDs without writing back or invalidating the L2); a barrier then needs
// only each wave's stores to have completed before the arrival and relaxed
// counter traffic.  Measured on MI355X at 12 / 83 / 170 workgroups: 1.6 /
// 2.3 / 3.3 us per barrier, against 3.8 / 13.9 / 26.5 us for the fenced
// form above (whose per-workgroup L2 write-back serialises).  Only data
// accessed through agent_store / agent_load may cross a *_rx barrier.
template <typename T>
__device__ __forceinline__ void agent_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T agent_load(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Ordering.  The handed-over values and the counters are all agent-scope
// atomics, which live at the device's coherence point.  The hardware issues
// a wave's memory operations in program order and does not speculate past
// the spin's branch; what the compiler could reorder (relaxed atomics to
// different addresses) is pinned by compiler-only fences: no store moves
// below the arrival's s_waitcnt, no load above the end of the spin.
__device__ __forceinline__ bool barrier_wait_rx(unsigned int* ctr, unsigned int target, int* err) {
  unsigned int spins = 0;
  bool ok = true;
  while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    __builtin_amdgcn_s_sleep(2);
    ++spins;
    if ((spins & 255u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      ok = false;
      break;
    }
    if (spins > (1u << 20)) { atomicOr(err, 1); ok = false; break; }
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return ok;
}
__device__ __forceinline__ void group_barrier_rx(unsigned int* ctr, unsigned int target, int* err) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);  // this wave's agent stores have completed
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    barrier_wait_rx(ctr, target, err);
  }
  __syncthreads();
}
// Two-level form (fan-in counters as in group_barrier2).
__device__ __forceinline__ void group_barrier2_rx(unsigned int* sync, int g, int nwg, unsigned int phase, int* err) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int grp = g / kFanIn, ngrp = (nwg + kFanIn - 1) / kFanIn;
    const unsigned int sz = (unsigned int)min(kFanIn, nwg - grp * kFanIn);
    const unsigned int old =
        __hip_atomic_fetch_add(sync + 16 * (1 + grp), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1 == phase * sz) __hip_atomic_fetch_add(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    barrier_wait_rx(sync, phase * (unsigned int)ngrp, err);
  }
  __syncthreads();
}

// Find the descriptor owning flattened tile `t` given exclusive prefix sums
// `start[0..n)` (start[0] == 0, nondecreasing).
// A 64-ary search by the whole wave: each round every lane reads one sample
// of the current range (one load round trip instead of one per halving), so
// n <= 4096 takes two rounds.  Must be called by full, converged waves with
// the same t (every caller passes its block's tile index at kernel entry).
__device__ __forceinline__ int find_owner(const int* start, int n, int t) {
  const int lane = threadIdx.x & (kWave - 1);
  int lo = 0, hi = n;  // start[lo] <= t, and hi == n or start[hi] > t
  while (hi - lo > 1) {
    const int step = (hi - lo + kWave - 1) / kWave;
    const int idx = lo + lane * step;
    const bool le = idx < hi && start[idx] <= t;
    const unsigned long long m = __ballot(le);
    const int last = kWave - 1 - __clzll((long long)m);  // lane 0 always qualifies
    const int nlo = lo + last * step;
    hi = min(hi, nlo + step);
    lo = nlo;
  }
  return lo;
}

}  // namespace spq
