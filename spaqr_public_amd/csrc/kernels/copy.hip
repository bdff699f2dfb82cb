// Batched 2-D block copies (gather / scatter / merge / transpose / precision
// conversion) over descriptors of uneven size: one 32x32 tile per workgroup,
// tiles flattened across descriptors with an exclusive prefix-sum table.
#include <algorithm>

#include "kcommon.hpp"
#include "descs.hpp"
#include "launch.hpp"

namespace spq {

template <typename Ts, typename Td>
__global__ __launch_bounds__(256) void k_copy(const CopyDesc* __restrict__ d, const int* __restrict__ tstart,
                                              int nd) {
  const int t = blockIdx.x;
  const int p = find_owner(tstart, nd, t);
  const CopyDesc D = d[p];
  const int local = t - tstart[p];
  const int ntr = (D.rows + 31) >> 5;
  const int r0 = (local % ntr) * 32, c0 = (local / ntr) * 32;
  const Ts* __restrict__ src = (const Ts*)D.src;
  Td* __restrict__ dst = (Td*)D.dst;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (!D.trans) {
    for (int c = ty; c < 32; c += 8) {
      const int i = r0 + tx, j = c0 + c;
      if (i < D.rows && j < D.cols) dst[i + (int64_t)j * D.ldd] = (Td)src[i + (int64_t)j * D.lds];
    }
  } else {
    __shared__ Ts tile[32][33];
    for (int c = ty; c < 32; c += 8) {
      const int i = r0 + tx, j = c0 + c;
      if (i < D.rows && j < D.cols) tile[c][tx] = src[i + (int64_t)j * D.lds];
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
      const int i = r0 + c, j = c0 + tx;
      if (i < D.rows && j < D.cols) dst[j + (int64_t)i * D.ldd] = (Td)tile[tx][c];
    }
  }
}

// Initialise a list of blocks: [I; 0] (desc.trans == 1) or zero (desc.trans == 0).
template <typename T>
__global__ __launch_bounds__(256) void k_init_blocks(const CopyDesc* __restrict__ d, const int* __restrict__ tstart,
                                                      int nd) {
  const int t = blockIdx.x;
  const int p = find_owner(tstart, nd, t);
  const CopyDesc D = d[p];
  const int local = t - tstart[p];
  const int ntr = (D.rows + 31) >> 5;
  const int r0 = (local % ntr) * 32, c0 = (local / ntr) * 32;
  T* __restrict__ dst = (T*)D.dst;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int c = ty; c < 32; c += 8) {
    const int i = r0 + tx, j = c0 + c;
    if (i < D.rows && j < D.cols) dst[i + (int64_t)j * D.ldd] = (D.trans && i == j) ? T(1) : T(0);
  }
}

void launch_copy(hipStream_t s, int elem_src, int elem_dst, const CopyDesc* d, const int* tstart, int nd,
                 int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  dim3 g(ntiles), b(256);
  if (elem_src == 8 && elem_dst == 8) hipLaunchKernelGGL((k_copy<double, double>), g, b, 0, s, d, tstart, nd);
  else if (elem_src == 4 && elem_dst == 4) hipLaunchKernelGGL((k_copy<float, float>), g, b, 0, s, d, tstart, nd);
  else if (elem_src == 8 && elem_dst == 4) hipLaunchKernelGGL((k_copy<double, float>), g, b, 0, s, d, tstart, nd);
  else hipLaunchKernelGGL((k_copy<float, double>), g, b, 0, s, d, tstart, nd);
}

void launch_init_blocks(hipStream_t s, int elem, const CopyDesc* d, const int* tstart, int nd, int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  dim3 g(ntiles), b(256);
  if (elem == 8) hipLaunchKernelGGL((k_init_blocks<double>), g, b, 0, s, d, tstart, nd);
  else hipLaunchKernelGGL((k_init_blocks<float>), g, b, 0, s, d, tstart, nd);
}

// Row norms of row ranges of blocks: one thread per row, a tile of 256 rows
// per workgroup (tiles: exclusive prefix sums over descriptors).
template <typename T>
__global__ __launch_bounds__(256) void k_rownorm(const RowNormDesc<T>* __restrict__ d, const int* __restrict__ tstart,
                                                 int nd) {
  const int p = find_owner(tstart, nd, blockIdx.x);
  const RowNormDesc<T> q = d[p];
  const int i = (blockIdx.x - tstart[p]) * 256 + threadIdx.x;
  if (i >= q.nrows) return;
  double s = 0;
  for (int j = 0; j < q.ncols; ++j) {
    const double v = (double)q.A[(q.row0 + i) + (int64_t)j * q.ld];
    s += v * v;
  }
  q.out[i] = sqrt(s);
}
// Row gather: one tile = 32 rows x 32 columns, one thread per (row, column-lane).
template <typename T>
__global__ __launch_bounds__(256) void k_rowgather(const RowGatherDesc<T>* __restrict__ d,
                                                   const int* __restrict__ tstart, int nd) {
  const int p = find_owner(tstart, nd, blockIdx.x);
  const RowGatherDesc<T> q = d[p];
  const int ntc = (q.cols + 31) / 32;
  const int t = blockIdx.x - tstart[p];
  const int r0 = (t / ntc) * 32, c0 = (t % ntc) * 32;
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int r = r0 + (e & 31), c = c0 + (e >> 5);
    if (r < q.nrows && c < q.cols)
      q.dst[(q.row0 + r) + (int64_t)c * q.ldd] = q.src[q.idx[r] + (int64_t)c * q.lds];
  }
}
int rowgather_tiles(int nrows, int cols) { return ((nrows + 31) / 32) * ((cols + 31) / 32); }
template <typename T>
void launch_rownorm(hipStream_t s, const RowNormDesc<T>* d, const int* tstart, int nd, int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL((k_rownorm<T>), dim3(ntiles), dim3(256), 0, s, d, tstart, nd);
}
template <typename T>
void launch_rowgather(hipStream_t s, const RowGatherDesc<T>* d, const int* tstart, int nd, int ntiles) {
  if (nd <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL((k_rowgather<T>), dim3(ntiles), dim3(256), 0, s, d, tstart, nd);
}
template void launch_rownorm<double>(hipStream_t, const RowNormDesc<double>*, const int*, int, int);
template void launch_rownorm<float>(hipStream_t, const RowNormDesc<float>*, const int*, int, int);
template void launch_rowgather<double>(hipStream_t, const RowGatherDesc<double>*, const int*, int, int);
template void launch_rowgather<float>(hipStream_t, const RowGatherDesc<float>*, const int*, int, int);

template <typename T>
__global__ void k_zero(T* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = i; q < n; q += stride) p[q] = T(0);
}
void launch_zero(hipStream_t s, int elem, void* p, int64_t n) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  if (elem == 8) hipLaunchKernelGGL((k_zero<double>), dim3(blocks), dim3(256), 0, s, (double*)p, n);
  else hipLaunchKernelGGL((k_zero<float>), dim3(blocks), dim3(256), 0, s, (float*)p, n);
}

}  // namespace spq
