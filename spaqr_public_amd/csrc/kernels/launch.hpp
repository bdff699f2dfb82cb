// Host-side launchers of the spaQR gfx950 kernels.  All take device pointers
// and enqueue on the given stream; descriptor arrays live in device memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "descs.hpp"

namespace spq {

// copy.hip
void launch_copy(hipStream_t s, int elem_src, int elem_dst, const CopyDesc* d, const int* tstart, int nd, int ntiles);
void launch_init_blocks(hipStream_t s, int elem, const CopyDesc* d, const int* tstart, int nd, int ntiles);
void launch_zero(hipStream_t s, int elem, void* p, int64_t n);
int rowgather_tiles(int nrows, int cols);
template <typename T>
void launch_rownorm(hipStream_t s, const RowNormDesc<T>* d, const int* tstart, int nd, int ntiles);
template <typename T>
void launch_rowgather(hipStream_t s, const RowGatherDesc<T>* d, const int* tstart, int nd, int ntiles);

// gemm.hip (tiles per descriptor: gemm_tiles(m, n) * ksplit, descs.hpp)
int gemm_tile_k();
template <typename T>
void launch_gemm_reduce(hipStream_t s, const GemmDesc<T>* d, const int* tstart, int nd, int ntiles);
template <typename T>
void launch_gemm(hipStream_t s, const GemmDesc<T>* d, const int* tstart, int nd, int ntiles);

// panel.hip
// Register form of the panel step: rows per workgroup, resident capacity,
// launch (multi-workgroup problems need coop = 1 or a grid within capacity).
int panel_rows_per_wg();
template <typename T>
int panel_r_capacity();
template <typename T>
int launch_panel_r(hipStream_t s, const PanelDesc<T>* d, const int* wstart, int nd, int nwg_total, int jb, int coop);
int panel_nb();
int trsm_rows_per_tile();
template <typename T>
void launch_larft(hipStream_t s, const LarftDesc<T>* d, const int* tstart, int nd, int ntiles);
template <typename T>
void launch_trsm_block(hipStream_t s, const TrsmDesc<T>* d, const int* tstart, int nd, int ntiles, int jb);

// qp3.hip
// Register form of the truncated pivoted QR: rows held per lane, resident
// workgroups for lane-group size G and nt threads per workgroup (256 or
// 1024), launch (gmax = largest G, nmax = largest N of the launch).
// register form: lanes per column for k rows at rc rows per lane (rc in
// {16, 20, 24, 28, 32}; 0 if k is beyond 64 lanes), capacity and launch
int qp3r_lanes(int k, int rc);
template <typename T>
int qp3r_capacity(int G, int nt, int rc);
template <typename T>
int launch_qp3r(hipStream_t s, const Qp3Desc<T>* d, const int* wstart, int nd, int nwg_total, int gmax, int nmax,
                int nt, int rc, int coop);
int qp3_threads();
template <typename T>
int qp3_capacity(int kmax);
// wstart: exclusive prefix sums of the descriptors' nwg; coop != 0 launches
// cooperatively (required when any descriptor has nwg > 1).  Returns a
// hipError_t value.
template <typename T>
int launch_qp3(hipStream_t s, const Qp3Desc<T>* d, const int* wstart, int nd, int nwg_total, int kmax, int coop);

// solve.hip
// register form of the reflector apply: nrw row waves per workgroup (1, 2
// or 7), so <= 64 nrw rows per workgroup; reflr_rows() = 448 (nrw = 7) and
// reflr_row_waves(m) the smallest form for ops of at most m rows
int reflr_rows();
int reflr_row_waves(int mmax);
size_t reflr_lds(int nseg_max);
template <typename T>
void launch_reflr(hipStream_t s, const ReflDesc<T>* d, const int* wstart, int nd, int nwg_total, const Seg* segs, T* xbuf,
                  int nseg_max, int nrw);
// latency form of the triangular sweep: one workgroup per op, no in-kernel
// R12 products, k <= bwdr_max_k()
int bwdr_max_k();
template <typename T>
void launch_bwdr(hipStream_t s, const BwdDesc<T>* d, int nd, int kmax, T* xbuf);
int refl_lds_cap(int elem);
template <typename T>
void launch_refl(hipStream_t s, const ReflDesc<T>* d, const int* wstart, int nd, int nwg_total, const Seg* segs,
                 T* xbuf, T* work, int mmax);
template <typename T>
void launch_bwd(hipStream_t s, const BwdDesc<T>* d, const int* wstart, int nd, int nwg_total, const R12Desc<T>* r12,
                T* xbuf);
int r12_chunk_cols();
template <typename T>
void launch_r12(hipStream_t s, const R12Chunk<T>* d, int nd, T* xbuf);
template <typename T>
void launch_move(hipStream_t s, const MoveDesc* d, int nd, T* dst, const T* src);
template <typename T>
void launch_gather_in(hipStream_t s, const int64_t* dst, const int* src, int n, const double* b, T* xbuf);
template <typename T>
void launch_scatter_out(hipStream_t s, const int64_t* src, const int* dst, int n, const T* xbuf, double* x);

// smin.hip
int smin_max_k();
size_t smin_work_elems(int k);
template <typename T>
void launch_smin(hipStream_t s, const SminDesc<T>* d, int nd, int kmax);
template <typename T>
void launch_scale_inv(hipStream_t s, const ScaleInvDesc<T>* d, const int* tstart, int nd, int ntiles);

// krylov.hip
void launch_spmv(hipStream_t s, int n, const int* rp, const int* ci, const double* v, const double* x, double* y,
                 double alpha, double beta);
// part: dots_part_elems(ncol) doubles of workspace
size_t dots_part_elems(int ncol);
void launch_dots(hipStream_t s, int n, int ncol, const double* V, int64_t ld, const double* w, double* out,
                 double* part);
void launch_gemv_sub(hipStream_t s, int n, int ncol, const double* V, int64_t ld, const double* h, double* w);
void launch_gemv(hipStream_t s, int n, int ncol, const double* V, int64_t ld, const double* h, double* y);
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y);
void launch_csc_outer(hipStream_t s, int n, int nnz, const int* colptr, const int* rowind, double alpha,
                      const double* u, const double* v, double* out);

}  // namespace spq
