// Descriptor records shared by the host planner and the batched kernels.
// Every batched launch takes a device array of descriptors (one per problem)
// plus, for tiled kernels, exclusive prefix sums of per-problem tile counts,
// so one launch covers problems of uneven size.
#pragma once
#include <stdint.h>

namespace spq {

// Inter-workgroup barriers: groups of more than kFanIn workgroups use a
// two-level counter tree (fan-in counters on separate 64-byte lines).
// barrier_words(nwg) = unsigned ints of counter storage a group needs.
constexpr int kFanIn = 16;
inline int barrier_words(int nwg) { return nwg <= kFanIn ? 16 : 16 * (1 + (nwg + kFanIn - 1) / kFanIn); }

// 2-D block copy: dst(i,j) = src(i,j) (trans=0) or dst(j,i) = src(i,j)
// (trans=1) for i < rows, j < cols of src.  Col-major, element offsets.
// Source and destination may have different element types (double/float).
struct CopyDesc {
  const void* src;
  void* dst;
  int lds, ldd, rows, cols;
  int trans;
  int pad_;
};

// Batched GEMM: C = alpha * op(A) * op(B) + beta * C, C is m x n.
// flags: bit0 transA, bit1 transB, bits 4-5 A structure, bits 6-7 B structure
// (structure applies to the stored matrix: 1 = unit lower trapezoidal,
// 2 = upper triangular incl. diagonal, 3 = lower triangular incl. diagonal).
enum : int { kTransA = 1, kTransB = 2 };
enum : int { kStructNone = 0, kStructUnitLower = 1, kStructUpper = 2, kStructLower = 3 };
constexpr int gemm_flags(bool ta, bool tb, int sa = 0, int sb = 0) {
  return (ta ? kTransA : 0) | (tb ? kTransB : 0) | (sa << 4) | (sb << 6);
}
// Tile shape of a descriptor (k_gemm): 64 x 64, 32 x 128 (m <= 32) or
// 128 x 32 (n <= 32); gemm_tiles = workgroups of one K chunk.  Inline: the
// planner counts tiles for every descriptor of every launch.
constexpr int gemm_shape(int m, int n) { return m <= 32 ? 1 : (n <= 32 ? 2 : 0); }
constexpr int gemm_shape_m(int shape) { return shape == 1 ? 32 : (shape == 2 ? 128 : 64); }
constexpr int gemm_shape_n(int shape) { return shape == 1 ? 128 : (shape == 2 ? 32 : 64); }
constexpr int gemm_tiles(int m, int n) {
  const int sh = gemm_shape(m, n), tm = gemm_shape_m(sh), tn = gemm_shape_n(sh);
  return ((m + tm - 1) / tm) * ((n + tn - 1) / tn);
}
// Split-K (ksplit > 1): the K range is cut into ksplit chunks whose partial
// products alpha * op(A) op(B) go to work (ksplit x m x n, ld m); a reduce
// pass sums them in chunk order and applies beta (deterministic).
template <typename T>
struct GemmDesc {
  const T* A;
  const T* B;
  T* C;
  int lda, ldb, ldc;
  int m, n, k;
  int flags;
  T alpha, beta;
  int ksplit = 1;
  T* work = nullptr;
};

// One block step of the Householder update of a panel's targets,
// C <- (I - V_b T_b V_b^T)^T C = C - (V_b T_b^T)(V_b^T C): the larfb of
// util.cpp:421-443 (side 'L', trans 'T', forward, columnwise) as batched
// GEMM descriptors.  W = V_b T_b^T (rows x w) is formed once per panel, with
// T_b entering as a dense factor (the panel kernel stores its strict lower
// part as zeros); then per target Y = V_b^T C (w x cols) goes to g1 and
// C -= W Y to g3.  g1 runs before g3.  alloc(elems) returns device scratch of
// at least elems elements of U.
template <typename U>
struct LarfbTarget {
  U* C;
  int ld, cols;
};
template <typename U, typename Alloc, typename Vec>
void larfb_step_descs(const U* Vb, int ldv, const U* Tb, int ldt, int rows, int w, const LarfbTarget<U>* tg,
                      int ntg, Alloc&& alloc, Vec& g1, Vec& g3) {
  if (ntg <= 0) return;
  const int ldw = rows > 1 ? rows : 1;
  U* W = alloc((int64_t)ldw * ldt);
  GemmDesc<U> d{};
  d.A = Vb; d.lda = ldv; d.B = Tb; d.ldb = ldt; d.C = W; d.ldc = ldw;
  d.m = rows; d.n = w; d.k = w;
  d.flags = gemm_flags(false, true, kStructUnitLower, kStructNone);
  d.alpha = U(1); d.beta = U(0);
  g1.push_back(d);
  for (int t = 0; t < ntg; ++t) {
    U* Y = alloc((int64_t)ldt * tg[t].cols);
    GemmDesc<U> y{};
    y.A = Vb; y.lda = ldv; y.B = tg[t].C; y.ldb = tg[t].ld; y.C = Y; y.ldc = ldt;
    y.m = w; y.n = tg[t].cols; y.k = rows;
    y.flags = gemm_flags(true, false, kStructUnitLower);
    y.alpha = U(1); y.beta = U(0);
    g1.push_back(y);
    GemmDesc<U> c{};
    c.A = W; c.lda = ldw; c.B = Y; c.ldb = ldt; c.C = tg[t].C; c.ldc = tg[t].ld;
    c.m = rows; c.n = tg[t].cols; c.k = w;
    c.flags = gemm_flags(false, false);
    c.alpha = U(-1); c.beta = U(1);
    g3.push_back(c);
  }
}

// Householder panel block: factor columns [jb, jb+w) of P (m x k, ld), writing
// reflectors below the diagonal, R on/above it, tau[j] and the nb x nb block
// T factor of this block into Tb + jb*nb (ld nb).  The register form splits
// rows [jb, m) over nwg workgroups (part: 2 * nwg * 64 partial sums and the
// published pivot row; sync: barrier_words(nwg) counters, zero on entry).
template <typename T>
struct PanelDesc {
  T* P;
  T* tau;
  T* Tb;
  int ld, m, k;
  int nwg;
  T* part;
  unsigned int* sync;
  int* err;
};

// Block T factors of an existing reflector set (V m x r unit lower, tau).
template <typename T>
struct LarftDesc {
  const T* V;
  const T* tau;
  T* Tb;
  int ldv, m, r;
  int pad_;
};

// Row-block triangular solve: B[:, jb:jb+w) <- B[:, jb:jb+w) * R[jb:jb+w, jb:jb+w]^{-1}
template <typename T>
struct TrsmDesc {
  T* B;
  const T* R;
  int ldb, ldr, rows, k;
};

// Truncated column-pivoted QR on C (k x N) stored transposed: Ct is N x k
// col-major (ld = ldc >= N), so row r of C is column r of Ct.  A problem is
// processed by `nwg` workgroups that own contiguous column slices and meet at
// one inter-workgroup barrier per pivot step (nwg == 1: no barrier).
//
// Two kernels share the descriptor.  The streaming form (G == 0) keeps C in
// memory and uses vn, pos (nwg * 2N: per-workgroup perm / ipos copies).  The
// register form (G > 0) holds each column in the registers of G lanes and uses
// pos (N: position -> column table written by workgroup 0) and part_col
// (2 * nwg * k: every workgroup's local-best column, double-buffered).
template <typename T>
struct Qp3Desc {
  T* Ct;
  T* V;               // k x k (ld ldv): reflector i in column i (unit lower)
  T* tau;             // k
  double* vn;         // 2*N partial norms (FP64), streaming form
  int* pos;
  double* part_v;     // 2 * nwg partial argmax values (double-buffered by step parity)
  int* part_p;        // 2 * nwg * 2: (position, column) of the partial argmax
  double* beta;       // k: R(i,i) of each step
  unsigned int* sync; // barrier counter (zero on entry)
  int* err;           // set when a barrier wait gives up
  int* rank_out;      // device int
  T* part_col;        // register form: published local-best columns
  int ldc, k, N, ldv;
  double tol;
  int nwg, G;
  int rc = 0;         // register form: rows per lane (the kernel's RC)
};

// Solve-side vector segment (cluster x buffer offset, length).
struct Seg {
  int64_t off;
  int len;
  int pad_;
};

// Apply a blocked reflector product to a gathered vector:
// x = concat(segs) ; x <- H^T x (trans=1) or H x (trans=0) ; scatter back.
template <typename T>
struct ReflDesc {
  const T* V;
  const T* Tb;
  int ldv, m, k, nb;
  int seg0, nseg;  // segment range in the segment table
  int trans;
  int nwg;         // workgroups (row slices); > 1 uses work/part/sync
  T* work;         // m: gathered vector (multi-workgroup form)
  T* part;         // 2 * nwg * 32 partial V_b^T x sums
  unsigned int* sync;
  int* err;
};

// x_c[0:k] -= sum_n A_n[0:k, 0:s_n] x_n[0:s_n] ; x_c[0:k] <- R^{-1} x_c[0:k]
template <typename T>
struct BwdDesc {
  const T* R;
  int ldr, k;
  int64_t xoff;
  int blk0, nblk;  // range in the R12 table
  int trans;       // 1: R^{-T} (transposed solve)
  int nwg;         // workgroups (row slices of x_c); > 1 uses sync
  unsigned int* sync;
  int* err;
};
template <typename T>
struct R12Desc {
  const T* A;
  int ld, s;
  int64_t xoff;
};

// Least-squares row surgery (tree.cpp:319-411).
// RowNormDesc: out[i] = ||A(row0 + i, 0:ncols)||_2 for i < nrows (FP64).
template <typename T>
struct RowNormDesc {
  const T* A;
  double* out;
  int ld, row0, nrows, ncols;
};
// RowGatherDesc: dst(row0 + i, 0:cols) = src(idx[i], 0:cols) for i < nrows.
template <typename T>
struct RowGatherDesc {
  const T* src;
  T* dst;
  const int* idx;
  int lds, ldd, row0, nrows, cols;
  int pad_;
};

// One column chunk of an R12 product in the solve (x-buffer offsets):
// trans = 0: x[xc + r] -= sum_{j<nc} A[r + j*ld] * x[xn + j], r < k
// trans = 1: x[xn + j] -= sum_{r<k} A[r + j*ld] * x[xc + r],  j < nc
template <typename T>
struct R12Chunk {
  const T* A;
  int64_t xc, xn;
  int ld, k, nc, trans;
};

// Vector segment move dst[0:len] = src[0:len] (x-buffer offsets).
struct MoveDesc {
  int64_t src, dst;
  int len;
  int pad_;
};

// sigma_min of an upper-triangular k x k factor (k <= kSminMaxK); work holds
// k * (kSminSteps + 1) doubles of Lanczos basis.
constexpr int kSminSteps = 64;
constexpr int kSminMaxK = 4096;
template <typename T>
struct SminDesc {
  const T* R;
  int ld, k;
  double* work;
  double* out;
};
// C[r0:r1, 0:cols] *= 1 / *s
template <typename T>
struct ScaleInvDesc {
  T* C;
  const double* s;
  int ld, r0, r1, cols;
};

}  // namespace spq
