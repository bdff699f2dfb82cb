/*
 * spaqr.h — C ABI of the MI355X spaQR library (libspaqr.so).
 *
 * The reference (Abeynaya/spaQR_public) exposes its hot path as the C++ class
 * `Tree` (include/tree.h:36-152) plus the duck-typed preconditioner contract
 * used by its Krylov templates (src/is.cpp:95-346).  Every function below
 * replaces one member or free function of that surface; the citation is given
 * per declaration.  Conventions:
 *   - sparse matrices are 0-based compressed sparse column (CSC) arrays with
 *     sorted row indices, the layout of Eigen::SparseMatrix<double,ColMajor,int>
 *     after makeCompressed() (the reference's SpMat, tree.h:31);
 *   - host pointers are borrowed for the duration of the call only;
 *   - device pointers (`*_device`) are HIP device addresses on the tree's GPU;
 *   - every function returns 0 on success or a negative SPAQR_E_* code, and
 *     spaqr_last_error() describes the last failure of the calling thread;
 *   - one tree per host thread; calls on one tree are not re-entrant (the
 *     reference's solve() is const but mutates per-cluster buffers too).
 */
#ifndef SPAQR_H
#define SPAQR_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPAQR_OK 0
#define SPAQR_E_HIP (-1)
#define SPAQR_E_ARG (-2)
#define SPAQR_E_NUMERIC (-3)
#define SPAQR_E_OOM (-4)
#define SPAQR_E_STATE (-5)

#define SPAQR_FP64 8 /* factor storage and arithmetic in double */
#define SPAQR_FP32 4 /* factor in float (mixed form: FP64 Krylov on top)   */

typedef struct spaqr_tree spaqr_tree;
typedef struct spaqr_matrix spaqr_matrix;

/* Last error message of the calling thread (never NULL). */
const char* spaqr_last_error(void);

/* Tree(int lvls, int skip_)                               tree.h:113-120
 * precision: SPAQR_FP64 or SPAQR_FP32; device: HIP device ordinal.        */
spaqr_tree* spaqr_tree_create(int nlevels, int skip, int precision, int device);
/* ~Tree()                                                  tree.h:150     */
void spaqr_tree_destroy(spaqr_tree* t);

/* Tree::set_tol / set_skip / set_scale / set_square / set_order / set_hsl
 *                                                          tree.cpp:14-19 */
int spaqr_set_tol(spaqr_tree* t, double tol);
int spaqr_set_skip(spaqr_tree* t, int skip);
int spaqr_set_scale(spaqr_tree* t, int scale);
int spaqr_set_square(spaqr_tree* t, int square);
int spaqr_set_order(spaqr_tree* t, float order);
int spaqr_set_hsl(spaqr_tree* t, int hsl); /* only 0 (no matching) is available */
/* Tree::set_Xcoo(Eigen::MatrixXd*)                         tree.cpp:20
 * X is dim x n column-major (coordinate d of dof i at X[i*dim + d]).      */
int spaqr_set_coordinates(spaqr_tree* t, const double* X, int dim, int n);
/* linspace_nd(n, dim) coordinates (util.cpp:852-881), as spaQR.cpp:159.  */
int spaqr_set_tensor_coordinates(spaqr_tree* t, int n, int dim);
/* Row partition for the non-geometric path (trees without coordinates,
 * tree.cpp:608-617): parts[i] in [0, 2^(nlevels-1)) for each of the m rows,
 * parts 2g and 2g+1 being the halves of part g one level up -- the output of
 * partition_patoh (partition.cpp:422-448), brought by the caller.  Without
 * it, partition uses spaqr_row_parts.                                       */
int spaqr_set_row_parts(spaqr_tree* t, const int* parts, int m);
/* The library's own row partitioner in place of partition_patoh
 * (partition.cpp:422-448, PaToH): recursive bisection of the column-net
 * hypergraph of the m x n pattern with Fiduccia-Mattheyses refinement of
 * the cut nets.  Host only (no device needed).  Writes m parts.            */
int spaqr_row_parts(int m, int n, const int* colptr, const int* rowind, int nlevels, int* parts);
/* Per-level progress lines on stdout (tree.cpp:1417-1434).               */
int spaqr_set_verbose(spaqr_tree* t, int verbose);

/* Tree::rows() / cols() / levels()                         tree.cpp:23-25 */
int spaqr_rows(const spaqr_tree* t);
int spaqr_cols(const spaqr_tree* t);
int spaqr_levels(const spaqr_tree* t);

/* Tree::partition(SpMat& A)                                tree.cpp:595-723 */
int spaqr_partition(spaqr_tree* t, int m, int n, const int* colptr, const int* rowind, const double* val);
/* Tree::assemble(SpMat& A)                                 tree.cpp:725-778 */
int spaqr_assemble(spaqr_tree* t, int m, int n, const int* colptr, const int* rowind, const double* val);
/* Tree::factorize()                                        tree.cpp:1309-1437 */
int spaqr_factorize(spaqr_tree* t);
/* Tree::solve(VectorXd b, VectorXd& x), host vectors       tree.cpp:1440-1466 */
int spaqr_solve(spaqr_tree* t, const double* b, double* x);
/* Tree::solve on device vectors (b may alias x).                          */
int spaqr_solve_device(spaqr_tree* t, const double* b, double* x);
/* Tree::solve_nrml(VectorXd b, VectorXd& x)                tree.cpp:1471-1517
 * x = (W^T W)^{-1} b with the R-only factors of a least-squares tree
 * (set_square(0)); b and x have length cols.                              */
int spaqr_solve_nrml(spaqr_tree* t, const double* b, double* x);
/* Same on device vectors (b may alias x).                                 */
int spaqr_solve_nrml_device(spaqr_tree* t, const double* b, double* x);

/* Sparse operator for the Krylov drivers (the SpMat `mat` argument of
 * gmres/cgls, is.cpp:95/275), uploaded once to the tree's device.         */
spaqr_matrix* spaqr_matrix_create(spaqr_tree* t, int m, int n, const int* colptr, const int* rowind, const double* val);
void spaqr_matrix_destroy(spaqr_matrix* A);

/* gmres(mat, rhs, x, precond, iters, restart, tol, verb)   is.cpp:95-272
 * Left-preconditioned restarted GMRES with the tree as preconditioner,
 * FP64.  Host vectors; x holds the initial guess on entry.
 * iters_out: iterations; est_out: the stopping quantity |g_k|/||M^{-1}r_0||.
 * For a tree factored in FP32 (the mixed form) the call runs FP64 iterative
 * refinement around this GMRES: true residual r = b - Ax by the FP64 SpMV,
 * GMRES on A d = r to a relative estimate of max(1e-6, tol / (2 ||r||/||b||))
 * with the FP32 factors as
 * preconditioner, x += d, until ||b - Ax||/||b|| <= tol (or a step stops
 * halving it, or `iters` inner iterations are spent); then iters_out is
 * the total of inner iterations and est_out the final ||b - Ax||/||b||.   */
int spaqr_gmres(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, int restart,
                double tol, int verbose, int* iters_out, double* est_out);
/* Same on device vectors.                                                 */
int spaqr_gmres_device(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, int restart,
                       double tol, int verbose, int* iters_out, double* est_out);

/* spaqr_gmres_device without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * iterations and results as spaqr_gmres_device; the Krylov basis and partial
 * sums live in a fixed 64 MiB device array that every GPU the library runs
 * on holds until the process ends, so n * (restart + 1) plus O(n) work must
 * fit in 2^23 doubles and restart <= 1024 (else SPAQR_E_ARG before any GPU
 * work).  Calls that overlap on one GPU race on that array.               */
int spaqr_gmres_device_noalloc(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters,
                               int restart, double tol, int verbose, int* iters_out, double* est_out);

/* Adjoint of x = A^{-1} b (differentiable solve).  No counterpart in the
 * reference, whose tOperations (toperations.h:14-20) transpose only the
 * R-only least-squares factors: these transpose the full square factors.
 * x = M^{-T} b: the op tape's backward steps transposed in reverse, then its
 * forward steps transposed in reverse (Householder products as Q instead of
 * Q^T, triangular sweeps with R^T, R12 products as x_n -= R12^T x_c,
 * permutations inverted).  Square trees; host vectors of length rows.     */
int spaqr_solve_transpose(spaqr_tree* t, const double* b, double* x);
/* Same on device vectors (b may alias x).                                 */
int spaqr_solve_transpose_device(spaqr_tree* t, const double* b, double* x);
/* GMRES on A^T x = b left-preconditioned with M^{-T}, otherwise exactly
 * spaqr_gmres_device (same stopping rule; iterative refinement for FP32
 * trees).  A must be the tree's square matrix.                             */
int spaqr_gmres_transpose_device(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters,
                                 int restart, double tol, int verbose, int* iters_out, double* est_out);
/* out[k] = alpha * u[i_k] * v[j_k] for the nonzeros (i_k, j_k) of A in CSC
 * order; device vectors.  With u = lambda (A^T lambda = dL/dx), v = x and
 * alpha = -1 it is dL/dA on A's sparsity pattern.                          */
int spaqr_csc_outer_device(const spaqr_matrix* A, double alpha, const double* u, const double* v, double* out);

/* cgls(mat, rhs, x, precond, iters, tol, verb)            is.cpp:275-346
 * CGLS on A^T A x = A^T b preconditioned by the tree's normal-equation
 * apply; stops when ||A^T r||^2 < tol^2 ||A^T b||^2.  A is m x n (uploaded
 * with its transpose), b has length m, x length n (initial guess on entry).
 * iters_out: iterations; est_out: ||A^T r|| / ||A^T b||.                 */
int spaqr_cgls(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, double tol, int verbose,
               int* iters_out, double* est_out);
/* Same on device vectors.                                                 */
int spaqr_cgls_device(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, double tol,
                      int verbose, int* iters_out, double* est_out);

/* spaQR.cpp:131-147: D_jj = 1/||A(:,j)||_2, A <- (A D) * 10 in place; D may be NULL. */
int spaqr_colscale(int m, int n, const int* colptr, const int* rowind, double* val, double* D);
/* random(size, seed): mt19937 + uniform_real_distribution(-1,1)  util.cpp:943-952 */
int spaqr_random(int n, int seed, double* out);

/* ---- inspection (counters the reference prints, plus the tree itself) ---- */
/* cperm (n) and rperm (m) of Tree (tree.h:53-54).                          */
int spaqr_perms(const spaqr_tree* t, int* cperm, int* rperm);
/* Per level (8 values): t_fill, t_elmn, t_shift, t_scale, t_sprsfy, t_merge,
 * top_rows, top_cols — the line of tree.cpp:1417-1426.                     */
int spaqr_level_stats(const spaqr_tree* t, double* out);
/* nnz4: nnzA, nnzR, nnzH, nnzQ (tree.h:98-101); dbl8: factor flops (LAWN 41
 * ledger), t_partition, t_assemble, t_factorize, top separator size,
 * device bytes reserved, factor storage bytes, bytes streamed per apply.   */
int spaqr_counters(const spaqr_tree* t, unsigned long long* nnz4, double* dbl8);
/* Sparsified interfaces in call order: ilvl, sep lvl, sep index, cols before, cols after. */
int spaqr_num_sparsified(const spaqr_tree* t);
int spaqr_sparsified(const spaqr_tree* t, int* out5);
/* Least squares: every sparsify_extra compression (tree.cpp:1260-1307) in
 * call order: ilvl, sep lvl, sep index, extra rows before, extra rows kept. */
int spaqr_num_extra_sparsified(const spaqr_tree* t);
int spaqr_extra_sparsified(const spaqr_tree* t, int* out5);
/* The factorization's profile (the reference's Profile, profile.h:13-106,
 * which it collects but never prints) as JSON text: per level the phase
 * times and top separator, per sparsified interface its columns before and
 * after and its RRQR columns (neighbours), per merged cluster its size and
 * aspect ratio, the least-squares extra-row compressions and, when kernel
 * timing ran, the kernel families.  Writes at most buflen - 1 bytes and a
 * NUL into buf (buf may be NULL) and returns the full length, or a negative
 * SPAQR_E_* code.                                                          */
int spaqr_profile_json(spaqr_tree* t, char* buf, int buflen);
/* Cluster table: per cluster of bottoms[l] (12 ints): l, self(lvl,sep),
 * left(lvl,sep), right(lvl,sep), cstart, csize, rstart, rsize, order.      */
int spaqr_num_clusters(const spaqr_tree* t);
int spaqr_cluster_table(const spaqr_tree* t, int* out12);

/* ---- measurement (no counterpart in the reference, which only prints
 * wall-clock phase times) ---------------------------------------------- */
/* phase_sync = 1 synchronises after every phase so the per-level times are
 * exact (the default); 0 lets phases overlap on the stream.                */
int spaqr_set_phase_sync(spaqr_tree* t, int phase_sync);
/* HIP-event timing of every batched launch, per kernel family.            */
int spaqr_kernel_timing(spaqr_tree* t, int on);
int spaqr_kernel_stat_count(spaqr_tree* t);
/* family i: name (NUL-terminated, at most namelen bytes) and
 * out4 = launches, total device ms, algorithmic flops, algorithmic bytes.  */
int spaqr_kernel_stat(spaqr_tree* t, int i, char* name, int namelen, double* out4);
int spaqr_kernel_stats_reset(spaqr_tree* t);

#ifdef __cplusplus
}
#endif
#endif /* SPAQR_H */
