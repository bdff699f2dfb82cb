// ORACLE — test infrastructure only. This is the CPU checker for the spaQR hot
// path; the product (spaqr_public_amd) never links, loads or calls it.
//
// Dense LAPACK/BLAS semantics used by the reference's wrapper layer
// (util.cpp:233-477, 785-816), restated from the published LAPACK algorithms
// (dgeqr2/dlarfg, dlarft 'F','C', dlarfb 'L', dgeqp3 via dlaqp2 norm
// downdating, dormqr, dtrsv, dtrsm).  Two backends:
//   * "own"      : the restatement in oracle_lapack.cpp (always available);
//   * "openblas" : the same entry points resolved at run time (dlopen) from an
//                  OpenBLAS build that exports LAPACK symbols; the reference
//                  links OpenBLAS (Makefile-confs/Makefile-mac.conf:11-12).
// Column-major storage, leading dimension = rows (as the reference's Eigen
// MatrixXd wrappers pass `ld = rows`).
#pragma once
#include <string>
#include <vector>

namespace orc {

struct Mat {
  int r = 0, c = 0;
  std::vector<double> a;
  Mat() {}
  Mat(int r_, int c_) : r(r_), c(c_), a((size_t)r_ * c_, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)j * r + i]; }
  double operator()(int i, int j) const { return a[(size_t)j * r + i]; }
  double* data() { return a.data(); }
  const double* data() const { return a.data(); }
  int rows() const { return r; }
  int cols() const { return c; }
  void setZero() { std::fill(a.begin(), a.end(), 0.0); }
  // Eigen conservativeResize semantics: keeps the overlapping top-left part.
  void conservativeResize(int nr, int nc);
  Mat topRows(int k) const;
  Mat middleRows(int s, int k) const;
  Mat bottomRows(int k) const;
  Mat leftCols(int k) const;
  Mat middleCols(int s, int k) const;
  Mat block(int i, int j, int nr, int nc) const;
  void setBlock(int i, int j, const Mat& B);
  Mat transpose() const;
};

// Backend selection. Returns the name of the active backend.
// which: "own" or "openblas" (path of the shared library in `lib`).
bool use_openblas(const std::string& lib, std::string* err);
void use_own();
const char* backend_name();
void set_blas_threads(int n);

// y = A x style helpers (plain loops, used everywhere in the oracle).
void gemm(const Mat& A, bool tA, const Mat& B, bool tB, double alpha, double beta, Mat& C);

// --- LAPACK-semantics routines (in place, col-major, ld = rows) ---
void geqrf(Mat& A, std::vector<double>& tau);                        // util.cpp:785-793
void larft(const Mat& V, const std::vector<double>& tau, Mat& T);    // util.cpp:403-411
// C <- H^T C (trans='T') or C <- H C (trans='N'), H = I - V T V^T, V unit lower.
void larfb(const Mat& V, const Mat& T, Mat& C, char trans);          // util.cpp:421-462
void larfb_vec(const Mat& V, const Mat& T, double* x, int m, char trans);
// AP = QR. jpvt 0-based on output (util.cpp:466-477).
void geqp3(Mat& A, std::vector<int>& jpvt, std::vector<double>& tau);
// x <- Q x / Q^T x with Q from (v, tau) (util.cpp:306-329)
void ormqr_vec(const Mat& v, const std::vector<double>& tau, double* x, int m, char trans);
// A <- Q^T A (util.cpp:332-342), A <- A Q (util.cpp:358-369)
void ormqr_left(const Mat& v, const std::vector<double>& tau, Mat& A, char trans);
void ormqr_right_notrans(const Mat& v, const std::vector<double>& tau, Mat& A);
// B <- B U^{-1}, U upper non-unit (util.cpp:233-241)
void trsm_right_upper(const Mat& U, Mat& B);
// x <- U^{-1} x or U^{-T} x, U = top-left n x n of `U` (util.cpp:253-260)
void trsv_upper(const Mat& U, double* x, int n, char trans);
// Smallest singular values helper for --scale 0 (util.cpp:767-782): returns
// all singular values in descending order (one-sided Jacobi).
std::vector<double> singular_values(const Mat& A);
// choose_rank (util.cpp:795-816)
int choose_rank(const std::vector<double>& s, double tol, bool rel = true);

}  // namespace orc
