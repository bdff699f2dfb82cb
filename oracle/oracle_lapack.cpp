// ORACLE — test infrastructure only (see oracle_lapack.h).
//
// Own restatement of the LAPACK algorithms the reference calls through its
// wrappers (util.cpp:233-477, 785-816), following the published reference
// LAPACK routines: dlarfg, dgeqr2, dlarft (forward, columnwise), dlarfb (left),
// dgeqp3 with the dlaqp2 partial-norm downdating rule (tol3z = sqrt(eps)),
// dormqr (reflector by reflector), dtrsv, dtrsm.  An optional backend resolves
// the same LAPACK/BLAS entry points from an OpenBLAS shared library at run time.
#include "oracle_lapack.h"

#include <dlfcn.h>

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace orc {

// ------------------------------------------------------------------ Mat utils
void Mat::conservativeResize(int nr, int nc) {
  Mat B(nr, nc);
  int mr = std::min(nr, r), mc = std::min(nc, c);
  for (int j = 0; j < mc; ++j)
    for (int i = 0; i < mr; ++i) B(i, j) = (*this)(i, j);
  *this = std::move(B);
}
Mat Mat::block(int i0, int j0, int nr, int nc) const {
  Mat B(nr, nc);
  for (int j = 0; j < nc; ++j)
    for (int i = 0; i < nr; ++i) B(i, j) = (*this)(i0 + i, j0 + j);
  return B;
}
Mat Mat::topRows(int k) const { return block(0, 0, k, c); }
Mat Mat::middleRows(int s, int k) const { return block(s, 0, k, c); }
Mat Mat::bottomRows(int k) const { return block(r - k, 0, k, c); }
Mat Mat::leftCols(int k) const { return block(0, 0, r, k); }
Mat Mat::middleCols(int s, int k) const { return block(0, s, r, k); }
void Mat::setBlock(int i0, int j0, const Mat& B) {
  for (int j = 0; j < B.c; ++j)
    for (int i = 0; i < B.r; ++i) (*this)(i0 + i, j0 + j) = B(i, j);
}
Mat Mat::transpose() const {
  Mat B(c, r);
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i) B(j, i) = (*this)(i, j);
  return B;
}

void gemm(const Mat& A, bool tA, const Mat& B, bool tB, double alpha, double beta, Mat& C) {
  int m = tA ? A.c : A.r;
  int k = tA ? A.r : A.c;
  int n = tB ? B.r : B.c;
  assert((tB ? B.c : B.r) == k);
  if (C.r != m || C.c != n) { C = Mat(m, n); beta = 0.0; }
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int l = 0; l < k; ++l) {
        double a = tA ? A(l, i) : A(i, l);
        double b = tB ? B(j, l) : B(l, j);
        s += a * b;
      }
      C(i, j) = alpha * s + (beta == 0.0 ? 0.0 : beta * C(i, j));
    }
  }
}

// ------------------------------------------------------------ backend table
namespace {
typedef void (*fn_dgeqrf)(const int*, const int*, double*, const int*, double*, double*, const int*, int*);
typedef void (*fn_dgeqp3)(const int*, const int*, double*, const int*, int*, double*, double*, const int*, int*);
typedef void (*fn_dlarft)(const char*, const char*, const int*, const int*, const double*, const int*,
                          const double*, double*, const int*, size_t, size_t);
typedef void (*fn_dlarfb)(const char*, const char*, const char*, const char*, const int*, const int*,
                          const int*, const double*, const int*, const double*, const int*, double*,
                          const int*, double*, const int*, size_t, size_t, size_t, size_t);
typedef void (*fn_dormqr)(const char*, const char*, const int*, const int*, const int*, const double*,
                          const int*, const double*, double*, const int*, double*, const int*, int*,
                          size_t, size_t);
typedef void (*fn_dtrsm)(const char*, const char*, const char*, const char*, const int*, const int*,
                         const double*, const double*, const int*, double*, const int*, size_t, size_t,
                         size_t, size_t);
typedef void (*fn_dtrsv)(const char*, const char*, const char*, const int*, const double*, const int*,
                         double*, const int*, size_t, size_t, size_t);
typedef void (*fn_threads)(int);

struct Backend {
  bool ob = false;
  void* h = nullptr;
  fn_dgeqrf dgeqrf = nullptr;
  fn_dgeqp3 dgeqp3 = nullptr;
  fn_dlarft dlarft = nullptr;
  fn_dlarfb dlarfb = nullptr;
  fn_dormqr dormqr = nullptr;
  fn_dtrsm dtrsm = nullptr;
  fn_dtrsv dtrsv = nullptr;
  fn_threads threads = nullptr;
} g_be;

void* sym(void* h, const char* a, const char* b) {
  void* p = dlsym(h, a);
  if (!p && b) p = dlsym(h, b);
  return p;
}
}  // namespace

bool use_openblas(const std::string& lib, std::string* err) {
  void* h = dlopen(lib.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!h) { if (err) *err = dlerror(); return false; }
  Backend b;
  b.h = h;
  b.dgeqrf = (fn_dgeqrf)sym(h, "scipy_dgeqrf_", "dgeqrf_");
  b.dgeqp3 = (fn_dgeqp3)sym(h, "scipy_dgeqp3_", "dgeqp3_");
  b.dlarft = (fn_dlarft)sym(h, "scipy_dlarft_", "dlarft_");
  b.dlarfb = (fn_dlarfb)sym(h, "scipy_dlarfb_", "dlarfb_");
  b.dormqr = (fn_dormqr)sym(h, "scipy_dormqr_", "dormqr_");
  b.dtrsm = (fn_dtrsm)sym(h, "scipy_dtrsm_", "dtrsm_");
  b.dtrsv = (fn_dtrsv)sym(h, "scipy_dtrsv_", "dtrsv_");
  b.threads = (fn_threads)sym(h, "scipy_openblas_set_num_threads", "openblas_set_num_threads");
  if (!b.dgeqrf || !b.dgeqp3 || !b.dlarft || !b.dlarfb || !b.dormqr || !b.dtrsm || !b.dtrsv) {
    if (err) *err = "missing LAPACK symbols in " + lib;
    dlclose(h);
    return false;
  }
  b.ob = true;
  g_be = b;
  return true;
}
void use_own() { g_be.ob = false; }
const char* backend_name() { return g_be.ob ? "openblas" : "own"; }
void set_blas_threads(int n) { if (g_be.ob && g_be.threads) g_be.threads(n); }

// ------------------------------------------------------------- own routines
namespace {
const double kEps = std::ldexp(1.0, -53);  // dlamch('E')

double nrm2(int n, const double* x) {
  // Scaled sum of squares (dnrm2-style) so tiny/huge entries do not flush.
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; ++i) {
    if (x[i] != 0.0) {
      double ax = std::fabs(x[i]);
      if (scale < ax) { ssq = 1.0 + ssq * (scale / ax) * (scale / ax); scale = ax; }
      else ssq += (ax / scale) * (ax / scale);
    }
  }
  return scale * std::sqrt(ssq);
}
double lapy2(double x, double y) {
  double xa = std::fabs(x), ya = std::fabs(y);
  double w = std::max(xa, ya), z = std::min(xa, ya);
  if (z == 0.0) return w;
  return w * std::sqrt(1.0 + (z / w) * (z / w));
}
// dlarfg: on exit alpha = beta, x = v(2:n), returns tau.
double larfg(int n, double& alpha, double* x) {
  if (n <= 1) return 0.0;
  double xnorm = nrm2(n - 1, x);
  if (xnorm == 0.0) return 0.0;
  double beta = -std::copysign(lapy2(alpha, xnorm), alpha);
  double tau = (beta - alpha) / beta;
  double sc = 1.0 / (alpha - beta);
  for (int i = 0; i < n - 1; ++i) x[i] *= sc;
  alpha = beta;
  return tau;
}
// Apply H = I - tau v v^T (v(0) = 1 implicit via v0) from the left to A(r0:, c0:c1).
void apply_h_left(Mat& A, int r0, int c0, int c1, const double* v, double tau) {
  if (tau == 0.0) return;
  int m = A.r - r0;
  for (int j = c0; j < c1; ++j) {
    double* col = &A(r0, j);
    double w = 0.0;
    for (int i = 0; i < m; ++i) w += v[i] * col[i];
    w *= tau;
    for (int i = 0; i < m; ++i) col[i] -= w * v[i];
  }
}
}  // namespace

void geqrf(Mat& A, std::vector<double>& tau) {
  int m = A.r, n = A.c, k = std::min(m, n);
  tau.assign(k, 0.0);
  if (m == 0 || n == 0) return;
  if (g_be.ob) {
    int lda = m, info = 0, lwork = -1;
    double wq = 0;
    g_be.dgeqrf(&m, &n, A.data(), &lda, tau.data(), &wq, &lwork, &info);
    lwork = std::max(1, (int)wq);
    std::vector<double> w(lwork);
    g_be.dgeqrf(&m, &n, A.data(), &lda, tau.data(), w.data(), &lwork, &info);
    if (info != 0) throw std::runtime_error("dgeqrf failed");
    return;
  }
  std::vector<double> v(m);
  for (int i = 0; i < k; ++i) {
    double alpha = A(i, i);
    tau[i] = larfg(m - i, alpha, (i + 1 < m) ? &A(i + 1, i) : nullptr);
    A(i, i) = alpha;
    if (i + 1 < n) {
      v[0] = 1.0;
      for (int r = i + 1; r < m; ++r) v[r - i] = A(r, i);
      apply_h_left(A, i, i + 1, n, v.data(), tau[i]);
    }
  }
}

void larft(const Mat& V, const std::vector<double>& tau, Mat& T) {
  int m = V.r, k = V.c;
  T = Mat(k, k);
  if (k == 0) return;
  if (g_be.ob) {
    const char d = 'F', s = 'C';
    int ldv = std::max(1, m), ldt = k;
    g_be.dlarft(&d, &s, &m, &k, V.data(), &ldv, tau.data(), T.data(), &ldt, 1, 1);
    return;
  }
  // T(i,i) = tau_i ; T(0:i,i) = -tau_i * T(0:i,0:i) * V(:,0:i)^T v_i
  std::vector<double> w(k);
  for (int i = 0; i < k; ++i) {
    if (tau[i] == 0.0) {
      for (int j = 0; j <= i; ++j) T(j, i) = 0.0;
      continue;
    }
    for (int j = 0; j < i; ++j) {
      // v_i has v_i(i)=1 and zeros above i; V(:,j) unit at j, zeros above j.
      double s = V(i, j);  // row i: v_i(i)=1 times V(i,j)
      for (int r = i + 1; r < m; ++r) s += V(r, j) * V(r, i);
      w[j] = -tau[i] * s;
    }
    for (int j = 0; j < i; ++j) {
      double s = 0.0;
      for (int l = j; l < i; ++l) s += T(j, l) * w[l];
      T(j, i) = s;
    }
    T(i, i) = tau[i];
  }
}

namespace {
// W = V^T C (k x n) with V unit lower trapezoidal (m x k).
void vt_c(const Mat& V, const double* C, int ldc, int m, int n, std::vector<double>& W) {
  int k = V.c;
  W.assign((size_t)k * n, 0.0);
  for (int j = 0; j < n; ++j) {
    const double* c = C + (size_t)j * ldc;
    for (int l = 0; l < k; ++l) {
      double s = (l < m) ? c[l] : 0.0;
      for (int r = l + 1; r < m; ++r) s += V(r, l) * c[r];
      W[(size_t)j * k + l] = s;
    }
  }
}
}  // namespace

static void larfb_raw(const Mat& V, const Mat& T, double* C, int ldc, int m, int n, char trans) {
  int k = V.c;
  if (m == 0 || n == 0 || k == 0) return;
  if (g_be.ob) {
    const char side = 'L', d = 'F', s = 'C';
    int ldv = std::max(1, V.r), ldt = k, ldw = n;
    std::vector<double> work((size_t)n * k);
    g_be.dlarfb(&side, &trans, &d, &s, &m, &n, &k, V.data(), &ldv, T.data(), &ldt, C, &ldc,
                work.data(), &ldw, 1, 1, 1, 1);
    return;
  }
  // H^T C = C - V T^T (V^T C) ; H C = C - V T (V^T C)
  std::vector<double> W;
  vt_c(V, C, ldc, m, n, W);
  std::vector<double> Z((size_t)k * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < k; ++i) {
      double s = 0.0;
      if (trans == 'T') { for (int l = 0; l <= i; ++l) s += T(l, i) * W[(size_t)j * k + l]; }
      else { for (int l = i; l < k; ++l) s += T(i, l) * W[(size_t)j * k + l]; }
      Z[(size_t)j * k + i] = s;
    }
  for (int j = 0; j < n; ++j) {
    double* c = C + (size_t)j * ldc;
    for (int l = 0; l < k; ++l) {
      double z = Z[(size_t)j * k + l];
      if (z == 0.0) continue;
      if (l < m) c[l] -= z;
      for (int r = l + 1; r < m; ++r) c[r] -= V(r, l) * z;
    }
  }
}

void larfb(const Mat& V, const Mat& T, Mat& C, char trans) {
  assert(C.r == V.r);
  larfb_raw(V, T, C.data(), std::max(1, C.r), C.r, C.c, trans);
}
void larfb_vec(const Mat& V, const Mat& T, double* x, int m, char trans) {
  assert(m == V.r);
  larfb_raw(V, T, x, std::max(1, m), m, 1, trans);
}

void geqp3(Mat& A, std::vector<int>& jpvt, std::vector<double>& tau) {
  int m = A.r, n = A.c, k = std::min(m, n);
  jpvt.assign(n, 0);
  tau.assign(k, 0.0);
  if (m == 0 || n == 0) return;
  if (g_be.ob) {
    int lda = m, info = 0, lwork = -1;
    double wq = 0;
    g_be.dgeqp3(&m, &n, A.data(), &lda, jpvt.data(), tau.data(), &wq, &lwork, &info);
    lwork = std::max(1, (int)wq);
    std::vector<double> w(lwork);
    g_be.dgeqp3(&m, &n, A.data(), &lda, jpvt.data(), tau.data(), w.data(), &lwork, &info);
    if (info != 0) throw std::runtime_error("dgeqp3 failed");
    for (auto& p : jpvt) p -= 1;
    return;
  }
  const double tol3z = std::sqrt(kEps);
  for (int j = 0; j < n; ++j) jpvt[j] = j;
  std::vector<double> vn1(n), vn2(n), v(m);
  for (int j = 0; j < n; ++j) { vn1[j] = nrm2(m, &A(0, j)); vn2[j] = vn1[j]; }
  for (int i = 0; i < k; ++i) {
    // pivot: first index of max vn1 over i..n-1 (idamax)
    int pvt = i;
    double best = -1.0;
    for (int j = i; j < n; ++j) if (std::fabs(vn1[j]) > best) { best = std::fabs(vn1[j]); pvt = j; }
    if (pvt != i) {
      for (int r = 0; r < m; ++r) std::swap(A(r, pvt), A(r, i));
      std::swap(jpvt[pvt], jpvt[i]);
      vn1[pvt] = vn1[i];
      vn2[pvt] = vn2[i];
    }
    double alpha = A(i, i);
    tau[i] = (i < m - 1) ? larfg(m - i, alpha, &A(i + 1, i)) : 0.0;
    A(i, i) = alpha;
    if (i < n - 1) {
      v[0] = 1.0;
      for (int r = i + 1; r < m; ++r) v[r - i] = A(r, i);
      apply_h_left(A, i, i + 1, n, v.data(), tau[i]);
    }
    for (int j = i + 1; j < n; ++j) {
      if (vn1[j] != 0.0) {
        double t = std::fabs(A(i, j)) / vn1[j];
        t = 1.0 - t * t;
        t = std::max(t, 0.0);
        double t2 = t * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
        if (t2 <= tol3z) {
          if (i < m - 1) { vn1[j] = nrm2(m - i - 1, &A(i + 1, j)); vn2[j] = vn1[j]; }
          else { vn1[j] = 0.0; vn2[j] = 0.0; }
        } else {
          vn1[j] *= std::sqrt(t);
        }
      }
    }
  }
}

static void ormqr_raw(const Mat& v, const std::vector<double>& tau, double* A, int lda, int m, int n,
                      char side, char trans) {
  int k = v.c;
  if (m == 0 || n == 0 || k == 0) return;
  if (g_be.ob) {
    int ldv = std::max(1, v.r), info = 0, lwork = -1;
    double wq = 0;
    g_be.dormqr(&side, &trans, &m, &n, &k, v.data(), &ldv, tau.data(), A, &lda, &wq, &lwork, &info, 1, 1);
    lwork = std::max(1, (int)wq);
    std::vector<double> w(lwork);
    g_be.dormqr(&side, &trans, &m, &n, &k, v.data(), &ldv, tau.data(), A, &lda, w.data(), &lwork, &info, 1, 1);
    if (info != 0) throw std::runtime_error("dormqr failed");
    return;
  }
  int nq = (side == 'L') ? m : n;
  std::vector<double> h(nq);
  // Q = H_0 H_1 ... H_{k-1}
  bool forward = (side == 'L') == (trans == 'T');
  for (int t = 0; t < k; ++t) {
    int i = forward ? t : k - 1 - t;
    if (tau[i] == 0.0) continue;
    for (int r = 0; r < nq; ++r) h[r] = (r < i) ? 0.0 : (r == i ? 1.0 : v(r, i));
    if (side == 'L') {
      for (int j = 0; j < n; ++j) {
        double* col = A + (size_t)j * lda;
        double w = 0.0;
        for (int r = i; r < m; ++r) w += h[r] * col[r];
        w *= tau[i];
        for (int r = i; r < m; ++r) col[r] -= w * h[r];
      }
    } else {
      for (int r = 0; r < m; ++r) {
        double w = 0.0;
        for (int c = i; c < n; ++c) w += A[(size_t)c * lda + r] * h[c];
        w *= tau[i];
        for (int c = i; c < n; ++c) A[(size_t)c * lda + r] -= w * h[c];
      }
    }
  }
}

void ormqr_vec(const Mat& v, const std::vector<double>& tau, double* x, int m, char trans) {
  assert(v.r == m);
  ormqr_raw(v, tau, x, std::max(1, m), m, 1, 'L', trans);
}
void ormqr_left(const Mat& v, const std::vector<double>& tau, Mat& A, char trans) {
  assert(v.r == A.r);
  ormqr_raw(v, tau, A.data(), std::max(1, A.r), A.r, A.c, 'L', trans);
}
void ormqr_right_notrans(const Mat& v, const std::vector<double>& tau, Mat& A) {
  assert(v.r == A.c);
  ormqr_raw(v, tau, A.data(), std::max(1, A.r), A.r, A.c, 'R', 'N');
}

void trsm_right_upper(const Mat& U, Mat& B) {
  int m = B.r, n = B.c;
  assert(U.r == n && U.c == n);
  if (m == 0 || n == 0) return;
  if (g_be.ob) {
    const char s = 'R', u = 'U', t = 'N', d = 'N';
    const double one = 1.0;
    int ldu = n, ldb = m;
    g_be.dtrsm(&s, &u, &t, &d, &m, &n, &one, U.data(), &ldu, B.data(), &ldb, 1, 1, 1, 1);
    return;
  }
  for (int j = 0; j < n; ++j) {
    for (int l = 0; l < j; ++l) {
      double u = U(l, j);
      if (u == 0.0) continue;
      for (int i = 0; i < m; ++i) B(i, j) -= B(i, l) * u;
    }
    double d = 1.0 / U(j, j);
    for (int i = 0; i < m; ++i) B(i, j) *= d;
  }
}

void trsv_upper(const Mat& U, double* x, int n, char trans) {
  if (n == 0) return;
  if (g_be.ob) {
    const char u = 'U', d = 'N';
    int ldu = std::max(1, U.r), inc = 1;
    g_be.dtrsv(&u, &trans, &d, &n, U.data(), &ldu, x, &inc, 1, 1, 1);
    return;
  }
  if (trans == 'N') {
    for (int i = n - 1; i >= 0; --i) {
      x[i] /= U(i, i);
      double xi = x[i];
      for (int r = 0; r < i; ++r) x[r] -= U(r, i) * xi;
    }
  } else {
    for (int i = 0; i < n; ++i) {
      double s = x[i];
      for (int r = 0; r < i; ++r) s -= U(r, i) * x[r];
      x[i] = s / U(i, i);
    }
  }
}

std::vector<double> singular_values(const Mat& A0) {
  // One-sided Jacobi on the taller orientation; converges to full accuracy.
  Mat A = A0.r >= A0.c ? A0 : A0.transpose();
  int m = A.r, n = A.c;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double a = 0, b = 0, g = 0;
        for (int i = 0; i < m; ++i) { a += A(i, p) * A(i, p); b += A(i, q) * A(i, q); g += A(i, p) * A(i, q); }
        if (std::fabs(g) <= 1e-300 || std::fabs(g) <= kEps * std::sqrt(a * b)) continue;
        off = std::max(off, std::fabs(g) / std::sqrt(a * b));
        double zeta = (b - a) / (2.0 * g);
        double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < m; ++i) {
          double x = A(i, p), y = A(i, q);
          A(i, p) = c * x - s * y;
          A(i, q) = s * x + c * y;
        }
      }
    if (off < 1e-15) break;
  }
  std::vector<double> s(n);
  for (int j = 0; j < n; ++j) s[j] = nrm2(m, &A(0, j));
  std::sort(s.begin(), s.end(), [](double x, double y) { return x > y; });
  return s;
}

int choose_rank(const std::vector<double>& s, double tol, bool rel) {
  int n = (int)s.size();
  if (tol == 0) return n;
  if (tol >= 1.0) return 0;
  if (n <= 1) return n;
  double sref = rel ? std::fabs(s[0]) : 1.0;
  int rank = 1;
  while (rank < n && std::fabs(s[rank]) / sref >= tol) rank++;
  return rank;
}

}  // namespace orc
