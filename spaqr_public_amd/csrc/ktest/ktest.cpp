// Kernel-level test entry points (libspaqr_ktest.so, loaded only by tests/).
// Each function uploads host arrays, runs one batched kernel family on the
// device and downloads the result, so single kernels can be checked against
// the oracle's LAPACK restatement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../kernels/descs.hpp"
#include "../kernels/launch.hpp"

using namespace spq;

namespace {
thread_local std::string g_err;
#define KT(call)                                                     \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) { g_err = hipGetErrorString(e_); return -1; } \
  } while (0)

template <typename X>
X* dup(const X* h, size_t n) {
  X* d = nullptr;
  if (hipMalloc((void**)&d, sizeof(X) * (n ? n : 1)) != hipSuccess) return nullptr;
  if (h && n) (void)hipMemcpy(d, h, sizeof(X) * n, hipMemcpyHostToDevice);
  else (void)hipMemset(d, 0, sizeof(X) * (n ? n : 1));
  return d;
}
template <typename X>
void back(X* h, const X* d, size_t n) {
  (void)hipMemcpy(h, d, sizeof(X) * n, hipMemcpyDeviceToHost);
}
}  // namespace

extern "C" {

const char* ktest_last_error() { return g_err.c_str(); }

}  // extern "C"

namespace {
// C = alpha op(A) op(B) + beta C (single problem)
template <typename X>
int gemm_one(int m, int n, int k, const X* A, int lda, int acols, const X* B, int ldb, int bcols, X* Cm, int ldc,
             int flags, X alpha, X beta) {
  X *dA = dup(A, (size_t)lda * acols), *dB = dup(B, (size_t)ldb * bcols), *dC = dup(Cm, (size_t)ldc * n);
  GemmDesc<X> g{dA, dB, dC, lda, ldb, ldc, m, n, k, flags, alpha, beta};
  GemmDesc<X>* dg = dup(&g, 1);
  int tiles = gemm_tiles(m, n), zero = 0;
  int* ds = dup(&zero, 1);
  launch_gemm<X>(0, dg, ds, 1, tiles);
  KT(hipDeviceSynchronize());
  back(Cm, dC, (size_t)ldc * n);
  (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(dg); (void)hipFree(ds);
  return 0;
}
}  // namespace

extern "C" {

int ktest_gemm(int m, int n, int k, const double* A, int lda, int acols, const double* B, int ldb, int bcols,
               double* Cm, int ldc, int flags, double alpha, double beta) {
  return gemm_one<double>(m, n, k, A, lda, acols, B, ldb, bcols, Cm, ldc, flags, alpha, beta);
}
int ktest_gemm_f32(int m, int n, int k, const float* A, int lda, int acols, const float* B, int ldb, int bcols,
                   float* Cm, int ldc, int flags, float alpha, float beta) {
  return gemm_one<float>(m, n, k, A, lda, acols, B, ldb, bcols, Cm, ldc, flags, alpha, beta);
}

// Same product split along K into ksplit chunks (partials + ordered reduce).
int ktest_gemm_split(int m, int n, int k, const double* A, int lda, int acols, const double* B, int ldb, int bcols,
                     double* Cm, int ldc, int flags, double alpha, double beta, int ksplit) {
  if (ksplit < 1 || m < 1 || n < 1) { g_err = "bad split"; return -1; }
  double *dA = dup(A, (size_t)lda * acols), *dB = dup(B, (size_t)ldb * bcols), *dC = dup(Cm, (size_t)ldc * n);
  double* dW = dup<double>(nullptr, (size_t)ksplit * m * n);
  GemmDesc<double> g{dA, dB, dC, lda, ldb, ldc, m, n, k, flags, alpha, beta, ksplit, dW};
  GemmDesc<double>* dg = dup(&g, 1);
  int zero = 0;
  int* ds = dup(&zero, 1);
  launch_gemm<double>(0, dg, ds, 1, gemm_tiles(m, n) * ksplit);
  if (ksplit > 1) launch_gemm_reduce<double>(0, dg, ds, 1, (m * n + 255) / 256);
  KT(hipDeviceSynchronize());
  back(Cm, dC, (size_t)ldc * n);
  (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC); (void)hipFree(dW); (void)hipFree(dg); (void)hipFree(ds);
  return 0;
}

// sigma_min of upper-triangular R (k x k, ld) by the Lanczos kernel.
int ktest_smin(int k, const double* R, int ld, double* out) {
  if (k < 1 || k > smin_max_k() || ld < k) { g_err = "bad smin shape"; return -1; }
  double* dR = dup(R, (size_t)ld * k);
  double* dW = dup<double>(nullptr, smin_work_elems(k));
  double* dO = dup<double>(nullptr, 1);
  SminDesc<double> sd{dR, ld, k, dW, dO};
  SminDesc<double>* dd = dup(&sd, 1);
  launch_smin<double>(0, dd, 1, k);
  KT(hipDeviceSynchronize());
  back(out, dO, 1);
  (void)hipFree(dR); (void)hipFree(dW); (void)hipFree(dO); (void)hipFree(dd);
  return 0;
}

// Blocked Householder QR of one m x k panel (col-major, ld m): A <- V\R,
// tau (min(m,k)), Tb (nb x roundup(k,nb)), and C (m x nc) <- Q^T C.
int ktest_qr(int m, int k, double* A, double* tau, double* Tb, int nc, double* Cm) {
  const int nb = panel_nb(), kk = m < k ? m : k, nblk = (k + nb - 1) / nb;
  double *dA = dup(A, (size_t)m * k), *dt = dup<double>(nullptr, k), *dT = dup<double>(nullptr, (size_t)nb * nb * nblk);
  double* dC = dup(Cm, (size_t)m * (nc > 0 ? nc : 1));
  int zero = 0;
  int* ds = dup(&zero, 1);
  int* err = dup<int>(nullptr, 1);
  for (int jb = 0; jb < kk; jb += nb) {
    // register-form panel step over ceil((m - jb) / rows-per-workgroup) workgroups
    const int nwg = (m - jb + panel_rows_per_wg() - 1) / panel_rows_per_wg();
    if (nwg > panel_r_capacity<double>()) { g_err = "nwg exceeds resident capacity"; return -1; }
    double* part = dup<double>(nullptr, 2 * (size_t)nwg * 2 * nb);
    unsigned int* sync = dup<unsigned int>(nullptr, barrier_words(nwg));
    PanelDesc<double> pd{dA, dt, dT, m, m, k, nwg, part, sync, err};
    PanelDesc<double>* dpd = dup(&pd, 1);
    const int rc = launch_panel_r<double>(0, dpd, ds, 1, nwg, jb, nwg > 1 ? 1 : 0);
    if (rc != 0) { g_err = hipGetErrorString((hipError_t)rc); return -1; }
    KT(hipDeviceSynchronize());
    for (void* q : {(void*)part, (void*)sync, (void*)dpd}) (void)hipFree(q);
    const int w = (nb < kk - jb) ? nb : kk - jb, rows = m - jb;
    const double* Vb = dA + jb + (size_t)jb * m;
    const double* Tbb = dT + (size_t)jb * nb;
    // the engine's descriptors for this step (W = V_b T_b^T, Y = V_b^T C,
    // C -= W Y), each list run as one batched launch as the engine does
    std::vector<LarfbTarget<double>> tg;
    const int c0 = jb + ((nb < k - jb) ? nb : k - jb);
    if (k > c0) tg.push_back({dA + jb + (size_t)c0 * m, m, k - c0});
    if (nc > 0) tg.push_back({dC + jb, m, nc});
    std::vector<void*> scratch;
    auto alloc = [&](int64_t elems) {
      double* p = dup<double>(nullptr, (size_t)(elems > 0 ? elems : 1));
      scratch.push_back(p);
      return p;
    };
    std::vector<GemmDesc<double>> g1, g3;
    larfb_step_descs<double>(Vb, m, Tbb, nb, rows, w, tg.data(), (int)tg.size(), alloc, g1, g3);
    for (auto* gl : {&g1, &g3}) {
      if (gl->empty()) continue;
      std::vector<int> st(gl->size());
      int nt = 0;
      for (size_t i = 0; i < gl->size(); ++i) { st[i] = nt; nt += gemm_tiles((*gl)[i].m, (*gl)[i].n); }
      GemmDesc<double>* dg = dup(gl->data(), gl->size());
      int* dst = dup(st.data(), st.size());
      launch_gemm<double>(0, dg, dst, (int)gl->size(), nt);
      KT(hipDeviceSynchronize());
      (void)hipFree(dg);
      (void)hipFree(dst);
    }
    for (void* p : scratch) (void)hipFree(p);
  }
  KT(hipDeviceSynchronize());
  back(A, dA, (size_t)m * k);
  back(tau, dt, kk);
  back(Tb, dT, (size_t)nb * nb * nblk);
  if (nc > 0) back(Cm, dC, (size_t)m * nc);
  int e = 0;
  back(&e, err, 1);
  for (void* p : {(void*)dA, (void*)dt, (void*)dT, (void*)dC, (void*)ds, (void*)err}) (void)hipFree(p);
  if (e) { g_err = "barrier wait timed out"; return -1; }
  return 0;
}

// Block T factors of V (m x r) and tau.
int ktest_larft(int m, int r, const double* V, const double* tau, double* Tb) {
  const int nb = panel_nb(), nblk = (r + nb - 1) / nb;
  double *dV = dup(V, (size_t)m * r), *dt = dup(tau, r), *dT = dup<double>(nullptr, (size_t)nb * nb * nblk);
  LarftDesc<double> d{dV, dt, dT, m, m, r, 0};
  LarftDesc<double>* dd = dup(&d, 1);
  int zero = 0;
  int* ds = dup(&zero, 1);
  launch_larft<double>(0, dd, ds, 1, nblk);
  KT(hipDeviceSynchronize());
  back(Tb, dT, (size_t)nb * nb * nblk);
  for (void* p : {(void*)dV, (void*)dt, (void*)dT, (void*)dd, (void*)ds}) (void)hipFree(p);
  return 0;
}

// Truncated QP3 of C (k x N): Ct is N x k (ld N) in/out; V k x k, tau k.
// form 0: streaming kernel split over nwg workgroups; form 1 / 2: register
// kernel with 256 / 1024-thread workgroups and 32 rows per lane; form
// 10 * rc + 1 / 2: the same with rc rows per lane (group size and workgroup
// count follow from k and N).  Multi-
// workgroup problems are launched cooperatively.  Returns the rank, or -1 on
// a launch/run error.
int ktest_qp3(int k, int N, double* Ct, double* V, double* tau, double tol, int nwg, int form) {
  int G = 0, nt = 256;
  const int rcv = form >= 10 ? form / 10 : 32;  // register form rows per lane
  if (form >= 1) {
    G = qp3r_lanes(k, rcv);
    if (G == 0) { g_err = "k too large for the register form"; return -1; }
    nt = (form % 10) == 2 ? 1024 : 256;
    nwg = (N + nt / G - 1) / (nt / G);
    if (nwg > qp3r_capacity<double>(G, nt, rcv)) { g_err = "nwg exceeds resident capacity"; return -1; }
  } else if (nwg > qp3_capacity<double>(k)) {
    g_err = "nwg exceeds resident capacity";
    return -1;
  }
  double *dC = dup(Ct, (size_t)N * k), *dV = dup<double>(nullptr, (size_t)k * k), *dt = dup<double>(nullptr, k);
  double* vn = dup<double>(nullptr, 2 * (size_t)N);
  int* pos = dup<int>(nullptr, 2 * (size_t)N * nwg);
  double* pv = dup<double>(nullptr, 4 * (size_t)nwg);
  int* pp = dup<int>(nullptr, 4 * (size_t)nwg);
  double* beta = dup<double>(nullptr, k);
  unsigned int* sync = dup<unsigned int>(nullptr, barrier_words(nwg));
  int* err = dup<int>(nullptr, 1);
  int* rk = dup<int>(nullptr, 1);
  double* pcol = dup<double>(nullptr, 2 * (size_t)nwg * std::max(k, G * rcv));
  Qp3Desc<double> d{dC, dV, dt, vn, pos, pv, pp, beta, sync, err, rk, pcol, N, k, N, k, tol, nwg, G, rcv};
  Qp3Desc<double>* dd = dup(&d, 1);
  int zero = 0;
  int* ws = dup(&zero, 1);
  int rc = form >= 1 ? launch_qp3r<double>(0, dd, ws, 1, nwg, G, N, nt, rcv, nwg > 1 ? 1 : 0)
                     : launch_qp3<double>(0, dd, ws, 1, nwg, k, nwg > 1 ? 1 : 0);
  if (rc != 0) { g_err = hipGetErrorString((hipError_t)rc); return -1; }
  KT(hipDeviceSynchronize());
  int rank = 0, e = 0;
  back(&rank, rk, 1);
  back(&e, err, 1);
  back(Ct, dC, (size_t)N * k);
  back(V, dV, (size_t)k * k);
  back(tau, dt, k);
  for (void* q : {(void*)dC, (void*)dV, (void*)dt, (void*)vn, (void*)pos, (void*)pv, (void*)pp, (void*)beta,
                  (void*)sync, (void*)err, (void*)rk, (void*)dd, (void*)ws, (void*)pcol})
    (void)hipFree(q);
  if (e) { g_err = "barrier wait timed out"; return -1; }
  return rank;
}

// Timing of one truncated QP3 problem (either form): `reps` launches on a
// fresh copy of C each, the average kernel time in *ms.  Returns the rank.
int ktest_qp3_bench(int k, int N, const double* Ct, double tol, int form, int nwg, int reps, double* ms) {
  int G = 0, nt = 256;
  const int rcv = form >= 10 ? form / 10 : 32;  // register form rows per lane
  if (form >= 1) {
    G = qp3r_lanes(k, rcv);
    if (G == 0) { g_err = "k too large for the register form"; return -1; }
    nt = (form % 10) == 2 ? 1024 : 256;
    nwg = (N + nt / G - 1) / (nt / G);
    if (nwg > qp3r_capacity<double>(G, nt, rcv)) { g_err = "nwg exceeds resident capacity"; return -1; }
  } else if (nwg > qp3_capacity<double>(k)) {
    g_err = "nwg exceeds resident capacity";
    return -1;
  }
  double *d0 = dup(Ct, (size_t)N * k), *dC = dup<double>(nullptr, (size_t)N * k);
  double *dV = dup<double>(nullptr, (size_t)k * k), *dt = dup<double>(nullptr, k);
  double* vn = dup<double>(nullptr, 2 * (size_t)N);
  int* pos = dup<int>(nullptr, 2 * (size_t)N * nwg);
  double* pv = dup<double>(nullptr, 4 * (size_t)nwg);
  int* pp = dup<int>(nullptr, 4 * (size_t)nwg);
  double* beta = dup<double>(nullptr, k);
  unsigned int* sync = dup<unsigned int>(nullptr, barrier_words(nwg));
  int* err = dup<int>(nullptr, 1);
  int* rk = dup<int>(nullptr, 1);
  double* pcol = dup<double>(nullptr, 2 * (size_t)nwg * std::max(k, G * rcv));
  Qp3Desc<double> d{dC, dV, dt, vn, pos, pv, pp, beta, sync, err, rk, pcol, N, k, N, k, tol, nwg, G, rcv};
  Qp3Desc<double>* dd = dup(&d, 1);
  int zero = 0;
  int* ws = dup(&zero, 1);
  hipEvent_t a, b;
  KT(hipEventCreate(&a));
  KT(hipEventCreate(&b));
  double tot = 0;
  int rc = 0;
  for (int r = 0; r < reps && rc == 0; ++r) {
    KT(hipMemcpy(dC, d0, sizeof(double) * (size_t)N * k, hipMemcpyDeviceToDevice));
    KT(hipMemset(sync, 0, sizeof(unsigned int) * barrier_words(nwg)));
    KT(hipEventRecord(a, 0));
    rc = form >= 1 ? launch_qp3r<double>(0, dd, ws, 1, nwg, G, N, nt, rcv, nwg > 1 ? 1 : 0)
                   : launch_qp3<double>(0, dd, ws, 1, nwg, k, nwg > 1 ? 1 : 0);
    KT(hipEventRecord(b, 0));
    KT(hipEventSynchronize(b));
    float t = 0;
    KT(hipEventElapsedTime(&t, a, b));
    tot += t;
  }
  if (rc != 0) { g_err = hipGetErrorString((hipError_t)rc); return -1; }
  *ms = tot / reps;
  int rank = 0, e = 0;
  back(&rank, rk, 1);
  back(&e, err, 1);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  for (void* q : {(void*)d0, (void*)dC, (void*)dV, (void*)dt, (void*)vn, (void*)pos, (void*)pv, (void*)pp,
                  (void*)beta, (void*)sync, (void*)err, (void*)rk, (void*)dd, (void*)ws, (void*)pcol})
    (void)hipFree(q);
  if (e) { g_err = "barrier wait timed out"; return -1; }
  return rank;
}

// x <- H^T x (trans=1) or H x for blocked reflectors (V m x k, Tb), with x
// split into `nseg` equal-length segments stored at stride `gap` in a buffer.
// form 0: LDS form (4 row-slice workgroups for m >= 2048); form 1: register
// form, 7 row waves (ceil(m / 448) workgroups); form 2: register form with
// the row waves reflr_row_waves(m) gives (1 or 2 for m <= 128, else 7)
int ktest_refl(int m, int k, const double* V, const double* Tb, double* x, int trans, int nseg, int gap, int form) {
  const int nb = panel_nb(), nblk = (k + nb - 1) / nb;
  const int seglen = m / nseg;
  const size_t xlen = (size_t)nseg * (seglen + gap);
  double *dV = dup(V, (size_t)m * k), *dT = dup(Tb, (size_t)nb * nb * nblk);
  std::vector<double> xb(xlen, -7.0);
  std::vector<Seg> segs;
  for (int s = 0; s < nseg; ++s) {
    segs.push_back(Seg{(int64_t)s * (seglen + gap), seglen, 0});
    std::memcpy(&xb[(size_t)s * (seglen + gap)], x + (size_t)s * seglen, sizeof(double) * seglen);
  }
  double* dx = dup(xb.data(), xlen);
  Seg* dsg = dup(segs.data(), segs.size());
  const int nrw = form == 2 ? reflr_row_waves(m) : 7;
  const int nwg = form ? (m + 64 * nrw - 1) / (64 * nrw) : (m >= 2048 ? 4 : 1);
  double* wk = dup<double>(nullptr, m);
  double* part = dup<double>(nullptr, 2 * 32 * (size_t)nwg);
  unsigned int* sync = dup<unsigned int>(nullptr, 1);
  int* err = dup<int>(nullptr, 1);
  ReflDesc<double> d{dV, dT, m, m, k, nb, 0, nseg, trans, nwg, wk, part, sync, err};
  ReflDesc<double>* dd = dup(&d, 1);
  double* work = dup<double>(nullptr, m);
  int zero = 0;
  int* ws = dup(&zero, 1);
  if (form) launch_reflr<double>(0, dd, ws, 1, nwg, dsg, dx, nseg, nrw);
  else launch_refl<double>(0, dd, ws, 1, nwg, dsg, dx, work, nwg > 1 ? 0 : m);
  KT(hipDeviceSynchronize());
  back(xb.data(), dx, xlen);
  for (int s = 0; s < nseg; ++s)
    std::memcpy(x + (size_t)s * seglen, &xb[(size_t)s * (seglen + gap)], sizeof(double) * seglen);
  int e = 0;
  back(&e, err, 1);
  for (void* p : {(void*)dV, (void*)dT, (void*)dx, (void*)dsg, (void*)dd, (void*)work, (void*)wk, (void*)part,
                  (void*)sync, (void*)err, (void*)ws})
    (void)hipFree(p);
  if (e) { g_err = "barrier wait timed out"; return -1; }
  return 0;
}

// xs <- R^{-1}(xs - A1 x1) (trans=0) ; or xs <- R^{-T} xs, x1 -= A1^T xs (trans=1)
// form 1: the latency form (one workgroup, no R12 products: s must be 0)
int ktest_bwd(int k, const double* R, int s, const double* A1, double* xs, double* x1, int trans, int form) {
  double *dR = dup(R, (size_t)k * k), *dA = dup(A1, (size_t)k * (s > 0 ? s : 1));
  std::vector<double> xb((size_t)k + s);
  std::memcpy(xb.data(), xs, sizeof(double) * k);
  if (s > 0) std::memcpy(xb.data() + k, x1, sizeof(double) * s);
  double* dx = dup(xb.data(), xb.size());
  R12Desc<double> q{dA, k, s, (int64_t)k};
  R12Desc<double>* dq = dup(&q, 1);
  const int nwg = k >= 512 ? 4 : 1;
  unsigned int* sync = dup<unsigned int>(nullptr, 1);
  int* err = dup<int>(nullptr, 1);
  BwdDesc<double> d{dR, k, k, 0, 0, s > 0 ? 1 : 0, trans, nwg, sync, err};
  BwdDesc<double>* dd = dup(&d, 1);
  int zero = 0;
  int* ws = dup(&zero, 1);
  if (form) {
    BwdDesc<double> d1 = d;
    d1.nwg = 1;
    KT(hipMemcpy(dd, &d1, sizeof(d1), hipMemcpyHostToDevice));
    launch_bwdr<double>(0, dd, 1, k, dx);
  } else {
    launch_bwd<double>(0, dd, ws, 1, nwg, dq, dx);
  }
  KT(hipDeviceSynchronize());
  back(xb.data(), dx, xb.size());
  std::memcpy(xs, xb.data(), sizeof(double) * k);
  if (s > 0) std::memcpy(x1, xb.data() + k, sizeof(double) * s);
  int e = 0;
  back(&e, err, 1);
  for (void* p : {(void*)dR, (void*)dA, (void*)dx, (void*)dq, (void*)dd, (void*)sync, (void*)err, (void*)ws})
    (void)hipFree(p);
  if (e) { g_err = "barrier wait timed out"; return -1; }
  return 0;
}

// B (rows x k) <- B R^{-1} with the blocked GEMM + row-block solve.
int ktest_trsm(int rows, int k, const double* R, double* B) {
  const int nb = panel_nb();
  double *dR = dup(R, (size_t)k * k), *dB = dup(B, (size_t)rows * k);
  int zero = 0;
  int* ds = dup(&zero, 1);
  for (int jb = 0; jb < k; jb += nb) {
    const int w = (nb < k - jb) ? nb : k - jb;
    if (jb > 0) {
      GemmDesc<double> g{dB, dR + (size_t)jb * k, dB + (size_t)jb * rows, rows, k, rows, rows, w, jb,
                         gemm_flags(false, false), -1.0, 1.0};
      GemmDesc<double>* dg = dup(&g, 1);
      launch_gemm<double>(0, dg, ds, 1, gemm_tiles(rows, w));
      KT(hipDeviceSynchronize());
      (void)hipFree(dg);
    }
    TrsmDesc<double> t{dB, dR, rows, k, rows, k};
    TrsmDesc<double>* dt = dup(&t, 1);
    launch_trsm_block<double>(0, dt, ds, 1, (rows + trsm_rows_per_tile() - 1) / trsm_rows_per_tile(), jb);
    KT(hipDeviceSynchronize());
    (void)hipFree(dt);
  }
  back(B, dB, (size_t)rows * k);
  for (void* p : {(void*)dR, (void*)dB, (void*)ds}) (void)hipFree(p);
  return 0;
}

// Batched copy round trip: dst = src (trans 0/1) for one block.
int ktest_copy(int rows, int cols, const double* src, int lds, double* dst, int ldd, int trans) {
  const int drows = trans ? cols : rows, dcols = trans ? rows : cols;
  double *ds_ = dup(src, (size_t)lds * cols), *dd = dup(dst, (size_t)ldd * dcols);
  CopyDesc c{ds_, dd, lds, ldd, rows, cols, trans, 0};
  CopyDesc* dc = dup(&c, 1);
  int zero = 0;
  int* st = dup(&zero, 1);
  launch_copy(0, 8, 8, dc, st, 1, ((rows + 31) / 32) * ((cols + 31) / 32));
  KT(hipDeviceSynchronize());
  back(dst, dd, (size_t)ldd * dcols);
  (void)drows;
  for (void* p : {(void*)ds_, (void*)dd, (void*)dc, (void*)st}) (void)hipFree(p);
  return 0;
}

}  // extern "C"
