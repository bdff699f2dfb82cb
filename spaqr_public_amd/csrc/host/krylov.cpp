#include "krylov.hpp"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../kernels/launch.hpp"

namespace spq {

EngineBase::~EngineBase() {
  for (auto& w : work_dev_) dev_free(w.p);
  for (auto& w : work_host_) host_free(w.p);
}

double* EngineBase::work_dev(int slot, size_t elems) {
  if ((int)work_dev_.size() <= slot) work_dev_.resize(slot + 1);
  Work& w = work_dev_[slot];
  elems = std::max<size_t>(elems, 1);
  if (w.cap < elems) {
    dev_free(w.p);
    w.p = nullptr;
    w.cap = 0;
    SPQ_HIP(dev_malloc((void**)&w.p, sizeof(double) * elems));
    w.cap = elems;
  }
  return w.p;
}

double* EngineBase::work_host(int slot, size_t elems) {
  if ((int)work_host_.size() <= slot) work_host_.resize(slot + 1);
  Work& w = work_host_[slot];
  elems = std::max<size_t>(elems, 1);
  if (w.cap < elems) {
    host_free(w.p);
    w.p = nullptr;
    w.cap = 0;
    SPQ_HIP(host_malloc((void**)&w.p, sizeof(double) * elems));
    w.cap = elems;
  }
  return w.p;
}

DevMatrix::~DevMatrix() {
  for (void* p : {(void*)rp, (void*)ci, (void*)v, (void*)rpt, (void*)cit, (void*)vt}) dev_free(p);
}

namespace {
void up(void** d, const void* h, size_t bytes) {
  SPQ_HIP(dev_malloc(d, bytes > 0 ? bytes : 8));
  if (bytes) SPQ_HIP(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
}
}  // namespace

DevMatrix* upload_matrix(const CscMat& A, bool with_transpose) {
  DevMatrix* M = new DevMatrix;
  M->m = A.m; M->n = A.n; M->nnz = A.nnz();
  // CSR of A == CSC of A^T
  CscMat At = csc_transpose(A);
  up((void**)&M->rp, At.p.data(), sizeof(int) * At.p.size());
  up((void**)&M->ci, At.i.data(), sizeof(int) * At.i.size());
  up((void**)&M->v, At.x.data(), sizeof(double) * At.x.size());
  if (with_transpose) {
    up((void**)&M->rpt, A.p.data(), sizeof(int) * A.p.size());
    up((void**)&M->cit, A.i.data(), sizeof(int) * A.i.size());
    up((void**)&M->vt, A.x.data(), sizeof(double) * A.x.size());
  }
  return M;
}

GmresWork gmres_work_tree(EngineBase& M, int n, int restart, bool refine) {
  // slots 0-4 and host slot 0: gmres_dev; slots 5-8 and host slot 1: the
  // refinement loop around it
  GmresWork w{};
  n = std::max(n, 1);
  restart = std::max(1, std::min(restart, n));
  w.V = M.work_dev(0, (size_t)n * (restart + 1));
  w.w = M.work_dev(1, n);
  w.p0 = M.work_dev(2, n);
  w.h = M.work_dev(3, (size_t)(restart + 2) * 2);
  w.dpart = M.work_dev(4, dots_part_elems(restart + 1));
  w.hh = M.work_host(0, (size_t)(restart + 2) * 2);
  if (refine) {
    w.r = M.work_dev(5, n);
    w.d = M.work_dev(6, n);
    w.h_ir = M.work_dev(7, 2);
    w.dpart_ir = M.work_dev(8, dots_part_elems(1));
    w.hh_ir = M.work_host(1, 2);
  }
  return w;
}

KrylovResult gmres_dev(EngineBase& M, const DevMatrix& A, const double* b, double* x, int iters, int restart,
                       double tol, int verbose, bool trans, const GmresWork* wk) {
  KrylovResult res;
  hipStream_t s = M.stream();
  const int n = A.m;
  if (trans && !A.rpt) throw Error(kErrArg, "gmres: the matrix was uploaded without its transpose");
  const int* rp = trans ? A.rpt : A.rp;
  const int* ci = trans ? A.cit : A.ci;
  const double* av = trans ? A.vt : A.v;
  restart = std::max(1, std::min(restart, n));
  const double t0 = wall();
  // work vectors: the caller's, or the tree's own (kept across calls, so
  // repeated calls allocate nothing)
  const GmresWork own = wk ? GmresWork{} : gmres_work_tree(M, n, restart, false);
  if (!wk) wk = &own;
  double* V = wk->V;
  double* w = wk->w;
  double* p0 = wk->p0;
  double* h = wk->h;
  double* dpart = wk->dpart;
  double* hh = wk->hh;
  auto norm_dev = [&](const double* v) {
    launch_dots(s, n, 1, v, n, v, h, dpart);
    SPQ_HIP(hipMemcpyAsync(hh, h, sizeof(double), hipMemcpyDeviceToHost, s));
    SPQ_HIP(hipStreamSynchronize(s));
    return std::sqrt(hh[0]);
  };
  auto precond = [&](const double* in, double* out) {
    const double a = wall();
    if (trans) M.solve_t_dev(in, out);
    else M.solve_dev(in, out);
    if (verbose) { SPQ_HIP(hipStreamSynchronize(s)); res.t_precond += wall() - a; }
    res.n_precond++;
  };
  if (norm_dev(b) <= std::numeric_limits<double>::min()) {
    SPQ_HIP(hipMemsetAsync(x, 0, sizeof(double) * A.n, s));
    SPQ_HIP(hipStreamSynchronize(s));
    return res;
  }
  std::vector<double> H((size_t)(restart + 1) * restart), g(restart + 1), cs(restart), sn(restart), y(restart);
  int it = 0;
  // the stopping quantity divides by the norm of the first preconditioned
  // residual for the whole run, restarts included (is.cpp:133, 217)
  double r0norm = -1.0;
  for (;;) {
    // r0 = M^{-1} (b - A x)
    SPQ_HIP(hipMemcpyAsync(p0, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    launch_spmv(s, n, rp, ci, av, x, p0, -1.0, 1.0);
    precond(p0, V);
    const double beta = norm_dev(V);
    if (r0norm < 0.0) r0norm = beta;
    if (beta == 0.0) break;
    launch_axpby(s, n, 0.0, V, 1.0 / beta, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    bool stop = false;
    while (k < restart) {
      ++it;
      ++k;
      double* vk = V + (size_t)n * (k - 1);
      double* vn = V + (size_t)n * k;
      launch_spmv(s, n, rp, ci, av, vk, w, 1.0, 0.0);
      precond(w, vn);
      // CGS2 against V(:, 0:k)
      launch_dots(s, n, k, V, n, vn, h, dpart);
      launch_gemv_sub(s, n, k, V, n, h, vn);
      launch_dots(s, n, k, V, n, vn, h + k, dpart);
      launch_gemv_sub(s, n, k, V, n, h + k, vn);
      launch_dots(s, n, 1, vn, n, vn, h + 2 * k, dpart);
      SPQ_HIP(hipMemcpyAsync(hh, h, sizeof(double) * (2 * k + 1), hipMemcpyDeviceToHost, s));
      SPQ_HIP(hipStreamSynchronize(s));
      double* col = &H[(size_t)(k - 1) * (restart + 1)];
      for (int i = 0; i < k; ++i) col[i] = hh[i] + hh[k + i];
      const double hk = std::sqrt(hh[2 * k]);
      col[k] = hk;
      if (hk != 0.0) launch_axpby(s, n, 0.0, vn, 1.0 / hk, vn);
      for (int i = 0; i < k - 1; ++i) {
        const double a = col[i], bb = col[i + 1];
        col[i] = cs[i] * a + sn[i] * bb;
        col[i + 1] = -sn[i] * a + cs[i] * bb;
      }
      const double a = col[k - 1], bb = col[k];
      const double r = std::hypot(a, bb);
      cs[k - 1] = (r == 0.0) ? 1.0 : a / r;
      sn[k - 1] = (r == 0.0) ? 0.0 : bb / r;
      col[k - 1] = r;
      col[k] = 0.0;
      g[k] = -sn[k - 1] * g[k - 1];
      g[k - 1] = cs[k - 1] * g[k - 1];
      res.est = std::fabs(g[k]) / r0norm;
      if (verbose > 1) printf("%d: |Ax-b|/|b| = %3.2e <? %3.2e\n", it, res.est, tol);
      stop = (k == n || res.est < tol || it >= iters || hk == 0.0);
      if (stop) break;
    }
    // x += V(:, 0:k) y, H(0:k,0:k) y = g(0:k)
    for (int i = k - 1; i >= 0; --i) {
      double t = g[i];
      for (int j = i + 1; j < k; ++j) t -= H[(size_t)j * (restart + 1) + i] * y[j];
      y[i] = t / H[(size_t)i * (restart + 1) + i];
    }
    SPQ_HIP(hipMemcpyAsync(h, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    launch_gemv(s, n, k, V, n, h, w);
    launch_axpby(s, n, 1.0, w, 1.0, x);
    if (stop) break;
  }
  SPQ_HIP(hipStreamSynchronize(s));
  res.iters = it;
  res.t_total = wall() - t0;
  return res;
}

KrylovResult gmres_ir_dev(EngineBase& M, const DevMatrix& A, const double* b, double* x, int iters, int restart,
                          double tol, int verbose, bool trans, const GmresWork* wk) {
  KrylovResult res;
  if (trans && !A.rpt) throw Error(kErrArg, "gmres: the matrix was uploaded without its transpose");
  hipStream_t s = M.stream();
  const int n = A.m;
  const double t0 = wall();
  const GmresWork own = wk ? GmresWork{} : gmres_work_tree(M, n, restart, true);
  if (!wk) wk = &own;
  double* r = wk->r;
  double* d = wk->d;
  double* h = wk->h_ir;
  double* dpart = wk->dpart_ir;
  double* hh = wk->hh_ir;
  auto norm = [&](const double* v) {
    launch_dots(s, n, 1, v, n, v, h, dpart);
    SPQ_HIP(hipMemcpyAsync(hh, h, sizeof(double), hipMemcpyDeviceToHost, s));
    SPQ_HIP(hipStreamSynchronize(s));
    return std::sqrt(hh[0]);
  };
  const double bn = norm(b);
  if (bn <= std::numeric_limits<double>::min()) {
    SPQ_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    SPQ_HIP(hipStreamSynchronize(s));
    res.t_total = wall() - t0;
    return res;
  }
  double prev = 0;
  for (;;) {
    SPQ_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (trans) launch_spmv(s, n, A.rpt, A.cit, A.vt, x, r, -1.0, 1.0);
    else launch_spmv(s, n, A.rp, A.ci, A.v, x, r, -1.0, 1.0);
    const double rel = norm(r) / bn;
    res.est = rel;
    if (verbose) printf("refinement %d: |b - Ax|/|b| = %3.2e (%d inner iterations)\n", res.outer, rel, res.iters);
    if (rel <= tol || res.iters >= iters || res.outer >= kIrMaxOuter) break;
    if (res.outer > 0 && rel > 0.5 * prev) break;  // the FP32 factors gain no more
    prev = rel;
    SPQ_HIP(hipMemsetAsync(d, 0, sizeof(double) * n, s));
    const int left = iters - res.iters;
    // no tighter than this step needs to reach tol (with a factor 2 margin)
    const double inner_tol = std::max(kIrInnerTol, 0.5 * tol / rel);
    KrylovResult in = gmres_dev(M, A, r, d, left, std::min(restart, left), inner_tol, verbose > 1 ? verbose : 0, trans,
                                wk);
    res.iters += in.iters;
    res.n_precond += in.n_precond;
    res.t_precond += in.t_precond;
    launch_axpby(s, n, 1.0, d, 1.0, x);
    ++res.outer;
  }
  SPQ_HIP(hipStreamSynchronize(s));
  res.t_total = wall() - t0;
  return res;
}

// Preconditioned CGLS on the normal equations (is.cpp:275-346): the
// preconditioner is the normal-equation apply of the tree, the stop test is
// ||A^T r||^2 < tol^2 ||A^T b||^2.  FP64 vectors and recurrences; with FP32
// factors the direction update uses the Polak-Ribiere beta.  b (m), x (n) on
// the device.
KrylovResult cgls_dev(EngineBase& M, const DevMatrix& A, const double* b, double* x, int iters, double tol,
                      int verbose) {
  KrylovResult res;
  hipStream_t s = M.stream();
  const int m = A.m, n = A.n;
  if (!A.rpt) throw Error(kErrArg, "cgls: the matrix was uploaded without its transpose");
  const double t0 = wall();
  // slots 9-16 and host slot 2
  double* r = M.work_dev(9, std::max(m, 1));
  double* tmp = M.work_dev(10, std::max(m, 1));
  double* nr = M.work_dev(11, std::max(n, 1));
  double* p = M.work_dev(12, std::max(n, 1));
  double* z = M.work_dev(13, std::max(n, 1));
  const bool flexible = M.precision_bytes() == 4;
  double* zold = M.work_dev(14, flexible ? std::max(n, 1) : 1);
  double* h = M.work_dev(15, 4);
  double* dpart = M.work_dev(16, dots_part_elems(1));
  double* hh = M.work_host(2, 4);
  auto dot = [&](int len, const double* u, const double* v) {
    launch_dots(s, len, 1, u, len, v, h, dpart);
    SPQ_HIP(hipMemcpyAsync(hh, h, sizeof(double), hipMemcpyDeviceToHost, s));
    SPQ_HIP(hipStreamSynchronize(s));
    return hh[0];
  };
  auto At = [&](const double* in, double* out) { launch_spmv(s, n, A.rpt, A.cit, A.vt, in, out, 1.0, 0.0); };
  auto precond = [&](const double* in, double* out) {
    const double a = wall();
    M.solve_nrml_dev(in, out);
    if (verbose) { SPQ_HIP(hipStreamSynchronize(s)); res.t_precond += wall() - a; }
    res.n_precond++;
  };
  // r = b - A x ; nres = A^T r ; A^T b
  SPQ_HIP(hipMemcpyAsync(r, b, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
  launch_spmv(s, m, A.rp, A.ci, A.v, x, r, -1.0, 1.0);
  At(r, nr);
  At(b, z);
  const double rhsn2 = dot(n, z, z);
  if (rhsn2 == 0.0) {
    SPQ_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    SPQ_HIP(hipStreamSynchronize(s));
    res.t_total = wall() - t0;
    return res;
  }
  const double thr = tol * tol * rhsn2;
  double rn2 = dot(n, nr, nr);
  int i = 0;
  if (rn2 >= thr) {
    precond(nr, p);
    double absNew = dot(n, nr, p);
    while (i < iters) {
      launch_spmv(s, m, A.rp, A.ci, A.v, p, tmp, 1.0, 0.0);
      const double tn = dot(m, tmp, tmp);
      const double alpha = absNew / tn;
      launch_axpby(s, n, alpha, p, 1.0, x);
      launch_axpby(s, m, -alpha, tmp, 1.0, r);
      At(r, nr);
      rn2 = dot(n, nr, nr);
      if (verbose > 1) printf("%d %3.2e \n", i, std::sqrt(rn2 / rhsn2));
      if (rn2 < thr) break;
      if (flexible) SPQ_HIP(hipMemcpyAsync(zold, z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      precond(nr, z);
      const double absOld = absNew;
      absNew = dot(n, nr, z);
      // FP32 factors make the preconditioner slightly non-linear: the
      // Polak-Ribiere form (nr . (z - z_old)) keeps the directions conjugate
      const double beta = (flexible ? absNew - dot(n, nr, zold) : absNew) / absOld;
      launch_axpby(s, n, 1.0, z, beta, p);
      ++i;
    }
  }
  SPQ_HIP(hipStreamSynchronize(s));
  res.iters = i;
  res.est = std::sqrt(rn2 / rhsn2);
  res.t_total = wall() - t0;
  return res;
}

}  // namespace spq
