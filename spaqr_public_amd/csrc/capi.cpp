// extern "C" boundary of libspaqr.so (declared in include/spaqr.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>

#include "capi_impl.hpp"
#include "kernels/launch.hpp"

using namespace spq;

namespace spq {
std::string& last_error() {
  thread_local std::string msg;
  return msg;
}
}  // namespace spq

namespace {
CscMat make_csc(int m, int n, const int* colptr, const int* rowind, const double* val) {
  if (m < 0 || n < 0 || !colptr || (colptr[n] > 0 && (!rowind || !val)))
    throw Error(kErrArg, "invalid CSC arrays");
  CscMat A;
  A.m = m;
  A.n = n;
  A.p.assign(colptr, colptr + n + 1);
  const int nnz = colptr[n];
  A.i.assign(rowind, rowind + nnz);
  A.x.assign(val, val + nnz);
  for (int j = 0; j < n; ++j) {
    if (A.p[j] > A.p[j + 1]) throw Error(kErrArg, "colptr not nondecreasing");
    for (int q = A.p[j]; q < A.p[j + 1]; ++q) {
      if (A.i[q] < 0 || A.i[q] >= m) throw Error(kErrArg, "row index out of range");
      if (q > A.p[j] && A.i[q] <= A.i[q - 1]) throw Error(kErrArg, "row indices not sorted/unique");
    }
  }
  return A;
}

}  // namespace

extern "C" {

const char* spaqr_last_error(void) { return last_error().c_str(); }

spaqr_tree* spaqr_tree_create(int nlevels, int skip, int precision, int device) {
  spaqr_tree* t = nullptr;
  int rc = guard([&] {
    if (nlevels < 1) throw Error(kErrArg, "nlevels must be >= 1");
    if (precision != SPAQR_FP64 && precision != SPAQR_FP32) throw Error(kErrArg, "precision must be 8 or 4");
    t = new spaqr_tree;
    t->device = device;
    t->e.reset(make_engine(nlevels, skip, precision, device));
  });
  if (rc != SPAQR_OK) { delete t; return nullptr; }
  return t;
}
void spaqr_tree_destroy(spaqr_tree* t) { delete t; }

int spaqr_set_tol(spaqr_tree* t, double v) { return guard([&] { eng(t).set_tol(v); }); }
int spaqr_set_skip(spaqr_tree* t, int v) { return guard([&] { eng(t).set_skip(v); }); }
int spaqr_set_scale(spaqr_tree* t, int v) { return guard([&] { eng(t).set_scale(v); }); }
int spaqr_set_square(spaqr_tree* t, int v) { return guard([&] { eng(t).set_square(v); }); }
int spaqr_set_order(spaqr_tree* t, float v) { return guard([&] { eng(t).set_order(v); }); }
int spaqr_set_hsl(spaqr_tree* t, int v) {
  return guard([&] {
    eng(t);
    if (v != 0) throw Error(kErrArg, "HSL MC64 bipartite matching is not available (licensed library absent)");
  });
}
int spaqr_set_coordinates(spaqr_tree* t, const double* X, int dim, int n) {
  return guard([&] {
    if (!X || dim < 1 || n < 0) throw Error(kErrArg, "invalid coordinates");
    eng(t).set_coords(std::vector<double>(X, X + (size_t)dim * n), dim);
  });
}
int spaqr_set_tensor_coordinates(spaqr_tree* t, int n, int dim) {
  return guard([&] {
    if (n < 1 || dim < 1 || dim > 3) throw Error(kErrArg, "tensor coordinates need n >= 1 and 1 <= dim <= 3");
    eng(t).set_coords(tensor_coordinates(n, dim), dim);
  });
}
int spaqr_set_row_parts(spaqr_tree* t, const int* parts, int m) {
  return guard([&] {
    if (!parts || m < 0) throw Error(kErrArg, "invalid row partition");
    eng(t).set_row_parts(std::vector<int>(parts, parts + m));
  });
}
int spaqr_row_parts(int m, int n, const int* colptr, const int* rowind, int nlevels, int* parts) {
  return guard([&] {
    if (!parts || !colptr || n < 0) throw Error(kErrArg, "row_parts: null array");
    const std::vector<double> pattern(std::max(colptr[n], 0), 1.0);
    const std::vector<int> p = row_parts(make_csc(m, n, colptr, rowind, pattern.data()), nlevels);
    std::copy(p.begin(), p.end(), parts);
  });
}
int spaqr_set_verbose(spaqr_tree* t, int v) { return guard([&] { eng(t).set_verbose(v); }); }
int spaqr_rows(const spaqr_tree* t) { return t && t->e ? t->e->rows() : SPAQR_E_ARG; }
int spaqr_cols(const spaqr_tree* t) { return t && t->e ? t->e->cols() : SPAQR_E_ARG; }
int spaqr_levels(const spaqr_tree* t) { return t && t->e ? t->e->levels() : SPAQR_E_ARG; }

int spaqr_partition(spaqr_tree* t, int m, int n, const int* colptr, const int* rowind, const double* val) {
  return guard([&] {
    if (m < n) throw Error(kErrArg, "partition: rows < cols (pass the transpose, as spaQR.cpp:88-92 does)");
    eng(t).partition(make_csc(m, n, colptr, rowind, val));
  });
}
int spaqr_assemble(spaqr_tree* t, int m, int n, const int* colptr, const int* rowind, const double* val) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (m != e.rows() || n != e.cols()) throw Error(kErrArg, "assemble: matrix shape differs from partition");
    e.assemble(make_csc(m, n, colptr, rowind, val));
  });
}
int spaqr_factorize(spaqr_tree* t) { return guard([&] { eng(t).factorize(); }); }

int spaqr_solve_device(spaqr_tree* t, const double* b, double* x) {
  return guard([&] {
    EngineBase& e = eng(t);
    e.solve_dev(b, x);
    SPQ_HIP(hipStreamSynchronize(e.stream()));
    e.check_solve();
  });
}
int spaqr_solve(spaqr_tree* t, const double* b, double* x) {
  return guard([&] {
    EngineBase& e = eng(t);
    double *db = nullptr, *dx = nullptr;
    SPQ_HIP(dev_malloc((void**)&db, sizeof(double) * std::max(e.rows(), 1)));
    SPQ_HIP(dev_malloc((void**)&dx, sizeof(double) * std::max(e.cols(), 1)));
    struct F { double* p; ~F() { dev_free(p); } } fb{db}, fx{dx};
    SPQ_HIP(hipMemcpy(db, b, sizeof(double) * e.rows(), hipMemcpyHostToDevice));
    e.solve_dev(db, dx);
    SPQ_HIP(hipStreamSynchronize(e.stream()));
    e.check_solve();
    SPQ_HIP(hipMemcpy(x, dx, sizeof(double) * e.cols(), hipMemcpyDeviceToHost));
  });
}
int spaqr_solve_nrml(spaqr_tree* t, const double* b, double* x) {
  return guard([&] {
    EngineBase& e = eng(t);
    double *db = nullptr, *dx = nullptr;
    SPQ_HIP(dev_malloc((void**)&db, sizeof(double) * std::max(e.cols(), 1)));
    SPQ_HIP(dev_malloc((void**)&dx, sizeof(double) * std::max(e.cols(), 1)));
    struct F { double* p; ~F() { dev_free(p); } } fb{db}, fx{dx};
    SPQ_HIP(hipMemcpy(db, b, sizeof(double) * e.cols(), hipMemcpyHostToDevice));
    e.solve_nrml_dev(db, dx);
    SPQ_HIP(hipStreamSynchronize(e.stream()));
    e.check_solve();
    SPQ_HIP(hipMemcpy(x, dx, sizeof(double) * e.cols(), hipMemcpyDeviceToHost));
  });
}

int spaqr_solve_nrml_device(spaqr_tree* t, const double* b, double* x) {
  return guard([&] {
    EngineBase& e = eng(t);
    e.solve_nrml_dev(b, x);
    SPQ_HIP(hipStreamSynchronize(e.stream()));
    e.check_solve();
  });
}

spaqr_matrix* spaqr_matrix_create(spaqr_tree* t, int m, int n, const int* colptr, const int* rowind,
                                  const double* val) {
  spaqr_matrix* M = nullptr;
  int rc = guard([&] {
    eng(t);
    SPQ_HIP(hipSetDevice(t->device));
    M = new spaqr_matrix;
    M->d.reset(upload_matrix(make_csc(m, n, colptr, rowind, val), true));
  });
  if (rc != SPAQR_OK) { delete M; return nullptr; }
  return M;
}
void spaqr_matrix_destroy(spaqr_matrix* A) { delete A; }

int spaqr_gmres_device(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, int restart,
                       double tol, int verbose, int* iters_out, double* est_out) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!A || !A->d) throw Error(kErrArg, "null matrix");
    if (A->d->m != A->d->n || A->d->m != e.rows()) throw Error(kErrArg, "gmres needs the square matrix of the tree");
    KrylovResult r = e.precision_bytes() == 4 ? gmres_ir_dev(e, *A->d, b, x, iters, restart, tol, verbose)
                                               : gmres_dev(e, *A->d, b, x, iters, restart, tol, verbose);
    e.check_solve();
    if (iters_out) *iters_out = r.iters;
    if (est_out) *est_out = r.est;
  });
}
int spaqr_gmres_transpose_device(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters,
                                 int restart, double tol, int verbose, int* iters_out, double* est_out) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!A || !A->d) throw Error(kErrArg, "null matrix");
    if (A->d->m != A->d->n || A->d->m != e.rows()) throw Error(kErrArg, "gmres needs the square matrix of the tree");
    KrylovResult r = e.precision_bytes() == 4 ? gmres_ir_dev(e, *A->d, b, x, iters, restart, tol, verbose, true)
                                               : gmres_dev(e, *A->d, b, x, iters, restart, tol, verbose, true);
    e.check_solve();
    if (iters_out) *iters_out = r.iters;
    if (est_out) *est_out = r.est;
  });
}
int spaqr_solve_transpose_device(spaqr_tree* t, const double* b, double* x) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!b || !x) throw Error(kErrArg, "null vector");
    e.solve_t_dev(b, x);
    SPQ_HIP(hipStreamSynchronize(e.stream()));
    e.check_solve();
  });
}
int spaqr_solve_transpose(spaqr_tree* t, const double* b, double* x) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!b || !x) throw Error(kErrArg, "null vector");
    const int n = e.rows();
    double* d = nullptr;
    SPQ_HIP(dev_malloc((void**)&d, sizeof(double) * std::max(n, 1)));
    struct F { double* p; ~F() { dev_free(p); } } fd{d};
    SPQ_HIP(hipMemcpy(d, b, sizeof(double) * n, hipMemcpyHostToDevice));
    e.solve_t_dev(d, d);
    SPQ_HIP(hipStreamSynchronize(e.stream()));
    e.check_solve();
    SPQ_HIP(hipMemcpy(x, d, sizeof(double) * n, hipMemcpyDeviceToHost));
  });
}
int spaqr_csc_outer_device(const spaqr_matrix* A, double alpha, const double* u, const double* v, double* out) {
  return guard([&] {
    if (!A || !A->d) throw Error(kErrArg, "null matrix");
    if (!u || !v || !out) throw Error(kErrArg, "null vector");
    const DevMatrix& d = *A->d;
    launch_csc_outer(nullptr, d.n, d.nnz, d.rpt, d.cit, alpha, u, v, out);
    SPQ_HIP(hipGetLastError());
    SPQ_HIP(hipStreamSynchronize(nullptr));
  });
}
int spaqr_cgls_device(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, double tol,
                      int verbose, int* iters_out, double* est_out) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!A || !A->d) throw Error(kErrArg, "null matrix");
    if (A->d->m != e.rows() || A->d->n != e.cols()) throw Error(kErrArg, "cgls needs the matrix of the tree");
    KrylovResult r = cgls_dev(e, *A->d, b, x, iters, tol, verbose);
    e.check_solve();
    if (iters_out) *iters_out = r.iters;
    if (est_out) *est_out = r.est;
  });
}
int spaqr_cgls(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, double tol, int verbose,
               int* iters_out, double* est_out) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!A || !A->d) throw Error(kErrArg, "null matrix");
    const int n = A->d->n, m = A->d->m;
    if (m != e.rows() || n != e.cols()) throw Error(kErrArg, "cgls needs the matrix of the tree");
    double *db = nullptr, *dx = nullptr;
    SPQ_HIP(dev_malloc((void**)&db, sizeof(double) * std::max(m, 1)));
    SPQ_HIP(dev_malloc((void**)&dx, sizeof(double) * std::max(n, 1)));
    struct F { double* p; ~F() { dev_free(p); } } fb{db}, fx{dx};
    SPQ_HIP(hipMemcpy(db, b, sizeof(double) * m, hipMemcpyHostToDevice));
    SPQ_HIP(hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice));
    KrylovResult r = cgls_dev(e, *A->d, db, dx, iters, tol, verbose);
    e.check_solve();
    SPQ_HIP(hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (iters_out) *iters_out = r.iters;
    if (est_out) *est_out = r.est;
  });
}
int spaqr_gmres(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x, int iters, int restart,
                double tol, int verbose, int* iters_out, double* est_out) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!A || !A->d) throw Error(kErrArg, "null matrix");
    const int n = A->d->n, m = A->d->m;
    double *db = nullptr, *dx = nullptr;
    SPQ_HIP(dev_malloc((void**)&db, sizeof(double) * std::max(m, 1)));
    SPQ_HIP(dev_malloc((void**)&dx, sizeof(double) * std::max(n, 1)));
    struct F { double* p; ~F() { dev_free(p); } } fb{db}, fx{dx};
    SPQ_HIP(hipMemcpy(db, b, sizeof(double) * m, hipMemcpyHostToDevice));
    SPQ_HIP(hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice));
    if (A->d->m != A->d->n || A->d->m != e.rows()) throw Error(kErrArg, "gmres needs the square matrix of the tree");
    KrylovResult r = e.precision_bytes() == 4 ? gmres_ir_dev(e, *A->d, db, dx, iters, restart, tol, verbose)
                                               : gmres_dev(e, *A->d, db, dx, iters, restart, tol, verbose);
    e.check_solve();
    SPQ_HIP(hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (iters_out) *iters_out = r.iters;
    if (est_out) *est_out = r.est;
  });
}

int spaqr_colscale(int m, int n, const int* colptr, const int* rowind, double* val, double* D) {
  return guard([&] {
    if (!colptr || !val) throw Error(kErrArg, "null arrays");
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int q = colptr[j]; q < colptr[j + 1]; ++q) s += val[q] * val[q];
      const double d = 1.0 / std::sqrt(s);
      if (D) D[j] = d;
      for (int q = colptr[j]; q < colptr[j + 1]; ++q) val[q] = (val[q] * d) * 10.0;
    }
    (void)m;
    (void)rowind;
  });
}
int spaqr_random(int n, int seed, double* out) {
  return guard([&] {
    std::mt19937 rng;
    rng.seed(seed);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    for (int i = 0; i < n; ++i) out[i] = dist(rng);
  });
}

int spaqr_perms(const spaqr_tree* t, int* cperm, int* rperm) {
  return guard([&] {
    const EngineBase& e = eng(t);
    std::memcpy(cperm, e.cperm().data(), sizeof(int) * e.cperm().size());
    std::memcpy(rperm, e.rperm().data(), sizeof(int) * e.rperm().size());
  });
}
int spaqr_level_stats(const spaqr_tree* t, double* out) {
  return guard([&] {
    const EngineBase& e = eng(t);
    int l = 0;
    for (auto& s : e.level_info()) {
      double v[8] = {s.t_fill, s.t_elmn, s.t_shift, s.t_scale, s.t_sprsfy, s.t_merge, (double)s.top_rows,
                     (double)s.top_cols};
      std::memcpy(out + 8 * l++, v, sizeof(v));
    }
  });
}
int spaqr_counters(const spaqr_tree* t, unsigned long long* nnz4, double* dbl8) {
  return guard([&] { eng(t).counters(nnz4, dbl8); });
}
int spaqr_num_sparsified(const spaqr_tree* t) { return t && t->e ? (int)t->e->sparsified().size() : SPAQR_E_ARG; }
int spaqr_sparsified(const spaqr_tree* t, int* out) {
  return guard([&] {
    int i = 0;
    for (auto& s : eng(t).sparsified()) {
      out[5 * i] = s.ilvl; out[5 * i + 1] = s.lvl; out[5 * i + 2] = s.sep;
      out[5 * i + 3] = s.before; out[5 * i + 4] = s.after;
      ++i;
    }
  });
}
int spaqr_num_extra_sparsified(const spaqr_tree* t) {
  return t && t->e ? (int)t->e->extra_sparsified().size() : SPAQR_E_ARG;
}
int spaqr_profile_json(spaqr_tree* t, char* buf, int buflen) {
  int n = 0;
  const int rc = guard([&] {
    const std::string s = eng(t).profile_json();
    n = (int)s.size();
    if (buf && buflen > 0) {
      const int w = std::min(n, buflen - 1);
      std::memcpy(buf, s.data(), (size_t)w);
      buf[w] = '\0';
    }
  });
  return rc != SPAQR_OK ? rc : n;
}
int spaqr_extra_sparsified(const spaqr_tree* t, int* out) {
  return guard([&] {
    int i = 0;
    for (auto& s : eng(t).extra_sparsified()) {
      out[5 * i] = s.ilvl; out[5 * i + 1] = s.lvl; out[5 * i + 2] = s.sep;
      out[5 * i + 3] = s.before; out[5 * i + 4] = s.after;
      ++i;
    }
  });
}
int spaqr_num_clusters(const spaqr_tree* t) { return t && t->e ? (int)t->e->cluster_table().size() : SPAQR_E_ARG; }
int spaqr_cluster_table(const spaqr_tree* t, int* out) {
  return guard([&] {
    int i = 0;
    for (auto& r : eng(t).cluster_table()) std::memcpy(out + 12 * i++, r.data(), sizeof(int) * 12);
  });
}

int spaqr_set_phase_sync(spaqr_tree* t, int v) { return guard([&] { eng(t).set_timing(v); }); }
int spaqr_kernel_timing(spaqr_tree* t, int on) { return guard([&] { eng(t).set_kernel_timing(on); }); }
int spaqr_kernel_stat_count(spaqr_tree* t) {
  int n = 0;
  int rc = guard([&] { n = (int)eng(t).kernel_stats().size(); });
  return rc == SPAQR_OK ? n : rc;
}
int spaqr_kernel_stat(spaqr_tree* t, int i, char* name, int namelen, double* out4) {
  return guard([&] {
    auto v = eng(t).kernel_stats();
    if (i < 0 || i >= (int)v.size()) throw Error(kErrArg, "kernel stat index out of range");
    if (name && namelen > 0) {
      std::strncpy(name, v[i].name.c_str(), namelen - 1);
      name[namelen - 1] = 0;
    }
    out4[0] = v[i].launches; out4[1] = v[i].ms; out4[2] = v[i].flops; out4[3] = v[i].bytes;
  });
}
int spaqr_kernel_stats_reset(spaqr_tree* t) { return guard([&] { eng(t).reset_kernel_stats(); }); }

}  // extern "C"
