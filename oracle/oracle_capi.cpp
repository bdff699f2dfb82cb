// ORACLE — test infrastructure only.  extern "C" surface of the CPU spaQR
// restatement, loaded by tests/ (ctypes), __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  The product never loads this library.
#include <cstring>
#include <stdexcept>
#include <string>

#include "oracle_spaqr.h"

using namespace orc;

namespace {
struct Handle {
  Tree* t = nullptr;
  Csc A;
};
Csc make_csc(int m, int n, const int* colptr, const int* rowind, const double* val) {
  Csc A;
  A.m = m; A.n = n;
  A.colptr.assign(colptr, colptr + n + 1);
  int nnz = colptr[n];
  A.rowind.assign(rowind, rowind + nnz);
  A.val.assign(val, val + nnz);
  return A;
}
Mat make_mat(int m, int n, const double* a) {
  Mat M(m, n);
  if (a) std::memcpy(M.data(), a, sizeof(double) * (size_t)m * n);
  return M;
}
thread_local std::string g_err;
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_use_openblas(const char* path) {
  std::string err;
  if (!use_openblas(path, &err)) { g_err = err; return -1; }
  return 0;
}
void orc_use_own() { use_own(); }
const char* orc_backend() { return backend_name(); }
void orc_set_threads(int n) { set_blas_threads(n); }

// ---------------------------------------------------------- dense kernels
void orc_geqrf(int m, int n, double* A, double* tau) {
  Mat M = make_mat(m, n, A);
  std::vector<double> t;
  geqrf(M, t);
  std::memcpy(A, M.data(), sizeof(double) * (size_t)m * n);
  std::memcpy(tau, t.data(), sizeof(double) * t.size());
}
// singular values of an m x n matrix, descending (min(m, n) of them)
void orc_singular_values(int m, int n, const double* A, double* S) {
  std::vector<double> s = singular_values(make_mat(m, n, A));
  std::memcpy(S, s.data(), sizeof(double) * s.size());
}
void orc_larft(int m, int k, const double* V, const double* tau, double* T) {
  Mat Vm = make_mat(m, k, V), Tm;
  std::vector<double> t(tau, tau + k);
  larft(Vm, t, Tm);
  std::memcpy(T, Tm.data(), sizeof(double) * (size_t)k * k);
}
void orc_larfb(int m, int n, int k, const double* V, const double* T, double* C, char trans) {
  Mat Vm = make_mat(m, k, V), Tm = make_mat(k, k, T), Cm = make_mat(m, n, C);
  larfb(Vm, Tm, Cm, trans);
  std::memcpy(C, Cm.data(), sizeof(double) * (size_t)m * n);
}
void orc_geqp3(int m, int n, double* A, int* jpvt, double* tau) {
  Mat M = make_mat(m, n, A);
  std::vector<int> p;
  std::vector<double> t;
  geqp3(M, p, t);
  std::memcpy(A, M.data(), sizeof(double) * (size_t)m * n);
  std::memcpy(jpvt, p.data(), sizeof(int) * n);
  std::memcpy(tau, t.data(), sizeof(double) * t.size());
}
void orc_ormqr_left(int m, int n, int k, const double* V, const double* tau, double* A, char trans) {
  Mat Vm = make_mat(m, k, V), Am = make_mat(m, n, A);
  std::vector<double> t(tau, tau + k);
  ormqr_left(Vm, t, Am, trans);
  std::memcpy(A, Am.data(), sizeof(double) * (size_t)m * n);
}
void orc_trsm_right_upper(int m, int n, const double* U, double* B) {
  Mat Um = make_mat(n, n, U), Bm = make_mat(m, n, B);
  trsm_right_upper(Um, Bm);
  std::memcpy(B, Bm.data(), sizeof(double) * (size_t)m * n);
}
void orc_trsv_upper(int n, const double* U, double* x, char trans) {
  Mat Um = make_mat(n, n, U);
  trsv_upper(Um, x, n, trans);
}
int orc_choose_rank(int n, const double* s, double tol) {
  return choose_rank(std::vector<double>(s, s + n), tol);
}

// ---------------------------------------------------------- driver helpers
void orc_colscale(int m, int n, const int* colptr, const int* rowind, double* val, double* D) {
  Csc A = make_csc(m, n, colptr, rowind, val);
  std::vector<double> d = colscale(A);
  std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
  if (D) std::memcpy(D, d.data(), sizeof(double) * n);
}
void orc_random(int n, int seed, double* out) {
  std::vector<double> r = random_vec(n, seed);
  std::memcpy(out, r.data(), sizeof(double) * n);
}

// ---------------------------------------------------------- Tree surface
void* orc_tree_new(int nlevels, int skip) {
  Handle* h = new Handle;
  h->t = new Tree(nlevels, skip);
  return h;
}
void orc_tree_free(void* p) {
  Handle* h = (Handle*)p;
  if (!h) return;
  delete h->t;
  delete h;
}
void orc_tree_set(void* p, double tol, int scale, float order, int square, int verbose) {
  Tree* t = ((Handle*)p)->t;
  t->set_tol(tol); t->set_scale(scale); t->set_order(order); t->set_square(square); t->set_verbose(verbose);
}
void orc_tree_set_tensor_coords(void* p, int n, int dim) {
  ((Handle*)p)->t->set_Xcoo(linspace_nd(n, dim), dim);
}
void orc_tree_set_row_parts(void* p, const int* parts, int m) {
  ((Handle*)p)->t->set_row_parts(std::vector<int>(parts, parts + m));
}
// Column labels of the non-geometric path from a row partition: 6 ints per
// column (self, left, right as lvl, sep).
int orc_hypergraph_labels(int m, int n, const int* colptr, const int* rowind, int nlevels, const int* parts,
                          int* out) {
  try {
    std::vector<double> ones(colptr[n], 1.0);
    Csc A = make_csc(m, n, colptr, rowind, ones.data());
    std::vector<ClusterID> cmap = hypergraph_partition(A, nlevels, std::vector<int>(parts, parts + m));
    for (int i = 0; i < n; ++i) {
      const ClusterID& c = cmap[i];
      int* o = out + 6 * i;
      o[0] = c.self.lvl; o[1] = c.self.sep; o[2] = c.l.lvl; o[3] = c.l.sep; o[4] = c.r.lvl; o[5] = c.r.sep;
    }
  } catch (std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}
int orc_partition(void* p, int m, int n, const int* colptr, const int* rowind, const double* val) {
  Handle* h = (Handle*)p;
  try {
    h->A = make_csc(m, n, colptr, rowind, val);
    h->t->partition(h->A);
  } catch (std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}
int orc_assemble(void* p) {
  Handle* h = (Handle*)p;
  try { h->t->assemble(h->A); } catch (std::exception& e) { g_err = e.what(); return -1; }
  return 0;
}
int orc_factorize(void* p) {
  try { return ((Handle*)p)->t->factorize(); } catch (std::exception& e) { g_err = e.what(); return -1; }
}
void orc_solve(void* p, const double* b, double* x) {
  Tree* t = ((Handle*)p)->t;
  std::vector<double> bb(b, b + t->rows()), xx;
  t->solve(bb, xx);
  std::memcpy(x, xx.data(), sizeof(double) * t->cols());
}
void orc_solve_nrml(void* p, const double* b, double* x) {
  Tree* t = ((Handle*)p)->t;
  std::vector<double> bb(b, b + t->cols()), xx;
  t->solve_nrml(bb, xx);
  std::memcpy(x, xx.data(), sizeof(double) * t->cols());
}
int orc_gmres(void* p, const double* b, double* x, int iters, int restart, double tol, double* tol_out, int verbose) {
  Handle* h = (Handle*)p;
  std::vector<double> bb(b, b + h->A.m), xx(x, x + h->A.n);
  int it = gmres(h->A, bb, xx, *h->t, iters, restart > 0 ? restart : iters, tol, tol_out, verbose);
  std::memcpy(x, xx.data(), sizeof(double) * h->A.n);
  return it;
}
int orc_cgls(void* p, const double* b, double* x, int iters, double tol, double* tol_out, int verbose) {
  Handle* h = (Handle*)p;
  std::vector<double> bb(b, b + h->A.m), xx(x, x + h->A.n);
  int it = cgls(h->A, bb, xx, *h->t, iters, tol, tol_out, verbose);
  std::memcpy(x, xx.data(), sizeof(double) * h->A.n);
  return it;
}

// ---------------------------------------------------------- inspection
void orc_perms(void* p, int* cperm, int* rperm) {
  Tree* t = ((Handle*)p)->t;
  std::memcpy(cperm, t->cperm.data(), sizeof(int) * t->cperm.size());
  std::memcpy(rperm, t->rperm.data(), sizeof(int) * t->rperm.size());
}
// per level: t_fill, t_elmn, t_shift, t_scale, t_sprsfy, t_merge, top_rows, top_cols
void orc_level_stats(void* p, double* out) {
  Tree* t = ((Handle*)p)->t;
  for (int l = 0; l < t->levels(); ++l) {
    const LevelStat& s = t->stats[l];
    double v[8] = {s.t_fill, s.t_elmn, s.t_shift, s.t_scale, s.t_sprsfy, s.t_merge, (double)s.top_rows,
                   (double)s.top_cols};
    std::memcpy(out + 8 * l, v, sizeof(v));
  }
}
// nnzA, nnzR, nnzH, nnzQ ; then flops, t_partition, t_assemble, t_factorize, top_size
void orc_counters(void* p, unsigned long long* nnz4, double* dbl5) {
  Tree* t = ((Handle*)p)->t;
  nnz4[0] = t->nnzA; nnz4[1] = t->nnzR; nnz4[2] = t->nnzH; nnz4[3] = t->nnzQ;
  dbl5[0] = t->flops; dbl5[1] = t->t_partition; dbl5[2] = t->t_assemble; dbl5[3] = t->t_factorize;
  dbl5[4] = t->top_size();
}
int orc_num_sparsified(void* p) { return (int)((Handle*)p)->t->sparsified.size(); }
void orc_sparsified(void* p, int* out) {
  auto& s = ((Handle*)p)->t->sparsified;
  for (size_t i = 0; i < s.size(); ++i) {
    out[5 * i] = s[i].ilvl; out[5 * i + 1] = s[i].lvl; out[5 * i + 2] = s[i].sep;
    out[5 * i + 3] = s[i].before; out[5 * i + 4] = s[i].after;
  }
}
int orc_num_extra_sparsified(void* p) { return (int)((Handle*)p)->t->extra_sparsified.size(); }
void orc_extra_sparsified(void* p, int* out) {
  auto& s = ((Handle*)p)->t->extra_sparsified;
  for (size_t i = 0; i < s.size(); ++i) {
    out[5 * i] = s[i].ilvl; out[5 * i + 1] = s[i].lvl; out[5 * i + 2] = s[i].sep;
    out[5 * i + 3] = s[i].before; out[5 * i + 4] = s[i].after;
  }
}
int orc_num_clusters(void* p) { return (int)((Handle*)p)->t->cluster_table().size(); }
void orc_cluster_table(void* p, int* out) {
  auto tab = ((Handle*)p)->t->cluster_table();
  for (size_t i = 0; i < tab.size(); ++i)
    for (int j = 0; j < 12; ++j) out[12 * i + j] = tab[i][j];
}

}  // extern "C"
