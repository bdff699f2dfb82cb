// ORACLE — test infrastructure only (see oracle_spaqr.h).
//
// CPU restatement of the reference spaQR, function by function.  Each
// function names the reference lines it follows.  Dense arithmetic goes
// through oracle_lapack (own LAPACK restatement or a dlopen'd OpenBLAS).
//
// Iteration over std::set<Cluster*> in the reference is by pointer value; the
// oracle orders those sets by creation sequence (uid), which is the address
// order glibc gives the reference's consecutive `new Cluster` calls.
#include "oracle_spaqr.h"

#include <algorithm>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <limits>
#include <random>

namespace orc {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool ByUid::operator()(const Cluster* a, const Cluster* b) const { return a->uid < b->uid; }

// --------------------------------------------------------------- sparse utils
Csc csc_transpose(const Csc& A) {
  Csc T;
  T.m = A.n; T.n = A.m;
  T.colptr.assign(A.m + 1, 0);
  for (int r : A.rowind) T.colptr[r + 1]++;
  for (int i = 0; i < A.m; ++i) T.colptr[i + 1] += T.colptr[i];
  T.rowind.resize(A.nnz());
  T.val.resize(A.nnz());
  std::vector<int> pos(T.colptr.begin(), T.colptr.end() - 1);
  for (int j = 0; j < A.n; ++j)
    for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
      int q = pos[A.rowind[p]]++;
      T.rowind[q] = j;
      T.val[q] = A.val[p];
    }
  return T;
}

Csc csc_colperm(const Csc& A, const std::vector<int>& p) {
  // util.cpp:912-940: column k of the result is column p[k] of A.
  Csc B;
  B.m = A.m; B.n = A.n;
  B.colptr.assign(A.n + 1, 0);
  for (int k = 0; k < A.n; ++k) B.colptr[k + 1] = B.colptr[k] + (A.colptr[p[k] + 1] - A.colptr[p[k]]);
  B.rowind.resize(A.nnz());
  B.val.resize(A.nnz());
  for (int k = 0; k < A.n; ++k) {
    int q = B.colptr[k];
    for (int s = A.colptr[p[k]]; s < A.colptr[p[k] + 1]; ++s, ++q) {
      B.rowind[q] = A.rowind[s];
      B.val[q] = A.val[s];
    }
  }
  return B;
}

static Csc csc_rowperm(const Csc& A, const std::vector<int>& p) {
  // Row i of the result is row p[i] of A (rows re-sorted per column).
  std::vector<int> pinv(A.m);
  for (int i = 0; i < A.m; ++i) pinv[p[i]] = i;
  Csc B = A;
  std::vector<std::pair<int, double>> tmp;
  for (int j = 0; j < A.n; ++j) {
    tmp.clear();
    for (int s = A.colptr[j]; s < A.colptr[j + 1]; ++s) tmp.push_back({pinv[A.rowind[s]], A.val[s]});
    std::sort(tmp.begin(), tmp.end());
    int q = A.colptr[j];
    for (auto& t : tmp) { B.rowind[q] = t.first; B.val[q] = t.second; ++q; }
  }
  return B;
}

Csc csc_ata_pattern(const Csc& A) {
  Csc At = csc_transpose(A);
  Csc G;
  G.m = A.n; G.n = A.n;
  G.colptr.assign(A.n + 1, 0);
  std::vector<int> mark(A.n, -1), cols;
  for (int j = 0; j < A.n; ++j) {
    cols.clear();
    for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) {
      int i = A.rowind[p];
      for (int q = At.colptr[i]; q < At.colptr[i + 1]; ++q) {
        int jj = At.rowind[q];
        if (mark[jj] != j) { mark[jj] = j; cols.push_back(jj); }
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int c : cols) { G.rowind.push_back(c); G.val.push_back(1.0); }
    G.colptr[j + 1] = (int)G.rowind.size();
  }
  return G;
}

void csc_matvec(const Csc& A, const double* x, double* y) {
  std::fill(y, y + A.m, 0.0);
  for (int j = 0; j < A.n; ++j) {
    double xj = x[j];
    for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) y[A.rowind[p]] += A.val[p] * xj;
  }
}
void csc_matvec_t(const Csc& A, const double* x, double* y) {
  for (int j = 0; j < A.n; ++j) {
    double s = 0.0;
    for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) s += A.val[p] * x[A.rowind[p]];
    y[j] = s;
  }
}

std::vector<double> colscale(Csc& A) {
  std::vector<double> D(A.n);
  for (int j = 0; j < A.n; ++j) {
    double s = 0.0;
    for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) s += A.val[p] * A.val[p];
    D[j] = 1.0 / std::sqrt(s);
  }
  for (int j = 0; j < A.n; ++j)
    for (int p = A.colptr[j]; p < A.colptr[j + 1]; ++p) A.val[p] = (A.val[p] * D[j]) * 10.0;
  return D;
}

std::vector<double> linspace_nd(int n, int dim) {
  long N = 1;
  for (int d = 0; d < dim; ++d) N *= n;
  std::vector<double> X((size_t)dim * N, 0.0);
  long id = 0;
  if (dim == 1) for (int i = 0; i < n; ++i) X[i] = i;
  else if (dim == 2) for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j, ++id) { X[id * 2] = i; X[id * 2 + 1] = j; }
  else if (dim == 3)
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) for (int k = 0; k < n; ++k, ++id) {
      X[id * 3] = i; X[id * 3 + 1] = j; X[id * 3 + 2] = k;
    }
  return X;
}

std::vector<double> random_vec(int size, int seed) {
  std::mt19937 rng;
  rng.seed(seed);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  std::vector<double> x(size);
  for (int i = 0; i < size; ++i) x[i] = dist(rng);
  return x;
}

// ------------------------------------------------------------------ IDs
static SepID merge_sep(const SepID& s) { return SepID(s.lvl + 1, s.sep / 2); }  // cluster.cpp:19-21
ClusterID merge_if(const ClusterID& c, int lvl) {                               // cluster.cpp:23-32
  ClusterID o;
  o.self = c.self;
  o.l = c.l.lvl < lvl ? merge_sep(c.l) : c.l;
  o.r = c.r.lvl < lvl ? merge_sep(c.r) : c.r;
  o.section = lvl > 0 ? c.section / 2 : c.section;
  o.part = c.self.lvl == lvl ? 2 : c.part;
  return o;
}

// ------------------------------------------------------------------ partition
static SepID find_highest_common(SepID n1, SepID n2) {  // partition.cpp:63-86
  int l1 = n1.lvl, l2 = n2.lvl, s1 = n1.sep, s2 = n2.sep;
  while (l1 < l2) { l1++; s1 /= 2; }
  while (l2 < l1) { l2++; s2 /= 2; }
  while (s1 != s2) { l1++; l2++; s1 /= 2; s2 /= 2; }
  return SepID(l1, s1);
}
static SepID find_lowest_common(SepID n1, SepID n2) { return n1.lvl < n2.lvl ? n1 : n2; }  // :88-100

static void bissect_geo(const std::vector<int>& dofs, std::vector<int>& parts, const std::vector<double>& X,
                        int dim, int start) {  // partition.cpp:241-279
  int N = (int)dofs.size();
  if (N == 0) return;
  int bestdim = -1;
  double maxrange = -1.0;
  for (int d = 0; d < dim; ++d) {
    double mx = std::numeric_limits<double>::lowest(), mn = std::numeric_limits<double>::max();
    for (int i : dofs) { mx = std::max(mx, X[(size_t)i * dim + d]); mn = std::min(mn, X[(size_t)i * dim + d]); }
    if (mx - mn > maxrange) { bestdim = d; maxrange = mx - mn; }
  }
  // Only the value at sorted position N/2 is used (:267-269).
  std::vector<double> v(N);
  for (int i = 0; i < N; ++i) v[i] = X[(size_t)dofs[i] * dim + bestdim];
  std::nth_element(v.begin(), v.begin() + N / 2, v.end());
  double midv = v[N / 2];
  for (int j : dofs) parts[j] = (X[(size_t)j * dim + bestdim] <= midv) ? start : start + 1;
}

// form_rmap #1 (tree.cpp:57-98): rows to clusters by max squared group norm.
static void form_rmap1(const Csc& A, std::vector<ClusterID>& rmap, const std::vector<ClusterID>& cpermed,
                       const std::vector<int>& cperm) {
  Csc At = csc_transpose(csc_colperm(A, cperm));
  ClusterID last = cpermed[A.n - 1];
  for (int i = 0; i < A.m; ++i) {
    if (rmap[i] != ClusterID()) continue;
    double maxv = -1;
    ClusterID rid = rmap[i];
    int p = At.colptr[i], e = At.colptr[i + 1];
    while (p < e) {
      ClusterID cid = cpermed[At.rowind[p]];
      double val = At.val[p] * At.val[p];
      int q = p + 1;
      while (q < e && cid == cpermed[At.rowind[q]]) { val += At.val[q] * At.val[q]; ++q; }
      if (val > maxv) { maxv = val; rid = cid; }
      p = q;
    }
    if (rid == ClusterID()) rid = last;
    rmap[i] = rid;
  }
}

// getInterfacesUsingA (partition.cpp:506-578) with the no-HSL rmap2cmap
// (:149-156: cmap = rmap).
static void get_interfaces_using_a(const Csc& A, std::vector<ClusterID>& cmap, std::vector<ClusterID>& rmap) {
  int nrows = A.m;
  Csc At = csc_transpose(A);
  for (int i = 0; i < nrows; ++i) {
    SepID self = rmap[i].self;
    if (self.lvl > 0) {
      SepID left(-1, 0), right(-1, 0);
      for (int p = At.colptr[i]; p < At.colptr[i + 1]; ++p) {
        int j = At.rowind[p];
        SepID nbr = cmap[j].self;
        if (nbr.lvl < self.lvl) {
          int pj = nbr.sep;
          pj = (int)(pj / std::pow(2.0, self.lvl - nbr.lvl - 1));
          if (pj % 2 == 0) left = (left.lvl == -1) ? nbr : find_lowest_common(nbr, left);
          else right = (right.lvl == -1) ? nbr : find_lowest_common(nbr, right);
        }
      }
      rmap[i].l = left;
      rmap[i].r = right;
    }
  }
  cmap = rmap;
  for (int i = 0; i < nrows; ++i) {
    SepID self = rmap[i].self, left = rmap[i].l, right = rmap[i].r;
    if (right == SepID()) {
      for (int p = At.colptr[i]; p < At.colptr[i + 1]; ++p) {
        int j = At.rowind[p];
        if (self == cmap[j].self && cmap[j].r != SepID())
          right = (right == SepID()) ? cmap[j].r : find_highest_common(right, cmap[j].r);
      }
    } else if (left == SepID()) {
      for (int p = At.colptr[i]; p < At.colptr[i + 1]; ++p) {
        int j = At.rowind[p];
        if (self == cmap[j].self && cmap[j].l != SepID())
          left = (left == SepID()) ? cmap[j].l : find_highest_common(left, cmap[j].l);
      }
    }
    rmap[i].l = left;
    rmap[i].r = right;
  }
  for (int i = 0; i < nrows; ++i) {
    if (rmap[i].l == SepID()) rmap[i].l = rmap[i].self;
    if (rmap[i].r == SepID()) rmap[i].r = rmap[i].self;
  }
  cmap = rmap;
}

std::vector<ClusterID> geometric_partition_ata(const Csc& A, int nlevels, const std::vector<double>& X, int dim) {
  // partition.cpp:282-348
  Csc AtA = csc_ata_pattern(A);
  int c = AtA.n;
  std::vector<int> parts(c, 0);
  for (int depth = 0; depth < nlevels - 1; ++depth) {
    std::vector<std::vector<int>> dofs(1 << depth);
    for (int i = 0; i < c; ++i) dofs[parts[i]].push_back(i);
    for (int k = 0; k < (1 << depth); ++k) bissect_geo(dofs[k], parts, X, dim, 2 * k);
  }
  std::vector<ClusterID> cmap(c);
  for (int depth = 0; depth < nlevels - 1; ++depth) {
    int l = nlevels - depth - 1;
    for (int i = 0; i < c; ++i) {
      if (cmap[i].self.lvl != -1) continue;
      int pi = parts[i] >> (l - 1);
      if (pi % 2 != 0) continue;
      bool sep = false;
      for (int p = AtA.colptr[i]; p < AtA.colptr[i + 1]; ++p) {
        int j = AtA.rowind[p];
        if (cmap[j].self.lvl == -1 && (parts[j] >> (l - 1)) == pi + 1) { sep = true; break; }
      }
      if (sep) cmap[i].self = SepID(l, pi / 2);
    }
  }
  for (int i = 0; i < c; ++i)
    if (cmap[i].self.lvl == -1) {
      cmap[i].self = SepID(0, parts[i]);
      cmap[i].l = SepID(0, parts[i]);
      cmap[i].r = SepID(0, parts[i]);
    }
  std::vector<ClusterID> rmap(A.m);
  if (A.m != A.n) {  // rows2clusters (partition.cpp:103-134), no HSL
    std::vector<int> cperm(A.n);
    for (int i = 0; i < A.n; ++i) cperm[i] = i;
    std::stable_sort(cperm.begin(), cperm.end(), [&](int i, int j) { return cmap[i] < cmap[j]; });
    std::vector<ClusterID> cpermed(A.n);
    for (int i = 0; i < A.n; ++i) cpermed[i] = cmap[cperm[i]];
    form_rmap1(A, rmap, cpermed, cperm);
  } else {
    rmap = cmap;
  }
  get_interfaces_using_a(A, cmap, rmap);
  return cmap;
}

static int part_at_lvl(int part_leaf, int lvl) {  // partition.cpp:54-61
  for (int l = 0; l < lvl - 1; ++l) part_leaf /= 2;
  return part_leaf;
}

// partition.cpp:582-623 (getInterfacesUsingA_HUND)
static void get_interfaces_hund(const Csc& A, std::vector<ClusterID>& cmap, const std::vector<int>& parts) {
  Csc At = csc_transpose(A);
  for (int i = 0; i < A.n; ++i) {
    SepID self = cmap[i].self;
    if (self.lvl <= 0) continue;
    SepID left(-1, 0), right(-1, 0);
    for (int p = A.colptr[i]; p < A.colptr[i + 1]; ++p) {
      int j = A.rowind[p];
      int pj = part_at_lvl(parts[j], self.lvl);
      for (int q = At.colptr[j]; q < At.colptr[j + 1]; ++q) {
        SepID nbr = cmap[At.rowind[q]].self;
        if (nbr.lvl >= self.lvl) continue;
        if (pj % 2 == 0) left = left.lvl == -1 ? nbr : find_highest_common(nbr, left);
        else right = right.lvl == -1 ? nbr : find_highest_common(nbr, right);
      }
    }
    cmap[i].l = left;
    cmap[i].r = right;
  }
  for (auto& c : cmap) {
    if (c.l == SepID()) c.l = c.self;
    if (c.r == SepID()) c.r = c.self;
  }
}

// partition.cpp:450-499 on a given row partition (partition_patoh's output,
// partition.cpp:422-448: PaToH itself is a binary-only library)
std::vector<ClusterID> hypergraph_partition(const Csc& A, int nlevels, const std::vector<int>& parts) {
  if ((int)parts.size() != A.m) throw std::runtime_error("oracle: row partition of the wrong length");
  std::vector<ClusterID> cmap(A.n);
  for (int depth = 0; depth < nlevels - 1; ++depth) {
    int l = nlevels - depth - 1;
    for (int i = 0; i < A.n; ++i) {
      if (cmap[i].self.lvl != -1) continue;
      bool sep = false;
      int pi = -1;
      for (int p = A.colptr[i]; p < A.colptr[i + 1]; ++p) {
        int pj = part_at_lvl(parts[A.rowind[p]], l);
        if (pi == -1) pi = pj;
        else if (pi != pj) { sep = true; break; }
      }
      if (sep) cmap[i].self = SepID(l, pi / 2);
    }
  }
  for (int i = 0; i < A.n; ++i) {
    if (cmap[i].self.lvl != -1) continue;
    if (A.colptr[i] == A.colptr[i + 1]) throw std::runtime_error("oracle: empty column");
    int j = A.rowind[A.colptr[i]];
    cmap[i].self = SepID(0, parts[j]);
    cmap[i].l = SepID(0, parts[j]);
    cmap[i].r = SepID(0, parts[j]);
  }
  get_interfaces_hund(A, cmap, parts);
  return cmap;
}

void init_permutation(int nlevels, const std::vector<ClusterID>& cmap, std::vector<int>& cperm) {
  // tree.cpp:46-55
  std::vector<ClusterID> pm = cmap;
  auto cmp = [&pm](int i, int j) { return pm[i] < pm[j]; };
  std::stable_sort(cperm.begin(), cperm.end(), cmp);
  for (int lvl = 1; lvl < nlevels; ++lvl) {
    for (auto& s : pm) s = merge_if(s, lvl);
    std::stable_sort(cperm.begin(), cperm.end(), cmp);
  }
}

// form_rmap #2 (tree.cpp:100-163), matched = false.
static void form_rmap2(const Csc& Ap, const std::vector<ClusterID>& cpermed, std::vector<int>& rperm,
                       std::vector<int>& c2r_count) {
  int r = Ap.m, c = Ap.n;
  Csc At = csc_transpose(Ap);
  std::vector<std::vector<int>> c2r(c);
  for (int i = 0; i < r; ++i) {
    double maxv = -1;
    int cidx = -1;
    int p = At.colptr[i], e = At.colptr[i + 1];
    while (p < e) {
      ClusterID cid = cpermed[At.rowind[p]];
      double val = At.val[p] * At.val[p];
      int q = p + 1;
      while (q < e && cid == cpermed[At.rowind[q]]) { val += At.val[q] * At.val[q]; ++q; }
      if (val > maxv) { maxv = val; cidx = At.rowind[p]; }
      p = q;
    }
    if (cidx == -1) cidx = c - 1;
    c2r_count[cidx]++;
    c2r[cidx].push_back(i);
  }
  int start = 0;
  for (int i = 0; i < c; ++i)
    for (int j = 0; j < c2r_count[i]; ++j) rperm[start++] = c2r[i][j];
}

// ------------------------------------------------------------------ Tree
void Cluster::sort_edgesOut(bool reverse) {  // cluster.cpp:58-65
  if (reverse) edgesOut.sort([](Edge* a, Edge* b) { return a->n2->order > b->n2->order; });
  else edgesOut.sort([](Edge* a, Edge* b) { return a->n2->order < b->n2->order; });
}

Tree::Tree(int lvls, int skip_) : nlevels(lvls), skip(skip_) {
  bottoms.resize(nlevels);
  stats.resize(nlevels);
}
Tree::~Tree() {}

Cluster* Tree::new_cluster(int cs, int cn, int rs, int rn, ClusterID id) {
  owned.emplace_back(new Cluster(uid_counter++, cs, cn, rs, rn, id, get_new_order()));
  return owned.back().get();
}
Edge* Tree::new_edge(Cluster* a, Cluster* b, Mat* m) {
  edges_owned.emplace_back(new Edge(a, b, m));
  return edges_owned.back().get();
}

std::pair<int, int> Tree::topsize() const {  // tree.cpp:27-37
  int c = 0, r = 0;
  for (auto s : bottom_current())
    if (s->level() == nlevels - 1) { c += s->cols(); r += s->rows(); }
  return {r, c};
}
int Tree::top_size() const { return (*bottoms[nlevels - 1].begin())->cols(); }

std::vector<std::vector<int>> Tree::cluster_table() const {
  std::vector<std::vector<int>> t;
  for (int l = 0; l < nlevels; ++l)
    for (auto s : bottoms[l])
      t.push_back({l, s->id.self.lvl, s->id.self.sep, s->id.l.lvl, s->id.l.sep, s->id.r.lvl, s->id.r.sep,
                   s->cstart, s->csize_org, s->rstart, s->rsize_org, s->order});
  return t;
}

void Tree::partition(const Csc& A) {  // tree.cpp:595-723
  double t0 = now_s();
  int r = A.m, c = A.n;
  nrows = r; ncols = c;
  std::vector<ClusterID> cmap;
  if (geo) cmap = geometric_partition_ata(A, nlevels, Xcoo, xdim);
  else if (!row_parts.empty()) cmap = hypergraph_partition(A, nlevels, row_parts);
  else throw std::runtime_error("oracle: no coordinates and no row partition");
  rperm.resize(r); cperm.resize(c);
  for (int i = 0; i < r; ++i) rperm[i] = i;
  for (int i = 0; i < c; ++i) cperm[i] = i;
  init_permutation(nlevels, cmap, cperm);
  std::vector<ClusterID> cpermed(c);
  for (int k = 0; k < c; ++k) cpermed[k] = cmap[cperm[k]];
  std::vector<int> c2r_count(c, 0);
  if (!square) {
    Csc Ap = csc_colperm(A, cperm);
    form_rmap2(Ap, cpermed, rperm, c2r_count);
  } else {
    std::fill(c2r_count.begin(), c2r_count.end(), 1);
    rperm = cperm;
  }
  int l = 0;
  for (int k = 0; k < c;) {
    int kn = k + 1;
    ClusterID cid = cpermed[k];
    int nr2c = c2r_count[k];
    while (kn < c && cid == cpermed[kn]) { nr2c += c2r_count[kn]; kn++; }
    if (nr2c == 0) throw std::runtime_error("oracle: cluster without rows");
    bottoms[0].push_back(new_cluster(k, kn - k, l, nr2c, cid));
    k = kn;
    l += nr2c;
  }
  for (int lvl = 1; lvl < nlevels; ++lvl) {
    auto begin = std::find_if(bottoms[lvl - 1].begin(), bottoms[lvl - 1].end(),
                              [lvl](const Cluster* s) { return s->level() >= lvl; });
    auto end = bottoms[lvl - 1].end();
    for (auto it = begin; it != end; ++it) (*it)->parentid = merge_if((*it)->id, lvl);
    for (auto k = begin; k != end;) {
      ClusterID idp = (*k)->parentid;
      std::vector<Cluster*> ch{*k};
      int cs = (*k)->cols(), rs = (*k)->rows(), c0 = (*k)->cstart, r0 = (*k)->rstart;
      ++k;
      while (k != end && idp == (*k)->parentid) { ch.push_back(*k); cs += (*k)->cols(); rs += (*k)->rows(); ++k; }
      Cluster* p = new_cluster(c0, cs, r0, rs, idp);
      for (auto cc : ch) { cc->parent = p; p->children.push_back(cc); }
      bottoms[lvl].push_back(p);
    }
  }
  t_partition = now_s() - t0;
  if (verbose && verbose < 3) printf("Time to partition: %g\n", t_partition);
}

void Tree::assemble(const Csc& A) {  // tree.cpp:725-778
  double t0 = now_s();
  nnzA = A.nnz();
  Csc App = csc_rowperm(csc_colperm(A, cperm), rperm);
  std::vector<Cluster*> rmap(nrows);
  for (auto s : bottom_original())
    for (int i = s->rstart; i < s->rstart + s->rows(); ++i) rmap[i] = s;
  for (auto self : bottom_original()) {
    CSet cn;
    for (int j = self->cstart; j < self->cstart + self->cols(); ++j)
      for (int p = App.colptr[j]; p < App.colptr[j + 1]; ++p) cn.insert(rmap[App.rowind[p]]);
    cn.insert(self);
    for (auto nbr : cn) {
      Mat* sA = new Mat(nbr->rows(), self->cols());
      for (int col = 0; col < self->cols(); ++col) {  // block2dense (util.cpp:818-850)
        int j = self->cstart + col;
        for (int p = App.colptr[j]; p < App.colptr[j + 1]; ++p) {
          int row = App.rowind[p];
          if (row >= nbr->rstart && row < nbr->rstart + nbr->rows()) (*sA)(row - nbr->rstart, col) = App.val[p];
        }
      }
      Edge* e = new_edge(self, nbr, sA);
      self->edgesOut.push_back(e);
      if (self != nbr) nbr->edgesIn.push_back(e);
    }
    self->sort_edgesOut();
    self->cnbrs = cn;
  }
  t_assemble = now_s() - t0;
  if (verbose && verbose < 3) {
    printf("Time to assemble: %g\n", t_assemble);
    auto ts = topsize();
    printf("Aspect ratio of top separator: %g\n", (double)ts.first / (double)ts.second);
  }
}

bool Tree::check_if_valid(ClusterID cid, ClusterID nid, const std::set<SepID>& cs) const {  // tree.cpp:168-224
  SepID cself = cid.self, nself = nid.self;
  if (cself.lvl > nself.lvl) return false;
  int d = nself.lvl - cself.lvl;
  int csep = cself.sep / (int)std::pow(2, d);
  if (csep != nself.sep) return false;
  SepID nl = nid.l, nr = nid.r;
  int dl = nl.lvl - cself.lvl, dr = nr.lvl - cself.lvl;
  bool ly, ry;
  if (nl.lvl >= cself.lvl) ly = (cself.sep / (int)std::pow(2, dl)) == nl.sep;
  else ly = (nl.sep / (int)std::pow(2, -dl)) == cself.sep;
  if (nr.lvl >= cself.lvl) ry = (cself.sep / (int)std::pow(2, dr)) == nr.sep;
  else ry = (nr.sep / (int)std::pow(2, -dr)) == cself.sep;
  if (!ly && !ry) return false;
  if (geo) {
    if (!cs.count(nself)) return false;
    if (!cs.count(nl) && !cs.count(nr)) return false;
  }
  return true;
}

Edge* Tree::add_edge(Cluster* c1, Cluster* c2) {  // tree.cpp:251-261
  Edge* e = new_edge(c1, c2, new Mat(c2->rows(), c1->cols()));
  c1->cnbrs.insert(c2);
  c1->edgesOut.push_back(e);
  if (c1 != c2) c2->edgesIn.push_back(e);
  return e;
}

// |A^T B|_F for blocks of one row space (A: r x p, B: r x q)
static double norm_atb(const Mat& A, const Mat& B) {
  double s = 0;
  for (int i = 0; i < A.cols(); ++i)
    for (int j = 0; j < B.cols(); ++j) {
      double d = 0;
      for (int r = 0; r < A.rows(); ++r) d += A(r, i) * B(r, j);
      s += d * d;
    }
  return std::sqrt(s);
}

void Tree::get_sparsity(Cluster* c) {  // tree.cpp:834-870
  CSet rs;
  std::set<SepID> cs;
  for (auto e : c->edgesOut) cs.insert(e->n2->id.self);
  for (auto e : c->edgesOut) {
    if (!e->A21) continue;
    Cluster* n = e->n2;
    rs.insert(n);
    for (auto ein : n->edgesIn) {
      if (ein->A21 && ein->n1 != c && !ein->n1->eliminated) {
        Cluster* nin = ein->n1;
        bool valid = check_if_valid(c->id, nin->id, cs);
        // no coordinates: a cluster that does not name c as an interface
        // must couple to it numerically through n (tree.cpp:855-858)
        if (!geo && valid && !(nin->id.l == c->id.self || nin->id.r == c->id.self))
          valid = norm_atb(*e->A21, *ein->A21) > 1e-14;
        if (valid) rs.insert(nin);
      }
    }
  }
  c->rsparsity = rs;
}

Op Tree::make_qr(Cluster* c) const {  // operations.h:33-45
  Op o;
  o.kind = Op::QR;
  o.c = c;
  o.ncols = c->cols();
  o.ccols = c->cols();
  for (auto e : c->edgesOut)
    if (e->A21) { o.a21_sizes.push_back(e->n2->rows()); o.nrows += e->A21->rows(); }
  for (auto e : c->edgesIn) o.a12_sizes.push_back(e->A21->cols());
  return o;
}

void Tree::update_cluster(Cluster* c, Cluster* n, const Mat& Q) {  // tree.cpp:280-316
  std::vector<Mat*> N;
  Mat V(Q.rows(), n->cols());
  int cr = 0;
  for (auto edge : c->edgesOut) {
    if (!edge->A21) continue;
    Cluster* nb = edge->n2;
    auto f = std::find_if(n->edgesOut.begin(), n->edgesOut.end(), [nb](Edge* e) { return e->n2 == nb; });
    Edge* ecn = (f == n->edgesOut.end()) ? add_edge(n, nb) : *f;
    N.push_back(ecn->A21.get());
    V.setBlock(cr, 0, *ecn->A21);
    cr += ecn->A21->rows();
  }
  larfb(Q, c->T, V, 'T');
  int k = c->T.cols();
  flops += 4.0 * V.rows() * V.cols() * k - (double)V.cols() * k * k;
  cr = 0;
  for (auto m : N) { *m = V.middleRows(cr, m->rows()); cr += m->rows(); }
}

void Tree::eliminate_cluster(Cluster* c) {  // tree.cpp:780-832 (+ householder 263-277)
  if (c->cols() == 0) { c->eliminated = true; return; }
  std::vector<Mat*> H;
  int nr = 0, nc = 0;
  for (auto e : c->edgesOut)
    if (e->A21) { H.push_back(e->A21.get()); nr += e->A21->rows(); nc = e->A21->cols(); }
  Mat Q(nr, nc);
  int cr = 0;
  for (auto h : H) { Q.setBlock(cr, 0, *h); cr += h->rows(); }
  int k = std::min(nr, nc);
  geqrf(Q, c->tau);
  flops += 4.0 * nr * nc * k - 2.0 * (nr + nc) * k * k + 4.0 / 3.0 * k * k * k;
  if (square) {
    cr = 0;
    for (auto h : H) { *h = Q.middleRows(cr, h->rows()); cr += h->rows(); }
  } else {
    *H[0] = Q.topRows(H[0]->cols());
  }
  larft(Q, c->tau, c->T);
  flops += (double)nr * k * k - (double)k * k * k / 3.0;
  nnzR += (unsigned long long)(0.5) * c->cols() * c->cols();  // tree.cpp:805 adds 0
  nnzH += (unsigned long long)(nr * nc);
  for (auto n : c->rsparsity) {
    if (!n->eliminated && c != n) {
      update_cluster(c, n, Q);
      nnzR += (unsigned long long)(c->cols() * n->cols());
    }
  }
  if (!square)
    for (auto e : c->edgesOut)
      if (e->n2 != c) e->A21.reset();
  ops.push_back(make_qr(c));
  if (!square) tops.push_back(make_qr(c));
  c->eliminated = true;
}

void Tree::reassign_rows(Cluster* c) {  // tree.cpp:319-411
  if (c->rows() == c->cols()) return;
  int rsize = c->rows() - c->cols();
  std::vector<std::pair<Cluster*, double>> r2c(rsize, {c, 0.0});
  for (auto edge : c->edgesIn) {
    Cluster* n = edge->n1;
    bool is_cnbr = std::find_if(c->edgesOut.begin(), c->edgesOut.end(), [n](Edge* e) { return e->n2 == n; }) !=
                   c->edgesOut.end();
    if (!is_cnbr) continue;
    for (int i = 0; i < rsize; ++i) {
      double s = 0.0;
      for (int j = 0; j < edge->A21->cols(); ++j) s += (*edge->A21)(c->cols() + i, j) * (*edge->A21)(c->cols() + i, j);
      double rn = std::sqrt(s);
      if (rn > r2c[i].second) r2c[i] = {n, rn};
    }
  }
  std::vector<int> rp(rsize);
  for (int i = 0; i < rsize; ++i) rp[i] = i;
  std::stable_sort(rp.begin(), rp.end(), [&](int i, int j) { return r2c[i].first->order < r2c[j].first->order; });
  for (auto edge : c->edgesIn) {
    Cluster* nold = edge->n1;
    // A12t = P^T * A21.middleRows(cols, rsize): row i of A12t is row rp[i].
    Mat A12t(rsize, edge->A21->cols());
    for (int i = 0; i < rsize; ++i)
      for (int j = 0; j < edge->A21->cols(); ++j) A12t(i, j) = (*edge->A21)(c->cols() + rp[i], j);
    for (int i = 0; i < rsize;) {
      Cluster* nnew = r2c[rp[i]].first;
      int inext = i + 1;
      while (inext < rsize && nnew == r2c[rp[inext]].first) inext++;
      int rinc = inext - i;
      auto f = std::find_if(nold->edgesOut.begin(), nold->edgesOut.end(), [nnew](Edge* e) { return e->n2 == nnew; });
      Edge* e = nullptr;
      if (nold == nnew) {
        e = *f;
      } else {
        e = (f == nold->edgesOut.end()) ? add_edge(nold, nnew) : *f;
      }
      int old_r = e->A21->rows();
      e->A21->conservativeResize(old_r + rinc, e->A21->cols());
      for (int t = 0; t < rinc; ++t)
        for (int j = 0; j < e->A21->cols(); ++j) (*e->A21)(old_r + t, j) = A12t(i + t, j);
      if (nold == nnew) {
        nnew->resize_x(rinc);
        Op o;
        o.kind = Op::Reassign;
        o.c = c; o.n = nnew; o.nstart = nnew->rows(); o.nrows = rinc;
        for (int t = 0; t < rinc; ++t) o.indices.push_back(c->cols() + rp[i + t]);
        ops.push_back(o);
        if (!square) tops.push_back(o);
      }
      i = inext;
    }
  }
  for (auto edge : c->edgesOut) {
    if (edge->n2 == c) continue;
    Cluster* n = edge->n2;
    auto f = std::find_if(n->edgesOut.begin(), n->edgesOut.end(), [n](Edge* e) { return e->n2 == n; });
    Mat* ms = (*f)->A21.get();
    n->reset_size(ms->rows(), ms->cols());
    for (auto e : n->edgesIn) {
      if (!e->n1->eliminated && e->A21 && e->A21->rows() != n->rows()) {
        e->A21->conservativeResize(n->rows(), e->A21->cols());
      }
    }
  }
}

bool Tree::want_sparsify(Cluster* c) const {  // tree.cpp:583-591
  return c->level() > ilvl && c->id.l.lvl <= ilvl && c->id.r.lvl <= ilvl;
}

void Tree::diagScale(Cluster* c) {  // tree.cpp:533-580
  auto f = std::find_if(c->edgesOut.begin(), c->edgesOut.end(), [c](Edge* e) { return e->n2 == c; });
  Edge* e = *f;
  auto Q = std::make_shared<Mat>(*e->A21);
  auto t = std::make_shared<std::vector<double>>();
  geqrf(*Q, *t);
  int rr = Q->rows(), cc = c->cols(), k = std::min(rr, cc);
  flops += 4.0 * rr * cc * k - 2.0 * (rr + cc) * k * k + 4.0 / 3.0 * k * k * k;
  Mat T;
  larft(*Q, *t, T);
  flops += (double)rr * k * k - (double)k * k * k / 3.0;
  auto R = std::make_shared<Mat>(cc, cc);
  for (int j = 0; j < cc; ++j)
    for (int i = 0; i <= j; ++i) (*R)(i, j) = (*Q)(i, j);
  e->A21->setZero();
  for (int i = 0; i < cc; ++i) (*e->A21)(i, i) = 1.0;
  nnzR = (unsigned long long)((double)nnzR + 0.5 * c->cols() * c->cols());
  nnzH += (unsigned long long)(c->rows() * c->cols());
  for (auto ei : c->edgesIn)
    if (!ei->n1->eliminated) {
      larfb(*Q, T, *ei->A21, 'T');
      flops += 4.0 * rr * ei->A21->cols() * k - (double)ei->A21->cols() * k * k;
    }
  for (auto eo : c->edgesOut)
    if (!eo->n2->eliminated && eo->n2 != c) {
      trsm_right_upper(*R, *eo->A21);
      flops += (double)eo->A21->rows() * cc * cc;
    }
  Op o;
  o.kind = Op::ScaleD;
  o.c = c;
  o.xs_len = c->cols();
  o.xsf_len = c->rows();
  if (square) { o.Q = Q; o.tau = t; ops.push_back(o); }
  else { o.Q = R; ops.push_back(o); tops.push_back(o); }
}

void Tree::split_edges(Cluster* c, Cluster* f) {  // tree.cpp:480-530
  for (auto e : c->edgesIn) {
    if (e->n1->eliminated) continue;
    Mat* AT = new Mat(e->A21->bottomRows(f->rows()));
    Edge* ef = new_edge(e->n1, f, AT);
    *e->A21 = e->A21->topRows(c->rows());
    f->edgesIn.push_back(ef);
    f->rsparsity.insert(e->n1);
    e->n1->edgesOut.push_back(ef);
  }
  auto fe = std::find_if(c->edgesOut.begin(), c->edgesOut.end(), [c](Edge* e) { return e->n2 == c; });
  Edge* es = *fe;
  Mat& S = *es->A21;
  Mat* Aff = new Mat(S.block(S.rows() - f->rows(), S.cols() - f->cols(), f->rows(), f->cols()));
  Mat* Afc = new Mat(S.block(S.rows() - f->rows(), 0, f->rows(), c->cols()));
  Mat* Acf = new Mat(S.block(0, S.cols() - f->cols(), c->rows(), f->cols()));
  S = S.block(0, 0, c->rows(), c->cols());
  Edge* eff = new_edge(f, f, Aff);
  Edge* efc = new_edge(f, c, Acf);
  Edge* ecf = new_edge(c, f, Afc);
  f->edgesOut.push_back(eff);
  f->edgesOut.push_back(efc);
  f->edgesIn.push_back(ecf);
  c->edgesOut.push_back(ecf);
  f->cnbrs.insert(c);
  f->cnbrs.insert(f);
  f->rsparsity.insert(c);
  f->sort_edgesOut(true);
}

void Tree::sparsify(Cluster* c) {  // tree.cpp:887-989 (--scale 0)
  if (c->cols() == 0) return;
  std::vector<Mat*> Apm, Amp;
  Mat* App = nullptr;
  for (auto e : c->edgesIn) if (!e->n1->eliminated) Apm.push_back(e->A21.get());
  for (auto e : c->edgesOut)
    if (!e->n2->eliminated) { if (e->n2 != c) Amp.push_back(e->A21.get()); else App = e->A21.get(); }
  int ampr = 0, apmc = 0;
  for (auto m : Amp) ampr += m->rows();
  for (auto m : Apm) apmc += m->cols();
  int kc = App->cols();
  Mat Ampc(ampr, kc), Apmc(App->rows(), apmc);
  { int r = 0; for (auto m : Amp) { Ampc.setBlock(r, 0, *m); r += m->rows(); } }
  { int q = 0; for (auto m : Apm) { Apmc.setBlock(0, q, *m); q += m->cols(); } }
  Mat C(kc, ampr + apmc);
  C.setBlock(0, 0, Ampc.transpose());
  Mat Ap(App->rows() + ampr, kc);
  Ap.setBlock(0, 0, *App);
  Ap.setBlock(App->rows(), 0, Ampc);
  std::vector<double> S = singular_values(Ap);
  double s = S[std::min(Ap.rows(), Ap.cols()) - 1];
  Mat P;
  gemm(*App, true, Apmc, false, 1.0 / s, 0.0, P);
  C.setBlock(0, ampr, P);
  std::vector<int> jpvt;
  std::vector<double> t;
  geqp3(C, jpvt, t);
  int kk = std::min(C.rows(), C.cols());
  flops += 4.0 * C.rows() * C.cols() * kk - 2.0 * (C.rows() + C.cols()) * kk * kk + 4.0 / 3.0 * kk * kk * kk;
  std::vector<double> rii(kk);
  for (int i = 0; i < kk; ++i) rii[i] = C(i, i);
  int rank = choose_rank(rii, tol);
  if (rank >= C.rows()) return;
  auto v = std::make_shared<Mat>(C.leftCols(rank));
  auto h = std::make_shared<std::vector<double>>(t.begin(), t.begin() + rank);
  ormqr_right_notrans(*v, *h, *App);
  Op oo;
  oo.kind = Op::Orthogonal; oo.c = c; oo.V = v; oo.tau = h; oo.xs_len = c->cols();
  ops.push_back(oo);
  if (!square) tops.push_back(oo);
  int cf_csize = c->cols() - rank;
  Cluster* cf = new_cluster(c->cstart + rank, cf_csize, c->rstart + c->rows() - cf_csize, cf_csize, c->id);
  Op os;
  os.kind = Op::Split; os.c = c; os.n = cf; os.xs_len = c->cols(); os.xsf_len = c->rows(); os.crows = cf->rows();
  ops.push_back(os);
  if (!square) tops.push_back(os);
  c->reset_size(c->rows() - cf_csize, rank);
  // Amp[i] <- (triu(R[0:rank]) P^T)[:, its columns]^T
  std::vector<int> pos(C.cols());
  for (int p = 0; p < C.cols(); ++p) pos[jpvt[p]] = p;
  int q = 0;
  for (auto m : Amp) {
    *m = Mat(m->rows(), rank);
    for (int i = 0; i < m->rows(); ++i) {
      int p = pos[q + i];
      for (int r = 0; r < rank; ++r) (*m)(i, r) = (r <= p) ? C(r, p) : 0.0;
    }
    q += m->rows();
  }
  split_edges(c, cf);
  eliminate_cluster(cf);
}

void Tree::sparsifyD(Cluster* c, bool imp) {  // tree.cpp:992-1114 ; imp: 1117-1257
  if (c->cols() == 0) return;
  std::vector<Mat*> Apm, Amp;
  Edge* es = nullptr;
  int spm = 0, smp = 0;
  for (auto e : c->edgesIn)
    if (!e->n1->eliminated) { Apm.push_back(e->A21.get()); spm += e->A21->cols(); }
  for (auto e : c->edgesOut) {
    if (!e->n2->eliminated && e->n2 != c) { Amp.push_back(e->A21.get()); smp += e->A21->rows(); }
    else if (e->n2 == c) es = e;
  }
  int kc = c->cols();
  Mat C(kc, spm + smp);
  Mat Apmc(c->rows(), spm);
  { int q = 0; for (auto m : Apm) { Apmc.setBlock(0, q, *m); q += m->cols(); } }
  Mat Apmc_rest;
  if (!square) Apmc_rest = Apmc.bottomRows(c->rows() - c->cols());
  Mat Apmc_top = Apmc.topRows(kc);
  { int r = 0;
    for (auto m : Amp) { for (int i = 0; i < m->rows(); ++i) for (int j = 0; j < kc; ++j) C(j, r + i) = (*m)(i, j); r += m->rows(); } }
  for (int j = 0; j < spm; ++j) for (int i = 0; i < kc; ++i) C(i, smp + j) = Apmc(i, j);
  std::vector<int> jpvt;
  std::vector<double> t;
  geqp3(C, jpvt, t);
  int kk = std::min(C.rows(), C.cols());
  flops += 4.0 * C.rows() * C.cols() * kk - 2.0 * (C.rows() + C.cols()) * kk * kk + 4.0 / 3.0 * kk * kk * kk;
  std::vector<double> rii(kk);
  for (int i = 0; i < kk; ++i) rii[i] = C(i, i);
  int rank = choose_rank(rii, tol);
  if (rank >= C.rows()) return;
  auto v = std::make_shared<Mat>(C.leftCols(rank));
  auto h = std::make_shared<std::vector<double>>(t.begin(), t.begin() + rank);
  Op oo;
  oo.kind = Op::OrthogonalD; oo.c = c; oo.V = v; oo.tau = h; oo.xs_len = c->cols();
  ops.push_back(oo);
  if (!square) tops.push_back(oo);
  if (!imp) nnzQ += (unsigned long long)(C.rows() * rank);
  int cf_csize = c->cols() - rank;
  Cluster* cf = new_cluster(c->cstart + rank, cf_csize, c->rstart + c->cols() - cf_csize, cf_csize, c->id);
  Op os;
  os.kind = Op::SplitD; os.c = c; os.n = cf; os.crows = c->rows(); os.ccols = c->cols(); os.rank = rank;
  os.xsf_len = cf->rows();
  ops.push_back(os);
  if (!square) tops.push_back(os);
  if (imp) nnzQ += (unsigned long long)(C.rows() * rank);
  // Crank = triu(R[0:rank]) P^T, column j <- pivot position pos[j]
  std::vector<int> pos(C.cols());
  for (int p = 0; p < C.cols(); ++p) pos[jpvt[p]] = p;
  Mat Crank(rank, C.cols());
  for (int j = 0; j < C.cols(); ++j) {
    int p = pos[j];
    for (int r = 0; r < rank; ++r) Crank(r, j) = (r <= p) ? C(r, p) : 0.0;
  }
  if (!imp) {
    Mat self(c->rows() - cf->rows(), rank);
    for (int i = 0; i < rank; ++i) for (int j = 0; j < rank; ++j) self(i, j) = (*es->A21)(i, j);
    int rank_rows = rank + Apmc_rest.rows();
    c->reset_size(rank_rows, rank);
    int cc = 0;
    for (auto m : Amp) { *m = Crank.middleCols(cc, m->rows()).transpose(); cc += m->rows(); }
    int count = cc;
    for (auto m : Apm) {
      int mc = m->cols();
      m->conservativeResize(c->rows(), mc);
      for (int j = 0; j < mc; ++j) for (int i = 0; i < c->cols(); ++i) (*m)(i, j) = Crank(i, cc + j);
      if (!square)
        for (int j = 0; j < mc; ++j)
          for (int i = 0; i < c->rows() - c->cols(); ++i) (*m)(c->cols() + i, j) = Apmc_rest(i, cc - count + j);
      cc += mc;
    }
    *es->A21 = self.topRows(c->rows());
    cf->eliminated = true;
  } else {
    c->reset_size(c->rows() - cf_csize, rank);
    int cc = 0;
    for (auto m : Amp) { *m = Crank.middleCols(cc, m->rows()).transpose(); cc += m->rows(); }
    int count = cc;
    for (auto m : Apm) {
      int mc = m->cols();
      m->conservativeResize(c->rows(), mc);
      for (int j = 0; j < mc; ++j) for (int i = 0; i < c->cols(); ++i) (*m)(i, j) = Crank(i, cc + j);
      if (!square)
        for (int j = 0; j < mc; ++j)
          for (int i = 0; i < c->rows() - c->cols(); ++i) (*m)(c->cols() + i, j) = Apmc_rest(i, cc - count + j);
      cc += mc;
    }
    Mat* Aff = new Mat(es->A21->block(0, 0, cf->cols(), cf->cols()));
    Mat selfb = es->A21->block(es->A21->rows() - c->rows(), es->A21->cols() - c->cols(), c->rows(), c->cols());
    *es->A21 = selfb;
    Edge* eff = new_edge(cf, cf, Aff);
    cf->edgesOut.push_back(eff);
    ormqr_left(*v, *h, Apmc_top, 'T');
    int q = 0;
    for (auto e : c->edgesIn) {
      if (e->n1->eliminated) continue;
      Cluster* m = e->n1;
      Mat* Afm = new Mat(Apmc_top.block(rank, q, cf->rows(), m->cols()));
      Edge* ne = new_edge(m, cf, Afm);
      cf->edgesIn.push_back(ne);
      m->edgesOut.push_back(ne);
      q += m->cols();
    }
    cf->tau.assign(cf->cols(), 0.0);
    cf->T = Mat(cf->cols(), cf->cols());
    ops.push_back(make_qr(cf));
    if (!square) tops.push_back(make_qr(cf));
    cf->eliminated = true;
  }
}

void Tree::sparsify_extra(Cluster* c) {  // tree.cpp:1260-1307 (+ sparsify_extra_rows 226-247)
  if (c->cols() == 0 || c->rows() <= c->cols()) return;
  std::vector<Mat*> Apm;
  int spm = 0;
  for (auto e : c->edgesIn)
    if (!e->n1->eliminated) { Apm.push_back(e->A21.get()); spm += e->A21->cols(); }
  Mat Apmc(c->rows(), spm);
  { int q = 0; for (auto m : Apm) { Apmc.setBlock(0, q, *m); q += m->cols(); } }
  Mat rest = Apmc.bottomRows(c->rows() - c->cols());
  int rank = c->cols(), rank_rows = 0;
  if (rest.rows() > 0) {
    std::vector<int> jpvt;
    std::vector<double> t;
    Mat C = rest;
    geqp3(C, jpvt, t);
    int kk = std::min(C.rows(), C.cols());
    flops += 4.0 * C.rows() * C.cols() * kk - 2.0 * (C.rows() + C.cols()) * kk * kk + 4.0 / 3.0 * kk * kk * kk;
    std::vector<double> rii(kk);
    for (int i = 0; i < kk; ++i) rii[i] = C(i, i);
    rank_rows = choose_rank(rii, tol);
    std::vector<int> pos(C.cols());
    for (int p = 0; p < C.cols(); ++p) pos[jpvt[p]] = p;
    Mat Cr(rank_rows, C.cols());
    for (int j = 0; j < C.cols(); ++j)
      for (int r = 0; r < rank_rows; ++r) Cr(r, j) = (r <= pos[j]) ? C(r, pos[j]) : 0.0;
    rest = Cr;
    extra_sparsified.push_back({ilvl, c->id.self.lvl, c->id.self.sep, (int)C.rows(), rank_rows});
    rank_rows += rank;
  }
  c->reset_size(rank_rows, rank);
  auto f = std::find_if(c->edgesOut.begin(), c->edgesOut.end(), [c](Edge* e) { return e->n2 == c; });
  *(*f)->A21 = (*f)->A21->topRows(rank_rows);
  int q = 0;
  for (auto m : Apm) {
    int mc = m->cols();
    m->conservativeResize(c->rows(), mc);
    for (int j = 0; j < mc; ++j) {
      for (int i = 0; i < c->cols(); ++i) (*m)(i, j) = Apmc(i, q + j);
      for (int i = 0; i < c->rows() - c->cols(); ++i) (*m)(c->cols() + i, j) = rest(i, q + j);
    }
    q += mc;
  }
}

void Tree::update_size(Cluster* snew) {  // tree.cpp:414-434
  int rs = 0, cs = 0;
  for (auto s : snew->children) { s->rposparent = rs; s->cposparent = cs; rs += s->rows(); cs += s->cols(); }
  snew->set_size(rs, cs);
  snew->rsize_org = rs; snew->csize_org = cs;
  Op o;
  o.kind = Op::Merge;
  o.c = snew;
  ops.push_back(o);
  if (!square) tops.push_back(o);
}

void Tree::update_edges(Cluster* snew) {  // tree.cpp:436-475
  CSet cn;
  for (auto s : snew->children)
    for (auto e : s->edgesOut)
      if (!e->n2->eliminated) cn.insert(e->n2->parent);
  snew->cnbrs = cn;
  for (auto n : cn) add_edge(snew, n);
  snew->sort_edgesOut();
  for (auto s : snew->children)
    for (auto e : s->edgesOut) {
      Cluster* nold = e->n2;
      if (nold->eliminated) continue;
      Cluster* nnew = nold->parent;
      auto f = std::find_if(snew->edgesOut.begin(), snew->edgesOut.end(), [nnew](Edge* x) { return x->n2 == nnew; });
      (*f)->A21->setBlock(nold->rposparent, s->cposparent, *e->A21);
      e->A21.reset();
    }
}

void Tree::merge_all() {  // tree.cpp:872-881
  current_bottom++;
  for (auto s : bottom_current()) update_size(s);
  for (auto s : bottom_current()) update_edges(s);
}

int Tree::factorize() {  // tree.cpp:1309-1437
  double fstart = now_s();
  for (ilvl = 0; ilvl < nlevels; ++ilvl) {
    LevelStat& st = stats[ilvl];
    double t0 = now_s();
    for (auto s : bottom_current()) if (s->level() == ilvl) get_sparsity(s);
    double t1 = now_s();
    for (auto s : bottom_current()) if (s->level() == ilvl) eliminate_cluster(s);
    double t2 = now_s();
    if (ilvl < nlevels - 1)
      for (auto s : bottom_current()) if (s->eliminated && !square) reassign_rows(s);
    double t3 = now_s();
    if (scale && ilvl < nlevels - 2 && ilvl >= skip && tol != 0)
      for (auto s : bottom_current()) if (want_sparsify(s)) diagScale(s);
    double t4 = now_s();
    if (ilvl >= skip && ilvl < nlevels - 2 && tol != 0) {
      for (auto s : bottom_current()) if (want_sparsify(s) && scale && !square) sparsify_extra(s);
      for (auto s : bottom_current()) {
        if (!want_sparsify(s)) continue;
        int before = s->cols();
        if (scale) sparsifyD(s, order == 1.5f);
        else sparsify(s);
        sparsified.push_back({ilvl, s->id.self.lvl, s->id.self.sep, before, s->cols()});
      }
    }
    double t5 = now_s();
    if (ilvl < nlevels - 1) merge_all();
    double t6 = now_s();
    st.t_fill = t1 - t0; st.t_elmn = t2 - t1; st.t_shift = t3 - t2; st.t_scale = t4 - t3;
    st.t_sprsfy = t5 - t4; st.t_merge = t6 - t5;
    auto ts = topsize();
    st.top_rows = ts.first; st.top_cols = ts.second;
    // verbose >= 3: the level lines go to stderr (progress of a long run
    // whose stdout is kept for a result line)
    if (verbose)
      fprintf(verbose >= 3 ? stderr : stdout,
             "lvl: %d    update fill-in: %.3f  elmn: %.3f  shift: %.3f  scale: %.3f  sprsfy: %.3f  merge: %.3f  "
             "size top_sep: %d, %d   a.r top_sep: %.3f\n",
             ilvl, st.t_fill, st.t_elmn, st.t_shift, st.t_scale, st.t_sprsfy, st.t_merge, ts.first, ts.second,
             (double)ts.first / (double)ts.second);
  }
  t_factorize = now_s() - fstart;
  if (verbose && verbose < 3) {
    printf("Tolerance set: %.3e\n", tol);
    printf("Time to factorize:  %.3e\n", t_factorize);
    printf("Size of top separator: %d\n", top_size());
    printf("nnzA: %llu nnzR: %llu\n", nnzA, nnzR);
    printf("nnzH: %llu nnzQ: %llu\n", nnzH, nnzQ);
  }
  return 0;
}

// ------------------------------------------------------------------ op tape
void Tree::op_fwd(const Op& o) const {  // operations.cpp
  switch (o.kind) {
    case Op::QR: {  // :12-38
      Cluster* c = o.c;
      Mat Q(o.nrows, o.ncols);
      std::vector<double> xc(o.nrows);
      int cr = 0, i = 0;
      for (auto e : c->edgesOut) {
        if (!e->A21) continue;
        int s = o.a21_sizes[i++];
        for (int r = 0; r < s; ++r) xc[cr + r] = e->n2->x[r];
        Q.setBlock(cr, 0, *e->A21);
        cr += s;
      }
      larfb_vec(Q, c->T, xc.data(), o.nrows, 'T');
      cr = 0; i = 0;
      for (auto e : c->edgesOut) {
        if (!e->A21) continue;
        int s = o.a21_sizes[i++];
        for (int r = 0; r < s; ++r) e->n2->x[r] = xc[cr + r];
        cr += s;
      }
      break;
    }
    case Op::Reassign:  // :59-68
      for (int i = o.nstart; i < o.nstart + o.nrows; ++i) o.n->x[i] = o.c->x[o.indices[i - o.nstart]];
      break;
    case Op::ScaleD:  // :75-78
      ormqr_vec(*o.Q, *o.tau, o.c->x.data(), o.xsf_len, 'T');
      break;
    case Op::OrthogonalD:  // :88
      ormqr_vec(*o.V, *o.tau, o.c->x.data(), o.xs_len, 'T');
      break;
    case Op::Merge: {  // :93-102
      int k = 0;
      for (auto ch : o.c->children) for (int i = 0; i < ch->rows(); ++i) o.c->x[k++] = ch->x[i];
      break;
    }
    case Op::Split: {  // :117
      int fl = o.n->rows();
      for (int i = 0; i < fl; ++i) o.n->x[i] = o.c->x[o.xsf_len - fl + i];
      break;
    }
    case Op::SplitD: {  // :120-125
      std::vector<double>& x = o.c->x;
      int cr = o.crows, cc = o.ccols, rk = o.rank;
      std::vector<double> xf(x.begin() + rk, x.begin() + cc);
      for (int i = 0; i < cc - rk; ++i) o.n->x[i] = xf[i];
      std::vector<double> rem(x.begin() + cc, x.begin() + cr);
      for (int i = 0; i < cr - cc; ++i) x[rk + i] = rem[i];
      for (int i = 0; i < cc - rk; ++i) x[cr - (cc - rk) + i] = xf[i];
      break;
    }
    default: break;
  }
}

void Tree::op_bwd(const Op& o) const {
  switch (o.kind) {
    case Op::QR: {  // :40-56
      Cluster* c = o.c;
      double* xs = c->x.data();
      int i = 0;
      for (auto e : c->edgesIn) {
        if (i >= (int)o.a12_sizes.size()) break;
        int s = o.a12_sizes[i++];
        const Mat& A = *e->A21;
        const std::vector<double>& xn = e->n1->x;
        for (int j = 0; j < s; ++j) {
          double v = xn[j];
          for (int r = 0; r < o.ccols; ++r) xs[r] -= A(r, j) * v;
        }
      }
      const Mat& S = *(*c->edgesOut.begin())->A21;
      trsv_upper(S, xs, o.ccols, 'N');
      break;
    }
    case Op::ScaleD:  // :80-83
      trsv_upper(*o.Q, o.c->x.data(), o.xs_len, 'N');
      break;
    case Op::Orthogonal:  // :86
    case Op::OrthogonalD:  // :89
      ormqr_vec(*o.V, *o.tau, o.c->x.data(), o.xs_len, 'N');
      break;
    case Op::Merge: {  // :104-114
      int k = 0;
      for (auto ch : o.c->children) for (int i = 0; i < ch->cols(); ++i) ch->x[i] = o.c->x[k++];
      break;
    }
    case Op::Split: {  // :118 (xsc_head = coarse head at creation)
      int fl = o.n->rows();
      for (int i = 0; i < fl; ++i) o.c->x[o.xs_len - fl + i] = o.n->x[i];
      break;
    }
    case Op::SplitD: {  // :126-130
      std::vector<double>& x = o.c->x;
      int cr = o.crows, cc = o.ccols, rk = o.rank;
      std::vector<double> rem(x.begin() + rk, x.begin() + rk + (cr - cc));
      for (int i = 0; i < cc - rk; ++i) x[rk + i] = o.n->x[i];
      for (int i = 0; i < cr - cc; ++i) x[cc + i] = rem[i];
      break;
    }
    default: break;
  }
}

void Tree::top_fwd(const Op& o) const {  // toperations.cpp
  switch (o.kind) {
    case Op::QR: {  // :12-24
      Cluster* c = o.c;
      double* xs = c->xt.data();
      const Mat& S = *(*c->edgesOut.begin())->A21;
      trsv_upper(S, xs, o.ccols, 'T');
      int i = 0;
      for (auto e : c->edgesIn) {
        if (i >= (int)o.a12_sizes.size()) break;
        int s = o.a12_sizes[i++];
        const Mat& A = *e->A21;
        std::vector<double>& xn = e->n1->xt;
        for (int j = 0; j < s; ++j) {
          double acc = 0.0;
          for (int r = 0; r < o.ccols; ++r) acc += A(r, j) * xs[r];
          xn[j] -= acc;
        }
      }
      break;
    }
    case Op::ScaleD:  // :71-74
      trsv_upper(*o.Q, o.c->xt.data(), o.xs_len, 'T');
      break;
    case Op::Orthogonal:
    case Op::OrthogonalD:  // :77, :79-81
      ormqr_vec(*o.V, *o.tau, o.c->xt.data(), o.xs_len, 'T');
      break;
    case Op::Merge: {  // :87-95
      int k = 0;
      for (auto ch : o.c->children) for (int i = 0; i < ch->cols(); ++i) o.c->xt[k++] = ch->xt[i];
      break;
    }
    case Op::Split: {  // :109 (xsc_head)
      int fl = o.n->rows();
      for (int i = 0; i < fl; ++i) o.n->xt[i] = o.c->xt[o.xs_len - fl + i];
      break;
    }
    case Op::SplitD: {
      std::vector<double>& x = o.c->xt;
      int cr = o.crows, cc = o.ccols, rk = o.rank;
      std::vector<double> xf(x.begin() + rk, x.begin() + cc);
      for (int i = 0; i < cc - rk; ++i) o.n->xt[i] = xf[i];
      std::vector<double> rem(x.begin() + cc, x.begin() + cr);
      for (int i = 0; i < cr - cc; ++i) x[rk + i] = rem[i];
      for (int i = 0; i < cc - rk; ++i) x[cr - (cc - rk) + i] = xf[i];
      break;
    }
    default: break;
  }
}

void Tree::top_bwd(const Op& o) const {
  switch (o.kind) {
    case Op::Reassign:  // toperations.cpp:56-63
      for (int i = o.nstart; i < o.nstart + o.nrows; ++i) o.c->xt[o.indices[i - o.nstart]] = o.n->xt[i];
      break;
    case Op::OrthogonalD:
      ormqr_vec(*o.V, *o.tau, o.c->xt.data(), o.xs_len, 'N');
      break;
    case Op::Merge: {  // :97-106
      int k = 0;
      for (auto ch : o.c->children) for (int i = 0; i < ch->rows(); ++i) ch->xt[i] = o.c->xt[k++];
      break;
    }
    case Op::Split: {  // :110 (xsc = coarse full at creation)
      int fl = o.n->rows();
      for (int i = 0; i < fl; ++i) o.c->xt[o.xsf_len - fl + i] = o.n->xt[i];
      break;
    }
    case Op::SplitD: {
      std::vector<double>& x = o.c->xt;
      int cr = o.crows, cc = o.ccols, rk = o.rank;
      std::vector<double> rem(x.begin() + rk, x.begin() + rk + (cr - cc));
      for (int i = 0; i < cc - rk; ++i) x[rk + i] = o.n->xt[i];
      for (int i = 0; i < cr - cc; ++i) x[cc + i] = rem[i];
      break;
    }
    default: break;
  }
}

void Tree::solve(const std::vector<double>& b0, std::vector<double>& x) const {  // tree.cpp:1440-1466
  std::vector<double> b(nrows);
  for (int i = 0; i < nrows; ++i) b[i] = b0[rperm[i]];
  for (auto c : bottom_original()) for (int i = 0; i < c->rsize_org; ++i) c->x[i] = b[c->rstart + i];
  for (auto& o : ops) op_fwd(o);
  for (auto it = ops.rbegin(); it != ops.rend(); ++it) op_bwd(*it);
  std::vector<double> xp(ncols);
  for (auto c : bottom_original()) for (int i = 0; i < c->csize_org; ++i) xp[c->cstart + i] = c->x[i];
  x.assign(ncols, 0.0);
  for (int i = 0; i < ncols; ++i) x[cperm[i]] = xp[i];
}

static bool skip_in_nrml(Op::Kind k) { return k == Op::QR || k == Op::ScaleD || k == Op::OrthogonalD; }

void Tree::solve_nrml(const std::vector<double>& b0, std::vector<double>& x) const {  // tree.cpp:1471-1517
  std::vector<double> b(ncols);
  for (int i = 0; i < ncols; ++i) b[i] = b0[cperm[i]];
  for (auto c : bottom_original()) for (int i = 0; i < c->csize_org; ++i) c->xt[i] = b[c->cstart + i];
  for (auto& o : tops) top_fwd(o);
  for (auto it = tops.rbegin(); it != tops.rend(); ++it) if (!skip_in_nrml(it->kind)) top_bwd(*it);
  for (auto c : bottom_original()) for (int i = 0; i < c->rsize_org; ++i) c->x[i] = c->xt[i];
  for (auto& o : ops) if (!skip_in_nrml(o.kind)) op_fwd(o);
  for (auto it = ops.rbegin(); it != ops.rend(); ++it) op_bwd(*it);
  std::vector<double> xp(ncols);
  for (auto c : bottom_original()) for (int i = 0; i < c->csize_org; ++i) xp[c->cstart + i] = c->x[i];
  x.assign(ncols, 0.0);
  for (int i = 0; i < ncols; ++i) x[cperm[i]] = xp[i];
}

// ------------------------------------------------------------------ Krylov
namespace {
double norm(const std::vector<double>& v) {
  double s = 0;
  for (double a : v) s += a * a;
  return std::sqrt(s);
}
// Eigen makeHouseholder on x[off..m): essential stored in ess[0..m-off-1).
void make_householder(const double* x, int n, double* ess, double& tau, double& beta) {
  double c0 = x[0], tsq = 0.0;
  for (int i = 1; i < n; ++i) tsq += x[i] * x[i];
  const double tiny = std::numeric_limits<double>::min();
  if (tsq <= tiny) {
    tau = 0.0; beta = c0;
    for (int i = 1; i < n; ++i) ess[i - 1] = 0.0;
  } else {
    beta = std::sqrt(c0 * c0 + tsq);
    if (c0 >= 0) beta = -beta;
    for (int i = 1; i < n; ++i) ess[i - 1] = x[i] / (c0 - beta);
    tau = (beta - c0) / beta;
  }
}
// Eigen applyHouseholderOnTheLeft on v[0..n) with essential ess (n-1), tau.
void apply_householder(double* v, int n, const double* ess, double tau) {
  if (n == 1) { v[0] *= (1.0 - tau); return; }
  if (tau == 0.0) return;
  double tmp = v[0];
  for (int i = 1; i < n; ++i) tmp += ess[i - 1] * v[i];
  v[0] -= tau * tmp;
  for (int i = 1; i < n; ++i) v[i] -= tau * ess[i - 1] * tmp;
}
struct Givens { double c = 1, s = 0; };
Givens make_givens(double p, double q) {  // Eigen JacobiRotation::makeGivens (real)
  Givens g;
  if (q == 0.0) { g.c = p < 0 ? -1 : 1; g.s = 0; }
  else if (p == 0.0) { g.c = 0; g.s = q < 0 ? 1 : -1; }
  else if (std::fabs(p) > std::fabs(q)) {
    double t = q / p, u = std::sqrt(1 + t * t);
    if (p < 0) u = -u;
    g.c = 1 / u; g.s = -t * g.c;
  } else {
    double t = p / q, u = std::sqrt(1 + t * t);
    if (q < 0) u = -u;
    g.s = -1 / u; g.c = -t * g.s;
  }
  return g;
}
// v.applyOnTheLeft(i, j, G.adjoint()): x' = c x - s y ; y' = s x + c y
void apply_givens_adj(double& x, double& y, const Givens& g) {
  double xi = x, yi = y;
  x = g.c * xi - g.s * yi;
  y = g.s * xi + g.c * yi;
}
}  // namespace

int gmres(const Csc& A, const std::vector<double>& rhs, std::vector<double>& x, const Tree& M, int iters,
          int restart, double tol, double* tol_out, int verbose) {  // is.cpp:95-272
  int m = A.m;
  if (norm(rhs) <= std::numeric_limits<double>::min()) { x.assign(m, 0.0); if (tol_out) *tol_out = 0; return 0; }
  std::vector<double> p0(m), r0(m), t(m), v(m), xn(m), Ax(m);
  csc_matvec(A, x.data(), Ax.data());
  for (int i = 0; i < m; ++i) p0[i] = rhs[i] - Ax[i];
  M.solve(p0, r0);
  double r0n = norm(r0);
  if (verbose && verbose < 3) {
    csc_matvec(A, r0.data(), Ax.data());
    double s = 0;
    for (int i = 0; i < m; ++i) s += (Ax[i] - p0[i]) * (Ax[i] - p0[i]);
    printf("initial residual %.3e\n", std::sqrt(s) / norm(p0));
  }
  if (r0n == 0) return 0;
  int maxIters = iters, it = 0;
  Mat H(m, restart + 1);
  std::vector<double> w(restart + 1, 0.0), tau(restart + 1, 0.0);
  std::vector<Givens> G(restart);
  double beta;
  make_householder(r0.data(), m, &H(1, 0), tau[0], beta);
  w[0] = beta;
  double tol_error = 0;
  for (int k = 1; k <= restart; ++k) {
    ++it;
    std::fill(v.begin(), v.end(), 0.0);
    v[k - 1] = 1.0;
    for (int i = k - 1; i >= 0; --i) apply_householder(&v[i], m - i, &H(i + 1, i), tau[i]);
    csc_matvec(A, v.data(), t.data());
    M.solve(t, v);
    for (int i = 0; i < k; ++i) apply_householder(&v[i], m - i, &H(i + 1, i), tau[i]);
    double tn = 0;
    for (int i = k; i < m; ++i) tn += v[i] * v[i];
    if (std::sqrt(tn) != 0.0 && k <= restart) {
      make_householder(&v[k], m - k, &H(k + 1, k), tau[k], beta);
      apply_householder(&v[k], m - k, &H(k + 1, k), tau[k]);
    }
    if (k > 1) for (int i = 0; i < k - 1; ++i) apply_givens_adj(v[i], v[i + 1], G[i]);
    if (k < m && v[k] != 0.0) {
      G[k - 1] = make_givens(v[k - 1], v[k]);
      apply_givens_adj(v[k - 1], v[k], G[k - 1]);
      apply_givens_adj(w[k - 1], w[k], G[k - 1]);
    }
    for (int i = 0; i < k; ++i) H(i, k - 1) = v[i];
    tol_error = std::fabs(w[k]) / r0n;
    bool stop = (k == m || tol_error < tol || it == maxIters);
    if (verbose > 1) printf("%d: |Ax-b|/|b| = %3.2e <? %3.2e\n", it, tol_error, tol);
    if (stop || k == restart) {
      std::vector<double> y(w.begin(), w.begin() + k);
      for (int i = k - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < k; ++j) s -= H(i, j) * y[j];
        y[i] = s / H(i, i);
      }
      std::fill(xn.begin(), xn.end(), 0.0);
      for (int i = k - 1; i >= 0; --i) {
        xn[i] += y[i];
        apply_householder(&xn[i], m - i, &H(i + 1, i), tau[i]);
      }
      for (int i = 0; i < m; ++i) x[i] += xn[i];
      if (stop) { if (tol_out) *tol_out = tol_error; return it; }
      k = 0;
      csc_matvec(A, x.data(), Ax.data());
      for (int i = 0; i < m; ++i) p0[i] = rhs[i] - Ax[i];
      M.solve(p0, r0);
      H.setZero();
      std::fill(w.begin(), w.end(), 0.0);
      std::fill(tau.begin(), tau.end(), 0.0);
      make_householder(r0.data(), m, &H(1, 0), tau[0], beta);
      w[0] = beta;
    }
  }
  if (tol_out) *tol_out = tol_error;
  return it;
}

int cgls(const Csc& A, const std::vector<double>& rhs, std::vector<double>& x, const Tree& M, int iters,
         double tol, double* tol_out, int verbose) {  // is.cpp:275-346
  int m = A.m, n = A.n;
  std::vector<double> res(m), nres(n), tmp(m), p(n), z(n), Atb(n);
  csc_matvec(A, x.data(), tmp.data());
  for (int i = 0; i < m; ++i) res[i] = rhs[i] - tmp[i];
  csc_matvec_t(A, res.data(), nres.data());
  csc_matvec_t(A, rhs.data(), Atb.data());
  double rhsn2 = 0;
  for (double a : Atb) rhsn2 += a * a;
  if (rhsn2 == 0) { x.assign(n, 0.0); if (tol_out) *tol_out = 0; return 0; }
  double thr = tol * tol * rhsn2;
  double rn2 = 0;
  for (double a : nres) rn2 += a * a;
  if (rn2 < thr) { if (tol_out) *tol_out = std::sqrt(rn2 / rhsn2); return 0; }
  M.solve_nrml(nres, p);
  double absNew = 0;
  for (int i = 0; i < n; ++i) absNew += nres[i] * p[i];
  int i = 0;
  while (i < iters) {
    csc_matvec(A, p.data(), tmp.data());
    double tn = 0;
    for (double a : tmp) tn += a * a;
    double alpha = absNew / tn;
    for (int j = 0; j < n; ++j) x[j] += alpha * p[j];
    for (int j = 0; j < m; ++j) res[j] -= alpha * tmp[j];
    csc_matvec_t(A, res.data(), nres.data());
    rn2 = 0;
    for (double a : nres) rn2 += a * a;
    if (verbose > 1) printf("%d %3.2e \n", i, std::sqrt(rn2 / rhsn2));
    if (rn2 < thr) break;
    M.solve_nrml(nres, z);
    double absOld = absNew;
    absNew = 0;
    for (int j = 0; j < n; ++j) absNew += nres[j] * z[j];
    double beta = absNew / absOld;
    for (int j = 0; j < n; ++j) p[j] = z[j] + beta * p[j];
    i++;
  }
  if (tol_out) *tol_out = std::sqrt(rn2 / rhsn2);
  return i;
}

}  // namespace orc
