// Geometric nested dissection and interface labelling (host, integer work).
// Reference: partition.cpp:54-157 (helpers, rows2clusters, rmap2cmap),
// 241-348 (bissect_geo, GeometricPartitionAtA), 506-578 (getInterfacesUsingA);
// tree.cpp:46-163 (initPermutation, form_rmap).
#include "partition.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace spq {

CscMat csc_transpose(const CscMat& A) {
  CscMat T;
  T.m = A.n;
  T.n = A.m;
  T.p.assign(A.m + 1, 0);
  for (int r : A.i) T.p[r + 1]++;
  std::partial_sum(T.p.begin(), T.p.end(), T.p.begin());
  T.i.resize(A.nnz());
  T.x.resize(A.nnz());
  std::vector<int> fill(T.p.begin(), T.p.end() - 1);
  for (int c = 0; c < A.n; ++c)
    for (int q = A.p[c]; q < A.p[c + 1]; ++q) {
      const int dst = fill[A.i[q]]++;
      T.i[dst] = c;
      T.x[dst] = A.x[q];
    }
  return T;
}

// B(r, c) = A(rperm[r], cperm[c]) (either permutation may be null).
CscMat csc_permute(const CscMat& A, const std::vector<int>* rperm, const std::vector<int>* cperm) {
  CscMat B;
  B.m = A.m;
  B.n = A.n;
  B.p.assign(A.n + 1, 0);
  B.i.reserve(A.nnz());
  B.x.reserve(A.nnz());
  std::vector<int> rinv;
  if (rperm) {
    rinv.resize(A.m);
    for (int r = 0; r < A.m; ++r) rinv[(*rperm)[r]] = r;
  }
  std::vector<std::pair<int, double>> col;
  for (int c = 0; c < A.n; ++c) {
    const int src = cperm ? (*cperm)[c] : c;
    col.clear();
    for (int q = A.p[src]; q < A.p[src + 1]; ++q) col.push_back({rperm ? rinv[A.i[q]] : A.i[q], A.x[q]});
    if (rperm) std::sort(col.begin(), col.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& e : col) { B.i.push_back(e.first); B.x.push_back(e.second); }
    B.p[c + 1] = (int)B.i.size();
  }
  return B;
}

std::vector<double> tensor_coordinates(int n, int dim) {
  int64_t N = 1;
  for (int d = 0; d < dim; ++d) N *= n;
  std::vector<double> X((size_t)dim * N);
  for (int64_t id = 0; id < N; ++id) {
    int64_t rem = id;
    for (int d = dim - 1; d >= 0; --d) {  // coordinate 0 varies slowest
      X[(size_t)id * dim + d] = (double)(rem % n);
      rem /= n;
    }
  }
  return X;
}

namespace {

// Median split of `dofs` along the coordinate with the largest range (first
// such coordinate on ties); X <= median value goes to part `lo`, else lo + 1.
void split_geo(const std::vector<int>& dofs, std::vector<int>& part, const std::vector<double>& X, int dim, int lo) {
  if (dofs.empty()) return;
  int best = -1;
  double range = -1.0;
  for (int d = 0; d < dim; ++d) {
    double mx = std::numeric_limits<double>::lowest(), mn = std::numeric_limits<double>::max();
    for (int v : dofs) {
      const double c = X[(size_t)v * dim + d];
      mx = std::max(mx, c);
      mn = std::min(mn, c);
    }
    if (mx - mn > range) { range = mx - mn; best = d; }
  }
  std::vector<double> vals;
  vals.reserve(dofs.size());
  for (int v : dofs) vals.push_back(X[(size_t)v * dim + best]);
  const size_t mid = dofs.size() / 2;
  std::nth_element(vals.begin(), vals.begin() + mid, vals.end());
  const double cut = vals[mid];
  for (int v : dofs) part[v] = (X[(size_t)v * dim + best] <= cut) ? lo : lo + 1;
}

SepID climb_common(SepID a, SepID b) {  // partition.cpp:63-86
  while (a.lvl < b.lvl) { a.lvl++; a.sep /= 2; }
  while (b.lvl < a.lvl) { b.lvl++; b.sep /= 2; }
  while (a.sep != b.sep) { a.lvl++; a.sep /= 2; b.lvl++; b.sep /= 2; }
  return a;
}
SepID lower_of(SepID a, SepID b) { return a.lvl < b.lvl ? a : b; }  // partition.cpp:88-100

// Row labels of a rectangular matrix: each row joins the column cluster whose
// entries in that row carry the largest sum of squares (tree.cpp:57-98).
void rows_by_weight(const CscMat& A, const std::vector<CID>& cpermed, const std::vector<int>& cperm,
                    std::vector<CID>& rlab) {
  CscMat rowsP = csc_transpose(csc_permute(A, nullptr, &cperm));  // column r = row r, entries by permuted column
  const CID last = cpermed[A.n - 1];
  for (int r = 0; r < A.m; ++r) {
    if (rlab[r] != CID()) continue;
    double bestw = -1;
    CID pick = rlab[r];
    int q = rowsP.p[r];
    const int e = rowsP.p[r + 1];
    while (q < e) {
      const CID g = cpermed[rowsP.i[q]];
      double w = 0;
      int t = q;
      while (t < e && cpermed[rowsP.i[t]] == g) { w += rowsP.x[t] * rowsP.x[t]; ++t; }
      if (w > bestw) { bestw = w; pick = g; }
      q = t;
    }
    rlab[r] = (pick == CID()) ? last : pick;
  }
}

}  // namespace

std::vector<CID> geometric_labels(const CscMat& A, int nlevels, const std::vector<double>& X, int dim) {
  const int n = A.n;
  const CscMat R = csc_transpose(A);  // R.p/R.i: columns of each row, sorted
  // recursive bisection of all columns (nlevels - 1 times)
  std::vector<int> part(n, 0);
  for (int depth = 0; depth < nlevels - 1; ++depth) {
    const int nparts = 1 << depth;
    std::vector<std::vector<int>> groups(nparts);
    for (int v = 0; v < n; ++v) groups[part[v]].push_back(v);
    for (int g = 0; g < nparts; ++g) split_geo(groups[g], part, X, dim, 2 * g);
  }
  // separators: from the top level down, a still-unlabelled column in an even
  // subtree becomes a separator of level l when an unlabelled column of the
  // adjacent odd subtree is its neighbour in the graph of A^T A.
  std::vector<CID> lab(n);
  std::vector<int> stamp(n, -1);
  for (int depth = 0; depth < nlevels - 1; ++depth) {
    const int l = nlevels - depth - 1;
    for (int v = 0; v < n; ++v) {
      if (lab[v].self.lvl != -1) continue;
      const int pv = part[v] >> (l - 1);
      if (pv & 1) continue;
      bool sep = false;
      for (int q = A.p[v]; q < A.p[v + 1] && !sep; ++q) {
        const int row = A.i[q];
        for (int t = R.p[row]; t < R.p[row + 1]; ++t) {
          const int u = R.i[t];
          if (lab[u].self.lvl == -1 && (part[u] >> (l - 1)) == pv + 1) { sep = true; break; }
        }
      }
      if (sep) lab[v].self = SepID(l, pv / 2);
    }
  }
  for (int v = 0; v < n; ++v)
    if (lab[v].self.lvl == -1) {
      lab[v].self = SepID(0, part[v]);
      lab[v].l = lab[v].self;
      lab[v].r = lab[v].self;
    }
  // row labels
  std::vector<CID> rlab(A.m);
  if (A.m != A.n) {
    std::vector<int> cperm(n);
    std::iota(cperm.begin(), cperm.end(), 0);
    std::stable_sort(cperm.begin(), cperm.end(), [&](int a, int b) { return lab[a] < lab[b]; });
    std::vector<CID> cpermed(n);
    for (int c = 0; c < n; ++c) cpermed[c] = lab[cperm[c]];
    rows_by_weight(A, cpermed, cperm, rlab);
  } else {
    rlab = lab;
  }
  // interfaces (partition.cpp:506-578): pass 1 from lower-level neighbours
  for (int r = 0; r < A.m; ++r) {
    const SepID self = rlab[r].self;
    if (self.lvl <= 0) continue;
    SepID left, right;
    for (int t = R.p[r]; t < R.p[r + 1]; ++t) {
      const SepID nb = lab[R.i[t]].self;
      if (nb.lvl >= self.lvl) continue;
      const int side = (int)(nb.sep / std::pow(2.0, self.lvl - nb.lvl - 1));
      if (side % 2 == 0) left = (left.lvl == -1) ? nb : lower_of(nb, left);
      else right = (right.lvl == -1) ? nb : lower_of(nb, right);
    }
    rlab[r].l = left;
    rlab[r].r = right;
  }
  std::vector<CID> clab = rlab;  // row i matched to column i
  // pass 2: a missing side is borrowed from same-separator neighbours
  for (int r = 0; r < A.m; ++r) {
    const SepID self = rlab[r].self;
    SepID left = rlab[r].l, right = rlab[r].r;
    if (right == SepID()) {
      for (int t = R.p[r]; t < R.p[r + 1]; ++t) {
        const CID& c = clab[R.i[t]];
        if (c.self == self && c.r != SepID()) right = (right == SepID()) ? c.r : climb_common(right, c.r);
      }
    } else if (left == SepID()) {
      for (int t = R.p[r]; t < R.p[r + 1]; ++t) {
        const CID& c = clab[R.i[t]];
        if (c.self == self && c.l != SepID()) left = (left == SepID()) ? c.l : climb_common(left, c.l);
      }
    }
    rlab[r].l = left;
    rlab[r].r = right;
  }
  for (auto& c : rlab) {
    if (c.l == SepID()) c.l = c.self;
    if (c.r == SepID()) c.r = c.self;
  }
  return rlab;
}

void order_columns(int nlevels, const std::vector<CID>& labels, std::vector<int>& cperm) {
  std::vector<CID> key = labels;
  auto by_key = [&key](int a, int b) { return key[a] < key[b]; };
  std::stable_sort(cperm.begin(), cperm.end(), by_key);
  for (int lvl = 1; lvl < nlevels; ++lvl) {
    for (auto& c : key) c = cid_merge_if(c, lvl);
    std::stable_sort(cperm.begin(), cperm.end(), by_key);
  }
}

void order_rows_ls(const CscMat& Ap, const std::vector<CID>& cpermed, std::vector<int>& rperm,
                   std::vector<int>& rows_per_col) {
  const int m = Ap.m, n = Ap.n;
  const CscMat R = csc_transpose(Ap);
  std::vector<std::vector<int>> owned(n);
  rows_per_col.assign(n, 0);
  for (int r = 0; r < m; ++r) {
    double bestw = -1;
    int pick = -1;
    int q = R.p[r];
    const int e = R.p[r + 1];
    while (q < e) {
      const CID& g = cpermed[R.i[q]];
      double w = 0;
      int t = q;
      while (t < e && cpermed[R.i[t]] == g) { w += R.x[t] * R.x[t]; ++t; }
      if (w > bestw) { bestw = w; pick = R.i[q]; }
      q = t;
    }
    if (pick == -1) pick = n - 1;
    rows_per_col[pick]++;
    owned[pick].push_back(r);
  }
  rperm.clear();
  rperm.reserve(m);
  for (int c = 0; c < n; ++c)
    for (int r : owned[c]) rperm.push_back(r);
}

}  // namespace spq
