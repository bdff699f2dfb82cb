// ORACLE — test infrastructure only.  CPU restatement of the reference spaQR
// (Abeynaya/spaQR_public): geometric nested-dissection partition, cluster
// tree, per-level factorization phases, the solve op tape and the Krylov
// drivers.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it, and only as the checker.
#pragma once
#include <list>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "oracle_lapack.h"

namespace orc {

// Compressed sparse column matrix with sorted row indices (Eigen's compressed
// SparseMatrix<double, ColMajor, int> layout).
struct Csc {
  int m = 0, n = 0;
  std::vector<int> colptr, rowind;
  std::vector<double> val;
  int nnz() const { return (int)rowind.size(); }
};
Csc csc_transpose(const Csc& A);
Csc csc_colperm(const Csc& A, const std::vector<int>& p);  // util.cpp:912-940
Csc csc_ata_pattern(const Csc& A);                          // structure of A^T A
void csc_matvec(const Csc& A, const double* x, double* y);  // y = A x
void csc_matvec_t(const Csc& A, const double* x, double* y);  // y = A^T x
// spaQR.cpp:131-147: D_jj = 1/||A(:,j)||, A <- (A D) 10 ; returns D.
std::vector<double> colscale(Csc& A);
// util.cpp:852-881 ; X is dim x n^dim col-major.
std::vector<double> linspace_nd(int n, int dim);
// util.cpp:943-952
std::vector<double> random_vec(int size, int seed);

struct SepID {  // cluster.h:22-40
  int lvl = -1, sep = 0;
  SepID() {}
  SepID(int l, int s) : lvl(l), sep(s) {}
  bool operator==(const SepID& o) const { return lvl == o.lvl && sep == o.sep; }
  bool operator!=(const SepID& o) const { return !(*this == o); }
  bool operator<(const SepID& o) const { return lvl < o.lvl || (lvl == o.lvl && sep < o.sep); }
};
struct ClusterID {  // cluster.h:43-111
  SepID self, l, r;
  int part = 2, section = 0;
  bool operator==(const ClusterID& o) const {
    return self == o.self && l == o.l && r == o.r && part == o.part && section == o.section;
  }
  bool operator!=(const ClusterID& o) const { return !(*this == o); }
  bool operator<(const ClusterID& o) const {
    if (self != o.self) return self < o.self;
    if (l != o.l) return l < o.l;
    if (r != o.r) return r < o.r;
    if (section != o.section) return section < o.section;
    return part < o.part;
  }
};
ClusterID merge_if(const ClusterID& c, int lvl);  // cluster.cpp:23-32
// partition.cpp:450-499 + 582-623 from a row partition (parts[i] in
// [0, 2^(nlevels-1))): column labels of the non-geometric path.
std::vector<ClusterID> hypergraph_partition(const Csc& A, int nlevels, const std::vector<int>& parts);

// Partition (partition.cpp:241-348, 506-578 ; tree.cpp:46-163)
std::vector<ClusterID> geometric_partition_ata(const Csc& A, int nlevels, const std::vector<double>& X, int dim);
void init_permutation(int nlevels, const std::vector<ClusterID>& cmap, std::vector<int>& cperm);

struct Cluster;
struct Edge {  // edge.h:14-34 : A21 = A(rows of n2, cols of n1)
  Cluster* n1;
  Cluster* n2;
  std::unique_ptr<Mat> A21;
  Edge(Cluster* a, Cluster* b, Mat* m) : n1(a), n2(b), A21(m) {}
};
struct ByUid {
  bool operator()(const Cluster* a, const Cluster* b) const;
};
using CSet = std::set<Cluster*, ByUid>;

struct Cluster {  // cluster.h:118-248
  int uid;
  int cstart, csize, rstart, rsize;
  int csize_org, rsize_org;
  ClusterID id;
  int order;
  bool eliminated = false;
  Cluster* parent = nullptr;
  ClusterID parentid;
  std::vector<Cluster*> children;
  int rposparent = 0, cposparent = 0;
  CSet rsparsity, cnbrs;
  std::list<Edge*> edgesOut, edgesIn;
  std::vector<double> tau;
  Mat T;
  std::vector<double> x, xt;
  Cluster(int uid_, int cs, int cn, int rs, int rn, ClusterID id_, int ord)
      : uid(uid_), cstart(cs), csize(cn), rstart(rs), rsize(rn), csize_org(cn), rsize_org(rn), id(id_), order(ord) {
    set_size(rn, cn);
  }
  int level() const { return id.self.lvl; }
  int cols() const { return csize; }
  int rows() const { return rsize; }
  void set_size(int r, int c) { rsize = r; csize = c; x.assign(r, 0.0); xt.assign(r, 0.0); }
  void reset_size(int r, int c) { rsize = r; csize = c; }
  void resize_x(int r) { x.resize(x.size() + r, 0.0); xt.resize(xt.size() + r, 0.0); }
  void sort_edgesOut(bool reverse = false);
};

// Solve op tape (operations.h / toperations.h).  Segments are (cluster, start,
// length) captured when the op is created, exactly like Eigen Segment views.
struct Op {
  enum Kind { QR, Reassign, ScaleD, Orthogonal, OrthogonalD, Merge, Split, SplitD } kind;
  Cluster* c = nullptr;
  Cluster* n = nullptr;  // Reassign target / fine cluster
  int nrows = 0, ncols = 0, ccols = 0, crows = 0, rank = 0, nstart = 0;
  std::vector<int> a21_sizes, a12_sizes, indices;
  int xs_len = 0, xsf_len = 0;  // head()/full() lengths at creation
  std::shared_ptr<Mat> Q;        // ScaleD (square: Q with R on top; LS: R)
  std::shared_ptr<Mat> V;        // Orthogonal(D)
  std::shared_ptr<std::vector<double>> tau;
};

struct LevelStat {
  double t_fill = 0, t_elmn = 0, t_shift = 0, t_scale = 0, t_sprsfy = 0, t_merge = 0;
  int top_rows = 0, top_cols = 0;
};
struct SparsifyRec {  // one sparsified interface: level, sep, cols before/after
  int ilvl, lvl, sep, before, after;
};

class Tree {
 public:
  Tree(int lvls, int skip_);
  ~Tree();
  void set_tol(double t) { tol = t; }
  void set_skip(int s) { skip = s; }
  void set_scale(int s) { scale = s; }
  void set_square(int s) { square = s; }
  void set_order(float s) { order = s; }
  void set_Xcoo(const std::vector<double>& X, int dim) { Xcoo = X; xdim = dim; geo = true; }
  // row partition of the non-geometric path (tree.cpp:608-617)
  void set_row_parts(const std::vector<int>& p) { row_parts = p; }
  void set_verbose(int v) { verbose = v; }
  int rows() const { return nrows; }
  int cols() const { return ncols; }
  int levels() const { return nlevels; }
  void partition(const Csc& A);
  void assemble(const Csc& A);
  int factorize();
  void solve(const std::vector<double>& b, std::vector<double>& x) const;
  void solve_nrml(const std::vector<double>& b, std::vector<double>& x) const;

  // Inspection for tests
  std::vector<int> cperm, rperm;
  std::vector<LevelStat> stats;
  std::vector<SparsifyRec> sparsified;
  // sparsify_extra calls that ran an RRQR (tree.cpp:1283-1286): level, sep,
  // extra rows before and after the compression
  std::vector<SparsifyRec> extra_sparsified;
  unsigned long long nnzA = 0, nnzR = 0, nnzH = 0, nnzQ = 0;
  double flops = 0;  // LAWN-41 ledger of the dense kernels
  double t_partition = 0, t_assemble = 0, t_factorize = 0;
  int top_size() const;
  std::vector<std::vector<int>> cluster_table() const;  // per leaf+parent: lvl, sep, l, r, cstart, csize, rstart, rsize, order

 private:
  int ilvl = -1, nlevels, nrows = 0, ncols = 0;
  double tol = 0;
  int skip, scale = 0, square = 0, max_order = 0;
  float order = 1;
  int verbose = 0;
  bool geo = false;
  std::vector<double> Xcoo;
  int xdim = 0;
  std::vector<int> row_parts;
  int uid_counter = 0;
  std::vector<std::list<Cluster*>> bottoms;
  std::vector<std::unique_ptr<Cluster>> owned;
  std::vector<std::unique_ptr<Edge>> edges_owned;
  int current_bottom = 0;
  std::vector<Op> ops, tops;

  Cluster* new_cluster(int cs, int cn, int rs, int rn, ClusterID id);
  Edge* new_edge(Cluster* a, Cluster* b, Mat* m);
  int get_new_order() { return max_order++; }
  const std::list<Cluster*>& bottom_current() const { return bottoms[current_bottom]; }
  const std::list<Cluster*>& bottom_original() const { return bottoms[0]; }
  std::pair<int, int> topsize() const;
  Edge* add_edge(Cluster* c1, Cluster* c2);
  bool check_if_valid(ClusterID cid, ClusterID nid, const std::set<SepID>& col_sparsity) const;
  void get_sparsity(Cluster* c);
  void eliminate_cluster(Cluster* c);
  void update_cluster(Cluster* c, Cluster* n, const Mat& Q);
  void reassign_rows(Cluster* c);
  bool want_sparsify(Cluster* c) const;
  void diagScale(Cluster* c);
  void sparsify(Cluster* c);
  void sparsifyD(Cluster* c, bool imp);
  void sparsify_extra(Cluster* c);
  void split_edges(Cluster* c, Cluster* f);
  void merge_all();
  void update_size(Cluster* snew);
  void update_edges(Cluster* snew);
  void op_fwd(const Op& o) const;
  void op_bwd(const Op& o) const;
  void top_fwd(const Op& o) const;
  void top_bwd(const Op& o) const;
  Op make_qr(Cluster* c) const;
};

// is.cpp:95-272 (Householder GMRES, left preconditioned) ; returns iterations.
int gmres(const Csc& A, const std::vector<double>& b, std::vector<double>& x, const Tree& M, int iters,
          int restart, double tol, double* tol_out, int verbose);
// is.cpp:275-346 ; returns iterations.
int cgls(const Csc& A, const std::vector<double>& b, std::vector<double>& x, const Tree& M, int iters,
         double tol, double* tol_out, int verbose);

}  // namespace orc
