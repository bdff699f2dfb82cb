// spaQR engine implementation.  Reference behaviour cited per function
// (tree.cpp line numbers).  Cluster sets that the reference keeps as
// std::set<Cluster*> are kept sorted by creation sequence, which is the address
// order of the reference's consecutive `new Cluster` allocations.
#include "engine.hpp"
#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <type_traits>

#include "../kernels/launch.hpp"
#include "tuner.hpp"

namespace spq {

namespace {
inline int iceil(int a, int b) { return (a + b - 1) / b; }
}  // namespace

template <typename T>
class Engine final : public EngineBase {
 public:
  Engine(int nlevels, int skip, int device) : nlevels_(nlevels), skip_(skip), perm_(size_t(512) << 20),
                                               scratch_(size_t(256) << 20), ws_(size_t(64) << 20) {
    SPQ_HIP(hipSetDevice(device));
    SPQ_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    stager_.init(stream_, size_t(64) << 20);
    bottoms_.resize(nlevels);
    info_.resize(nlevels);
    nb_ = panel_nb();
    for (auto& row : qp3r_cap_)
      for (int& c : row) c = -1;
    SPQ_HIP(host_malloc((void**)&rank_host_, sizeof(int) * 65536));
    qp3_err_ = (int*)perm_.alloc(256);
    SPQ_HIP(hipMemset(qp3_err_, 0, 256));
    if (const char* e = std::getenv("SPAQR_QP3_MAX_WG")) qp3_max_wg_ = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("SPAQR_QP3_REGISTER")) qp3_reg_ = std::atoi(e) != 0;
    if (const char* e = std::getenv("SPAQR_QP3_LOG")) qp3_log_ = std::atoi(e) != 0;
    if (const char* e = std::getenv("SPAQR_REFL_REGISTER")) reflr_ = std::atoi(e) != 0;
    SPQ_HIP(hipDeviceGetAttribute(&num_cus_, hipDeviceAttributeMultiprocessorCount, device));
    {
      hipDeviceProp_t prop;
      SPQ_HIP(hipGetDeviceProperties(&prop, device));
      std::string arch = prop.gcnArchName;
      global_tuner().init(arch.substr(0, arch.find(':')));
    }
    if (const char* e = std::getenv("SPAQR_SOLVE_STAGES")) stage_prof_ = std::atoi(e) != 0;
    if (const char* e = std::getenv("SPAQR_SOLVE_GRAPH")) solve_graph_ = std::atoi(e) != 0;
    if (const char* e = std::getenv("SPAQR_ELIM_COLOR")) elim_color_ = std::atoi(e) != 0;
    if (const char* e = std::getenv("SPAQR_QP3_RC")) {
      const int r = std::atoi(e);
      for (int c : qp3r_rcs())
        if (r == c) qp3_rc_ = r;
    }
  }
  ~Engine() override {
    (void)hipStreamSynchronize(stream_);
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    std::lock_guard<std::recursive_mutex> lk(capture_lock());
    if (graph_exec_fwd_) hipGraphExecDestroy(graph_exec_fwd_);
    if (graph_exec_nrml_) hipGraphExecDestroy(graph_exec_nrml_);
    if (graph_exec_t_) hipGraphExecDestroy(graph_exec_t_);
    dev_free(xtbuf_);
    dev_free(x64_);
    host_free(rank_host_);
    dev_free(xbuf_);
    dev_free(bin_);
    dev_free(xout_);
    hipStreamDestroy(stream_);
  }

  void set_tol(double t) override { tol_ = t; }
  void set_skip(int s) override { skip_ = s; }
  void set_scale(int s) override { scale_ = s; }
  void set_square(int s) override { square_ = s; }
  void set_order(float o) override { order_ = o; }
  void set_coords(const std::vector<double>& X, int dim) override { X_ = X; xdim_ = dim; geo_ = true; }
  void set_row_parts(const std::vector<int>& parts) override { row_parts_ = parts; }
  void set_verbose(int v) override { verbose_ = v; }
  void set_timing(int t) override { timing_ = t; }
  int rows() const override { return nrows_; }
  int cols() const override { return ncols_; }
  int levels() const override { return nlevels_; }
  hipStream_t stream() const override { return stream_; }
  const std::vector<int>& cperm() const override { return cperm_; }
  const std::vector<int>& rperm() const override { return rperm_; }
  const std::vector<LevelInfo>& level_info() const override { return info_; }
  const std::vector<SparsifyInfo>& sparsified() const override { return sprs_; }
  const std::vector<SparsifyInfo>& extra_sparsified() const override { return extra_; }
  const std::vector<MergeInfo>& merged() const override { return merges_; }
  int skip() const override { return skip_; }
  int precision_bytes() const override { return (int)sizeof(T); }

  void set_kernel_timing(int on) override { kprof_ = on != 0; }
  void check_solve() override {
    if (!solve_err_) return;
    int e = 0;
    SPQ_HIP(hipMemcpyAsync(&e, solve_err_, sizeof(int), hipMemcpyDeviceToHost, stream_));
    SPQ_HIP(hipStreamSynchronize(stream_));
    if (e) throw Error(kErrNumeric, "solve: an inter-workgroup barrier wait timed out");
  }
  void reset_kernel_stats() override {
    SPQ_HIP(hipStreamSynchronize(stream_));
    kev_.clear();
    ev_used_ = 0;  // the pool's events are recorded again
  }
  std::vector<KernelStat> kernel_stats() override {
    SPQ_HIP(hipStreamSynchronize(stream_));
    std::vector<KernelStat> out;
    for (auto& kv : kev_) {
      KernelStat k;
      k.name = kv.first;
      k.launches = (int)kv.second.ev.size();
      k.flops = kv.second.flops;
      k.bytes = kv.second.bytes;
      for (auto& e : kv.second.ev) {
        float ms = 0;
        SPQ_HIP(hipEventElapsedTime(&ms, e.first, e.second));
        k.ms += ms;
      }
      out.push_back(k);
    }
    return out;
  }

  void partition(const CscMat& A) override;
  void assemble(const CscMat& A) override;
  int factorize() override;
  void solve_dev(const double* b, double* x) override;
  void solve_nrml_dev(const double* b, double* x) override;
  void solve_t_dev(const double* b, double* x) override;

  std::vector<std::array<int, 12>> cluster_table() const override {
    std::vector<std::array<int, 12>> t;
    for (int l = 0; l < nlevels_; ++l)
      for (int c : bottoms_[l]) {
        const Clu& s = cl_[c];
        t.push_back({l, s.id.self.lvl, s.id.self.sep, s.id.l.lvl, s.id.l.sep, s.id.r.lvl, s.id.r.sep, s.cstart,
                     s.csize_org, s.rstart, s.rsize_org, s.order});
      }
    return t;
  }
  void counters(unsigned long long* nnz4, double* d8) const override {
    nnz4[0] = nnzA_; nnz4[1] = nnzR_; nnz4[2] = nnzH_; nnz4[3] = nnzQ_;
    d8[0] = flops_; d8[1] = t_partition_; d8[2] = t_assemble_; d8[3] = t_factorize_;
    d8[4] = top_size_; d8[5] = (double)perm_.reserved(); d8[6] = factor_bytes_; d8[7] = solve_bytes_;
  }

 private:
  template <typename U>
  struct BlockT {
    U* p = nullptr;
    int rows = 0, cols = 0, ld = 0;
  };
  using Block = BlockT<T>;
  struct Edge {
    int n1, n2;
    Block a;
    bool alive = true;
  };
  struct Clu {
    int cstart = 0, csize = 0, rstart = 0, rsize = 0, csize_org = 0, rsize_org = 0;
    CID id;
    int order = 0;
    bool elim = false;
    int parent = -1;
    CID parentid;
    std::vector<int> children;
    int rpos = 0, cpos = 0;
    std::vector<int> out, in;  // edge ids in reference list order
    std::vector<int> outn;     // n2 of each out edge (same order as out; scanned by find_out)
    std::vector<int> inn;      // n1 of each in edge (same order as in)
    std::vector<int> rsp;      // rsparsity, creation order
    std::vector<int> cnbrs;    // sorted, creation order
    int64_t xoff = -1;
    int xcap = 0;
    int level() const { return id.self.lvl; }
  };
  // Solve op records (sizes captured at creation, operations.h / toperations.h)
  // OpFineQR: the QR op of an order-1.5 fine cluster (tree.cpp:1252-1253):
  // tau = 0, so only its backward R12 / R sweep does work
  // OpReassign: least-squares rows moved from an eliminated cluster to a
  // neighbour (operations.cpp:59-68, toperations.cpp:56-66)
  enum OpKind { OpQR, OpScaleD, OpOrthoD, OpSplitD, OpMerge, OpFineQR, OpReassign };
  struct Op {
    OpKind kind;
    int level, c = -1, f = -1;
    // reflector data
    const T* V = nullptr;
    const T* Tb = nullptr;
    int ldv = 0, m = 0, k = 0;
    std::vector<std::pair<int, int>> segs;  // (cluster, length)
    // backward data
    const T* R = nullptr;
    int ldr = 0, kr = 0;
    std::vector<std::pair<Block, std::pair<int, int>>> r12;  // block, (cluster, s)
    // SplitD
    int crows = 0, ccols = 0, rank = 0;
    int wave = 0;  // sparsify wave (OrthoD, SplitD, FineQR); elimination wave (QR)
    bool nofwd = false;     // OrthoD without a forward step (Orthogonal, operations.h:113-129)
    const double* V64 = nullptr;   // FP64 factors of the top separator of an FP32 tree (V and R
    const double* Tb64 = nullptr;  // share the panel, as V / R do)
    bool fwd_refl = false;  // FineQR with reflectors (the --scale 0 fine elimination)
    // Reassign: x_f[nstart + t] = x_c[idx[t]]
    std::vector<int> idx;
    int nstart = 0;
  };

  // ------------------------------------------------------------ helpers
  int new_cluster(int cs, int cn, int rs, int rn, const CID& id) {
    Clu c;
    c.cstart = cs; c.csize = cn; c.rstart = rs; c.rsize = rn; c.csize_org = cn; c.rsize_org = rn;
    c.id = id;
    c.order = max_order_++;
    cl_.push_back(c);
    return (int)cl_.size() - 1;
  }
  // A cluster's x segment needs room for the most rows it holds during its
  // one-level lifetime (reassigned least-squares rows grow it, tree.cpp:365);
  // offsets are laid out when the solve program is built.
  void alloc_x(int c, int cap) { cl_[c].xcap = std::max(cl_[c].xcap, cap); }
  void layout_x() {
    xsize_ = 0;
    for (auto& c : cl_) {
      if (c.xcap > 0) { c.xoff = xsize_; xsize_ += c.xcap; }
      else c.xoff = 0;
    }
  }
  T* alloc_perm(size_t n) { return (T*)perm_.alloc(sizeof(T) * std::max<size_t>(n, 1)); }
  T* alloc_scr(size_t n) { return (T*)scratch_.alloc(sizeof(T) * std::max<size_t>(n, 1)); }
  Block new_block(int rows, int cols) {
    Block b;
    b.rows = rows; b.cols = cols; b.ld = std::max(rows, 1);
    b.p = alloc_perm((size_t)b.ld * std::max(cols, 1));
    return b;
  }
  // link_nbr / link_in = false: the edge joins c1's out list only (the fine
  // cluster edges of split_edges, tree.cpp:480-530, which touch neither the
  // column neighbours of c1 nor the in list of c2 unless stated)
  int add_edge(int c1, int c2, std::vector<Block>* zero_list, bool link_nbr = true, bool link_in = true) {  // tree.cpp:251-261
    Edge e;
    e.n1 = c1; e.n2 = c2;
    e.a = new_block(cl_[c2].rsize, cl_[c1].csize);
    if (zero_list) zero_list->push_back(e.a);
    ed_.push_back(e);
    const int id = (int)ed_.size() - 1;
    if (link_nbr) {
      auto& cn = cl_[c1].cnbrs;
      auto it = std::lower_bound(cn.begin(), cn.end(), c2);
      if (it == cn.end() || *it != c2) cn.insert(it, c2);
    }
    cl_[c1].out.push_back(id);
    cl_[c1].outn.push_back(c2);
    if (c1 != c2 && link_in) {
      cl_[c2].in.push_back(id);
      cl_[c2].inn.push_back(c1);
    }
    return id;
  }
  int find_out(int c, int n2) const {
    const Clu& k = cl_[c];
    for (size_t q = 0; q < k.outn.size(); ++q)
      if (k.outn[q] == n2) return k.out[q];
    return -1;
  }
  void sort_out(int c) {
    // stable sort by the column cluster's order: (order, position) keys,
    // sorted without allocating for the short lists of the merge
    auto& o = cl_[c].out;
    static thread_local std::vector<std::pair<std::pair<int, int>, int>> key;
    key.resize(o.size());
    for (size_t q = 0; q < o.size(); ++q) key[q] = {{cl_[ed_[o[q]].n2].order, (int)q}, o[q]};
    std::sort(key.begin(), key.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    auto& on = cl_[c].outn;
    on.resize(o.size());
    for (size_t q = 0; q < o.size(); ++q) { o[q] = key[q].second; on[q] = ed_[o[q]].n2; }
  }
  std::pair<int, int> topsize() const {
    int r = 0, c = 0;
    for (int s : bottoms_[cur_])
      if (cl_[s].level() == nlevels_ - 1) { r += cl_[s].rsize; c += cl_[s].csize; }
    return {r, c};
  }
  bool want_sparsify(int c) const {  // tree.cpp:583-591
    const Clu& s = cl_[c];
    return s.level() > ilvl_ && s.id.l.lvl <= ilvl_ && s.id.r.lvl <= ilvl_;
  }
  void sync() {
    stager_.flush_uploads();
    const double a = wall();
    SPQ_HIP(hipStreamSynchronize(stream_));
    t_wait_ += wall() - a;
    stager_.reset();
  }
  double t_wait_ = 0;  // host time blocked in stream synchronisations (diagnostics)
  double t_flush_ = 0, t_launch_ = 0;  // host time in descriptor uploads / launch calls (verbose > 1)
  long n_launch_ = 0;

  // Records a HIP event pair around one launch of a kernel family when kernel
  // timing is on (events are recorded on the engine's stream, where the
  // kernels run).
  template <typename F>
  void timed(const char* name, double flops, double bytes, F&& launch) {
    const double ta = verbose_ > 1 ? wall() : 0.0;
    stager_.flush();
    if (verbose_ > 1) {
      const double tb = wall();
      t_flush_ += tb - ta;
      ++n_launch_;
    }
    if (!kprof_) {
      const double tb = verbose_ > 1 ? wall() : 0.0;
      launch();
      if (verbose_ > 1) t_launch_ += wall() - tb;
      return;
    }
    const double tb = verbose_ > 1 ? wall() : 0.0;
    hipEvent_t a = pooled_event(), b = pooled_event();
    SPQ_HIP(hipEventRecord(a, stream_));
    launch();
    SPQ_HIP(hipEventRecord(b, stream_));
    if (verbose_ > 1) t_launch_ += wall() - tb;
    auto& k = kev_[name];
    k.ev.push_back({a, b});
    k.flops += flops;
    k.bytes += bytes;
  }

  // --------------------------------------------------- batched launch helpers
  // esrc / edst: element bytes of source and destination (a conversion when
  // they differ)
  // Launches planned ahead: tables are pushed to the staging ring when a
  // launch is planned, and the launches of a phase are issued together
  // afterwards, so their tables reach the device in one copy.
  using Run = std::vector<std::function<void()>>;
  // cnt[i] = f(i, bytes) for i < n (cnt = tile_bufs().cnt, sized by the
  // caller) on the host pool; returns the bytes f adds up, summed per chunk
  // and then in chunk order
  template <typename F>
  double count_tiles(size_t n, F&& f) {
    std::vector<int>& cnt = tile_bufs().cnt;
    const int grain = 8192, nch = std::max(plan_chunks((int)n, grain), 1);
    std::vector<double> part(nch, 0.0);
    parallel_for((int)n, grain, [&](int ch, int b, int e) {
      double by = 0;
      for (int i = b; i < e; ++i) cnt[i] = f((size_t)i, by);
      part[ch] = by;
    });
    double bytes = 0;
    for (double v : part) bytes += v;
    return bytes;
  }
  void copy_blocks(const std::vector<CopyDesc>& ds, int esrc = (int)sizeof(T), int edst = (int)sizeof(T)) {
    copy_blocks(ds.data(), ds.size(), esrc, edst);
  }
  void copy_blocks(const CopyDesc* ds, size_t n, int esrc = (int)sizeof(T), int edst = (int)sizeof(T)) {
    if (auto f = copy_plan(ds, n, esrc, edst)) f();
  }
  std::function<void()> copy_plan(const CopyDesc* ds, size_t n, int esrc = (int)sizeof(T),
                                  int edst = (int)sizeof(T)) {
    if (n == 0) return {};
    std::vector<int>& cnt = tile_bufs().cnt;
    std::vector<int>& st = tile_bufs().st;
    cnt.resize(n);
    const double bytes = count_tiles(n, [&](size_t i, double& by) {
      by += (double)(esrc + edst) * ds[i].rows * ds[i].cols;
      return iceil(std::max(ds[i].rows, 1), 32) * iceil(std::max(ds[i].cols, 1), 32);
    });
    const int nt = prefix_tiles(cnt, st);
    const CopyDesc* dd = (const CopyDesc*)stager_.push_raw(ds, sizeof(CopyDesc) * n);
    const int* dst = stager_.push(st);
    const int nd = (int)n;
    return [this, bytes, esrc, edst, dd, dst, nd, nt] {
      timed("copy", 0.0, bytes, [&] { launch_copy(stream_, esrc, edst, dd, dst, nd, nt); });
    };
  }
  void init_blocks(const std::vector<Block>& bs, int identity) {
    if (bs.empty()) return;
    std::vector<CopyDesc> ds;
    ds.reserve(bs.size());
    for (auto& b : bs) {
      if (b.rows == 0 || b.cols == 0) continue;
      CopyDesc d{};
      d.dst = b.p; d.ldd = b.ld; d.rows = b.rows; d.cols = b.cols; d.trans = identity;
      ds.push_back(d);
    }
    if (ds.empty()) return;
    std::vector<int>& cnt = tile_bufs().cnt;
    std::vector<int>& st = tile_bufs().st;
    cnt.resize(ds.size());
    const double bytes = count_tiles(ds.size(), [&](size_t i, double& by) {
      by += (double)sizeof(T) * ds[i].rows * ds[i].cols;
      return iceil(ds[i].rows, 32) * iceil(ds[i].cols, 32);
    });
    const int nt = prefix_tiles(cnt, st);
    const CopyDesc* dd = stager_.push(ds);
    const int* dst = stager_.push(st);
    timed("init", 0.0, bytes, [&] { launch_init_blocks(stream_, sizeof(T), dd, dst, (int)ds.size(), nt); });
  }
  // Batched GEMM.  When the batch has few output tiles, descriptors with a
  // long K are split along K (partials in ws_, then an ordered reduce) so the
  // launch fills the device.
  // fam: timing family (every family runs k_gemm, plus k_gemm_reduce)
  template <typename U>
  void gemm_batch(std::vector<GemmDesc<U>>& ds, const char* fam = "gemm") {  // ds may get split-K fields
    gemm_batch(ds.data(), ds.size(), fam);
  }
  template <typename U>
  void gemm_batch(GemmDesc<U>* ds, size_t n, const char* fam = "gemm") {
    gemm_launch(gemm_plan(ds, n, fam));
  }
  // A batch's tables pushed to the staging ring, ready to launch (plan and
  // launch apart let a caller push the tables of several launches before the
  // first of them, so they reach the device in one copy)
  template <typename U>
  struct GemmLaunch {
    const char* fam = nullptr;
    double fl = 0, by = 0;
    const GemmDesc<U>* dd = nullptr;
    const int* dst = nullptr;
    int n = 0, nt = 0;
    const GemmDesc<U>* drd = nullptr;
    const int* drst = nullptr;
    int rn = 0, rnt = 0;
  };
  template <typename U>
  void gemm_launch(const GemmLaunch<U>& L) {
    if (L.n == 0) return;
    timed(L.fam, L.fl, L.by, [&] {
      launch_gemm<U>(stream_, L.dd, L.dst, L.n, L.nt);
      if (L.rn) launch_gemm_reduce<U>(stream_, L.drd, L.drst, L.rn, L.rnt);
    });
  }
  template <typename U>
  GemmLaunch<U> gemm_plan(GemmDesc<U>* ds, size_t n, const char* fam) {
    GemmLaunch<U> L;
    L.fam = fam;
    if (n == 0) return L;
    std::vector<int>& cnt = tile_bufs().cnt;
    cnt.resize(n);
    // tile counts and the ledger's flops / bytes, in chunks joined in order
    const int grain = 4096, nch = std::max(plan_chunks((int)n, grain), 1);
    std::vector<int> ctiles(nch, 0);
    std::vector<double> cfl(nch, 0.0), cby(nch, 0.0);
    parallel_for((int)n, grain, [&](int ch, int b, int e) {
      int t = 0;
      double f = 0, y = 0;
      for (int i = b; i < e; ++i) {
        const GemmDesc<U>& d = ds[i];
        t += cnt[i] = gemm_tiles(d.m, d.n);
        f += 2.0 * d.m * d.n * d.k;
        y += sizeof(U) * ((double)d.m * d.k + (double)d.k * d.n + (d.beta == U(0) ? 1.0 : 2.0) * d.m * d.n);
      }
      ctiles[ch] = t; cfl[ch] = f; cby[ch] = y;
    });
    int tiles = 0;
    double fl = 0, by = 0;
    for (int ch = 0; ch < nch; ++ch) { tiles += ctiles[ch]; fl += cfl[ch]; by += cby[ch]; }
    const int want = 4 * num_cus_;
    bool split = false;
    if (tiles < want) {
      // partials live in gws_ until this batch's reduce has run; the next
      // batch is ordered after it on the stream, so the arena is reused.
      gws_.rewind();
      size_t room = kSplitWsBytes;  // bytes, allocations rounded to 256
      for (size_t i = 0; i < n; ++i) {
        auto& d = ds[i];
        if (d.m <= 0 || d.n <= 0) continue;
        const size_t mn = (size_t)d.m * d.n;
        int ks = std::min(std::max(1, want / std::max(tiles, 1)), std::max(1, d.k / 256));
        ks = (int)std::min<size_t>(ks, room / (sizeof(U) * mn + 256));
        if (ks > 1) {
          const size_t bytes = (sizeof(U) * (size_t)ks * mn + 255) / 256 * 256;
          d.ksplit = ks;
          d.work = (U*)gws_.alloc(bytes);
          room -= bytes;
          split = true;
        }
      }
    }
    for (size_t i = 0; i < n; ++i)
      if (ds[i].ksplit > 1) cnt[i] *= ds[i].ksplit;
    std::vector<int>& st = tile_bufs().st;
    const int nt = prefix_tiles(cnt, st);
    const GemmDesc<U>* dd = (const GemmDesc<U>*)stager_.push_raw(ds, sizeof(GemmDesc<U>) * n);
    const int* dst = stager_.push(st);
    std::vector<GemmDesc<U>> rd;
    std::vector<int> rcnt, rst;
    if (split)
      for (size_t i = 0; i < n; ++i)
        if (ds[i].ksplit > 1) { rd.push_back(ds[i]); rcnt.push_back(iceil(ds[i].m * ds[i].n, 256)); }
    const int rnt = prefix_tiles(rcnt, rst);
    L.fl = fl; L.by = by;
    L.dd = dd; L.dst = dst; L.n = (int)n; L.nt = nt;
    if (!rd.empty()) {
      L.drd = stager_.push(rd);
      L.drst = stager_.push(rst);
      L.rn = (int)rd.size();
      L.rnt = rnt;
    }
    return L;
  }

  template <typename U>
  static GemmDesc<U> gd(const U* A, int lda, const U* B, int ldb, U* C, int ldc, int m, int n, int k, int flags,
                        U alpha, U beta) {
    GemmDesc<U> d;
    d.A = A; d.B = B; d.C = C; d.lda = lda; d.ldb = ldb; d.ldc = ldc; d.m = m; d.n = n; d.k = k;
    d.flags = flags; d.alpha = alpha; d.beta = beta;
    return d;
  }

  // Blocked Householder QR of a batch of panels.  For every block step the
  // panel kernel factors NB columns, then the block reflector is applied to the
  // panel's trailing columns and to every extra matrix `extra[p]` (same row
  // space as the panel, ld given) with three batched GEMMs:
  //   Y = V_b^T C ; Z = T_b^T Y ; C -= V_b Z.
  template <typename U>
  struct QrJobT {
    U* P; int ld, m, k;
    U* tau; U* Tb;
    std::vector<BlockT<U>> extra;  // matrices receiving H^T (rows = m)
  };
  using QrJob = QrJobT<T>;
  // U is the element type of the batch (T, or double for the FP64 top
  // separator of an FP32 tree)
  template <typename U>
  void blocked_qr(std::vector<QrJobT<U>>& jobs) {
    Run run;
    blocked_qr_plan(jobs, run);
    for (auto& f : run) f();
  }
  // Pushes every block step's tables and appends the steps' launches to
  // `run` (to be issued in order, after any launches already there; a large
  // step issues `run` itself).  Step jb + 1 reuses step jb's scratch in ws_
  // (the stream orders it after step jb's kernels).
  template <typename U>
  void blocked_qr_plan(std::vector<QrJobT<U>>& jobs, Run& run) {
    int kmax = 0;
    for (auto& j : jobs) kmax = std::max(kmax, std::min(j.m, j.k));
    // panel steps run in the register form: rows [jb, m) over nwg
    // workgroups; barrier counters of all multi-workgroup steps come from one
    // pool cleared by one launch
    const int rpw = panel_rows_per_wg();
    size_t words = 0;
    for (int jb = 0; jb < kmax; jb += nb_)
      for (auto& j : jobs)
        if (jb < std::min(j.m, j.k) && iceil(j.m - jb, rpw) > 1) words += barrier_words(iceil(j.m - jb, rpw));
    unsigned int* pool = words ? (unsigned int*)scratch_.alloc(sizeof(unsigned int) * words) : nullptr;
    if (words) run.push_back([this, pool, words] { launch_zero(stream_, 4, pool, words); });
    size_t pool_used = 0;
    const int cap = std::is_same<U, T>::value ? panel_r_cap() : panel_r_capacity<U>();
    for (int jb = 0; jb < kmax; jb += nb_) {
      ws_.rewind();
      std::vector<PanelDesc<U>> single, multi;
      std::vector<int> live;
      double pfl = 0, pby = 0;
      for (int q = 0; q < (int)jobs.size(); ++q) {
        auto& j = jobs[q];
        if (jb >= std::min(j.m, j.k)) continue;
        PanelDesc<U> d{};
        d.P = j.P; d.tau = j.tau; d.Tb = j.Tb; d.ld = j.ld; d.m = j.m; d.k = j.k;
        d.nwg = iceil(j.m - jb, rpw);
        d.err = qp3_err_;
        if (d.nwg > 1) {
          d.part = (U*)ws_.alloc(sizeof(U) * 2 * (size_t)d.nwg * 2 * nb_);
          d.sync = pool + pool_used;
          pool_used += barrier_words(d.nwg);
          multi.push_back(d);
        } else {
          single.push_back(d);
        }
        live.push_back(q);
        const int w = std::min(nb_, std::min(j.m, j.k) - jb), r = j.m - jb, c = std::min(nb_, j.k - jb);
        pfl += 4.0 * r * w * c + 2.0 * r * w * w;  // column steps: dots + updates, then T
        pby += 2.0 * sizeof(U) * (double)r * c;
      }
      if (live.empty()) break;
      // single-workgroup panels in one plain launch; the others cooperatively,
      // packed within the resident capacity
      std::vector<std::pair<std::vector<PanelDesc<U>>, int>> launches;
      if (!single.empty()) launches.push_back({single, 0});
      std::vector<PanelDesc<U>> cur;
      int used = 0;
      for (auto& d : multi) {
        if (d.nwg > cap) throw Error(kErrState, "panel: problem exceeds the resident workgroup capacity");
        if (used + d.nwg > cap && !cur.empty()) { launches.push_back({cur, 1}); cur.clear(); used = 0; }
        cur.push_back(d);
        used += d.nwg;
      }
      if (!cur.empty()) launches.push_back({cur, 1});
      for (size_t li = 0; li < launches.size(); ++li) {
        auto& L = launches[li];
        std::vector<int> cnt, st;
        for (auto& d : L.first) cnt.push_back(d.nwg);
        const int nt = prefix_tiles(cnt, st);
        const PanelDesc<U>* dpd = stager_.push(L.first);
        const int* dst = stager_.push(st);
        const double f = li == 0 ? pfl : 0.0, by = li == 0 ? pby : 0.0;
        const int nd = (int)L.first.size(), coop = L.second;
        run.push_back([this, f, by, dpd, dst, nd, nt, jb, coop] {
          int rc = 0;
          timed("panel", f, by, [&] { rc = launch_panel_r<U>(stream_, dpd, dst, nd, nt, jb, coop); });
          if (rc != 0) throw Error(kErrHip, std::string("panel launch failed: ") + hipGetErrorString((hipError_t)rc));
        });
      }
      // trailing matrices: C -= (V_b T_b^T)(V_b^T C).  W = V_b T_b^T is formed
      // once per problem in the same launch as the Y = V_b^T C of its
      // targets (W and Y live in ws_, rewound per step); T_b's strict lower
      // part is stored as zeros, so it enters as a dense factor
      auto& bufs = larfb_bufs<U>();
      auto &g1 = bufs.g1, &g3 = bufs.g3;
      // the step's targets of job q: its trailing columns, then its extra
      // matrices with columns, in that order (larfb_step_descs' order)
      auto targets = [&](int q, std::vector<LarfbTarget<U>>& tg) {
        auto& j = jobs[q];
        tg.clear();
        const int c0 = jb + std::min(nb_, j.k - jb);
        if (j.k > c0) tg.push_back({j.P + jb + (int64_t)c0 * j.ld, j.ld, j.k - c0});
        for (auto& b : j.extra)
          if (b.cols > 0) tg.push_back({b.p + jb, b.ld, b.cols});
      };
      // pass 1: per live job its descriptor counts and the bytes of its W
      // and Y products (each allocation 256-aligned, as the arena places
      // them), so every job's tables and scratch have fixed offsets
      const int nl = (int)live.size();
      std::vector<size_t> n1(nl + 1, 0), n3(nl + 1, 0), nby(nl + 1, 0);
      auto al = [](size_t b) { return (b + 255) / 256 * 256; };
      parallel_for(nl, 256, [&](int, int b, int e) {
        std::vector<LarfbTarget<U>> tg;
        for (int i = b; i < e; ++i) {
          targets(live[i], tg);
          if (tg.empty()) continue;
          const int rows = jobs[live[i]].m - jb;
          size_t by = al(sizeof(U) * (size_t)std::max<int64_t>((int64_t)std::max(rows, 1) * nb_, 1));
          for (auto& t : tg) by += al(sizeof(U) * (size_t)std::max<int64_t>((int64_t)nb_ * t.cols, 1));
          n1[i + 1] = 1 + tg.size();
          n3[i + 1] = tg.size();
          nby[i + 1] = by;
        }
      });
      for (int i = 0; i < nl; ++i) { n1[i + 1] += n1[i]; n3[i + 1] += n3[i]; nby[i + 1] += nby[i]; }
      if (g1.size() < n1[nl]) g1.resize(n1[nl]);
      if (g3.size() < n3[nl]) g3.resize(n3[nl]);
      char* wbase = nby[nl] ? (char*)ws_.alloc(nby[nl]) : nullptr;
      // pass 2: the descriptors, written at those offsets
      parallel_for(nl, 64, [&](int, int b, int e) {
        std::vector<LarfbTarget<U>> tg;
        struct Out {
          GemmDesc<U>* p;
          void push_back(const GemmDesc<U>& d) { *p++ = d; }
        };
        for (int i = b; i < e; ++i) {
          const int q = live[i];
          auto& j = jobs[q];
          targets(q, tg);
          if (tg.empty()) continue;
          const int w = std::min(nb_, std::min(j.m, j.k) - jb);
          const int rows = j.m - jb;
          const U* Vb = j.P + jb + (int64_t)jb * j.ld;
          const U* Tb = j.Tb + (int64_t)jb * nb_;
          char* cur = wbase + nby[i];
          auto alloc = [&](int64_t elems) {
            U* p = (U*)cur;
            cur += al(sizeof(U) * (size_t)std::max<int64_t>(elems, 1));
            return p;
          };
          Out o1{g1.data() + n1[i]}, o3{g3.data() + n3[i]};
          larfb_step_descs<U>(Vb, j.ld, Tb, nb_, rows, w, tg.data(), (int)tg.size(), alloc, o1, o3);
        }
      });
      const GemmLaunch<U> l1 = gemm_plan(g1.data(), n1[nl], "gemm_y");
      const GemmLaunch<U> l3 = gemm_plan(g3.data(), n3[nl], "gemm_c");
      run.push_back([this, l1, l3] {
        gemm_launch(l1);
        gemm_launch(l3);
      });
      // a step whose planning costs more host time than an upload (the
      // batches of many small problems) is launched at once, so the next
      // step is planned while it runs
      if (n1[nl] + n3[nl] >= 4096) {
        for (auto& f : run) f();
        run.clear();
      }
    }
  }

  // tables of the larfb block steps and gemm_batch's tile counts, kept per
  // host thread across block steps and trees: a level-0 step plans ~300k
  // descriptors, and fresh memory costs more host time (page faults) than
  // filling it.  The memory is held until the thread ends.
  template <typename U>
  struct LarfbBufs {
    std::vector<GemmDesc<U>> g1, g3;
    std::vector<LarfbTarget<U>> tg;
  };
  template <typename U>
  static LarfbBufs<U>& larfb_bufs() {
    static thread_local LarfbBufs<U> b;
    return b;
  }
  struct TileBufs {
    std::vector<int> cnt, st;
  };
  // copy tables of the elimination's gather and scatter (reused; only the
  // leading entries a call fills are valid)
  struct CopyBufs {
    std::vector<CopyDesc> a, b;
  };
  static CopyBufs& copy_bufs() {
    static thread_local CopyBufs b;
    return b;
  }
  static TileBufs& tile_bufs() {
    static thread_local TileBufs b;
    return b;
  }

  // ------------------------------------------------------------ phases
  // numeric row-sparsity tests of the non-geometric path (tree.cpp:855-858)
  struct SparsityTest { int c, nin, e, ein; };
  std::vector<SparsityTest> sparsity_tests_;
  // per-thread state of get_sparsity: cluster marks (== stamp: in the set
  // being built) and the numeric tests found, in cluster order
  struct SparsityScratch {
    std::vector<unsigned> mark;
    unsigned stamp = 0;
    std::vector<SparsityTest> tests;
  };
  void get_sparsity(int c, SparsityScratch& sc);
  // the row sparsity of every cluster of the level, on the host pool
  void fill_sparsity(const std::vector<int>& lev);
  void finish_sparsity();
  // fine_wave >= 0: the inline elimination of --scale 0 fine clusters of
  // that sparsify wave (ops recorded as FineQR with reflectors)
  void eliminate_level(const std::vector<int>& cs, int fine_wave = -1);
  void sparsify0_level(const std::vector<int>& cs);
  void scale_level(const std::vector<int>& cs);
  void sparsify_level(const std::vector<int>& cs);
  void merge_all();
  void reassign_level(const std::vector<int>& cs);
  struct RrqrJob {
    T* Ct;
    int k, N;
    int rank = 0, nwg = 1;
    T* V = nullptr;
    T* tau = nullptr;
  };
  // rc_force > 0: rows per lane of the register form (the autotuner's trial
  // runs); else the tuned or closed-form choice
  void rrqr_batch(std::vector<RrqrJob>& jobs, int rc_force = 0);
  bool trial_ = false;  // inside an autotuner trial: no ledger, no logging
  void sparsify_extra_level(const std::vector<int>& cs);
  void build_solve_program();

  // ------------------------------------------------------------ state
  int nlevels_, skip_, scale_ = 0, square_ = 0, verbose_ = 0, timing_ = 1;
  double tol_ = 0;
  float order_ = 1;
  bool geo_ = false;
  std::vector<double> X_;
  std::vector<int> row_parts_;  // caller's row partition (non-geometric path)
  bool elim_color_ = true;      // elimination waves by colouring (else creation-order longest path)
  int xdim_ = 0;
  int nrows_ = 0, ncols_ = 0, ilvl_ = -1, cur_ = 0, max_order_ = 0, nb_ = 32;
  std::vector<int> cperm_, rperm_;
  std::vector<Clu> cl_;
  std::vector<Edge> ed_;
  std::vector<std::vector<int>> bottoms_;
  std::vector<Op> ops_;
  std::vector<LevelInfo> info_;
  std::vector<SparsifyInfo> sprs_;
  std::vector<SparsifyInfo> extra_;  // sparsify_extra: extra rows before / after (least squares)
  std::vector<MergeInfo> merges_;    // clusters created by merge_all, in creation order
  unsigned long long nnzA_ = 0, nnzR_ = 0, nnzH_ = 0, nnzQ_ = 0;
  double flops_ = 0, t_partition_ = 0, t_assemble_ = 0, t_factorize_ = 0, top_size_ = 0;
  double factor_bytes_ = 0, solve_bytes_ = 0;
  int64_t xsize_ = 0;
  T* xbuf_ = nullptr;
  hipStream_t stream_ = nullptr;
  Arena perm_, scratch_, ws_, gws_{kSplitWsBytes + 4096};
  static constexpr size_t kSplitWsBytes = size_t(64) << 20;
  int* qp3_err_ = nullptr;
  int qp3_max_wg_ = 64;
  Stager stager_;
  int* rank_host_ = nullptr;
  bool partitioned_ = false, assembled_ = false, factored_ = false;
  bool factorizing_ = false;  // set for good once factorize starts (a failed run leaves the tree unusable)
  bool kprof_ = false;
  struct KEv { std::vector<std::pair<hipEvent_t, hipEvent_t>> ev; double flops = 0, bytes = 0; };
  std::map<const int*, const char*> qp3_family_;  // problem -> timing family of its launch
  std::vector<hipEvent_t> ev_pool_;               // reused by kernel timing
  size_t ev_used_ = 0;
  hipEvent_t pooled_event() {
    if (ev_used_ == ev_pool_.size()) {
      hipEvent_t e;
      SPQ_HIP(hipEventCreate(&e));
      ev_pool_.push_back(e);
    }
    return ev_pool_[ev_used_++];
  }
  std::map<std::string, KEv> kev_;

  // solve program (device descriptor tables, built once after factorize)
  struct Stage {
    int type;  // 0 refl, 1 bwd, 2 move
    int nd = 0, mmax = 0, nwg_total = 0;
    int kmax = 0, mall = 0;  // largest k and m of the stage's ops (diagnostics)
    int form = 0, nseg_max = 0;  // 1: register form (k_reflr) / latency form (k_bwdr)
    int nrw = 7;                 // register form: row waves per workgroup
    int buf = 0, sbuf = 0;  // x buffer used (0 = x, 1 = xt); moves copy sbuf -> buf
    const ReflDesc<T>* refl = nullptr;
    const BwdDesc<T>* bwd = nullptr;
    const MoveDesc* mv = nullptr;
    const R12Chunk<T>* r12c = nullptr;
    const int* wstart = nullptr;
    T* work = nullptr;
    // FP64 top separator of an FP32 tree (types 4-7)
    const ReflDesc<double>* refl64 = nullptr;
    const BwdDesc<double>* bwd64 = nullptr;
    const CopyDesc* cvt = nullptr;
    double* work64 = nullptr;
  };
  double* x64_ = nullptr;              // FP64 workspace of the top-separator ops
  const Seg* segs64_dev_ = nullptr;
  unsigned int* solve_sync_ = nullptr;  // barrier counters of multi-workgroup solve ops
  int n_solve_sync_ = 0;
  int* solve_err_ = nullptr;
  int num_cus_ = 256;
  std::vector<Stage> fwd_, bwd_;
  const Seg* segs_dev_ = nullptr;
  const R12Desc<T>* r12_dev_ = nullptr;
  int64_t* in_dst_ = nullptr;
  int* in_src_ = nullptr;
  int64_t* out_src_ = nullptr;
  int* out_dst_ = nullptr;
  int n_in_ = 0, n_out_ = 0;
  hipGraphExec_t graph_exec_fwd_ = nullptr;
  double* bin_ = nullptr;
  double* xout_ = nullptr;
  // mode 0: M^{-1}; 1: normal equations (W^T W)^{-1}; 2: M^{-T}
  void run_program(const double* b, double* x, bool prof = false, int mode = 0);
  hipGraphExec_t capture_program(int mode);
  // M^{-T} (square trees): transposes of the backward, then forward programs
  std::vector<Stage> afwd_, abwd_;
  hipGraphExec_t graph_exec_t_ = nullptr;
  // normal-equation programs (least squares) and their entry map
  std::vector<Stage> tfwd_, tbwd_, pfwd_;
  T* xtbuf_ = nullptr;
  int64_t* tin_dst_ = nullptr;
  int* tin_src_ = nullptr;
  int n_tin_ = 0;
  hipGraphExec_t graph_exec_nrml_ = nullptr;
  bool stage_prof_ = false;  // SPAQR_SOLVE_STAGES=1: time every solve stage (uncaptured)
  bool solve_graph_ = true;  // SPAQR_SOLVE_GRAPH=0: launch the solve program directly (no hipGraph)
  bool qp3_reg_ = true;      // SPAQR_QP3_REGISTER=0: streaming truncated QR only
  bool qp3_log_ = false;     // SPAQR_QP3_LOG=1: per-launch RRQR lines (with kernel timing)
  bool reflr_ = true;        // SPAQR_REFL_REGISTER=0: solve reflector stages in the LDS form only
  // per-cluster scratch of the host bookkeeping (sized to cl_, reused)
  std::vector<int> pos_;        // panel position of a cluster, -1 when none
  int panel_cap_ = -1;
  int panel_r_cap() {
    if (panel_cap_ < 0) panel_cap_ = panel_r_capacity<T>();
    return panel_cap_;
  }
  static constexpr int kRcSlots = (100 - 16) / 4 + 1;
  int qp3r_cap_[2][kRcSlots];
  int qp3_rc_ = 0;  // SPAQR_QP3_RC: force the register form's rows per lane
  // rows per lane of the register form's kernels: 16 to 32 (256 and 1024
  // threads), 40 to 100 (256 threads; from 64 on one lane holds a column of
  // up to rc rows, so a problem needs 4x fewer workgroups than at 4 lanes)
  static std::vector<int> qp3r_rcs() { return {16, 20, 24, 28, 32, 40, 48, 64, 80, 100}; }
  // resident workgroups of the register-form QR with nt threads per
  // workgroup (largest LDS footprint, G = 64)
  int qp3r_cap(int nt, int rc) {
    int& c = qp3r_cap_[nt == 1024 ? 1 : 0][std::min((rc - 16) / 4, kRcSlots - 1)];
    if (c < 0) c = qp3r_capacity<T>(64, nt, rc);
    return c;
  }
};

// ======================================================================
template <typename T>
void Engine<T>::partition(const CscMat& A) {  // tree.cpp:595-723
  if (partitioned_) throw Error(kErrState, "partition: the tree is already partitioned (create a new tree)");
  const double t0 = wall();
  nrows_ = A.m;
  ncols_ = A.n;
  std::vector<CID> lab;
  if (geo_) {
    if ((int64_t)X_.size() != (int64_t)xdim_ * A.n) throw Error(kErrArg, "partition: coordinate array has wrong size");
    if (verbose_) printf("Using GeometricPartition... \n");
    lab = geometric_labels(A, nlevels_, X_, xdim_);
  } else {
    // no coordinates: a row partition (the caller's, else the library's
    // own) and the labels it induces (tree.cpp:608-617)
    if (verbose_) printf("Using Hypergraph Partition... \n");
    lab = hypergraph_labels(A, nlevels_, row_parts_.empty() ? row_parts(A, nlevels_) : row_parts_);
  }
  rperm_.resize(A.m);
  cperm_.resize(A.n);
  std::iota(rperm_.begin(), rperm_.end(), 0);
  std::iota(cperm_.begin(), cperm_.end(), 0);
  order_columns(nlevels_, lab, cperm_);
  std::vector<CID> cpermed(A.n);
  for (int k = 0; k < A.n; ++k) cpermed[k] = lab[cperm_[k]];
  std::vector<int> rows_per_col;
  if (!square_) {
    CscMat Ap = csc_permute(A, nullptr, &cperm_);
    order_rows_ls(Ap, cpermed, rperm_, rows_per_col);
  } else {
    rows_per_col.assign(A.n, 1);
    rperm_ = cperm_;
  }
  int r0 = 0;
  for (int k = 0; k < A.n;) {
    int kn = k + 1, nr = rows_per_col[k];
    while (kn < A.n && cpermed[kn] == cpermed[k]) nr += rows_per_col[kn++];
    if (nr == 0) throw Error(kErrArg, "partition: a cluster received no rows");
    bottoms_[0].push_back(new_cluster(k, kn - k, r0, nr, cpermed[k]));
    r0 += nr;
    k = kn;
  }
  for (int lvl = 1; lvl < nlevels_; ++lvl) {
    auto& prev = bottoms_[lvl - 1];
    size_t b = 0;
    while (b < prev.size() && cl_[prev[b]].level() < lvl) ++b;
    for (size_t q = b; q < prev.size(); ++q) cl_[prev[q]].parentid = cid_merge_if(cl_[prev[q]].id, lvl);
    for (size_t q = b; q < prev.size();) {
      const CID pid = cl_[prev[q]].parentid;
      std::vector<int> ch{prev[q]};
      int cs = cl_[prev[q]].csize, rs = cl_[prev[q]].rsize;
      const int c0 = cl_[prev[q]].cstart, rr0 = cl_[prev[q]].rstart;
      ++q;
      while (q < prev.size() && cl_[prev[q]].parentid == pid) {
        ch.push_back(prev[q]);
        cs += cl_[prev[q]].csize;
        rs += cl_[prev[q]].rsize;
        ++q;
      }
      const int p = new_cluster(c0, cs, rr0, rs, pid);
      for (int c : ch) { cl_[c].parent = p; cl_[p].children.push_back(c); }
      bottoms_[lvl].push_back(p);
    }
  }
  partitioned_ = true;
  t_partition_ = wall() - t0;
  if (verbose_) printf("Time to partition: %g\n", t_partition_);
}

template <typename T>
void Engine<T>::assemble(const CscMat& A) {  // tree.cpp:725-778
  if (!partitioned_) throw Error(kErrState, "assemble: partition has not run");
  if (assembled_) throw Error(kErrState, "assemble: the tree is already assembled (create a new tree)");
  const double t0 = wall();
  nnzA_ = A.nnz();
  const CscMat App = csc_permute(A, &rperm_, &cperm_);
  std::vector<int> rmap(nrows_);
  for (int c : bottoms_[0])
    for (int i = cl_[c].rstart; i < cl_[c].rstart + cl_[c].rsize; ++i) rmap[i] = c;
  for (int c : bottoms_[0]) alloc_x(c, cl_[c].rsize);
  // host-side dense blocks, uploaded in one batch
  std::vector<T> host;
  struct Pending { int edge; size_t off; };
  std::vector<Pending> pend;
  for (int self : bottoms_[0]) {
    std::vector<int> nb;
    const Clu& s = cl_[self];
    for (int j = s.cstart; j < s.cstart + s.csize; ++j)
      for (int q = App.p[j]; q < App.p[j + 1]; ++q) nb.push_back(rmap[App.i[q]]);
    nb.push_back(self);
    std::sort(nb.begin(), nb.end());
    nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    for (int n : nb) {
      Edge e;
      e.n1 = self; e.n2 = n;
      e.a.rows = cl_[n].rsize; e.a.cols = s.csize; e.a.ld = std::max(1, e.a.rows);
      const size_t off = host.size();
      host.resize(off + (size_t)e.a.ld * e.a.cols, T(0));
      const int r0 = cl_[n].rstart, r1 = r0 + cl_[n].rsize;
      for (int col = 0; col < s.csize; ++col) {
        const int j = s.cstart + col;
        const int* beg = App.i.data() + App.p[j];
        const int* end = App.i.data() + App.p[j + 1];
        for (const int* it = std::lower_bound(beg, end, r0); it != end && *it < r1; ++it)
          host[off + (size_t)(*it - r0) + (size_t)col * e.a.ld] = (T)App.x[it - App.i.data()];
      }
      ed_.push_back(e);
      const int id = (int)ed_.size() - 1;
      pend.push_back({id, off});
      cl_[self].out.push_back(id);
      cl_[self].outn.push_back(n);
      if (self != n) {
        cl_[n].in.push_back(id);
        cl_[n].inn.push_back(self);
      }
    }
    sort_out(self);
    cl_[self].cnbrs = nb;
  }
  T* dev = alloc_perm(host.size());
  SPQ_HIP(hipMemcpyAsync(dev, host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice, stream_));
  for (auto& p : pend) ed_[p.edge].a.p = dev + p.off;
  sync();
  assembled_ = true;
  t_assemble_ = wall() - t0;
  if (verbose_) {
    auto ts = topsize();
    printf("Time to assemble: %g\n", t_assemble_);
    printf("Aspect ratio of top separator: %g\n", (double)ts.first / (double)ts.second);
  }
}

template <typename T>
void Engine<T>::get_sparsity(int c, SparsityScratch& sc) {  // tree.cpp:834-870 (+ check_if_valid 168-224)
  std::vector<SepID> colsp;
  for (int e : cl_[c].out) colsp.push_back(cl_[ed_[e].n2].id.self);
  std::sort(colsp.begin(), colsp.end());
  auto has = [&colsp](const SepID& s) { return std::binary_search(colsp.begin(), colsp.end(), s); };
  const SepID cs = cl_[c].id.self;
  // integer division by 2^d (the reference's (int)pow(2, d), exact here)
  auto valid = [&](const CID& nid) {
    const SepID ns = nid.self;
    if (cs.lvl > ns.lvl) return false;
    const int d = ns.lvl - cs.lvl;
    if (cs.sep / (1 << d) != ns.sep) return false;
    auto side_ok = [&](const SepID& s) {
      const int dd = s.lvl - cs.lvl;
      if (s.lvl >= cs.lvl) return cs.sep / (1 << dd) == s.sep;
      return s.sep / (1 << -dd) == cs.sep;
    };
    if (!side_ok(nid.l) && !side_ok(nid.r)) return false;
    if (geo_) {  // a separator may only connect through the column sparsity
      if (!has(ns)) return false;
      if (!has(nid.l) && !has(nid.r)) return false;
    }
    return true;
  };
  // the distinct clusters found, marked as they are first accepted (a
  // cluster already in the set needs no further test), then sorted
  std::vector<int> rs;
  std::vector<unsigned>& mark = sc.mark;
  if (mark.size() < cl_.size()) mark.resize(cl_.size(), 0);
  const unsigned stamp = ++sc.stamp;
  auto take = [&](int n) {
    if (mark[n] != stamp) { mark[n] = stamp; rs.push_back(n); }
  };
  for (int e : cl_[c].out) {
    if (!ed_[e].alive) continue;
    const int n = ed_[e].n2;
    take(n);
    for (int ein : cl_[n].in) {
      const int nin = ed_[ein].n1;
      if (mark[nin] == stamp) continue;
      if (!ed_[ein].alive || nin == c || cl_[nin].elim || !valid(cl_[nin].id)) continue;
      // without coordinates, a cluster that does not name c as an interface
      // must also couple to c numerically through n: |A(n,c)^T A(n,nin)| >
      // 1e-14 (tree.cpp:855-858); decided for the whole level at once
      if (!geo_ && !(cl_[nin].id.l == cs || cl_[nin].id.r == cs)) {
        sc.tests.push_back(SparsityTest{c, nin, e, ein});
        continue;
      }
      take(nin);
    }
  }
  std::sort(rs.begin(), rs.end());
  cl_[c].rsp = std::move(rs);
}

// Clusters of one level only read the tree and write their own row
// sparsity; the numeric tests are joined in cluster order.
template <typename T>
void Engine<T>::fill_sparsity(const std::vector<int>& lev) {
  const int n = (int)lev.size(), grain = 64;
  std::vector<SparsityScratch> sc(std::max(plan_chunks(n, grain), 1));
  parallel_for(n, grain, [&](int ch, int b, int e) {
    for (int i = b; i < e; ++i) get_sparsity(lev[i], sc[ch]);
  });
  for (auto& x : sc) sparsity_tests_.insert(sparsity_tests_.end(), x.tests.begin(), x.tests.end());
}

// The numeric tests get_sparsity left for this level: P = A(n,c)^T A(n,nin)
// by batched GEMMs, |P|_F per product, one read-back; a passing cluster joins
// c's row sparsity (kept sorted).
template <typename T>
void Engine<T>::finish_sparsity() {
  if (sparsity_tests_.empty()) return;
  scratch_.rewind();
  std::vector<GemmDesc<T>> gds;
  std::vector<RowNormDesc<T>> nds;
  std::vector<int> ncnt;
  double* norms = (double*)scratch_.alloc(sizeof(double) * sparsity_tests_.size());
  std::vector<char> empty(sparsity_tests_.size(), 0);
  for (size_t q = 0; q < sparsity_tests_.size(); ++q) {
    const SparsityTest& t = sparsity_tests_[q];
    const Block& a = ed_[t.e].a;    // rows of n x cols of c
    const Block& b = ed_[t.ein].a;  // rows of n x cols of nin
    if (a.rows == 0 || a.cols == 0 || b.cols == 0) { empty[q] = 1; continue; }  // norm 0: no coupling
    T* P = alloc_scr((size_t)std::max(a.cols * b.cols, 1));
    gds.push_back(gd<T>(a.p, a.ld, b.p, b.ld, P, std::max(a.cols, 1), a.cols, b.cols, a.rows,
                        gemm_flags(true, false), T(1), T(0)));
    // |P|_F as the norm of one row of a.cols * b.cols entries
    nds.push_back(RowNormDesc<T>{P, norms + q, 1, 0, 1, a.cols * b.cols});
    ncnt.push_back(1);
  }
  if (gds.empty()) { sparsity_tests_.clear(); return; }
  gemm_batch<T>(gds, "sparsity");
  std::vector<int> nst;
  const int nt = prefix_tiles(ncnt, nst);
  const RowNormDesc<T>* dn = stager_.push(nds);
  const int* dst = stager_.push(nst);
  stager_.flush();
  launch_rownorm<T>(stream_, dn, dst, (int)nds.size(), nt);
  std::vector<double> h(sparsity_tests_.size());
  SPQ_HIP(hipMemcpyAsync(h.data(), norms, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream_));
  sync();
  for (size_t q = 0; q < sparsity_tests_.size(); ++q) {
    const SparsityTest& t = sparsity_tests_[q];
    if (empty[q] || !(h[q] > 1e-14)) continue;
    auto& rs = cl_[t.c].rsp;
    if (!std::binary_search(rs.begin(), rs.end(), t.nin)) rs.insert(std::lower_bound(rs.begin(), rs.end(), t.nin), t.nin);
  }
  sparsity_tests_.clear();
}

// tree.cpp:780-832 (eliminate_cluster, householder, update_cluster) for every
// cluster of the level at once.
template <typename T>
void Engine<T>::eliminate_level(const std::vector<int>& cs, int fine_wave) {
  struct Plan {
    int c, m, k, Nw;
    std::vector<int> pe;                  // panel edges
    std::vector<int> tg;                  // target clusters
    std::vector<int> te;                  // target edges: te[t * |pe| + s] for target t, panel block s
    std::vector<int> touched;             // sorted target edge ids
    T *P, *W, *tau, *Tb;
    double *P64 = nullptr, *W64 = nullptr, *tau64 = nullptr, *Tb64 = nullptr;  // FP64 top separator
  };
  // An FP32 tree factors its top separator in FP64 (the last level's QR):
  // the gathered panel is converted, factored in double, and the op keeps
  // the double factors for the solve.
  // (square and least-squares trees: the R-only top of a least-squares tree
  // keeps its FP64 R for both triangular sweeps of the normal equations)
  const bool top64 = sizeof(T) == 4 && fine_wave < 0 && ilvl_ == nlevels_ - 1;
  std::vector<Plan> plans;
  std::vector<Block> zero;
  for (int c : cs) {
    if (cl_[c].csize == 0) { cl_[c].elim = true; continue; }
    Plan p;
    p.c = c; p.m = 0; p.k = cl_[c].csize; p.Nw = 0;
    for (int e : cl_[c].out)
      if (ed_[e].alive) { p.pe.push_back(e); p.m += ed_[e].a.rows; }
    if (ed_[p.pe[0]].n2 != c) throw Error(kErrState, "eliminate: self block is not first");
    // the target blocks of row n in the panel's columns: one scan of n's
    // out edges finds the existing ones (pe's clusters are marked with their
    // panel position), the missing ones are added in panel order
    if (pos_.size() < cl_.size()) pos_.resize(cl_.size(), -1);
    for (size_t q = 0; q < p.pe.size(); ++q) pos_[ed_[p.pe[q]].n2] = (int)q;
    const size_t npe = p.pe.size();
    p.te.reserve(cl_[c].rsp.size() * npe);
    p.touched.reserve(cl_[c].rsp.size() * npe);
    for (int n : cl_[c].rsp) {
      if (cl_[n].elim || n == c) continue;
      const size_t base = p.te.size();
      p.te.resize(base + npe, -1);
      int* row = p.te.data() + base;  // p.te is not resized until this row is done
      {
        const Clu& cn = cl_[n];
        for (size_t q = 0; q < cn.outn.size(); ++q) {
          const int kc = cn.outn[q];
          if (kc < (int)pos_.size() && pos_[kc] >= 0 && row[pos_[kc]] < 0) row[pos_[kc]] = cn.out[q];
        }
      }
      for (size_t q = 0; q < npe; ++q) {
        if (row[q] < 0) row[q] = add_edge(n, ed_[p.pe[q]].n2, &zero);
        p.touched.push_back(row[q]);
      }
      p.tg.push_back(n);
      p.Nw += cl_[n].csize;
      nnzR_ += (unsigned long long)(cl_[c].csize * cl_[n].csize);
    }
    for (int e : p.pe) pos_[ed_[e].n2] = -1;
    std::sort(p.touched.begin(), p.touched.end());
    nnzH_ += (unsigned long long)(p.m * p.k);
    const int kk = std::min(p.m, p.k);
    flops_ += 4.0 * p.m * p.k * kk - 2.0 * (p.m + p.k) * kk * kk + 4.0 / 3.0 * kk * kk * kk;  // geqrf
    flops_ += (double)p.m * kk * kk - (double)kk * kk * kk / 3.0;                           // larft
    flops_ += 4.0 * p.m * p.Nw * kk - (double)p.Nw * kk * kk;                               // larfb
    plans.push_back(std::move(p));
  }
  init_blocks(zero, 0);
  // waves: two clusters that update a common target block may not share a
  // launch (their read-modify-writes of the block would race).  At level 0
  // their updates commute exactly: the partition separates the leaves in
  // the graph of A^T A, so no row of A couples two leaves and their
  // reflectors act on disjoint rows of a shared block.  There the waves are
  // a greedy colouring in creation order: the smallest wave no earlier
  // cluster sharing a block holds (per-block masks of used waves).  Above
  // level 0 the rows of an interface have been mixed by its sparsification,
  // so clusters sharing one do not commute and the order of
  // tree.cpp:1320-1330 (creation order) is kept: wave = 1 + the largest wave
  // of an earlier cluster sharing a block.  So is the inline elimination of
  // --scale 0 fine clusters.
  std::vector<int> wave(plans.size(), 0);
  int nw = plans.empty() ? 0 : 1;
  if (fine_wave < 0 && ilvl_ == 0 && elim_color_) {
    constexpr int kWords = 4;  // up to 256 waves by colouring
    std::vector<std::array<uint64_t, kWords>> used(ed_.size(), std::array<uint64_t, kWords>{});
    std::vector<int> maxw(ed_.size(), -1);
    for (size_t i = 0; i < plans.size(); ++i) {
      std::array<uint64_t, kWords> m{};
      int wl = 0;
      for (int f : plans[i].touched) {
        for (int q = 0; q < kWords; ++q) m[q] |= used[f][q];
        wl = std::max(wl, maxw[f] + 1);
      }
      int w = -1;
      for (int q = 0; q < kWords && w < 0; ++q)
        if (~m[q]) w = 64 * q + __builtin_ctzll(~m[q]);
      if (w < 0) w = wl;  // more than 256 waves: after every conflicting one
      wave[i] = w;
      for (int f : plans[i].touched) {
        if (w < 64 * kWords) used[f][w / 64] |= 1ull << (w % 64);
        maxw[f] = std::max(maxw[f], w);
      }
      nw = std::max(nw, w + 1);
    }
  } else {
    std::vector<int> maxw(ed_.size(), -1);
    for (size_t i = 0; i < plans.size(); ++i) {
      int w = 0;
      for (int f : plans[i].touched) w = std::max(w, maxw[f] + 1);
      wave[i] = w;
      for (int f : plans[i].touched) maxw[f] = std::max(maxw[f], w);
      nw = std::max(nw, w + 1);
    }
  }
  if (fine_wave < 0) {
    info_[ilvl_].elim_waves = nw;
    info_[ilvl_].n_elim = (int)plans.size();
  }
  for (auto& p : plans) {
    p.P = alloc_perm((size_t)p.m * p.k);
    p.tau = alloc_perm(p.k);
    p.Tb = alloc_perm((size_t)nb_ * iceil(p.k, nb_) * nb_);
    if (top64) {
      p.P64 = (double*)perm_.alloc(sizeof(double) * std::max<size_t>((size_t)p.m * p.k, 1));
      p.tau64 = (double*)perm_.alloc(sizeof(double) * std::max(p.k, 1));
      p.Tb64 = (double*)perm_.alloc(sizeof(double) * (size_t)nb_ * iceil(p.k, nb_) * nb_);
      factor_bytes_ += sizeof(double) * ((double)p.m * p.k + (double)nb_ * iceil(p.k, nb_) * nb_);
    } else {
      factor_bytes_ += sizeof(T) * ((double)p.m * p.k + (double)nb_ * iceil(p.k, nb_) * nb_);
    }
  }
  std::vector<CopyDesc>& gather = copy_bufs().a;
  std::vector<CopyDesc>& scatter = copy_bufs().b;
  for (int w = 0; w < nw; ++w) {
    scratch_.rewind();
    std::vector<QrJob> jobs;
    // the wave's plans in creation order; each one's panel blocks and the
    // target blocks of every panel row go to the gather table (panel block s,
    // then the targets of row s), the target blocks alone to the scatter
    // table, at offsets fixed here so the tables are written in parallel
    std::vector<int> idx;
    std::vector<size_t> goff(1, 0), soff(1, 0);
    for (size_t i = 0; i < plans.size(); ++i) {
      if (wave[i] != w) continue;
      Plan& p = plans[i];
      p.W = alloc_scr((size_t)p.m * std::max(p.Nw, 1));
      idx.push_back((int)i);
      goff.push_back(goff.back() + p.pe.size() * (1 + p.tg.size()));
      soff.push_back(soff.back() + p.pe.size() * p.tg.size());
      QrJob j;
      j.P = p.P; j.ld = p.m; j.m = p.m; j.k = p.k; j.tau = p.tau; j.Tb = p.Tb;
      if (p.Nw > 0) j.extra.push_back(Block{p.W, p.m, p.Nw, p.m});
      jobs.push_back(std::move(j));
    }
    if (gather.size() < goff.back()) gather.resize(goff.back());
    if (scatter.size() < soff.back()) scatter.resize(soff.back());
    parallel_for((int)idx.size(), 8, [&](int, int qb, int qe) {
      for (int q = qb; q < qe; ++q) {
        const Plan& p = plans[idx[q]];
        CopyDesc* g = gather.data() + goff[q];
        CopyDesc* sc = scatter.data() + soff[q];
        int r = 0;
        for (size_t s = 0; s < p.pe.size(); ++s) {
          const Block& b = ed_[p.pe[s]].a;
          *g++ = CopyDesc{b.p, p.P + r, b.ld, p.m, b.rows, b.cols, 0, 0};
          int col = 0;
          for (size_t t = 0; t < p.tg.size(); ++t) {
            const Block& tb = ed_[p.te[t * p.pe.size() + s]].a;
            const CopyDesc d{tb.p, p.W + r + (int64_t)col * p.m, tb.ld, p.m, tb.rows, tb.cols, 0, 0};
            *g++ = d;
            *sc++ = CopyDesc{d.dst, (void*)tb.p, p.m, tb.ld, tb.rows, tb.cols, 0, 0};
            col += cl_[p.tg[t]].csize;
          }
          r += b.rows;
        }
      }
    });
    Run run;
    if (auto f = copy_plan(gather.data(), goff.back())) run.push_back(std::move(f));
    if (top64) {
      for (auto& f : run) f();
      run.clear();
      std::vector<QrJobT<double>> j64;
      std::vector<CopyDesc> up, down;
      for (size_t i = 0; i < plans.size(); ++i) {
        if (wave[i] != w) continue;
        Plan& p = plans[i];
        up.push_back(CopyDesc{p.P, p.P64, p.m, p.m, p.m, p.k, 0, 0});
        QrJobT<double> j;
        j.P = p.P64; j.ld = p.m; j.m = p.m; j.k = p.k; j.tau = p.tau64; j.Tb = p.Tb64;
        if (p.Nw > 0) {
          p.W64 = (double*)scratch_.alloc(sizeof(double) * (size_t)p.m * p.Nw);
          up.push_back(CopyDesc{p.W, p.W64, p.m, p.m, p.m, p.Nw, 0, 0});
          down.push_back(CopyDesc{p.W64, p.W, p.m, p.m, p.m, p.Nw, 0, 0});
          j.extra.push_back(BlockT<double>{p.W64, p.m, p.Nw, p.m});
        }
        j64.push_back(std::move(j));
      }
      copy_blocks(up, 4, 8);
      blocked_qr(j64);
      copy_blocks(down, 8, 4);
    } else {
      blocked_qr_plan(jobs, run);
    }
    if (auto f = copy_plan(scatter.data(), soff.back())) run.push_back(std::move(f));
    for (auto& f : run) f();
  }
  // ops and bookkeeping (tree.cpp:826-829): the op records are built on the
  // host pool (each reads only its own cluster's lists), then appended in
  // creation order
  const size_t o0 = ops_.size();
  ops_.resize(o0 + plans.size());
  parallel_for((int)plans.size(), 64, [&](int, int b, int e) {
    for (int i = b; i < e; ++i) {
      const Plan& p = plans[i];
      Op& o = ops_[o0 + i];
      o.kind = fine_wave < 0 ? OpQR : OpFineQR; o.level = ilvl_; o.c = p.c;
      o.wave = wave[i];
      if (fine_wave >= 0) { o.wave = fine_wave; o.fwd_refl = true; }
      o.V = p.P; o.ldv = p.m; o.m = p.m; o.k = p.k; o.Tb = p.Tb;
      for (int e2 : p.pe) o.segs.push_back({ed_[e2].n2, cl_[ed_[e2].n2].rsize});
      o.R = p.P; o.ldr = p.m; o.kr = p.k;
      if (top64) { o.V64 = p.P64; o.Tb64 = p.Tb64; }
      for (int e2 : cl_[p.c].in) o.r12.push_back({ed_[e2].a, {ed_[e2].n1, ed_[e2].a.cols}});
    }
  });
  for (auto& p : plans) cl_[p.c].elim = true;
}

// tree.cpp:533-580 (diagScale) for every interface of the level: (i) QR of
// the diagonal blocks with the in-edge blocks updated by Q^T in the same
// blocked sweep, (ii) out-edge blocks times R^{-1}.
template <typename T>
void Engine<T>::scale_level(const std::vector<int>& cs) {
  std::vector<QrJob> jobs;
  std::vector<Block> ident;
  struct TrsmJob { Block R; std::vector<Block> B; };
  // the blocks each interface's QR updates (in-edges) and its TRSM scales
  // (out-edges), found on the host pool: no interface's lists include
  // another interface's diagonal block, which is the only block replaced below
  const int nc = (int)cs.size();
  std::vector<int> eself(nc);
  jobs.resize(nc);
  std::vector<TrsmJob> tj(nc);
  parallel_for(nc, 64, [&](int, int b, int e) {
    for (int i = b; i < e; ++i) {
      const int c = cs[i];
      eself[i] = find_out(c, c);
      for (int e2 : cl_[c].in)
        if (!cl_[ed_[e2].n1].elim) jobs[i].extra.push_back(ed_[e2].a);
      for (int e2 : cl_[c].out)
        if (!cl_[ed_[e2].n2].elim && ed_[e2].n2 != c) tj[i].B.push_back(ed_[e2].a);
    }
  });
  for (int i = 0; i < nc; ++i) {
    const int c = cs[i];
    const int es = eself[i];
    const Block q = ed_[es].a;  // becomes the ScaleD factor storage
    const int rr = q.rows, cc = cl_[c].csize, kk = std::min(rr, cc);
    Block nb = new_block(rr, cc);
    ident.push_back(nb);
    ed_[es].a = nb;
    QrJob& j = jobs[i];
    j.P = q.p; j.ld = q.ld; j.m = rr; j.k = cc;
    j.tau = alloc_perm(cc);
    j.Tb = alloc_perm((size_t)nb_ * iceil(cc, nb_) * nb_);
    factor_bytes_ += sizeof(T) * ((double)rr * cc + (double)nb_ * iceil(cc, nb_) * nb_);
    TrsmJob& t = tj[i];
    t.R = q;
    nnzR_ = (unsigned long long)((double)nnzR_ + 0.5 * cc * cc);
    nnzH_ += (unsigned long long)(cl_[c].rsize * cl_[c].csize);
    flops_ += 4.0 * rr * cc * kk - 2.0 * (rr + cc) * kk * kk + 4.0 / 3.0 * kk * kk * kk + (double)rr * kk * kk -
              (double)kk * kk * kk / 3.0;
    for (auto& b : j.extra) flops_ += 4.0 * rr * b.cols * kk - (double)b.cols * kk * kk;
    for (auto& b : t.B) flops_ += (double)b.rows * cc * cc;
    Op o;
    o.kind = OpScaleD; o.level = ilvl_; o.c = c;
    o.V = q.p; o.ldv = q.ld; o.m = cl_[c].rsize; o.k = cc; o.Tb = j.Tb;
    o.segs.push_back({c, cl_[c].rsize});
    o.R = q.p; o.ldr = q.ld; o.kr = cc;
    ops_.push_back(std::move(o));
  }
  info_[ilvl_].n_scale = (int)cs.size();
  init_blocks(ident, 1);
  scratch_.rewind();
  Run run;  // the QR's steps, then the TRSM's column blocks
  blocked_qr_plan(jobs, run);
  // (ii) B <- B R^{-1}, column block by column block
  int kmax = 0;
  for (auto& t : tj) kmax = std::max(kmax, t.R.cols);
  for (int jb = 0; jb < kmax; jb += nb_) {
    std::vector<GemmDesc<T>> g;
    std::vector<TrsmDesc<T>> td;
    std::vector<int> cnt;
    for (auto& t : tj) {
      if (jb >= t.R.cols) continue;
      const int w = std::min(nb_, t.R.cols - jb);
      for (auto& b : t.B) {
        if (b.rows == 0) continue;
        if (jb > 0)
          g.push_back(gd(b.p, b.ld, t.R.p + (int64_t)jb * t.R.ld, t.R.ld, b.p + (int64_t)jb * b.ld, b.ld, b.rows, w, jb,
                         gemm_flags(false, false, kStructNone, kStructNone), T(-1), T(1)));
        td.push_back(TrsmDesc<T>{b.p, t.R.p, b.ld, t.R.ld, b.rows, t.R.cols});
        cnt.push_back(iceil(b.rows, trsm_rows_per_tile()));
      }
    }
    const GemmLaunch<T> lg = gemm_plan(g.data(), g.size(), "gemm_trsm");
    run.push_back([this, lg] { gemm_launch(lg); });
    if (!td.empty()) {
      std::vector<int> st;
      const int nt = prefix_tiles(cnt, st);
      double tfl = 0;
      for (auto& d : td) tfl += (double)d.rows * std::min(nb_, d.k - jb) * std::min(nb_, d.k - jb);
      const TrsmDesc<T>* dtd = stager_.push(td);
      const int* dst = stager_.push(st);
      const int nd = (int)td.size();
      run.push_back([this, tfl, dtd, dst, nd, nt, jb] {
        timed("trsm", tfl, 0.0, [&] { launch_trsm_block<T>(stream_, dtd, dst, nd, nt, jb); });
      });
    }
  }
  for (auto& f : run) f();
}

// Truncated column-pivoted QR (the geqp3 + choose_rank of util.cpp:466-477,
// 795-816) of a batch of independent problems whose transposed inputs Ct
// (N x k, ld N) are already gathered.  Per problem: the register form when its
// columns fit the resident workgroups, else the streaming form; single-
// workgroup problems in plain launches, multi-workgroup ones cooperatively
// within the resident capacity.  Fills rank, V (k x k, ld k) and tau; on
// exit Ct(0:rank, j) = triu(R(0:rank,:)) P^T for every column j.
template <typename T>
void Engine<T>::rrqr_batch(std::vector<RrqrJob>& jobs, int rc_force) {
  if (jobs.empty()) return;
  int kmax = 0, nmax = 0;
  for (auto& j : jobs) { kmax = std::max(kmax, j.k); nmax = std::max(nmax, j.N); }
  // rows per lane of the register form, one for the whole batch (one launch
  // per thread count): the autotuner's choice for this batch's size class
  // when it has one, else the rc whose lanes per column, over the resident
  // 256-thread workgroups at its register budget, cover the batch with the
  // fewest rounds (ties keep the earlier candidate)
  int rc = 32;
  if (rc_force > 0) {
    rc = rc_force;
  } else if (qp3_rc_ > 0) {
    rc = qp3_rc_;
  } else {
    double best = -1;
    for (int r : {32, 28, 24, 20, 16, 40, 48}) {
      const int cap = qp3r_cap(256, r);
      if (cap <= 0) continue;
      double cost = 0;
      for (auto& j : jobs) {
        const int g = qp3r_lanes(j.k, r);
        cost += g ? (double)g * std::max(j.N, 1) / cap : 1e30;
      }
      if (best < 0 || cost < best) { best = cost; rc = r; }
    }
    if (qp3_reg_ && tol_ < 1.0) {
      char key[96];
      snprintf(key, sizeof key, "qp3r/%s/k%d/n%d/p%d", sizeof(T) == 8 ? "f64" : "f32", log2_bucket(kmax),
               log2_bucket(nmax), log2_bucket((long)jobs.size()));
      Tuner& tu = global_tuner();
      const int cached = tu.lookup(key);
      if (cached > 0 && qp3r_cap(256, cached) > 0) {
        rc = cached;
      } else if (tu.tuning()) {
        // time every form on copies of the batch's inputs (the RRQR works
        // in place), keep the fastest
        std::vector<std::pair<int, double>> ms;
        double best_ms = -1;
        for (int r : qp3r_rcs()) {
          if (qp3r_cap(256, r) <= 0) continue;
          std::vector<RrqrJob> trial;
          std::vector<CopyDesc> cp;
          for (auto& j : jobs) {
            RrqrJob t{alloc_scr((size_t)std::max(j.N, 1) * j.k), j.k, j.N};
            cp.push_back(CopyDesc{j.Ct, t.Ct, std::max(j.N, 1), std::max(j.N, 1), j.N, j.k, 0, 0});
            trial.push_back(t);
          }
          const bool kp = kprof_;
          kprof_ = false;
          copy_blocks(cp);
          sync();
          const double t0 = wall();
          trial_ = true;
          rrqr_batch(trial, r);
          trial_ = false;
          const double t1 = 1e3 * (wall() - t0);
          kprof_ = kp;
          ms.push_back({r, t1});
          if (best_ms < 0 || t1 < best_ms) { best_ms = t1; rc = r; }
        }
        tu.record(key, rc, ms);
      }
    }
  }
  const bool lds_ok = sizeof(T) * (size_t)std::max(kmax, 1) <= (size_t)48 * 1024;
  int* ranks_dev = (int*)scratch_.alloc(sizeof(int) * jobs.size());
  std::vector<Qp3Desc<T>> qd;
  std::vector<int> sync_off, nts;
  int sync_words = 0;
  for (size_t q = 0; q < jobs.size(); ++q) {
    RrqrJob& jb = jobs[q];
    const int N = std::max(jb.N, 1);
    int G = 0, nwg = 1, nt = 256;
    if (qp3_reg_ && (G = qp3r_lanes(jb.k, rc)) > 0) {
      // very large problems: 1024-thread workgroups (one per CU, fewer to
      // meet); below ~256 workgroups the 256-thread form is faster
      nt = iceil(N, 256 / G) > 256 && rc <= 32 ? 1024 : 256;
      nwg = iceil(N, nt / G);
      if (nwg > qp3r_cap(nt, rc)) G = 0;
    } else {
      G = 0;
    }
    if (G == 0) nwg = lds_ok ? std::max(1, std::min(qp3_max_wg_, jb.N / 256)) : 1;
    jb.nwg = nwg;
    jb.V = alloc_scr((size_t)jb.k * jb.k);
    jb.tau = alloc_scr(jb.k);
    Qp3Desc<T> d{};
    d.Ct = jb.Ct; d.V = jb.V; d.tau = jb.tau;
    d.vn = G ? nullptr : (double*)scratch_.alloc(sizeof(double) * 2 * N);
    d.pos = (int*)scratch_.alloc(sizeof(int) * (G ? (size_t)N : 2 * (size_t)N * nwg));
    d.part_v = (double*)scratch_.alloc(sizeof(double) * 4 * nwg);  // (norm, squared norm below i) x 2 buffers
    d.part_p = (int*)scratch_.alloc(sizeof(int) * 4 * nwg);
    d.beta = (double*)scratch_.alloc(sizeof(double) * std::max(jb.k, 1));
    sync_off.push_back(sync_words);
    sync_words += barrier_words(nwg);
    nts.push_back(G ? nt : 0);
    d.err = qp3_err_;
    d.rank_out = ranks_dev + q;
    d.ldc = N; d.k = jb.k; d.N = jb.N; d.ldv = jb.k; d.tol = tol_; d.nwg = nwg; d.G = G; d.rc = G ? rc : 0;
    d.part_col = G ? alloc_scr((size_t)2 * nwg * G * rc) : nullptr;  // a column slot of G * rc rows per workgroup
    qd.push_back(d);
    const int kk = std::min(jb.k, jb.N);
    if (!trial_) flops_ += 4.0 * jb.k * jb.N * kk - 2.0 * (jb.k + jb.N) * kk * kk + 4.0 / 3.0 * kk * kk * kk;
  }
  if (tol_ >= 1.0) {
    for (auto& j : jobs) j.rank = 0;  // util.cpp:798-799
    return;
  }
  // barrier counters start at zero
  unsigned int* syncs = (unsigned int*)scratch_.alloc(sizeof(unsigned int) * std::max(sync_words, 1));
  for (size_t q = 0; q < qd.size(); ++q) qd[q].sync = syncs + sync_off[q];
  launch_zero(stream_, 4, syncs, sync_words);
  struct Group { int nt, rc, coop, gmax; std::vector<Qp3Desc<T>> d; int nmax = 0; };
  std::vector<Group> groups;
  std::vector<std::pair<int, int>> forms = {{0, 0}};  // (threads, rows per lane); 0: streaming form
  for (int nt : {256, 1024})
    for (int r : qp3r_rcs())
      if (nt == 256 || r <= 32) forms.push_back({nt, r});
  for (auto form : forms) {
    const int nt = form.first, rcf = form.second;
    bool any = false;
    for (size_t q = 0; q < qd.size() && !any; ++q) any = nts[q] == nt && qd[q].rc == rcf;
    if (!any) continue;
    const int cap = nt ? qp3r_cap(nt, rcf) : qp3_capacity<T>(kmax);
    Group single{nt, rcf, 0, 1, {}};
    Group cur{nt, rcf, 1, 1, {}};
    int used = 0;
    for (size_t q = 0; q < qd.size(); ++q) {
      if (nts[q] != nt || qd[q].rc != rcf) continue;
      Qp3Desc<T> d = qd[q];
      if (d.nwg == 1) {
        single.d.push_back(d);
        single.gmax = std::max(single.gmax, d.G);
        continue;
      }
      if (!nt && d.nwg > cap) d.nwg = std::max(1, cap);
      if (used + d.nwg > cap && !cur.d.empty()) { groups.push_back(cur); cur = Group{nt, rcf, 1, 1, {}}; used = 0; }
      cur.d.push_back(d);
      cur.gmax = std::max(cur.gmax, d.G);
      used += d.nwg;
    }
    if (!single.d.empty()) groups.push_back(single);
    if (!cur.d.empty()) groups.push_back(cur);
  }
  // SPAQR_QP3_LOG=1 with kernel timing: one stderr line per launch (problem
  // shapes, ranks and device time), for tuning the launch forms
  struct LaunchLog { const char* fam; size_t ev; std::vector<const int*> ranks; int nd, nwg, kmax, nmax, wgmax; };
  std::vector<LaunchLog> llog;
  for (auto& gq : groups) {
    std::vector<int> cnt, st;
    double qby = 0;
    for (auto& d : gq.d) {
      cnt.push_back(d.nwg);
      qby += 2.0 * sizeof(T) * (double)d.k * d.N;  // C read once, written once
    }
    const int nt = prefix_tiles(cnt, st);
    const Qp3Desc<T>* dqd = stager_.push(gq.d);
    const int* dst = stager_.push(st);
    int rc = 0;
    // one timing family per kernel: streaming, register 256 / 1024 threads;
    // flops are added once the ranks are known
    const char* fam = "qp3s";
    if (gq.nt) {
      static std::mutex mu;
      static std::set<std::string> names;  // interned family names (the map below keeps pointers)
      char buf[32];
      snprintf(buf, sizeof buf, "qp3r%d_%d", gq.nt, gq.rc);
      std::lock_guard<std::mutex> lk(mu);
      fam = names.insert(buf).first->c_str();
    }
    for (auto& d : gq.d) qp3_family_[d.rank_out] = fam;
    timed(fam, 0.0, qby, [&] {
      int nmax = 0;
      for (auto& d : gq.d) nmax = std::max(nmax, d.N);
      rc = gq.nt ? launch_qp3r<T>(stream_, dqd, dst, (int)gq.d.size(), nt, gq.gmax, nmax, gq.nt, gq.rc, gq.coop)
                 : launch_qp3<T>(stream_, dqd, dst, (int)gq.d.size(), nt, kmax, gq.coop);
    });
    if (rc != 0) {
      int nmax = 0;
      for (auto& d : gq.d) nmax = std::max(nmax, d.N);
      char shape[200];
      snprintf(shape, sizeof shape,
               " (threads %d, rows per lane %d, cooperative %d, problems %zu, workgroups %d, lanes %d, N max %d, "
               "k max %d, capacity %d)",
               gq.nt, gq.rc, gq.coop, gq.d.size(), nt, gq.gmax, nmax, kmax,
               gq.nt ? qp3r_cap(gq.nt, gq.rc) : qp3_capacity<T>(kmax));
      throw Error(kErrHip, std::string("qp3 launch failed: ") + hipGetErrorString((hipError_t)rc) + shape);
    }
    if (qp3_log_ && kprof_ && !trial_) {
      LaunchLog L{fam, kev_[fam].ev.size() - 1, {}, (int)gq.d.size(), nt, 0, 0, 0};
      for (auto& d : gq.d) {
        L.ranks.push_back(d.rank_out);
        L.kmax = std::max(L.kmax, d.k);
        L.nmax = std::max(L.nmax, d.N);
        L.wgmax = std::max(L.wgmax, d.nwg);
      }
      llog.push_back(L);
    }
  }
  SPQ_HIP(hipMemcpyAsync(rank_host_, ranks_dev, sizeof(int) * jobs.size(), hipMemcpyDeviceToHost, stream_));
  SPQ_HIP(hipMemcpyAsync(rank_host_ + 65535, qp3_err_, sizeof(int), hipMemcpyDeviceToHost, stream_));
  sync();
  if (rank_host_[65535] != 0) throw Error(kErrNumeric, "qp3: an inter-workgroup barrier wait timed out");
  for (size_t q = 0; q < jobs.size(); ++q) jobs[q].rank = rank_host_[q];
  for (auto& L : llog) {
    float ms = 0;
    auto& e = kev_[L.fam].ev[L.ev];
    SPQ_HIP(hipEventElapsedTime(&ms, e.first, e.second));
    long rsum = 0;
    int rmax = 0;
    for (const int* r : L.ranks) {
      const int v = rank_host_[r - ranks_dev];
      rsum += v;
      rmax = std::max(rmax, v);
    }
    fprintf(stderr, "qp3 lvl %d %-12s problems %5d wgs %5d max(k %5d N %6d nwg %4d rank %4d) sum rank %7ld ms %.3f\n",
            ilvl_, L.fam, L.nd, L.nwg, L.kmax, L.nmax, L.wgmax, rmax, rsum, ms);
  }
  if (kprof_) {
    // truncated RRQR: 4kNr - 2(k+N)r^2 + 4r^3/3 + 2kN (initial norms)
    for (size_t q = 0; q < jobs.size(); ++q) {
      const double k = jobs[q].k, N = jobs[q].N, r = jobs[q].rank < jobs[q].k ? jobs[q].rank + 1 : jobs[q].rank;
      auto it = qp3_family_.find(qd[q].rank_out);
      if (it != qp3_family_.end())
        kev_[it->second].flops += 4 * k * N * r - 2 * (k + N) * r * r + 4.0 / 3.0 * r * r * r + 2 * k * N;
    }
  }
  qp3_family_.clear();
}

// tree.cpp:1260-1307 (+ sparsify_extra_rows 226-247), least squares: the
// extra rows (rows - cols) of an interface's in-edge blocks are compressed by
// a truncated pivoted QR; they become triu(R(0:r,:)) P^T, r = choose_rank,
// and the interface keeps cols + r rows.  Interfaces are independent here
// (each rewrites only its own in-edge blocks), so one batch covers a level.
template <typename T>
void Engine<T>::sparsify_extra_level(const std::vector<int>& cs) {
  struct X { int c, cols, k, N; std::vector<int> apm; T* Ct; };
  std::vector<X> xs;
  std::vector<CopyDesc> gather, scatter;
  std::vector<RrqrJob> rj;
  scratch_.rewind();
  for (int c : cs) {
    if (cl_[c].csize == 0 || cl_[c].rsize <= cl_[c].csize) continue;
    X x;
    x.c = c; x.cols = cl_[c].csize; x.k = cl_[c].rsize - x.cols; x.N = 0;
    for (int e : cl_[c].in)
      if (!cl_[ed_[e].n1].elim) { x.apm.push_back(e); x.N += ed_[e].a.cols; }
    x.Ct = alloc_scr((size_t)std::max(x.N, 1) * x.k);
    int off = 0;
    for (int e : x.apm) {
      const Block& b = ed_[e].a;
      if (b.cols > 0) gather.push_back(CopyDesc{b.p + x.cols, x.Ct + off, b.ld, std::max(x.N, 1), x.k, b.cols, 1, 0});
      off += b.cols;
    }
    if (x.N > 0) rj.push_back(RrqrJob{x.Ct, x.k, x.N});
    xs.push_back(std::move(x));
  }
  if (xs.empty()) return;
  copy_blocks(gather);
  rrqr_batch(rj);
  size_t q = 0;
  for (auto& x : xs) {
    const int r = x.N > 0 ? rj[q++].rank : 0;
    extra_.push_back({ilvl_, cl_[x.c].id.self.lvl, cl_[x.c].id.self.sep, x.k, r});
    int off = 0;
    for (int e : x.apm) {
      Block& b = ed_[e].a;
      if (r > 0 && b.cols > 0) scatter.push_back(CopyDesc{x.Ct + off, b.p + x.cols, x.N, b.ld, b.cols, r, 1, 0});
      b.rows = x.cols + r;
      off += b.cols;
    }
    cl_[x.c].rsize = x.cols + r;
    ed_[find_out(x.c, x.c)].a.rows = x.cols + r;
  }
  copy_blocks(scatter);
}

// tree.cpp:992-1114 (sparsifyD, scale = 1, order = 1) in dependency waves:
// an interface waits for every earlier interface it shares a block with.
template <typename T>
void Engine<T>::sparsify_level(const std::vector<int>& cs) {
  std::vector<int> S;
  for (int c : cs)
    if (cl_[c].csize > 0) S.push_back(c);
  // waves: an interface waits for every earlier adjacent interface (an edge
  // either way); neighbours come from the out- and in-edge lists
  std::vector<int> slot(cl_.size(), -1);
  for (size_t i = 0; i < S.size(); ++i) slot[S[i]] = (int)i;
  std::vector<int> wave(S.size(), 0);
  int nw = S.empty() ? 0 : 1;
  for (size_t i = 0; i < S.size(); ++i) {
    int w = 0;
    auto see = [&](int n) {
      const int jj = n >= 0 ? slot[n] : -1;
      if (jj >= 0 && jj < (int)i) w = std::max(w, wave[jj] + 1);
    };
    for (int n : cl_[S[i]].outn) if (n != S[i]) see(n);
    for (int n : cl_[S[i]].inn) if (n != S[i]) see(n);
    wave[i] = w;
    nw = std::max(nw, w + 1);
  }
  info_[ilvl_].sprs_waves = nw;
  info_[ilvl_].n_sprs = (int)S.size();
  std::vector<SparsifyInfo> recs(S.size());
  for (int w = 0; w < nw; ++w) {
    scratch_.rewind();
    const bool imp = order_ == 1.5f;  // sparsifyD_imp (tree.cpp:1117-1257)
    struct Job {
      int c, k, N, smp, spm = 0;
      std::vector<int> amp, apm;
      T* W = nullptr;  // order 1.5: top k rows of the in-edge blocks, k x spm (ld k)
      T* Rest = nullptr;  // least squares: rows [k, rows) of the in-edge blocks, rest x spm
      int rest = 0;
      T *Ct, *V, *tau;
    };
    std::vector<CopyDesc> gather;
    std::vector<RrqrJob> rj;
    int kmax = 0;
    std::vector<int> wave_idx;
    for (size_t i = 0; i < S.size(); ++i)
      if (wave[i] == w) { wave_idx.push_back((int)i); kmax = std::max(kmax, cl_[S[i]].csize); }
    // per interface of the wave: its edge lists and sizes (read-only scans,
    // on the host pool), then the scratch in creation order, then the
    // gather tables at fixed offsets (on the pool)
    const int nj = (int)wave_idx.size();
    std::vector<Job> jobs(nj);
    std::vector<size_t> goff(nj + 1, 0);
    parallel_for(nj, 64, [&](int, int b, int e) {
      for (int q = b; q < e; ++q) {
        const int c = S[wave_idx[q]];
        Job& jb = jobs[q];
        jb.c = c; jb.k = cl_[c].csize; jb.smp = 0;
        int spm = 0;
        const Clu& k = cl_[c];
        for (size_t t = 0; t < k.in.size(); ++t)
          if (!cl_[k.inn[t]].elim) { jb.apm.push_back(k.in[t]); spm += ed_[k.in[t]].a.cols; }
        for (size_t t = 0; t < k.out.size(); ++t)
          if (!cl_[k.outn[t]].elim && k.outn[t] != c) { jb.amp.push_back(k.out[t]); jb.smp += ed_[k.out[t]].a.rows; }
        jb.N = jb.smp + spm;
        jb.spm = spm;
        jb.rest = cl_[c].rsize - jb.k;
        size_t ng = jb.amp.size() + jb.apm.size();
        if (jb.rest > 0 && spm > 0) ng += jb.apm.size();
        if (imp && spm > 0) ng += jb.apm.size();
        goff[q + 1] = ng;
      }
    });
    for (int q = 0; q < nj; ++q) {
      Job& jb = jobs[q];
      goff[q + 1] += goff[q];
      jb.Ct = alloc_scr((size_t)std::max(jb.N, 1) * jb.k);
      // Apmc_rest (tree.cpp:1016), moved up under the rank rows afterwards
      if (jb.rest > 0 && jb.spm > 0) jb.Rest = alloc_scr((size_t)jb.rest * jb.spm);
      // Apmc_top (tree.cpp:1130-1140), kept before the blocks are rewritten
      if (imp && jb.spm > 0) jb.W = alloc_scr((size_t)jb.k * jb.spm);
      rj.push_back(RrqrJob{jb.Ct, jb.k, jb.N});
    }
    gather.resize(goff[nj]);
    parallel_for(nj, 64, [&](int, int b, int e) {
      for (int q = b; q < e; ++q) {
        const Job& jb = jobs[q];
        const int N = std::max(jb.N, 1);
        CopyDesc* g = gather.data() + goff[q];
        int off = 0;
        for (int e2 : jb.amp) {
          const Block& bl = ed_[e2].a;
          *g++ = CopyDesc{bl.p, jb.Ct + off, bl.ld, N, bl.rows, jb.k, 0, 0};
          off += bl.rows;
        }
        for (int e2 : jb.apm) {
          const Block& bl = ed_[e2].a;
          *g++ = CopyDesc{bl.p, jb.Ct + off, bl.ld, N, jb.k, bl.cols, 1, 0};
          off += bl.cols;
        }
        if (jb.Rest) {
          int c2 = 0;
          for (int e2 : jb.apm) {
            const Block& bl = ed_[e2].a;
            *g++ = CopyDesc{bl.p + jb.k, jb.Rest + (int64_t)c2 * jb.rest, bl.ld, jb.rest, jb.rest, bl.cols, 0, 0};
            c2 += bl.cols;
          }
        }
        if (jb.W) {
          int c2 = 0;
          for (int e2 : jb.apm) {
            const Block& bl = ed_[e2].a;
            *g++ = CopyDesc{bl.p, jb.W + (int64_t)c2 * jb.k, bl.ld, jb.k, jb.k, bl.cols, 0, 0};
            c2 += bl.cols;
          }
        }
      }
    });
    copy_blocks(gather);
    rrqr_batch(rj);
    std::vector<int> ranks(jobs.size(), 0);
    for (size_t q = 0; q < jobs.size(); ++q) {
      ranks[q] = rj[q].rank;
      jobs[q].V = rj[q].V;
      jobs[q].tau = rj[q].tau;
    }
    if (verbose_ > 1) {
      int nmax = 0, nwgs = 0, rsum = 0, ksum = 0;
      double kn = 0;
      for (size_t q = 0; q < jobs.size(); ++q) {
        nmax = std::max(nmax, jobs[q].N);
        nwgs += rj[q].nwg;
        rsum += ranks[q];
        ksum += jobs[q].k;
        kn += (double)jobs[q].k * jobs[q].N;
      }
      fprintf(stderr, "sprs lvl %d wave %d: problems %zu kmax %d Nmax %d sum_kN %.3g workgroups %d ranks %d/%d\n", ilvl_,
              w, jobs.size(), kmax, nmax, kn, nwgs, rsum, ksum);
    }
    std::vector<CopyDesc> scatter, afm_copies;
    std::vector<LarftDesc<T>> lf;
    std::vector<int> lcnt;
    struct QtJob { const T* V; const T* Tb; T* W; int k, rank, spm; };
    std::vector<QtJob> qt;  // order 1.5: Apmc_top <- Q^T Apmc_top (ormqr_left, tree.cpp:1234)
    for (size_t q = 0; q < jobs.size(); ++q) {
      Job& jb = jobs[q];
      const int c = jb.c, k = jb.k, rank = ranks[q];
      const int before = cl_[c].csize;
      if (rank >= k) {  // tree.cpp:1050
        recs[slot[c]] = {ilvl_, cl_[c].id.self.lvl, cl_[c].id.self.sep, before, cl_[c].csize};
        continue;
      }
      const int N = std::max(jb.N, 1);
      // OrthogonalD op: V (k x rank) kept, block T factors from tau
      T* V = alloc_perm((size_t)k * std::max(rank, 1));
      T* tau = alloc_perm(std::max(rank, 1));
      T* Tb = alloc_perm((size_t)nb_ * iceil(std::max(rank, 1), nb_) * nb_);
      factor_bytes_ += sizeof(T) * ((double)k * rank + (double)nb_ * iceil(std::max(rank, 1), nb_) * nb_);
      if (rank > 0) {
        scatter.push_back(CopyDesc{jb.V, V, k, k, k, rank, 0, 0});
        scatter.push_back(CopyDesc{jb.tau, tau, rank, rank, rank, 1, 0, 0});
        lf.push_back(LarftDesc<T>{V, tau, Tb, k, k, rank, 0});
        lcnt.push_back(iceil(rank, nb_));
      }
      nnzQ_ += (unsigned long long)(k * rank);
      Op od;
      od.kind = OpOrthoD; od.level = ilvl_; od.c = c; od.wave = w;
      od.V = V; od.ldv = k; od.m = k; od.k = rank; od.Tb = Tb;
      od.segs.push_back({c, k});
      ops_.push_back(std::move(od));
      // fine cluster (tree.cpp:1068-1079)
      const int cfn = k - rank;
      const int cf = new_cluster(cl_[c].cstart + rank, cfn, cl_[c].rstart + k - cfn, cfn, cl_[c].id);
      alloc_x(cf, cfn);
      cl_[cf].elim = true;
      Op sp;
      sp.kind = OpSplitD; sp.level = ilvl_; sp.c = c; sp.f = cf; sp.wave = w;
      sp.crows = cl_[c].rsize; sp.ccols = k; sp.rank = rank;
      ops_.push_back(std::move(sp));
      // new blocks (tree.cpp:1082-1110)
      const int rest = cl_[c].rsize - k;
      const int rank_rows = rank + rest;
      cl_[c].rsize = rank_rows;
      cl_[c].csize = rank;
      int off = 0;
      for (int e : jb.amp) {
        Block& b = ed_[e].a;
        if (rank > 0) scatter.push_back(CopyDesc{jb.Ct + off, b.p, N, b.ld, b.rows, rank, 0, 0});
        b.cols = rank;
        off += b.rows;
      }
      int qr0 = 0;
      for (int e : jb.apm) {
        Block& b = ed_[e].a;
        if (rank > 0) scatter.push_back(CopyDesc{jb.Ct + off, b.p, N, b.ld, b.cols, rank, 1, 0});
        if (rest > 0 && b.cols > 0)
          scatter.push_back(CopyDesc{jb.Rest + (int64_t)qr0 * rest, b.p + rank, rest, b.ld, rest, b.cols, 0, 0});
        b.rows = rank_rows;
        off += b.cols;
        qr0 += b.cols;
      }
      const int es = find_out(c, c);
      if (imp) {
        // fine cluster of the order-1.5 scheme (tree.cpp:1196-1254): self block
        // Aff = top-left cf x cf of the self block, in-edges E2 = rows
        // [rank, k) of Q^T Apmc_top, and a QR op with tau = 0
        const Block self = ed_[es].a;
        const int eff = add_edge(cf, cf, nullptr);
        scatter.push_back(CopyDesc{self.p, ed_[eff].a.p, self.ld, ed_[eff].a.ld, cfn, cfn, 0, 0});
        Op fq;
        fq.kind = OpFineQR; fq.level = ilvl_; fq.c = cf; fq.wave = w;
        fq.R = ed_[eff].a.p; fq.ldr = ed_[eff].a.ld; fq.kr = cfn;
        int qc = 0;
        for (int e : jb.apm) {
          const int m = ed_[e].n1, mc = cl_[m].csize;
          const int ne = add_edge(m, cf, nullptr);
          const Block& nb2 = ed_[ne].a;
          if (mc > 0)
            afm_copies.push_back(CopyDesc{jb.W + rank + (int64_t)qc * k, nb2.p, k, nb2.ld, cfn, mc, 0, 0});
          fq.r12.push_back({nb2, {m, mc}});
          qc += mc;
        }
        ops_.push_back(std::move(fq));
        if (rank > 0 && jb.spm > 0) qt.push_back(QtJob{V, Tb, jb.W, k, rank, jb.spm});
        else if (jb.spm > 0 && rank == 0) {
          // Q = I: the in-edges are the top rows unchanged (copies above)
        }
      }
      ed_[es].a.rows = rank_rows;
      ed_[es].a.cols = rank;
      recs[slot[c]] = {ilvl_, cl_[c].id.self.lvl, cl_[c].id.self.sep, before, rank};
    }
    for (auto& jb : jobs) {  // profile: the RRQR's columns (neighbour columns) and wave
      recs[slot[jb.c]].nbrs = jb.N;
      recs[slot[jb.c]].wave = w;
    }
    copy_blocks(scatter);
    if (!lf.empty()) {
      std::vector<int> st;
      const int nt = prefix_tiles(lcnt, st);
      const LarftDesc<T>* dl = stager_.push(lf);
      const int* dst = stager_.push(st);
      timed("larft", 0.0, 0.0, [&] { launch_larft<T>(stream_, dl, dst, (int)lf.size(), nt); });
    }
    if (!qt.empty()) {
      // W <- H^T W block by block: Y = V_b^T W, Z = T_b^T Y, W -= V_b Z
      int rmax = 0;
      for (auto& j : qt) rmax = std::max(rmax, j.rank);
      for (int jb = 0; jb < rmax; jb += nb_) {
        ws_.rewind();
        std::vector<GemmDesc<T>> g1, g2, g3;
        for (auto& j : qt) {
          if (jb >= j.rank) continue;
          const int wdt = std::min(nb_, j.rank - jb), rows = j.k - jb;
          const T* Vb = j.V + jb + (int64_t)jb * j.k;
          const T* Tbb = j.Tb + (int64_t)jb * nb_;
          T* Y = (T*)ws_.alloc(sizeof(T) * (size_t)nb_ * j.spm);
          T* Z = (T*)ws_.alloc(sizeof(T) * (size_t)nb_ * j.spm);
          g1.push_back(gd(Vb, j.k, j.W + jb, j.k, Y, nb_, wdt, j.spm, rows, gemm_flags(true, false, kStructUnitLower), T(1), T(0)));
          g2.push_back(gd(Tbb, nb_, Y, nb_, Z, nb_, wdt, j.spm, wdt, gemm_flags(true, false, kStructUpper), T(1), T(0)));
          g3.push_back(gd(Vb, j.k, Z, nb_, j.W + jb, j.k, rows, j.spm, wdt, gemm_flags(false, false, kStructUnitLower), T(-1), T(1)));
        }
        gemm_batch(g1);
        gemm_batch(g2);
        gemm_batch(g3);
      }
    }
    copy_blocks(afm_copies);
  }
  sprs_.insert(sprs_.end(), recs.begin(), recs.end());
}

// tree.cpp:887-989 (sparsify, --scale 0) for every interface of the level,
// wave by wave as in sparsify_level: per interface
//   s = sigma_min([App; Ampc])             (QR, then k_smin on R)
//   C = [Ampc^T, (1/s) App^T Apmc]         (gathers, GEMMs, k_scale_inv)
//   geqp3 + choose_rank on C               (rrqr_batch)
//   App <- App Q                           (larft, three GEMMs per block)
//   Amp blocks <- triu(R[0:rank]) P^T      (from the RRQR output)
// then the fine cluster (the last k - rank columns and rows), split_edges
// (tree.cpp:480-530) and the inline eliminate_cluster(cf) (tree.cpp:988),
// batched over the wave.  Square trees only (least squares forces scale 1).
template <typename T>
void Engine<T>::sparsify0_level(const std::vector<int>& cs) {
  if (!square_) throw Error(kErrState, "--scale 0 sparsification needs a square tree");
  std::vector<int> S;
  for (int c : cs)
    if (cl_[c].csize > 0) S.push_back(c);
  std::vector<int> slot(cl_.size(), -1);
  for (size_t i = 0; i < S.size(); ++i) slot[S[i]] = (int)i;
  std::vector<int> wave(S.size(), 0);
  int nw = S.empty() ? 0 : 1;
  for (size_t i = 0; i < S.size(); ++i) {
    int w = 0;
    auto see = [&](int n) {
      const int jj = n >= 0 && n < (int)slot.size() ? slot[n] : -1;
      if (jj >= 0 && jj < (int)i) w = std::max(w, wave[jj] + 1);
    };
    for (int n : cl_[S[i]].outn) if (n != S[i]) see(n);
    for (int n : cl_[S[i]].inn) if (n != S[i]) see(n);
    wave[i] = w;
    nw = std::max(nw, w + 1);
  }
  info_[ilvl_].sprs_waves = nw;
  info_[ilvl_].n_sprs = (int)S.size();
  std::vector<SparsifyInfo> recs(S.size());
  for (int w = 0; w < nw; ++w) {
    scratch_.rewind();
    struct Job {
      int c, k, rows, N, smp, spm, es, map;
      std::vector<int> amp, apm;
      T *Ap, *Ct;
    };
    std::vector<Job> jobs;
    std::vector<CopyDesc> gather;
    std::vector<QrJob> qr;
    for (size_t i = 0; i < S.size(); ++i) {
      if (wave[i] != w) continue;
      const int c = S[i];
      Job jb;
      jb.c = c; jb.k = cl_[c].csize; jb.rows = cl_[c].rsize; jb.smp = 0; jb.spm = 0;
      jb.es = find_out(c, c);
      if (jb.es < 0) throw Error(kErrState, "sparsify: interface without a self block");
      if (jb.k > kSminMaxK) throw Error(kErrArg, "--scale 0: interface wider than the sigma_min kernel limit");
      for (int e : cl_[c].in)
        if (!cl_[ed_[e].n1].elim && ed_[e].n1 != c) { jb.apm.push_back(e); jb.spm += ed_[e].a.cols; }
      for (int e : cl_[c].out)
        if (!cl_[ed_[e].n2].elim && ed_[e].n2 != c) { jb.amp.push_back(e); jb.smp += ed_[e].a.rows; }
      jb.N = jb.smp + jb.spm;
      jb.map = jb.rows + jb.smp;
      // Ap = [App; Ampc] (tree.cpp:921-923), factorised for its R
      jb.Ap = alloc_scr((size_t)jb.map * jb.k);
      const Block& self = ed_[jb.es].a;
      gather.push_back(CopyDesc{self.p, jb.Ap, self.ld, jb.map, jb.rows, jb.k, 0, 0});
      // Ct = C^T: rows [0, smp) = Ampc, rows [smp, N) = (1/s) Apmc^T App
      const int N = std::max(jb.N, 1);
      jb.Ct = alloc_scr((size_t)N * jb.k);
      int off = 0;
      for (int e : jb.amp) {
        const Block& b = ed_[e].a;
        gather.push_back(CopyDesc{b.p, jb.Ap + jb.rows + off, b.ld, jb.map, b.rows, jb.k, 0, 0});
        gather.push_back(CopyDesc{b.p, jb.Ct + off, b.ld, N, b.rows, jb.k, 0, 0});
        off += b.rows;
      }
      QrJob q;
      q.P = jb.Ap; q.ld = jb.map; q.m = jb.map; q.k = jb.k;
      q.tau = alloc_scr(jb.k);
      q.Tb = alloc_scr((size_t)nb_ * iceil(jb.k, nb_) * nb_);
      qr.push_back(std::move(q));
      jobs.push_back(std::move(jb));
    }
    copy_blocks(gather);
    // P blocks: Ct[smp + q0 : ..., :] = A(c, n)^T App, before App changes
    std::vector<GemmDesc<T>> gp;
    for (auto& jb : jobs) {
      const Block& self = ed_[jb.es].a;
      const int N = std::max(jb.N, 1);
      int off = jb.smp;
      for (int e : jb.apm) {
        const Block& b = ed_[e].a;
        if (b.cols > 0 && jb.rows > 0)
          gp.push_back(gd(b.p, b.ld, self.p, self.ld, jb.Ct + off, N, b.cols, jb.k, jb.rows,
                          gemm_flags(true, false, kStructNone, kStructNone), T(1), T(0)));
        off += b.cols;
      }
    }
    gemm_batch(gp);
    blocked_qr(qr);
    // sigma_min of each R, then the (1/s) scaling of the P rows
    {
      std::vector<SminDesc<T>> sd;
      int kmax = 0;
      if (jobs.size() > 16000) throw Error(kErrArg, "--scale 0: too many interfaces in one sparsification wave");
      double* sdev = (double*)scratch_.alloc(sizeof(double) * std::max<size_t>(jobs.size(), 1));
      for (size_t q = 0; q < jobs.size(); ++q) {
        SminDesc<T> d;
        d.R = jobs[q].Ap; d.ld = jobs[q].map; d.k = jobs[q].k;
        d.work = (double*)scratch_.alloc(sizeof(double) * smin_work_elems(jobs[q].k));
        d.out = sdev + q;
        sd.push_back(d);
        kmax = std::max(kmax, jobs[q].k);
      }
      const SminDesc<T>* dsd = stager_.push(sd);
      timed("smin", 0.0, 0.0, [&] { launch_smin<T>(stream_, dsd, (int)sd.size(), kmax); });
      std::vector<ScaleInvDesc<T>> si;
      std::vector<int> cnt, st;
      for (size_t q = 0; q < jobs.size(); ++q) {
        const Job& jb = jobs[q];
        if (jb.spm == 0) continue;
        si.push_back(ScaleInvDesc<T>{jb.Ct, sdev + q, std::max(jb.N, 1), jb.smp, jb.N, jb.k});
        cnt.push_back(iceil(jb.spm * jb.k, 256));
      }
      if (!si.empty()) {
        const int nt = prefix_tiles(cnt, st);
        const ScaleInvDesc<T>* dsi = stager_.push(si);
        const int* dst = stager_.push(st);
        timed("copy", 0.0, 0.0, [&] { launch_scale_inv<T>(stream_, dsi, dst, (int)si.size(), nt); });
      }
      // an exactly singular Ap gives s = 0 (the reference then divides by 0)
      SPQ_HIP(hipMemcpyAsync(rank_host_ + 32768, sdev, sizeof(double) * jobs.size(), hipMemcpyDeviceToHost, stream_));
    }
    std::vector<RrqrJob> rj;
    for (auto& jb : jobs) rj.push_back(RrqrJob{jb.Ct, jb.k, jb.N});
    rrqr_batch(rj);  // synchronises (rank read-back)
    for (size_t q = 0; q < jobs.size(); ++q) {
      double sv;
      std::memcpy(&sv, (const char*)(rank_host_ + 32768) + sizeof(double) * q, sizeof(double));
      if (!(sv > 0.0)) throw Error(kErrNumeric, "--scale 0: [App; Amp] of an interface is numerically singular");
    }
    std::vector<CopyDesc> vcopy, bcopy;
    std::vector<LarftDesc<T>> lf;
    std::vector<int> lcnt;
    struct RqJob { T* App; int ld, rows, k, rank; const T* V; const T* Tb; };
    std::vector<RqJob> rq;
    std::vector<int> fines;
    for (size_t q = 0; q < jobs.size(); ++q) {
      Job& jb = jobs[q];
      const int c = jb.c, k = jb.k, rank = rj[q].rank;
      if (rank >= k) {  // tree.cpp:965
        recs[slot[c]] = {ilvl_, cl_[c].id.self.lvl, cl_[c].id.self.sep, k, k};
        continue;
      }
      const int N = std::max(jb.N, 1);
      T* V = alloc_perm((size_t)k * std::max(rank, 1));
      T* tau = alloc_perm(std::max(rank, 1));
      T* Tb = alloc_perm((size_t)nb_ * iceil(std::max(rank, 1), nb_) * nb_);
      factor_bytes_ += sizeof(T) * ((double)k * rank + (double)nb_ * iceil(std::max(rank, 1), nb_) * nb_);
      const Block self = ed_[jb.es].a;  // a copy: add_edge below may move ed_
      if (rank > 0) {
        vcopy.push_back(CopyDesc{rj[q].V, V, k, k, k, rank, 0, 0});
        vcopy.push_back(CopyDesc{rj[q].tau, tau, rank, rank, rank, 1, 0, 0});
        lf.push_back(LarftDesc<T>{V, tau, Tb, k, k, rank, 0});
        lcnt.push_back(iceil(rank, nb_));
        rq.push_back(RqJob{self.p, self.ld, jb.rows, k, rank, V, Tb});
      }
      nnzQ_ += (unsigned long long)(k * rank);
      // Orthogonal op (operations.h:113-129): backward x_c <- Q x_c only
      Op od;
      od.kind = OpOrthoD; od.level = ilvl_; od.c = c; od.wave = w; od.nofwd = true;
      od.V = V; od.ldv = k; od.m = k; od.k = rank; od.Tb = Tb;
      od.segs.push_back({c, k});
      ops_.push_back(std::move(od));
      // fine cluster: columns [rank, k), the last k - rank rows (tree.cpp:965-970)
      const int cfn = k - rank, rows0 = jb.rows, rows1 = jb.rows - cfn;
      const int cf = new_cluster(cl_[c].cstart + rank, cfn, cl_[c].rstart + rows0 - cfn, cfn, cl_[c].id);
      alloc_x(cf, cfn);
      Op sp;  // Split == SplitD on a square cluster (operations.cpp:118-131)
      sp.kind = OpSplitD; sp.level = ilvl_; sp.c = c; sp.f = cf; sp.wave = w;
      sp.crows = rows0; sp.ccols = k; sp.rank = rank;
      ops_.push_back(std::move(sp));
      cl_[c].rsize = rows1;
      cl_[c].csize = rank;
      // Amp blocks <- (triu(R[0:rank]) P^T)[:, their columns]^T (tree.cpp:976-983)
      int off = 0;
      for (int e : jb.amp) {
        Block& b = ed_[e].a;
        if (rank > 0) vcopy.push_back(CopyDesc{jb.Ct + off, b.p, N, b.ld, b.rows, rank, 0, 0});
        b.cols = rank;
        off += b.rows;
      }
      // split_edges (tree.cpp:480-530): in-edges of c give their last cfn rows
      // to new edges n1 -> cf; the self block App Q splits into four
      for (int e : jb.apm) {
        const int n1 = ed_[e].n1;
        const int ne = add_edge(n1, cf, nullptr, false, true);
        Block& b = ed_[e].a;
        const Block& nb2 = ed_[ne].a;
        if (b.cols > 0) bcopy.push_back(CopyDesc{b.p + rows1, nb2.p, b.ld, nb2.ld, cfn, b.cols, 0, 0});
        b.rows = rows1;
      }
      const int eff = add_edge(cf, cf, nullptr, true, true);
      const int efc = add_edge(cf, c, nullptr, true, false);   // A(c, f) = Acf (not an in-edge of c)
      const int ecf = add_edge(c, cf, nullptr, false, true);   // A(f, c) = Afc
      bcopy.push_back(CopyDesc{self.p + rows1 + (int64_t)rank * self.ld, ed_[eff].a.p, self.ld, ed_[eff].a.ld, cfn, cfn, 0, 0});
      if (rows1 > 0)
        bcopy.push_back(CopyDesc{self.p + (int64_t)rank * self.ld, ed_[efc].a.p, self.ld, ed_[efc].a.ld, rows1, cfn, 0, 0});
      if (rank > 0) bcopy.push_back(CopyDesc{self.p + rows1, ed_[ecf].a.p, self.ld, ed_[ecf].a.ld, cfn, rank, 0, 0});
      ed_[jb.es].a.rows = rows1;
      ed_[jb.es].a.cols = rank;
      auto& rsp = cl_[cf].rsp;
      for (int e : jb.apm) rsp.push_back(ed_[e].n1);
      rsp.push_back(c);
      cl_[cf].elim = false;
      fines.push_back(cf);
      recs[slot[c]] = {ilvl_, cl_[c].id.self.lvl, cl_[c].id.self.sep, k, rank};
    }
    for (auto& jb : jobs) {  // profile: the RRQR's columns and wave
      recs[slot[jb.c]].nbrs = jb.N;
      recs[slot[jb.c]].wave = w;
    }
    copy_blocks(vcopy);
    if (!lf.empty()) {
      std::vector<int> st;
      const int nt = prefix_tiles(lcnt, st);
      const LarftDesc<T>* dl = stager_.push(lf);
      const int* dst = stager_.push(st);
      timed("larft", 0.0, 0.0, [&] { launch_larft<T>(stream_, dl, dst, (int)lf.size(), nt); });
    }
    // App <- App Q = App H_1 ... H_rank, 32 reflectors at a time:
    // W = App[:, jb:] V_b, Z = W T_b, App[:, jb:] -= Z V_b^T (tree.cpp:957)
    {
      int rmax = 0;
      for (auto& j : rq) rmax = std::max(rmax, j.rank);
      for (int jb0 = 0; jb0 < rmax; jb0 += nb_) {
        ws_.rewind();
        std::vector<GemmDesc<T>> g1, g2, g3;
        for (auto& j : rq) {
          if (jb0 >= j.rank || j.rows == 0) continue;
          const int wdt = std::min(nb_, j.rank - jb0), kk = j.k - jb0;
          const T* Vb = j.V + jb0 + (int64_t)jb0 * j.k;
          const T* Tbb = j.Tb + (int64_t)jb0 * nb_;
          T* A0 = j.App + (int64_t)jb0 * j.ld;
          T* W = (T*)ws_.alloc(sizeof(T) * (size_t)j.rows * nb_);
          T* Z = (T*)ws_.alloc(sizeof(T) * (size_t)j.rows * nb_);
          g1.push_back(gd(A0, j.ld, Vb, j.k, W, j.rows, j.rows, wdt, kk, gemm_flags(false, false, kStructNone, kStructUnitLower), T(1), T(0)));
          g2.push_back(gd(W, j.rows, Tbb, nb_, Z, j.rows, j.rows, wdt, wdt, gemm_flags(false, false, kStructNone, kStructUpper), T(1), T(0)));
          g3.push_back(gd(Z, j.rows, Vb, j.k, A0, j.ld, j.rows, kk, wdt, gemm_flags(false, true, kStructNone, kStructUnitLower), T(-1), T(1)));
        }
        gemm_batch(g1);
        gemm_batch(g2);
        gemm_batch(g3);
      }
    }
    copy_blocks(bcopy);
    // inline elimination of the fine clusters (tree.cpp:988)
    if (!fines.empty()) eliminate_level(fines, w);
  }
  sprs_.insert(sprs_.end(), recs.begin(), recs.end());
}

// tree.cpp:319-411 (least squares), for every cluster eliminated at this
// level with more rows than columns: each extra row goes to the column
// neighbour whose in-edge block has the largest norm on that row (first one
// on ties, else the cluster itself); rows are ordered by target, appended to
// the target's blocks, and every other in-edge block of a grown cluster is
// padded with zero rows.  The growth is simulated on the host in the
// reference's order and each grown block is materialised once: zero, old
// rows, gathered rows.
template <typename T>
void Engine<T>::reassign_level(const std::vector<int>& cs) {
  std::vector<int> E;
  for (int c : cs)
    if (cl_[c].elim && cl_[c].rsize > cl_[c].csize) E.push_back(c);
  if (E.empty()) return;
  scratch_.rewind();
  // (1) norms of the extra rows of the column-neighbour in-edge blocks
  std::vector<RowNormDesc<T>> nd;
  std::vector<std::pair<int, int>> nkey;  // (cluster, edge)
  std::vector<int> ncnt, nstart_host;
  int total = 0;
  for (int c : E) {
    const int rsz = cl_[c].rsize - cl_[c].csize;
    for (int e : cl_[c].in) {
      const int n = ed_[e].n1;
      if (find_out(c, n) < 0) continue;
      const Block& b = ed_[e].a;
      nd.push_back(RowNormDesc<T>{b.p, nullptr, b.ld, cl_[c].csize, rsz, b.cols});
      nkey.push_back({c, e});
      nstart_host.push_back(total);
      ncnt.push_back(iceil(rsz, 256));
      total += rsz;
    }
  }
  std::vector<double> norms(total);
  if (total > 0) {
    double* dn = (double*)scratch_.alloc(sizeof(double) * total);
    for (size_t q = 0; q < nd.size(); ++q) nd[q].out = dn + nstart_host[q];
    std::vector<int> st;
    const int nt = prefix_tiles(ncnt, st);
    const RowNormDesc<T>* dd = stager_.push(nd);
    const int* dst = stager_.push(st);
    stager_.flush();
    launch_rownorm<T>(stream_, dd, dst, (int)nd.size(), nt);
    SPQ_HIP(hipMemcpyAsync(norms.data(), dn, sizeof(double) * total, hipMemcpyDeviceToHost, stream_));
    sync();
  }
  // (2) the reference's sequence on a symbolic copy of the block rows
  struct Grow {
    Block base;            // content before this level's reassignments
    int rows;              // current row count
    struct App { const T* src; int lds, row0; std::vector<int> idx; };
    std::vector<App> app;  // appended row runs
  };
  std::map<int, Grow> grow;  // edge -> growth
  auto rows_of = [&](int e) { auto it = grow.find(e); return it == grow.end() ? ed_[e].a.rows : it->second.rows; };
  auto touch = [&](int e) -> Grow& {
    auto it = grow.find(e);
    if (it == grow.end()) it = grow.emplace(e, Grow{ed_[e].a, ed_[e].a.rows, {}}).first;
    return it->second;
  };
  std::vector<Block> zero;
  size_t ni = 0;
  for (int c : E) {
    const int cols = cl_[c].csize, rsz = cl_[c].rsize - cols;
    std::vector<std::pair<int, double>> r2c(rsz, {c, 0.0});
    for (; ni < nkey.size() && nkey[ni].first == c; ++ni) {
      const int n = ed_[nkey[ni].second].n1;
      const double* nv = norms.data() + nstart_host[ni];
      for (int i = 0; i < rsz; ++i)
        if (nv[i] > r2c[i].second) r2c[i] = {n, nv[i]};
    }
    std::vector<int> rp(rsz);
    std::iota(rp.begin(), rp.end(), 0);
    std::stable_sort(rp.begin(), rp.end(), [&](int a, int b) { return cl_[r2c[a].first].order < cl_[r2c[b].first].order; });
    const std::vector<int> cin = cl_[c].in;  // add_edge below never adds in-edges to c
    for (int ein : cin) {
      const int nold = ed_[ein].n1;
      const Block src = ed_[ein].a;
      for (int i = 0; i < rsz;) {
        const int nnew = r2c[rp[i]].first;
        int inext = i + 1;
        while (inext < rsz && r2c[rp[inext]].first == nnew) ++inext;
        const int rinc = inext - i;
        int e = find_out(nold, nnew);
        if (e < 0) {
          if (nold == nnew) throw Error(kErrState, "reassign: missing self block");
          e = add_edge(nold, nnew, &zero);
        }
        Grow& g = touch(e);
        typename Grow::App a;
        a.src = src.p; a.lds = src.ld; a.row0 = g.rows;
        for (int t = i; t < inext; ++t) a.idx.push_back(cols + rp[t]);
        g.app.push_back(std::move(a));
        g.rows += rinc;
        if (nold == nnew) {
          Op o;
          o.kind = OpReassign; o.level = ilvl_; o.c = c; o.f = nnew;
          o.nstart = cl_[nnew].rsize;
          for (int t = i; t < inext; ++t) o.idx.push_back(cols + rp[t]);
          alloc_x(nnew, cl_[nnew].rsize + rinc);
          ops_.push_back(std::move(o));
        }
        i = inext;
      }
    }
    // sizes and zero padding (tree.cpp:394-410)
    for (int eo : cl_[c].out) {
      const int n = ed_[eo].n2;
      if (n == c) continue;
      const int es = find_out(n, n);
      cl_[n].rsize = rows_of(es);
      cl_[n].csize = ed_[es].a.cols;
      alloc_x(n, cl_[n].rsize);
      for (int e : cl_[n].in) {
        if (cl_[ed_[e].n1].elim) continue;
        if (rows_of(e) != cl_[n].rsize) touch(e).rows = cl_[n].rsize;
      }
    }
  }
  // (3) materialise: zero the new blocks, copy the old rows, gather the new
  std::vector<CopyDesc> copies;
  std::vector<RowGatherDesc<T>> rg;
  std::vector<int> rgcnt;
  for (auto& kv : grow) {
    Grow& g = kv.second;
    if (g.rows == g.base.rows && g.app.empty()) continue;
    Block nb = new_block(g.rows, g.base.cols);
    zero.push_back(nb);
    if (g.base.rows > 0 && g.base.cols > 0)
      copies.push_back(CopyDesc{g.base.p, nb.p, g.base.ld, nb.ld, g.base.rows, g.base.cols, 0, 0});
    for (auto& a : g.app) {
      const int* didx = stager_.upload(perm_, a.idx);
      rg.push_back(RowGatherDesc<T>{a.src, nb.p, didx, a.lds, nb.ld, a.row0, (int)a.idx.size(), nb.cols, 0});
      rgcnt.push_back(rowgather_tiles((int)a.idx.size(), nb.cols));
    }
    ed_[kv.first].a = nb;
  }
  init_blocks(zero, 0);
  copy_blocks(copies);
  if (!rg.empty()) {
    std::vector<int> st;
    const int nt = prefix_tiles(rgcnt, st);
    const RowGatherDesc<T>* dd = stager_.push(rg);
    const int* dst = stager_.push(st);
    timed("copy", 0.0, 0.0, [&] { launch_rowgather<T>(stream_, dd, dst, (int)rg.size(), nt); });
  }
}

template <typename T>
void Engine<T>::merge_all() {  // tree.cpp:872-881, 414-475
  cur_++;
  for (int snew : bottoms_[cur_]) {
    int rs = 0, cs = 0;
    for (int s : cl_[snew].children) { cl_[s].rpos = rs; cl_[s].cpos = cs; rs += cl_[s].rsize; cs += cl_[s].csize; }
    Clu& p = cl_[snew];
    p.rsize = rs; p.csize = cs; p.rsize_org = rs; p.csize_org = cs;
    merges_.push_back({ilvl_, p.id.self.lvl, p.id.self.sep, rs, cs});
    alloc_x(snew, rs);
    Op o;
    o.kind = OpMerge; o.level = ilvl_; o.c = snew;
    ops_.push_back(std::move(o));
  }
  std::vector<Block> zero;
  std::vector<CopyDesc> copies;
  // every new cluster's neighbours (the parents of its children's live
  // column neighbours): the children's edge lists do not change below, so
  // these are found on the host pool first
  const std::vector<int>& snews = bottoms_[cur_];
  std::vector<std::vector<int>> nbs(snews.size());
  parallel_for((int)snews.size(), 64, [&](int, int b, int e) {
    for (int i = b; i < e; ++i) {
      std::vector<int>& nb = nbs[i];
      for (int s : cl_[snews[i]].children)
        for (int e2 : cl_[s].out)
          if (!cl_[ed_[e2].n2].elim) nb.push_back(cl_[ed_[e2].n2].parent);
      std::sort(nb.begin(), nb.end());
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
    }
  });
  // the new clusters' edge lists get their final lengths up front (out:
  // one edge per neighbour, in: one per cluster naming it), so add_edge
  // appends without reallocating
  {
    size_t tot = 0;
    std::vector<int> indeg(cl_.size(), 0);
    for (size_t i = 0; i < snews.size(); ++i) {
      tot += nbs[i].size();
      for (int n : nbs[i])
        if (n != snews[i]) ++indeg[n];
    }
    ed_.reserve(ed_.size() + tot);
    zero.reserve(tot);
    for (size_t i = 0; i < snews.size(); ++i) {
      Clu& k = cl_[snews[i]];
      k.out.reserve(k.out.size() + nbs[i].size());
      k.outn.reserve(k.outn.size() + nbs[i].size());
      k.in.reserve(k.in.size() + indeg[snews[i]]);
      k.inn.reserve(k.inn.size() + indeg[snews[i]]);
    }
  }
  for (size_t i = 0; i < snews.size(); ++i) {
    const int snew = snews[i];
    const std::vector<int>& nb = nbs[i];
    cl_[snew].cnbrs = nb;
    for (int n : nb) add_edge(snew, n, &zero);
    sort_out(snew);
  }
  // every child block moves into its parent's block (rows at the child row
  // cluster's offset, columns at the child's); per new cluster in order, on
  // the host pool: the blocks of different new clusters are disjoint
  const int ns = (int)snews.size();
  std::vector<size_t> coff(ns + 1, 0);
  parallel_for(ns, 64, [&](int, int b, int e) {
    for (int i = b; i < e; ++i) {
      size_t n = 0;
      for (int s : cl_[snews[i]].children)
        for (int e2 : cl_[s].out) n += cl_[ed_[e2].n2].elim ? 0 : 1;
      coff[i + 1] = n;
    }
  });
  for (int i = 0; i < ns; ++i) coff[i + 1] += coff[i];
  copies.resize(coff[ns]);
  parallel_for(ns, 64, [&](int, int b, int e) {
    // snew's out edge towards each cluster (find_out's first match)
    std::vector<int> pos(cl_.size(), -1);
    for (int i = b; i < e; ++i) {
      const Clu& P = cl_[snews[i]];
      for (size_t q = P.outn.size(); q-- > 0;) pos[P.outn[q]] = P.out[q];
      CopyDesc* d = copies.data() + coff[i];
      for (int s : P.children)
        for (int e2 : cl_[s].out) {
          const int nold = ed_[e2].n2;
          if (cl_[nold].elim) continue;
          const Block& dst = ed_[pos[cl_[nold].parent]].a;
          const Block& src = ed_[e2].a;
          *d++ = CopyDesc{src.p, dst.p + cl_[nold].rpos + (int64_t)cl_[s].cpos * dst.ld, src.ld, dst.ld, src.rows,
                          src.cols, 0, 0};
          ed_[e2].alive = false;
        }
      for (int n2 : P.outn) pos[n2] = -1;
    }
  });
  init_blocks(zero, 0);
  copy_blocks(copies);
}

template <typename T>
int Engine<T>::factorize() {  // tree.cpp:1309-1437
  // the level loop consumes the assembled blocks and the cluster lists: one
  // factorization per assembled tree
  if (!assembled_) throw Error(kErrState, "factorize: assemble has not run");
  if (factored_ || factorizing_) throw Error(kErrState, "factorize: the tree is already factored (create a new tree)");
  factorizing_ = true;
  const double fstart = wall();
  for (ilvl_ = 0; ilvl_ < nlevels_; ++ilvl_) {
    LevelInfo& li = info_[ilvl_];
    const double t0 = wall(), w0 = t_wait_, f0 = t_flush_, l0 = t_launch_, a0 = Arena::t_chunks;
    const long n0 = n_launch_, c0 = Arena::n_chunks, rg0 = stager_.n_grow, rw0 = stager_.n_wrap;
    std::vector<int> lev;
    for (int s : bottoms_[cur_])
      if (cl_[s].level() == ilvl_) lev.push_back(s);
    fill_sparsity(lev);
    finish_sparsity();
    const double t1 = wall();
    eliminate_level(lev);
    if (timing_) sync();
    const double t2 = wall();
    if (!square_ && ilvl_ < nlevels_ - 1) reassign_level(lev);
    if (timing_) sync();
    const double t3 = wall();
    std::vector<int> itf;
    for (int s : bottoms_[cur_])
      if (want_sparsify(s)) itf.push_back(s);
    if (scale_ && ilvl_ < nlevels_ - 2 && ilvl_ >= skip_ && tol_ != 0) scale_level(itf);
    if (timing_) sync();
    const double t4 = wall();
    if (ilvl_ >= skip_ && ilvl_ < nlevels_ - 2 && tol_ != 0) {
      if (!scale_) {
        sparsify0_level(itf);
      } else {
        if (!square_) sparsify_extra_level(itf);
        sparsify_level(itf);
      }
    }
    if (timing_) sync();
    const double t5 = wall();
    if (ilvl_ < nlevels_ - 1) merge_all();
    // without phase synchronisation the next level's host planning (and,
    // after the last level, the solve program's) overlaps this level's last
    // kernels: the stream keeps the device work in order
    if (timing_) sync();
    const double t6 = wall();
    li.t_fill = t1 - t0; li.t_elmn = t2 - t1; li.t_shift = t3 - t2; li.t_scale = t4 - t3;
    li.t_sprsfy = t5 - t4; li.t_merge = t6 - t5;
    if (verbose_ > 1 && kprof_) {
      static thread_local std::map<std::string, double> prev;
      if (ilvl_ == 0) prev.clear();
      std::string line;
      for (auto& k : kernel_stats()) {
        char buf[96];
        snprintf(buf, sizeof buf, " %s %.1f", k.name.c_str(), k.ms - prev[k.name]);
        line += buf;
        prev[k.name] = k.ms;
      }
      fprintf(stderr,
              "lvl %d kernel ms:%s | wall %.1f (host busy %.1f: uploads %.1f, launches %.1f, %ld launches, %ld chunk "
              "allocations %.1f, staging ring %ld grown %ld wrapped) | elim %d in %d waves, sprs %d in %d waves\n",
              ilvl_, line.c_str(), 1e3 * (t6 - t0), 1e3 * ((t6 - t0) - (t_wait_ - w0)), 1e3 * (t_flush_ - f0),
              1e3 * (t_launch_ - l0), n_launch_ - n0, Arena::n_chunks - c0, 1e3 * (Arena::t_chunks - a0),
              stager_.n_grow - rg0, stager_.n_wrap - rw0, li.n_elim, li.elim_waves, li.n_sprs, li.sprs_waves);
    }
    auto ts = topsize();
    li.top_rows = ts.first; li.top_cols = ts.second;
    if (verbose_)
      printf("lvl: %d    update fill-in: %.3f  elmn: %.3f  shift: %.3f  scale: %.3f  sprsfy: %.3f  merge: %.3f  "
             "size top_sep: %d, %d   a.r top_sep: %.3f\n",
             ilvl_, li.t_fill, li.t_elmn, li.t_shift, li.t_scale, li.t_sprsfy, li.t_merge, ts.first, ts.second,
             (double)ts.first / (double)ts.second);
  }
  const double tbp = wall();
  build_solve_program();
  // the apply's graph is instantiated here, so solves (and the Krylov drivers
  // around them) allocate nothing on the device
  if (solve_graph_ && !stage_prof_) {
    if (square_) graph_exec_fwd_ = capture_program(0);
    else graph_exec_nrml_ = capture_program(1);
  }
  if (verbose_ > 1) fprintf(stderr, "solve program built in %.1f ms\n", 1e3 * (wall() - tbp));
  // a multi-workgroup kernel whose barrier wait gave up leaves the flag set
  SPQ_HIP(hipMemcpyAsync(rank_host_ + 65535, qp3_err_, sizeof(int), hipMemcpyDeviceToHost, stream_));
  sync();
  if (rank_host_[65535] != 0) throw Error(kErrNumeric, "factorize: an inter-workgroup barrier wait timed out");
  t_factorize_ = wall() - fstart;
  top_size_ = cl_[bottoms_[nlevels_ - 1][0]].csize;
  factored_ = true;
  if (verbose_) {
    printf("Tolerance set: %.3e\n", tol_);
    printf("Time to factorize:  %.3e\n", t_factorize_);
    printf("Size of top separator: %d\n", (int)top_size_);
    printf("nnzA: %llu nnzR: %llu\n", nnzA_, nnzR_);
    printf("nnzH: %llu nnzQ: %llu\n", nnzH_, nnzQ_);
  }
  return 0;
}

// Solve program: per level, forward stages QR (waves) -> ScaleD -> OrthogonalD
// -> SplitD -> Merge, backward stages in the reverse order (tree.cpp:1440-1466).
template <typename T>
void Engine<T>::build_solve_program() {
  layout_x();
  // temporary segments for SplitD moves whose ranges overlap (least-squares
  // clusters with extra rows)
  std::map<const Op*, int64_t> split_tmp;
  for (auto& o : ops_)
    if (o.kind == OpSplitD && o.crows > o.ccols) { split_tmp[&o] = xsize_; xsize_ += o.crows - o.ccols; }
  SPQ_HIP(dev_malloc((void**)&xbuf_, sizeof(T) * std::max<int64_t>(xsize_, 1)));
  if (!square_) SPQ_HIP(dev_malloc((void**)&xtbuf_, sizeof(T) * std::max<int64_t>(xsize_, 1)));
  SPQ_HIP(dev_malloc((void**)&bin_, sizeof(double) * std::max(nrows_, 1)));
  SPQ_HIP(dev_malloc((void**)&xout_, sizeof(double) * std::max(ncols_, 1)));
  std::vector<Seg> segs;
  std::vector<R12Desc<T>> r12;
  // buf: the x buffer a stage works on (0 = x, 1 = xt); moves copy sbuf -> buf
  // types 4 / 5: reflectors / triangular sweeps in FP64 on x64_ (the top
  // separator of an FP32 tree); 6 / 7: segment conversion x -> x64_ / back
  struct HStage {
    int type;  // 0 reflectors, 1 triangular sweeps, 2 moves, 3 R12 products
    std::vector<ReflDesc<T>> refl;
    std::vector<BwdDesc<T>> bwd;
    std::vector<MoveDesc> mv;
    std::vector<R12Chunk<T>> r12c;
    std::vector<ReflDesc<double>> refl64;
    std::vector<BwdDesc<double>> bwd64;
    std::vector<CopyDesc> cvt;
    int mmax = 0, buf = 0, sbuf = 0;
  };
  std::vector<HStage> hf, hb, htf, htb, hpf;
  auto push = [](std::vector<HStage>& v, HStage&& st) {
    if (!st.refl.empty() || !st.bwd.empty() || !st.mv.empty() || !st.r12c.empty() || !st.refl64.empty() ||
        !st.bwd64.empty() || !st.cvt.empty())
      v.push_back(std::move(st));
  };
  auto stage = [](int type, int buf, int sbuf = -1) {
    HStage st;
    st.type = type; st.buf = buf; st.sbuf = sbuf < 0 ? buf : sbuf;
    return st;
  };  // FP64 top-separator ops: their segments are converted into x64_ (one
  // contiguous range per op), applied in double, and converted back
  std::vector<Seg> segs64;
  int64_t x64_len = 0;
  for (auto& o : ops_)
    if (o.V64) {
      if (o.kind != OpQR || !o.r12.empty()) throw Error(kErrState, "FP64 top separator: unexpected op structure");
      x64_len = std::max<int64_t>(x64_len, o.m);
    }
  if (x64_len > 0) SPQ_HIP(dev_malloc((void**)&x64_, sizeof(double) * x64_len));
  // xb: the x buffer the op's segments live in (xbuf_, or xtbuf_ for the
  // tops program of the normal equations)
  auto qr64_stages = [&](const Op& o, int trans, bool bwd, std::vector<HStage>& out, T* xb) {
    HStage a = stage(6, 0), b = stage(bwd ? 5 : 4, 0), c = stage(7, 0);
    int64_t off = 0;
    T* const xbuf_ = xb;  // the conversions address this buffer
    if (!bwd) {
      for (auto& sg : o.segs) {
        a.cvt.push_back(CopyDesc{xbuf_ + cl_[sg.first].xoff, x64_ + off, sg.second, sg.second, sg.second, 1, 0, 0});
        c.cvt.push_back(CopyDesc{x64_ + off, xbuf_ + cl_[sg.first].xoff, sg.second, sg.second, sg.second, 1, 0, 0});
        off += sg.second;
      }
      ReflDesc<double> d{};
      d.V = o.V64; d.Tb = o.Tb64; d.ldv = o.ldv; d.m = o.m; d.k = o.k; d.nb = nb_;
      d.seg0 = (int)segs64.size(); d.nseg = 1; d.trans = trans; d.nwg = 1;
      segs64.push_back(Seg{0, o.m, 0});
      b.refl64.push_back(d);
      b.mmax = o.m;
      solve_bytes_ += sizeof(double) * ((double)o.m * o.k + (double)nb_ * o.k) + 2.0 * sizeof(double) * o.m;
    } else {
      const int64_t xo = cl_[o.c].xoff;
      a.cvt.push_back(CopyDesc{xbuf_ + xo, x64_, o.kr, o.kr, o.kr, 1, 0, 0});
      c.cvt.push_back(CopyDesc{x64_, xbuf_ + xo, o.kr, o.kr, o.kr, 1, 0, 0});
      BwdDesc<double> d{};
      d.R = o.V64; d.ldr = o.ldv; d.k = o.kr; d.xoff = 0; d.blk0 = 0; d.nblk = 0; d.trans = trans; d.nwg = 1;
      b.bwd64.push_back(d);
      solve_bytes_ += sizeof(double) * 0.5 * o.kr * o.kr + 2.0 * sizeof(double) * o.kr;
    }
    push(out, std::move(a));
    push(out, std::move(b));
    push(out, std::move(c));
  };

  auto refl_of = [&](const Op& o, int trans) {
    ReflDesc<T> d{};
    d.V = o.V; d.Tb = o.Tb; d.ldv = o.ldv; d.m = o.m; d.k = o.k; d.nb = nb_;
    d.seg0 = (int)segs.size(); d.nseg = (int)o.segs.size(); d.trans = trans;
    for (auto& s : o.segs) segs.push_back(Seg{cl_[s.first].xoff, s.second, 0});
    solve_bytes_ += sizeof(T) * ((double)o.m * o.k + (double)nb_ * o.k) + 2.0 * sizeof(T) * o.m;
    return d;
  };
  // triangular sweep of a QR op; its R12 products go to `rs` as column
  // chunks (run before the sweep when trans = 0, after it when trans = 1)
  auto bwd_of = [&](const Op& o, int trans, HStage& rs) {
    BwdDesc<T> d{};
    d.R = o.R; d.ldr = o.ldr; d.k = o.kr; d.xoff = cl_[o.c].xoff;
    d.blk0 = 0; d.nblk = 0; d.trans = trans;
    const int cc = r12_chunk_cols();
    for (auto& q : o.r12) {
      const int sn = q.second.second;
      for (int j0 = 0; j0 < sn; j0 += cc)
        rs.r12c.push_back(R12Chunk<T>{q.first.p + (int64_t)j0 * q.first.ld, cl_[o.c].xoff,
                                      cl_[q.second.first].xoff + j0, q.first.ld, o.kr, std::min(cc, sn - j0), trans});
      solve_bytes_ += sizeof(T) * (double)o.kr * sn;
    }
    solve_bytes_ += sizeof(T) * 0.5 * o.kr * o.kr + 2.0 * sizeof(T) * o.kr;
    return d;
  };
  auto scaled_of = [&](const Op& o, int trans) {
    BwdDesc<T> d{};
    d.R = o.R; d.ldr = o.ldr; d.k = o.k; d.xoff = cl_[o.c].xoff; d.trans = trans;
    solve_bytes_ += sizeof(T) * 0.5 * o.k * o.k + 2.0 * sizeof(T) * o.k;
    return d;
  };
  // SplitD (operations.cpp:120-131): forward x_f = x[rk:cc), x[rk:rk+rest) =
  // x[cc:cr); backward x[rk:cc) = x_f after x[cc:cr) = x[rk:rk+rest).  The
  // overlapping shift of the rest rows goes through a temporary segment, so
  // each direction is two stages (A then B).
  auto splitd_fwd = [&](const Op* o, HStage& A, HStage& B) {
    const int64_t xc = cl_[o->c].xoff, xf = cl_[o->f].xoff;
    const int n = o->ccols - o->rank, rest = o->crows - o->ccols;
    if (n > 0) A.mv.push_back(MoveDesc{xc + o->rank, xf, n, 0});
    if (rest > 0) {
      const int64_t t = split_tmp.at(o);
      A.mv.push_back(MoveDesc{xc + o->ccols, t, rest, 0});
      B.mv.push_back(MoveDesc{t, xc + o->rank, rest, 0});
    }
  };
  auto splitd_bwd = [&](const Op* o, HStage& A, HStage& B) {
    const int64_t xc = cl_[o->c].xoff, xf = cl_[o->f].xoff;
    const int n = o->ccols - o->rank, rest = o->crows - o->ccols;
    if (rest > 0) {
      const int64_t t = split_tmp.at(o);
      A.mv.push_back(MoveDesc{xc + o->rank, t, rest, 0});
      B.mv.push_back(MoveDesc{t, xc + o->ccols, rest, 0});
    }
    if (n > 0) B.mv.push_back(MoveDesc{xf, xc + o->rank, n, 0});
  };
  // Merge: children -> parent (rows on the way up, cols in the normal-
  // equation forward pass) and parent -> children
  auto merge_up = [&](const Op* o, HStage& S, bool rows) {
    int64_t k = 0;
    for (int ch : cl_[o->c].children) {
      const int len = rows ? cl_[ch].rsize : cl_[ch].csize;
      if (len > 0) S.mv.push_back(MoveDesc{cl_[ch].xoff, cl_[o->c].xoff + k, len, 0});
      k += len;
    }
  };
  auto merge_down = [&](const Op* o, HStage& S, bool rows) {
    int64_t k = 0;
    for (int ch : cl_[o->c].children) {
      const int len = rows ? cl_[ch].rsize : cl_[ch].csize;
      if (len > 0) S.mv.push_back(MoveDesc{cl_[o->c].xoff + k, cl_[ch].xoff, len, 0});
      k += len;
    }
  };
  // Reassign: x_f[nstart + t] = x_c[idx[t]] (forward) and back
  auto reassign = [&](const Op* o, HStage& S, bool fwd) {
    for (size_t t = 0; t < o->idx.size(); ++t) {
      const int64_t c = cl_[o->c].xoff + o->idx[t], f = cl_[o->f].xoff + o->nstart + (int64_t)t;
      S.mv.push_back(fwd ? MoveDesc{c, f, 1, 0} : MoveDesc{f, c, 1, 0});
    }
  };
  struct LevelOps { std::vector<const Op*> qr, sc, od, sd, mg, fq, ra; };
  std::vector<std::vector<const Op*>> by_level(nlevels_);
  for (auto& o : ops_) by_level[o.level].push_back(&o);
  auto level_ops = [&](int lvl) {
    LevelOps L;
    for (const Op* po : by_level[lvl]) {
      const Op& o = *po;
      switch (o.kind) {
        case OpQR: L.qr.push_back(&o); break;
        case OpScaleD: L.sc.push_back(&o); break;
        case OpOrthoD: L.od.push_back(&o); break;
        case OpSplitD: L.sd.push_back(&o); break;
        case OpMerge: L.mg.push_back(&o); break;
        case OpFineQR: L.fq.push_back(&o); break;
        case OpReassign: L.ra.push_back(&o); break;
      }
    }
    return L;
  };
  // ---- x <- M^{-1} x: ops forward, then backward (tree.cpp:1450-1457)
  std::vector<int> seg_wave(cl_.size(), -1);
  for (int lvl = 0; lvl < nlevels_; ++lvl) {
    LevelOps L = level_ops(lvl);
    // QR forward waves: the factorization's order (elimination wave, then
    // creation order), where ops sharing an x segment keep that order (wave
    // = 1 + the latest wave of an earlier op on one of its segments, tracked
    // per segment).  Same-wave eliminations share no block, hence (the
    // diagonal block of every panel row being a target) no segment, so the
    // solve has as many QR stages as the factorization had waves.
    std::vector<int> wv(L.qr.size(), 0);
    std::vector<int> ord(L.qr.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return L.qr[a]->wave < L.qr[b]->wave; });
    int nw = 0;
    for (int i : ord) {
      int w = 0;
      for (auto& a2 : L.qr[i]->segs) w = std::max(w, seg_wave[a2.first] + 1);
      wv[i] = w;
      for (auto& a2 : L.qr[i]->segs) seg_wave[a2.first] = std::max(seg_wave[a2.first], w);
      nw = std::max(nw, w + 1);
    }
    for (size_t i = 0; i < L.qr.size(); ++i)
      for (auto& a2 : L.qr[i]->segs) seg_wave[a2.first] = -1;
    for (int w = 0; w < nw; ++w) {
      HStage st = stage(0, 0);
      for (size_t i = 0; i < L.qr.size(); ++i)
        if (wv[i] == w && !L.qr[i]->V64) { st.refl.push_back(refl_of(*L.qr[i], 1)); st.mmax = std::max(st.mmax, L.qr[i]->m); }
      push(hf, std::move(st));
      for (size_t i = 0; i < L.qr.size(); ++i)
        if (wv[i] == w && L.qr[i]->V64) qr64_stages(*L.qr[i], 1, false, hf, xbuf_);
    }
    HStage sra = stage(2, 0);
    for (auto o : L.ra) reassign(o, sra, true);
    push(hf, std::move(sra));
    HStage ssc = stage(0, 0);
    for (auto o : L.sc) { ssc.refl.push_back(refl_of(*o, 1)); ssc.mmax = std::max(ssc.mmax, o->m); }
    push(hf, std::move(ssc));
    HStage sod = stage(0, 0);
    for (auto o : L.od)
      if (o->k > 0 && !o->nofwd) { sod.refl.push_back(refl_of(*o, 1)); sod.mmax = std::max(sod.mmax, o->m); }
    push(hf, std::move(sod));
    HStage sa = stage(2, 0), sb = stage(2, 0);
    for (auto o : L.sd) splitd_fwd(o, sa, sb);
    push(hf, std::move(sa));
    push(hf, std::move(sb));
    // --scale 0 fine clusters: their QR forward (segments {f, c} per
    // interface, disjoint across interfaces, so one stage)
    HStage sfq = stage(0, 0);
    for (auto o : L.fq)
      if (o->fwd_refl && o->k > 0) { sfq.refl.push_back(refl_of(*o, 1)); sfq.mmax = std::max(sfq.mmax, o->m); }
    push(hf, std::move(sfq));
    HStage smg = stage(2, 0);
    for (auto o : L.mg) merge_up(o, smg, true);
    push(hf, std::move(smg));
  }
  for (int lvl = nlevels_ - 1; lvl >= 0; --lvl) {
    LevelOps L = level_ops(lvl);
    HStage smg = stage(2, 0);
    for (auto o : L.mg) merge_down(o, smg, false);
    push(hb, std::move(smg));
    // sparsification ops in reverse: grouped by kind, or, when order-1.5
    // fine-cluster QRs read neighbouring interfaces, wave by wave (latest
    // wave first: FineQR, SplitD, OrthoD), which keeps the tape's order
    // between adjacent interfaces
    int wmax = 0;
    for (auto o : L.sd) wmax = std::max(wmax, o->wave);
    for (int w = L.fq.empty() ? 0 : wmax; w >= 0; --w) {
      const bool all = L.fq.empty();
      HStage sfr = stage(3, 0), sfq = stage(1, 0);
      for (auto o : L.fq) if (o->wave == w) sfq.bwd.push_back(bwd_of(*o, 0, sfr));
      push(hb, std::move(sfr));
      push(hb, std::move(sfq));
      HStage sa = stage(2, 0), sb = stage(2, 0);
      for (auto o : L.sd) if (all || o->wave == w) splitd_bwd(o, sa, sb);
      push(hb, std::move(sa));
      push(hb, std::move(sb));
      HStage sod = stage(0, 0);
      for (auto o : L.od) {
        if (!all && o->wave != w) continue;
        if (o->k > 0) { sod.refl.push_back(refl_of(*o, 0)); sod.mmax = std::max(sod.mmax, o->m); }
      }
      push(hb, std::move(sod));
    }
    HStage ssc = stage(1, 0);
    for (auto o : L.sc) ssc.bwd.push_back(scaled_of(*o, 0));
    push(hb, std::move(ssc));
    // Reassign has no backward step on x (operations.cpp:70)
    HStage sqrr = stage(3, 0), sqr = stage(1, 0);
    for (auto o : L.qr)
      if (!o->V64) sqr.bwd.push_back(bwd_of(*o, 0, sqrr));
    push(hb, std::move(sqrr));
    push(hb, std::move(sqr));
    for (auto o : L.qr)
      if (o->V64) qr64_stages(*o, 0, true, hb, xbuf_);
  }
  if (!square_) {
    // ---- normal equations (tree.cpp:1471-1517): tops forward on xt, tops
    // backward (permutations only) on xt, xt -> x for the leaves, ops
    // forward (permutations only) on x, then the backward program above
    for (int lvl = 0; lvl < nlevels_; ++lvl) {
      LevelOps L = level_ops(lvl);
      HStage sqr = stage(1, 1), sqrr = stage(3, 1);  // R^{-T}, then x_n -= A_n^T x_c (toperations.cpp:12-24)
      for (auto o : L.qr)
        if (!o->V64) sqr.bwd.push_back(bwd_of(*o, 1, sqrr));
      push(htf, std::move(sqr));
      push(htf, std::move(sqrr));
      for (auto o : L.qr)  // FP64 top of an FP32 tree: R^{-T} in double on xt
        if (o->V64) qr64_stages(*o, 1, true, htf, xtbuf_);
      HStage ssc = stage(1, 1);
      for (auto o : L.sc) ssc.bwd.push_back(scaled_of(*o, 1));
      push(htf, std::move(ssc));
      int wmax = 0;
      for (auto o : L.sd) wmax = std::max(wmax, o->wave);
      for (int w = 0; w <= (L.fq.empty() ? 0 : wmax); ++w) {
        const bool all = L.fq.empty();
        HStage sod = stage(0, 1);
        for (auto o : L.od) {
          if (!all && o->wave != w) continue;
          if (o->k > 0) { sod.refl.push_back(refl_of(*o, 1)); sod.mmax = std::max(sod.mmax, o->m); }
        }
        push(htf, std::move(sod));
        HStage sa = stage(2, 1), sb = stage(2, 1);
        for (auto o : L.sd) if (all || o->wave == w) splitd_fwd(o, sa, sb);
        push(htf, std::move(sa));
        push(htf, std::move(sb));
        HStage sfq = stage(1, 1), sfr = stage(3, 1);
        for (auto o : L.fq) if (o->wave == w) sfq.bwd.push_back(bwd_of(*o, 1, sfr));
        push(htf, std::move(sfq));
        push(htf, std::move(sfr));
      }
      HStage smg = stage(2, 1);
      for (auto o : L.mg) merge_up(o, smg, false);
      push(htf, std::move(smg));
    }
    for (int lvl = nlevels_ - 1; lvl >= 0; --lvl) {
      LevelOps L = level_ops(lvl);
      HStage smg = stage(2, 1);
      for (auto o : L.mg) merge_down(o, smg, true);
      push(htb, std::move(smg));
      HStage sa = stage(2, 1), sb = stage(2, 1);
      for (auto o : L.sd) splitd_bwd(o, sa, sb);
      push(htb, std::move(sa));
      push(htb, std::move(sb));
      HStage sra = stage(2, 1);
      for (auto o : L.ra) reassign(o, sra, false);
      push(htb, std::move(sra));
    }
    HStage cp = stage(2, 0, 1);
    for (int c : bottoms_[0])
      if (cl_[c].rsize_org > 0) cp.mv.push_back(MoveDesc{cl_[c].xoff, cl_[c].xoff, cl_[c].rsize_org, 0});
    push(htb, std::move(cp));
    for (int lvl = 0; lvl < nlevels_; ++lvl) {
      LevelOps L = level_ops(lvl);
      HStage sra = stage(2, 0);
      for (auto o : L.ra) reassign(o, sra, true);
      push(hpf, std::move(sra));
      HStage sa = stage(2, 0), sb = stage(2, 0);
      for (auto o : L.sd) splitd_fwd(o, sa, sb);
      push(hpf, std::move(sa));
      push(hpf, std::move(sb));
      HStage smg = stage(2, 0);
      for (auto o : L.mg) merge_up(o, smg, true);
      push(hpf, std::move(smg));
    }
  }
  // ---- x <- M^{-T} x (square): M^{-1} = S B F G with F, B the programs
  // above and G, S the entry and exit maps, so M^{-T} = G^T F^T B^T S^T.  A
  // program's transpose runs its stages in reverse, each transposed:
  // reflector and triangular stages flip `trans`, R12 products become
  // x_n -= R12^T x_c, and moves run dst -> src.  Every cluster owns its own
  // x segment, so a move's source is dead after it and its destination dead
  // before it, which makes the reversed move the exact transpose on the
  // values that are read.
  std::vector<HStage> haf, hab;
  if (square_) {
    auto transpose = [](const std::vector<HStage>& src, std::vector<HStage>& dst) {
      for (auto it = src.rbegin(); it != src.rend(); ++it) {
        HStage st = *it;
        for (auto& d : st.refl) d.trans ^= 1;
        for (auto& d : st.bwd) d.trans ^= 1;
        for (auto& d : st.r12c) d.trans ^= 1;
        for (auto& d : st.mv) std::swap(d.src, d.dst);
        for (auto& d : st.refl64) d.trans ^= 1;
        for (auto& d : st.bwd64) d.trans ^= 1;
        for (auto& d : st.cvt) {
          const void* a = d.src;
          d.src = d.dst;
          d.dst = const_cast<void*>(a);
          std::swap(d.lds, d.ldd);
        }
        if (st.type == 6) st.type = 7;
        else if (st.type == 7) st.type = 6;
        std::swap(st.buf, st.sbuf);
        dst.push_back(std::move(st));
      }
    };
    transpose(hb, haf);
    transpose(hf, hab);
  }
  segs_dev_ = stager_.upload(perm_, segs);
  segs64_dev_ = stager_.upload(perm_, segs64);
  r12_dev_ = stager_.upload(perm_, r12);
  // Multi-workgroup split of large ops: row slices of <= 256 rows for
  // reflector applies of >= 384 rows, or >= 4 blocks of 32 for triangular
  // sweeps of >= 512; the stage's multi-workgroup ops come first and use at
  // most one workgroup per CU.
  int nsync = 0;
  auto plan_nwg = [&](std::vector<int>& want) {
    int total = 0;
    for (int v : want) if (v > 1) total += v;
    if (total > num_cus_) {
      const double f = (double)num_cus_ / total;
      for (int& v : want) if (v > 1) v = std::max(2, (int)(v * f));
      total = 0;
      for (int v : want) if (v > 1) total += v;
      for (int& v : want) if (v > 1 && total > num_cus_) { total -= v - 1; v = 1; }
    }
  };
  auto finish = [&](std::vector<HStage>& hs, std::vector<Stage>& out) {
    for (auto& h : hs) {
      Stage s;
      s.type = h.type;
      s.mmax = h.mmax;
      s.buf = h.buf;
      s.sbuf = h.sbuf;
      std::vector<int> cnt, st;
      if (h.type == 0) {
        // register form for stages of few ops (one op's latency sets the
        // pace): every op gets ceil(m / rows) workgroups, all resident
        // stages of many small ops (every op within one workgroup of the
        // smallest register form that holds it) use the register form too
        std::vector<int> want;
        int reg_total = 0, mstage = 1;
        for (auto& d : h.refl) { reg_total += iceil(std::max(d.m, 1), reflr_rows()); mstage = std::max(mstage, d.m); }
        const bool few = (int)h.refl.size() < num_cus_ && reg_total <= num_cus_;
        const bool many = !few && mstage <= reflr_rows();
        const bool reg = reflr_ && (few || many);
        s.nrw = few ? 7 : reflr_row_waves(mstage);
        if (reg) {
          for (auto& d : h.refl) want.push_back(iceil(std::max(d.m, 1), 64 * s.nrw));
        } else {
          for (auto& d : h.refl) want.push_back(d.m >= 384 ? std::min(32, iceil(d.m, 256)) : 1);
          plan_nwg(want);
        }
        s.form = reg ? 1 : 0;
        std::vector<int> order(h.refl.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return (want[a] > 1) > (want[b] > 1); });
        std::vector<ReflDesc<T>> ds;
        int mmax1 = 0;
        for (int i : order) {
          ReflDesc<T> d = h.refl[i];
          s.kmax = std::max(s.kmax, d.k);
          s.mall = std::max(s.mall, d.m);
          d.nwg = want[i];
          s.nseg_max = std::max(s.nseg_max, d.nseg);
          if (d.nwg > 1) {
            d.work = reg ? nullptr : alloc_perm(d.m);
            d.part = alloc_perm((size_t)2 * d.nwg * 32);
            d.sync = (unsigned int*)(intptr_t)(nsync++);  // index, resolved below
          } else {
            mmax1 = std::max(mmax1, d.m);
          }
          ds.push_back(d);
          cnt.push_back(d.nwg);
        }
        s.mmax = mmax1;
        s.nwg_total = prefix_tiles(cnt, st);
        s.nd = (int)ds.size();
        if (!reg && mmax1 > refl_lds_cap(sizeof(T))) s.work = alloc_perm((size_t)s.nwg_total * mmax1);
        h.refl = ds;
      } else if (h.type == 1) {
        // latency form (one workgroup per op) for stages of few ops without
        // in-kernel R12 products
        bool lat = reflr_ && (int)h.bwd.size() < num_cus_;
        for (auto& d : h.bwd) lat = lat && d.nblk == 0 && d.k <= bwdr_max_k();
        s.form = lat ? 1 : 0;
        std::vector<int> want;
        for (auto& d : h.bwd) want.push_back(!lat && d.k >= 512 ? std::min(32, (d.k / 32) / 4) : 1);
        if (!lat) plan_nwg(want);
        std::vector<int> order(h.bwd.size());
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return (want[a] > 1) > (want[b] > 1); });
        std::vector<BwdDesc<T>> ds;
        for (int i : order) {
          BwdDesc<T> d = h.bwd[i];
          s.kmax = std::max(s.kmax, d.k);
          d.nwg = want[i];
          if (d.nwg > 1) d.sync = (unsigned int*)(intptr_t)(nsync++);
          ds.push_back(d);
          cnt.push_back(d.nwg);
        }
        s.nwg_total = prefix_tiles(cnt, st);
        s.nd = (int)ds.size();
        h.bwd = ds;
      } else if (h.type == 3) {
        s.nd = (int)h.r12c.size();
      } else if (h.type == 4 || h.type == 5) {
        // one workgroup per FP64 op
        s.nd = h.type == 4 ? (int)h.refl64.size() : (int)h.bwd64.size();
        if (h.type == 5) {
          bool lat = reflr_;
          for (auto& d : h.bwd64) { lat = lat && d.nblk == 0 && d.k <= bwdr_max_k(); s.kmax = std::max(s.kmax, d.k); }
          s.form = lat ? 1 : 0;
        }
        cnt.assign(s.nd, 1);
        s.nwg_total = prefix_tiles(cnt, st);
        if (h.type == 4 && h.mmax > refl_lds_cap(sizeof(double)))
          s.work64 = (double*)perm_.alloc(sizeof(double) * (size_t)s.nd * h.mmax);
      } else if (h.type == 6 || h.type == 7) {
        s.nd = (int)h.cvt.size();
        for (auto& d : h.cvt) cnt.push_back(iceil(std::max(d.rows, 1), 32) * iceil(std::max(d.cols, 1), 32));
        s.nwg_total = prefix_tiles(cnt, st);
      } else {
        s.nd = (int)h.mv.size();
      }
      if (h.type <= 1 || h.type >= 4) s.wstart = stager_.upload(perm_, st);
      out.push_back(s);
    }
  };
  finish(hf, fwd_);
  finish(hb, bwd_);
  finish(htf, tfwd_);
  finish(htb, tbwd_);
  finish(hpf, pfwd_);
  finish(haf, afwd_);
  finish(hab, abwd_);
  n_solve_sync_ = nsync;
  solve_sync_ = (unsigned int*)perm_.alloc(sizeof(unsigned int) * std::max(nsync, 1));
  solve_err_ = (int*)perm_.alloc(256);
  SPQ_HIP(hipMemsetAsync(solve_err_, 0, 256, stream_));
  auto upload_stage = [&](std::vector<HStage>& hs, std::vector<Stage>& st) {
    for (size_t i = 0; i < hs.size(); ++i) {
      auto& h = hs[i];
      if (h.type == 0) {
        for (auto& d : h.refl) {
          if (d.nwg > 1) d.sync = solve_sync_ + (intptr_t)d.sync;
          d.err = solve_err_;
        }
        st[i].refl = stager_.upload(perm_, h.refl);
      } else if (h.type == 1) {
        for (auto& d : h.bwd) {
          if (d.nwg > 1) d.sync = solve_sync_ + (intptr_t)d.sync;
          d.err = solve_err_;
        }
        st[i].bwd = stager_.upload(perm_, h.bwd);
      } else if (h.type == 3) {
        st[i].r12c = stager_.upload(perm_, h.r12c);
      } else if (h.type == 4) {
        for (auto& d : h.refl64) d.err = solve_err_;
        st[i].refl64 = stager_.upload(perm_, h.refl64);
      } else if (h.type == 5) {
        for (auto& d : h.bwd64) d.err = solve_err_;
        st[i].bwd64 = stager_.upload(perm_, h.bwd64);
      } else if (h.type == 6 || h.type == 7) {
        st[i].cvt = stager_.upload(perm_, h.cvt);
      } else {
        st[i].mv = stager_.upload(perm_, h.mv);
      }
    }
  };
  upload_stage(hf, fwd_);
  upload_stage(hb, bwd_);
  upload_stage(htf, tfwd_);
  upload_stage(htb, tbwd_);
  upload_stage(hpf, pfwd_);
  upload_stage(haf, afwd_);
  upload_stage(hab, abwd_);
  if (verbose_ > 1) {
    auto dump = [&](const char* dir, std::vector<HStage>& hs) {
      for (auto& h : hs) {
        int mm = 0, kk = 0;
        double bytes = 0;
        if (h.type == 0)
          for (auto& d : h.refl) { mm = std::max(mm, d.m); kk = std::max(kk, d.k); bytes += (double)d.m * d.k; }
        if (h.type == 1)
          for (auto& d : h.bwd) { kk = std::max(kk, d.k); bytes += 0.5 * d.k * d.k; }
        fprintf(stderr, "solve %s stage type %d ops %zu max_m %d max_k %d factor_entries %.3g\n", dir, h.type,
                h.type == 0 ? h.refl.size() : (h.type == 1 ? h.bwd.size() : h.mv.size()), mm, kk, bytes);
      }
    };
    dump("fwd", hf);
    dump("bwd", hb);
  }
  // entry / exit maps (set_vector / extract_vector, cluster.cpp:125-156)
  std::vector<int64_t> idst, osrc;
  std::vector<int> isrc, odst;
  for (int c : bottoms_[0]) {
    for (int i = 0; i < cl_[c].rsize_org; ++i) { idst.push_back(cl_[c].xoff + i); isrc.push_back(rperm_[cl_[c].rstart + i]); }
    for (int i = 0; i < cl_[c].csize_org; ++i) { osrc.push_back(cl_[c].xoff + i); odst.push_back(cperm_[cl_[c].cstart + i]); }
  }
  n_in_ = (int)idst.size();
  n_out_ = (int)osrc.size();
  in_dst_ = stager_.upload(perm_, idst);
  in_src_ = stager_.upload(perm_, isrc);
  out_src_ = stager_.upload(perm_, osrc);
  out_dst_ = stager_.upload(perm_, odst);
  if (!square_) {
    // normal equations: b (length cols) enters xt through the column
    // permutation (tree.cpp:1475-1480)
    std::vector<int64_t> tdst;
    std::vector<int> tsrc;
    for (int c : bottoms_[0])
      for (int i = 0; i < cl_[c].csize_org; ++i) { tdst.push_back(cl_[c].xoff + i); tsrc.push_back(cperm_[cl_[c].cstart + i]); }
    n_tin_ = (int)tdst.size();
    tin_dst_ = stager_.upload(perm_, tdst);
    tin_src_ = stager_.upload(perm_, tsrc);
  }
}

template <typename T>
void Engine<T>::run_program(const double* b, double* x, bool prof, int mode) {
  launch_zero(stream_, sizeof(T), xbuf_, xsize_);
  if (n_solve_sync_ > 0) launch_zero(stream_, 4, solve_sync_, n_solve_sync_);
  std::vector<hipEvent_t> ev;
  auto run = [&](const Stage& s) {
    if (prof) {
      ev.emplace_back();
      SPQ_HIP(hipEventCreate(&ev.back()));
      SPQ_HIP(hipEventRecord(ev.back(), stream_));
    }
    T* xb = s.buf ? xtbuf_ : xbuf_;
    if (s.type == 0 && s.form == 1)
      launch_reflr<T>(stream_, s.refl, s.wstart, s.nd, s.nwg_total, segs_dev_, xb, s.nseg_max, s.nrw);
    else if (s.type == 0) launch_refl<T>(stream_, s.refl, s.wstart, s.nd, s.nwg_total, segs_dev_, xb, s.work, s.mmax);
    else if (s.type == 1 && s.form == 1) launch_bwdr<T>(stream_, s.bwd, s.nd, s.kmax, xb);
    else if (s.type == 1) launch_bwd<T>(stream_, s.bwd, s.wstart, s.nd, s.nwg_total, r12_dev_, xb);
    else if (s.type == 3) launch_r12<T>(stream_, s.r12c, s.nd, xb);
    else if (s.type == 4)
      launch_refl<double>(stream_, s.refl64, s.wstart, s.nd, s.nwg_total, segs64_dev_, x64_, s.work64, s.mmax);
    else if (s.type == 5 && s.form == 1) launch_bwdr<double>(stream_, s.bwd64, s.nd, s.kmax, x64_);
    else if (s.type == 5) launch_bwd<double>(stream_, s.bwd64, s.wstart, s.nd, s.nwg_total, nullptr, x64_);
    else if (s.type == 6) launch_copy(stream_, (int)sizeof(T), 8, s.cvt, s.wstart, s.nd, s.nwg_total);
    else if (s.type == 7) launch_copy(stream_, 8, (int)sizeof(T), s.cvt, s.wstart, s.nd, s.nwg_total);
    else launch_move<T>(stream_, s.mv, s.nd, xb, s.sbuf ? xtbuf_ : xbuf_);
  };
  std::vector<const std::vector<Stage>*> progs;
  if (mode == 1) {
    launch_zero(stream_, sizeof(T), xtbuf_, xsize_);
    launch_gather_in<T>(stream_, tin_dst_, tin_src_, n_tin_, b, xtbuf_);
    progs = {&tfwd_, &tbwd_, &pfwd_, &bwd_};
  } else if (mode == 2) {
    // S^T: b enters through the exit map
    launch_gather_in<T>(stream_, out_src_, out_dst_, n_out_, b, xbuf_);
    progs = {&afwd_, &abwd_};
  } else {
    launch_gather_in<T>(stream_, in_dst_, in_src_, n_in_, b, xbuf_);
    progs = {&fwd_, &bwd_};
  }
  for (auto p : progs)
    for (auto& s : *p) run(s);
  if (prof) {
    ev.emplace_back();
    SPQ_HIP(hipEventCreate(&ev.back()));
    SPQ_HIP(hipEventRecord(ev.back(), stream_));
    SPQ_HIP(hipStreamSynchronize(stream_));
    double tot = 0;
    for (size_t i = 0; i + 1 < ev.size(); ++i) {
      float ms = 0;
      SPQ_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      tot += ms;
      size_t j = i, pi = 0;
      while (j >= progs[pi]->size()) j -= progs[pi++]->size();
      const Stage& s = (*progs[pi])[j];
      fprintf(stderr, "stage p%zu %3zu type %d form %d ops %5d wgs %5d mmax %6d mall %6d kmax %5d nrw %d ms %.4f\n", pi, j,
              s.type, s.form, s.nd, s.type == 2 ? s.nd : s.nwg_total, s.mmax, s.mall, s.kmax, s.nrw, ms);
    }
    fprintf(stderr, "stages total ms %.3f\n", tot);
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  if (mode == 2) launch_scatter_out<T>(stream_, in_dst_, in_src_, n_in_, xbuf_, x);  // G^T
  else launch_scatter_out<T>(stream_, out_src_, out_dst_, n_out_, xbuf_, x);
}

// One solve program (mode as in run_program) captured into an executable
// graph that reads bin_ and writes xout_.  The capture only enqueues kernels
// on this tree's stream; relaxed mode keeps other threads (other trees, the
// caller's allocations and synchronisations) free to run meanwhile.
template <typename T>
hipGraphExec_t Engine<T>::capture_program(int mode) {
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  std::lock_guard<std::recursive_mutex> lk(capture_lock());
  SPQ_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
  run_program(bin_, xout_, false, mode);
  SPQ_HIP(hipStreamEndCapture(stream_, &g));
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  SPQ_HIP(e);
  return ex;
}

template <typename T>
void Engine<T>::solve_dev(const double* b, double* x) {
  if (!factored_) throw Error(kErrState, "solve before factorize");
  if (!square_) throw Error(kErrState, "solve: rectangular (least-squares) tree; use solve_nrml");
  // b is taken by value (tree.cpp:1440): copy into the program's fixed input,
  // replay the captured apply, copy the fixed output out.
  SPQ_HIP(hipMemcpyAsync(bin_, b, sizeof(double) * nrows_, hipMemcpyDeviceToDevice, stream_));
  if (stage_prof_ || !solve_graph_) {
    timed("apply", 0.0, solve_bytes_, [&] { run_program(bin_, xout_, stage_prof_); });
    SPQ_HIP(hipMemcpyAsync(x, xout_, sizeof(double) * ncols_, hipMemcpyDeviceToDevice, stream_));
    return;
  }
  if (!graph_exec_fwd_) graph_exec_fwd_ = capture_program(0);
  timed("apply", 0.0, solve_bytes_, [&] { SPQ_HIP(hipGraphLaunch(graph_exec_fwd_, stream_)); });
  SPQ_HIP(hipMemcpyAsync(x, xout_, sizeof(double) * ncols_, hipMemcpyDeviceToDevice, stream_));
}

// x = (W^T W)^{-1} b with the R-only factors of a least-squares tree
// (tree.cpp:1471-1517); b and x have length cols, b may alias x.
template <typename T>
void Engine<T>::solve_nrml_dev(const double* b, double* x) {
  if (!factored_) throw Error(kErrState, "solve_nrml before factorize");
  if (square_) throw Error(kErrState, "solve_nrml: the tree was factored as a square system (set_square(0) first)");
  SPQ_HIP(hipMemcpyAsync(bin_, b, sizeof(double) * ncols_, hipMemcpyDeviceToDevice, stream_));
  if (stage_prof_ || !solve_graph_) {
    timed("apply", 0.0, solve_bytes_, [&] { run_program(bin_, xout_, stage_prof_, 1); });
  } else {
    if (!graph_exec_nrml_) graph_exec_nrml_ = capture_program(1);
    timed("apply", 0.0, solve_bytes_, [&] { SPQ_HIP(hipGraphLaunch(graph_exec_nrml_, stream_)); });
  }
  SPQ_HIP(hipMemcpyAsync(x, xout_, sizeof(double) * ncols_, hipMemcpyDeviceToDevice, stream_));
}

// x = M^{-T} b (square trees; the preconditioner of the adjoint solve with
// A^T); b may alias x.
template <typename T>
void Engine<T>::solve_t_dev(const double* b, double* x) {
  if (!factored_) throw Error(kErrState, "solve_transpose before factorize");
  if (!square_) throw Error(kErrState, "solve_transpose: rectangular (least-squares) tree");
  SPQ_HIP(hipMemcpyAsync(bin_, b, sizeof(double) * ncols_, hipMemcpyDeviceToDevice, stream_));
  if (stage_prof_ || !solve_graph_) {
    timed("apply", 0.0, solve_bytes_, [&] { run_program(bin_, xout_, stage_prof_, 2); });
  } else {
    if (!graph_exec_t_) graph_exec_t_ = capture_program(2);
    timed("apply", 0.0, solve_bytes_, [&] { SPQ_HIP(hipGraphLaunch(graph_exec_t_, stream_)); });
  }
  SPQ_HIP(hipMemcpyAsync(x, xout_, sizeof(double) * nrows_, hipMemcpyDeviceToDevice, stream_));
}

EngineBase* make_engine(int nlevels, int skip, int precision_bytes, int device) {
  host_prof_start();
  if (precision_bytes == 4) return new Engine<float>(nlevels, skip, device);
  return new Engine<double>(nlevels, skip, device);
}

}  // namespace spq
