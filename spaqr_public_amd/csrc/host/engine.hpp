// spaQR engine: the reference's Tree (tree.h:36-152) with the symbolic work
// (partition, cluster tree, fill, sizes, op bookkeeping) on the host and every
// dense numeric phase of Tree::factorize / Tree::solve as batched gfx950
// kernels over all clusters of a level.
#pragma once
#include <array>
#include <memory>
#include <string>
#include <vector>

#include "../kernels/descs.hpp"
#include "common.hpp"
#include "partition.hpp"

namespace spq {

struct LevelInfo {
  double t_fill = 0, t_elmn = 0, t_shift = 0, t_scale = 0, t_sprsfy = 0, t_merge = 0;
  int top_rows = 0, top_cols = 0;
  int n_elim = 0, n_scale = 0, n_sprs = 0, elim_waves = 0, sprs_waves = 0;
};

struct SparsifyInfo {
  int ilvl, lvl, sep, before, after;
  int nbrs = 0;  // columns of the interface's RRQR matrix (its neighbours' columns, tree.cpp:1052)
  int wave = 0;  // dependency wave within the level
};
// A cluster created by merge_all (tree.cpp:414-434): sizes after the merge.
struct MergeInfo {
  int ilvl, lvl, sep, rows, cols;
};

// Per kernel family: launches, device time (HIP events on the engine stream),
// algorithmic flops and bytes of those launches.
struct KernelStat {
  std::string name;
  int launches = 0;
  double ms = 0, flops = 0, bytes = 0;
};

// Interface the C ABI needs, independent of the factor precision.
class EngineBase {
 public:
  virtual ~EngineBase();
  // Work vectors of the Krylov drivers, kept with the tree across calls:
  // slot `slot` holds at least `elems` doubles (device, or pinned host for
  // work_host).  A slot grows only when asked for more than it holds; the
  // caller synchronises its stream before returning, so a slot is never
  // reallocated under work in flight.
  double* work_dev(int slot, size_t elems);
  double* work_host(int slot, size_t elems);
  // setters (tree.cpp:14-20)
  virtual void set_tol(double t) = 0;
  virtual void set_skip(int s) = 0;
  virtual void set_scale(int s) = 0;
  virtual void set_square(int s) = 0;
  virtual void set_order(float o) = 0;
  virtual void set_coords(const std::vector<double>& X, int dim) = 0;
  virtual void set_row_parts(const std::vector<int>& parts) = 0;
  virtual void set_verbose(int v) = 0;
  virtual void set_timing(int t) = 0;
  virtual int rows() const = 0;
  virtual int cols() const = 0;
  virtual int levels() const = 0;
  virtual void partition(const CscMat& A) = 0;
  virtual void assemble(const CscMat& A) = 0;
  virtual int factorize() = 0;
  // x = M^{-1} b with device FP64 vectors (tree.cpp:1440-1466)
  virtual void solve_dev(const double* b, double* x) = 0;
  virtual void solve_nrml_dev(const double* b, double* x) = 0;
  // x = M^{-T} b (square trees): the transposed ops, backward program first
  virtual void solve_t_dev(const double* b, double* x) = 0;
  virtual hipStream_t stream() const = 0;
  // inspection
  virtual const std::vector<int>& cperm() const = 0;
  virtual const std::vector<int>& rperm() const = 0;
  virtual const std::vector<LevelInfo>& level_info() const = 0;
  virtual const std::vector<SparsifyInfo>& sparsified() const = 0;
  virtual const std::vector<SparsifyInfo>& extra_sparsified() const = 0;
  virtual const std::vector<MergeInfo>& merged() const = 0;
  virtual int skip() const = 0;
  // The factorization's profile as JSON (profile.h:13-106): per level the
  // phase times and top separator, per sparsified interface its columns
  // before / after, RRQR columns (neighbours) and wave, per merged cluster
  // its size and aspect ratio, the least-squares extra-row compressions and,
  // when kernel timing ran, the kernel families' device time.
  std::string profile_json();
  virtual std::vector<std::array<int, 12>> cluster_table() const = 0;
  virtual void counters(unsigned long long* nnz4, double* dbl8) const = 0;
  virtual int precision_bytes() const = 0;
  // kernel-family timing with HIP events (off by default)
  virtual void set_kernel_timing(int on) = 0;
  // throws if a multi-workgroup solve op reported a barrier timeout
  virtual void check_solve() = 0;
  virtual std::vector<KernelStat> kernel_stats() = 0;
  virtual void reset_kernel_stats() = 0;

 private:
  struct Work { double* p = nullptr; size_t cap = 0; };
  std::vector<Work> work_dev_, work_host_;
};

EngineBase* make_engine(int nlevels, int skip, int precision_bytes, int device);

// SPAQR_HOST_PROF=1: sample the calling thread's host time (diagnostics;
// the report goes to stderr at process exit).  Called when a tree is made.
void host_prof_start();

}  // namespace spq
