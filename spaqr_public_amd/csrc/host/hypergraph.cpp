// Non-geometric ordering: a row partition of the column-net hypergraph of A
// and the cluster labels it induces.
//
// The reference partitions the rows of A with PaToH (partition.cpp:422-448:
// k = 2^(nlevels-1) parts, unit weights, connectivity metric, seed 42), a
// binary-only library; row_parts() is this library's own partitioner in its
// place (recursive bisection with Fiduccia-Mattheyses refinement of the cut
// nets), and a caller may bring any other row partition instead.  The
// labelling from a row partition follows HypergraphPartition
// (partition.cpp:450-499): a column whose rows lie in different parts at
// level l is a separator of level l, the others are leaves of their rows'
// part; and getInterfacesUsingA_HUND (partition.cpp:582-623): the left /
// right interface of a separator column is the highest common ancestor of
// the lower-level columns met through its rows (distance 2 in A), sided by
// the part of the row they are met through.
#include <algorithm>
#include <cstdint>
#include <set>
#include <stdexcept>
#include <utility>
#include <vector>

#include "partition.hpp"

namespace spq {

namespace {

SepID highest_common(SepID a, SepID b) {  // partition.cpp:63-86
  while (a.lvl < b.lvl) { a.lvl++; a.sep /= 2; }
  while (b.lvl < a.lvl) { b.lvl++; b.sep /= 2; }
  while (a.sep != b.sep) { a.lvl++; a.sep /= 2; b.lvl++; b.sep /= 2; }
  return a;
}

// Bisection of one part: rows `rows` (global ids) into side 0 / 1.
class Bisector {
 public:
  Bisector(const CscMat& A, const CscMat& R) : A_(A), R_(R), loc_(A.m, -1) {}

  // side[t] for rows[t]; side 0 gets ceil(r / 2) rows before refinement
  void run(const std::vector<int>& rows, std::vector<char>& side) {
    const int r = (int)rows.size();
    side.assign(r, 0);
    if (r < 2) return;
    for (int t = 0; t < r; ++t) loc_[rows[t]] = t;
    build_nets(rows);
    // initial split: breadth-first order from a pseudo-peripheral row (the
    // last row reached from the first one), first half to side 0
    std::vector<int> order;
    bfs(0, order);
    bfs(order.back(), order);
    for (int q = 0; q < (r + 1) / 2; ++q) side[order[q]] = 0;
    for (int q = (r + 1) / 2; q < r; ++q) side[order[q]] = 1;
    refine(side);
    for (int t = 0; t < r; ++t) loc_[rows[t]] = -1;
  }

 private:
  const CscMat& A_;
  const CscMat& R_;
  std::vector<int> loc_;                 // global row -> local index (-1 outside)
  std::vector<int> np_, pins_;           // nets (columns with >= 2 local pins): CSR of local rows
  std::vector<int> vp_, vnets_;          // local row -> its nets
  static constexpr int kBigNet = 1024;   // nets this large stay out of refinement

  void build_nets(const std::vector<int>& rows) {
    const int r = (int)rows.size();
    std::vector<int> cols;
    std::vector<char> seen(A_.n, 0);
    for (int g : rows)
      for (int t = R_.p[g]; t < R_.p[g + 1]; ++t)
        if (!seen[R_.i[t]]) { seen[R_.i[t]] = 1; cols.push_back(R_.i[t]); }
    std::sort(cols.begin(), cols.end());
    np_.assign(1, 0);
    pins_.clear();
    std::vector<int> deg(r, 0);
    for (int c : cols) {
      seen[c] = 0;
      int cnt = 0;
      for (int q = A_.p[c]; q < A_.p[c + 1]; ++q) cnt += loc_[A_.i[q]] >= 0;
      if (cnt < 2) continue;
      for (int q = A_.p[c]; q < A_.p[c + 1]; ++q)
        if (loc_[A_.i[q]] >= 0) { pins_.push_back(loc_[A_.i[q]]); deg[loc_[A_.i[q]]]++; }
      np_.push_back((int)pins_.size());
    }
    vp_.assign(r + 1, 0);
    for (int v = 0; v < r; ++v) vp_[v + 1] = vp_[v] + deg[v];
    vnets_.assign(vp_[r], 0);
    std::vector<int> fill(vp_.begin(), vp_.end() - 1);
    for (int e = 0; e + 1 < (int)np_.size(); ++e)
      for (int q = np_[e]; q < np_[e + 1]; ++q) vnets_[fill[pins_[q]]++] = e;
  }

  // breadth-first order of all local rows from `s` (unreached components
  // continue from their smallest row)
  void bfs(int s, std::vector<int>& order) {
    const int r = (int)vp_.size() - 1;
    std::vector<char> vis(r, 0);
    std::vector<char> nvis(np_.size(), 0);
    order.clear();
    int next = 0;
    while ((int)order.size() < r) {
      if (vis[s]) {
        while (vis[next]) ++next;
        s = next;
      }
      size_t head = order.size();
      vis[s] = 1;
      order.push_back(s);
      while (head < order.size()) {
        const int v = order[head++];
        for (int q = vp_[v]; q < vp_[v + 1]; ++q) {
          const int e = vnets_[q];
          if (nvis[e]) continue;
          nvis[e] = 1;
          for (int t = np_[e]; t < np_[e + 1]; ++t)
            if (!vis[pins_[t]]) { vis[pins_[t]] = 1; order.push_back(pins_[t]); }
        }
      }
    }
  }

  // Fiduccia-Mattheyses passes on the number of cut nets, side sizes kept
  // within 3 % of half; a pass moves each row at most once, keeps the best
  // prefix of its moves and stops after 256 moves without a new best.
  void refine(std::vector<char>& side) {
    const int r = (int)side.size(), ne = (int)np_.size() - 1;
    const int slack = std::max(1, r * 3 / 100);
    const int lo = r / 2 - slack, hi = (r + 1) / 2 + slack;  // bounds on side-0 size
    std::vector<int> cnt(2 * ne, 0);
    for (int e = 0; e < ne; ++e)
      for (int q = np_[e]; q < np_[e + 1]; ++q) cnt[2 * e + side[pins_[q]]]++;
    auto big = [&](int e) { return np_[e + 1] - np_[e] > kBigNet; };
    int n0 = 0;
    for (int v = 0; v < r; ++v) n0 += side[v] == 0;
    std::vector<int> gain(r);
    std::vector<char> locked(r);
    for (int pass = 0; pass < 8; ++pass) {
      std::set<std::pair<int, int>> bucket[2];  // (-gain, row)
      for (int v = 0; v < r; ++v) {
        int g = 0;
        const int s = side[v];
        for (int q = vp_[v]; q < vp_[v + 1]; ++q) {
          const int e = vnets_[q];
          if (big(e)) continue;
          if (cnt[2 * e + s] == 1 && cnt[2 * e + 1 - s] > 0) ++g;
          if (cnt[2 * e + 1 - s] == 0) --g;
        }
        gain[v] = g;
        locked[v] = 0;
        bucket[s].insert({-g, v});
      }
      auto bump = [&](int u, int d) {
        if (locked[u] || d == 0) return;
        auto& b = bucket[(int)side[u]];
        b.erase({-gain[u], u});
        gain[u] += d;
        b.insert({-gain[u], u});
      };
      std::vector<int> moves;
      int cur = 0, best = 0, best_len = 0, since = 0;
      while (since < 256) {
        // the best feasible move: from side 0 needs n0 > lo, from side 1 n0 < hi
        int pick = -1;
        for (int s = 0; s < 2; ++s) {
          if (bucket[s].empty()) continue;
          if (s == 0 ? n0 - 1 < lo : n0 + 1 > hi) continue;
          const auto& top = *bucket[s].begin();
          if (pick < 0 || top.first < -gain[pick] || (top.first == -gain[pick] && top.second < pick)) pick = top.second;
        }
        if (pick < 0) break;
        const int v = pick, from = side[v], to = 1 - from;
        bucket[from].erase({-gain[v], v});
        locked[v] = 1;
        cur -= gain[v];
        for (int q = vp_[v]; q < vp_[v + 1]; ++q) {
          const int e = vnets_[q];
          if (big(e)) {
            cnt[2 * e + from]--; cnt[2 * e + to]++;
            continue;
          }
          // standard FM gain updates for the cut-net metric
          if (cnt[2 * e + to] == 0) {
            for (int t = np_[e]; t < np_[e + 1]; ++t) bump(pins_[t], 1);
          } else if (cnt[2 * e + to] == 1) {
            for (int t = np_[e]; t < np_[e + 1]; ++t)
              if (side[pins_[t]] == to) bump(pins_[t], -1);
          }
          cnt[2 * e + from]--; cnt[2 * e + to]++;
          if (cnt[2 * e + from] == 0) {
            for (int t = np_[e]; t < np_[e + 1]; ++t) bump(pins_[t], -1);
          } else if (cnt[2 * e + from] == 1) {
            for (int t = np_[e]; t < np_[e + 1]; ++t)
              if (side[pins_[t]] == from && pins_[t] != v) bump(pins_[t], 1);
          }
        }
        side[v] = (char)to;
        n0 += from == 0 ? -1 : 1;
        moves.push_back(v);
        if (cur < best) { best = cur; best_len = (int)moves.size(); since = 0; }
        else ++since;
      }
      // undo the moves after the best prefix
      for (int q = (int)moves.size() - 1; q >= best_len; --q) {
        const int v = moves[q], from = side[v], to = 1 - from;
        for (int t = vp_[v]; t < vp_[v + 1]; ++t) { cnt[2 * vnets_[t] + from]--; cnt[2 * vnets_[t] + to]++; }
        side[v] = (char)to;
        n0 += from == 0 ? -1 : 1;
      }
      if (best == 0) break;
    }
  }
};

}  // namespace

std::vector<int> row_parts(const CscMat& A, int nlevels) {
  if (nlevels < 1 || nlevels > 30) throw Error(kErrArg, "row_parts: nlevels out of range");
  const CscMat R = csc_transpose(A);
  std::vector<int> part(A.m, 0);
  Bisector bis(A, R);
  std::vector<char> side;
  for (int depth = 0; depth < nlevels - 1; ++depth) {
    const int nparts = 1 << depth;
    std::vector<std::vector<int>> groups(nparts);
    for (int v = 0; v < A.m; ++v) groups[part[v]].push_back(v);
    for (int g = 0; g < nparts; ++g) {
      bis.run(groups[g], side);
      for (size_t t = 0; t < groups[g].size(); ++t) part[groups[g][t]] = 2 * g + side[t];
    }
  }
  return part;
}

std::vector<CID> hypergraph_labels(const CscMat& A, int nlevels, const std::vector<int>& parts) {
  const int n = A.n;
  if ((int)parts.size() != A.m) throw Error(kErrArg, "partition: the row partition has the wrong length");
  const int64_t np = int64_t(1) << (nlevels - 1);
  for (int p : parts)
    if (p < 0 || p >= np) throw Error(kErrArg, "partition: a row part lies outside [0, 2^(nlevels-1))");
  for (int c = 0; c < n; ++c)
    if (A.p[c] == A.p[c + 1]) throw Error(kErrArg, "partition: an empty column has no part (non-geometric partition)");
  // separators, top level first (partition.cpp:458-483): a column whose rows
  // fall in different parts at level l
  std::vector<CID> lab(n);
  for (int depth = 0; depth < nlevels - 1; ++depth) {
    const int l = nlevels - depth - 1;
    for (int c = 0; c < n; ++c) {
      if (lab[c].self.lvl != -1) continue;
      const int pi = parts[A.i[A.p[c]]] >> (l - 1);
      for (int q = A.p[c] + 1; q < A.p[c + 1]; ++q)
        if ((parts[A.i[q]] >> (l - 1)) != pi) { lab[c].self = SepID(l, pi / 2); break; }
    }
  }
  // leaves: the part of the first row (partition.cpp:486-495)
  for (int c = 0; c < n; ++c)
    if (lab[c].self.lvl == -1) {
      lab[c].self = SepID(0, parts[A.i[A.p[c]]]);
      lab[c].l = lab[c].self;
      lab[c].r = lab[c].self;
    }
  // interfaces (partition.cpp:589-621)
  const CscMat R = csc_transpose(A);
  for (int c = 0; c < n; ++c) {
    const SepID self = lab[c].self;
    if (self.lvl <= 0) continue;
    SepID left, right;
    for (int q = A.p[c]; q < A.p[c + 1]; ++q) {
      const int row = A.i[q];
      const bool even = ((parts[row] >> (self.lvl - 1)) & 1) == 0;
      for (int t = R.p[row]; t < R.p[row + 1]; ++t) {
        const SepID nb = lab[R.i[t]].self;
        if (nb.lvl >= self.lvl) continue;
        SepID& side = even ? left : right;
        side = side.lvl == -1 ? nb : highest_common(nb, side);
      }
    }
    lab[c].l = left;
    lab[c].r = right;
  }
  for (auto& c : lab) {
    if (c.l == SepID()) c.l = c.self;
    if (c.r == SepID()) c.r = c.self;
  }
  return lab;
}

}  // namespace spq
