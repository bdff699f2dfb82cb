// Factorization profile as JSON: the quantities the reference collects in
// its Profile (profile.h:13-106, filled at tree.cpp:427-430, 1052, 1176,
// 1332, 1365, 1387, 1399-1400) in the batched form this engine runs:
// per-level phase times instead of per-separator ones (a level's clusters
// share launches), and the kernel families' device time when kernel timing
// ran.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "engine.hpp"

namespace spq {

namespace {
void add(std::string& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void add(std::string& s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  const int n = vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (n > 0) s.append(buf, (size_t)std::min(n, (int)sizeof buf - 1));
}
}  // namespace

std::string EngineBase::profile_json() {
  std::string s;
  add(s, "{\"nlevels\": %d, \"skip\": %d, \"rows\": %d, \"cols\": %d, \"precision_bytes\": %d,\n", levels(), skip(),
      rows(), cols(), precision_bytes());
  s += " \"levels\": [";
  const auto& li = level_info();
  for (size_t l = 0; l < li.size(); ++l) {
    const LevelInfo& v = li[l];
    add(s,
        "%s\n  {\"ilvl\": %zu, \"t_fill\": %.6g, \"t_elmn\": %.6g, \"t_shift\": %.6g, \"t_scale\": %.6g, "
        "\"t_sprsfy\": %.6g, \"t_merge\": %.6g, \"top_rows\": %d, \"top_cols\": %d, \"eliminated\": %d, "
        "\"elim_waves\": %d, \"sparsified\": %d, \"sprs_waves\": %d}",
        l ? "," : "", l, v.t_fill, v.t_elmn, v.t_shift, v.t_scale, v.t_sprsfy, v.t_merge, v.top_rows, v.top_cols,
        v.n_elim, v.elim_waves, v.n_sprs, v.sprs_waves);
  }
  s += "],\n \"interfaces\": [";
  const auto& sp = sparsified();
  for (size_t i = 0; i < sp.size(); ++i)
    add(s, "%s\n  {\"ilvl\": %d, \"lvl\": %d, \"sep\": %d, \"cols_before\": %d, \"cols_after\": %d, "
           "\"neighbors\": %d, \"wave\": %d}",
        i ? "," : "", sp[i].ilvl, sp[i].lvl, sp[i].sep, sp[i].before, sp[i].after, sp[i].nbrs, sp[i].wave);
  s += "],\n \"extra_rows\": [";
  const auto& ex = extra_sparsified();
  for (size_t i = 0; i < ex.size(); ++i)
    add(s, "%s\n  {\"ilvl\": %d, \"lvl\": %d, \"sep\": %d, \"rows_before\": %d, \"rows_after\": %d}", i ? "," : "",
        ex[i].ilvl, ex[i].lvl, ex[i].sep, ex[i].before, ex[i].after);
  s += "],\n \"merged\": [";
  const auto& mg = merged();
  for (size_t i = 0; i < mg.size(); ++i)
    add(s, "%s\n  {\"ilvl\": %d, \"lvl\": %d, \"sep\": %d, \"rows\": %d, \"cols\": %d, \"aspect_ratio\": %.6g}",
        i ? "," : "", mg[i].ilvl, mg[i].lvl, mg[i].sep, mg[i].rows, mg[i].cols,
        mg[i].cols > 0 ? (double)mg[i].rows / mg[i].cols : 0.0);
  s += "],\n \"kernels\": [";
  const auto ks = kernel_stats();
  for (size_t i = 0; i < ks.size(); ++i)
    add(s, "%s\n  {\"family\": \"%s\", \"launches\": %d, \"ms\": %.6g, \"gflop\": %.6g, \"gbytes\": %.6g}",
        i ? "," : "", ks[i].name.c_str(), ks[i].launches, ks[i].ms, ks[i].flops * 1e-9, ks[i].bytes * 1e-9);
  s += "]}\n";
  return s;
}

}  // namespace spq
