#include "tuner.hpp"

#include <dlfcn.h>
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace spq {

int log2_bucket(long v) {
  int b = 0;
  while ((1L << b) < v) ++b;
  return b;
}

namespace {
// Directory of the package holding libspaqr.so (its _lib directory's parent).
std::string package_dir() {
  Dl_info info{};
  if (!dladdr((void*)&log2_bucket, &info) || !info.dli_fname) return ".";
  std::string p = info.dli_fname;
  for (int up = 0; up < 2; ++up) {
    const size_t s = p.rfind('/');
    if (s == std::string::npos) return ".";
    p = p.substr(0, s);
  }
  return p;
}
}  // namespace

void Tuner::init(const std::string& arch) {
  std::lock_guard<std::mutex> lk(mu_);
  if (loaded_) return;
  loaded_ = true;
  arch_ = arch;
  if (const char* e = std::getenv("SPAQR_TUNE")) tuning_ = std::atoi(e) != 0;
  if (const char* e = std::getenv("SPAQR_TUNE_FILE")) path_ = e;
  else path_ = package_dir() + "/tune/" + arch + ".json";
  std::ifstream f(path_);
  if (!f) return;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  // the file this class writes: one entry per line,
  //   {"key": "<key>", "choice": <int>, "ms": {"<cand>": <ms>, ...}}
  size_t pos = 0;
  while ((pos = s.find("{\"key\": \"", pos)) != std::string::npos) {
    pos += 9;
    const size_t q = s.find('"', pos);
    if (q == std::string::npos) break;
    const std::string key = s.substr(pos, q - pos);
    const size_t c = s.find("\"choice\": ", q);
    if (c == std::string::npos) break;
    Entry e;
    e.choice = std::atoi(s.c_str() + c + 10);
    size_t m = s.find("\"ms\": {", c);
    const size_t end = s.find('}', m == std::string::npos ? c : m + 7);
    if (m != std::string::npos && end != std::string::npos) {
      m += 7;
      while (true) {
        const size_t a = s.find('"', m);
        if (a == std::string::npos || a > end) break;
        const int cand = std::atoi(s.c_str() + a + 1);
        const size_t colon = s.find(':', a);
        if (colon == std::string::npos || colon > end) break;
        e.ms.push_back({cand, std::atof(s.c_str() + colon + 1)});
        m = colon + 1;
      }
    }
    entries_[key] = e;
    pos = q;
  }
}

int Tuner::lookup(const std::string& key) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = entries_.find(key);
  return it == entries_.end() ? -1 : it->second.choice;
}

void Tuner::record(const std::string& key, int choice, const std::vector<std::pair<int, double>>& ms) {
  std::lock_guard<std::mutex> lk(mu_);
  Entry& e = entries_[key];
  e.choice = choice;
  e.ms = ms;
  save();
}

Tuner& global_tuner() {
  static Tuner t;
  return t;
}

void Tuner::save() {
  const size_t slash = path_.rfind('/');
  if (slash != std::string::npos) (void)mkdir(path_.substr(0, slash).c_str(), 0755);
  std::ofstream f(path_);
  if (!f) {
    fprintf(stderr, "spaqr tuner: cannot write %s\n", path_.c_str());
    return;
  }
  f << "{\"arch\": \"" << arch_ << "\", \"entries\": [\n";
  size_t i = 0;
  for (auto& kv : entries_) {
    f << " {\"key\": \"" << kv.first << "\", \"choice\": " << kv.second.choice << ", \"ms\": {";
    for (size_t c = 0; c < kv.second.ms.size(); ++c) {
      char buf[64];
      snprintf(buf, sizeof buf, "%s\"%d\": %.4f", c ? ", " : "", kv.second.ms[c].first, kv.second.ms[c].second);
      f << buf;
    }
    f << "}}" << (++i < entries_.size() ? "," : "") << "\n";
  }
  f << "]}\n";
}

}  // namespace spq
