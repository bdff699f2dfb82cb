// Launch-form autotuner keyed on the cluster-size histogram of a batch.
//
// A batched launch (one level's RRQR wave, ...) is summarised by a key that
// buckets its problem sizes (log2 of the largest k and N, log2 of the
// problem count) per kernel family and precision.  For a key seen for the
// first time in tuning mode (SPAQR_TUNE=1) the engine times every candidate
// form on a copy of the batch's inputs and keeps the fastest; the choices and
// the measured times of all candidates are kept in a JSON file, which later
// runs read (SPAQR_TUNE_FILE, default tune/<gfx arch>.json beside the
// library's package).  Keys missing from the file fall back to the engine's
// closed-form rule.
#pragma once
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace spq {

class Tuner {
 public:
  // Loads the cache for this device (arch name, e.g. "gfx950").
  void init(const std::string& arch);
  bool tuning() const { return tuning_; }
  // Cached choice for key, or -1.
  int lookup(const std::string& key);
  // Records a tuned choice with the candidates' measured times (ms) and
  // rewrites the file.
  void record(const std::string& key, int choice, const std::vector<std::pair<int, double>>& ms);
  const std::string& path() const { return path_; }

 private:
  struct Entry {
    int choice = -1;
    std::vector<std::pair<int, double>> ms;
  };
  void save();
  std::string path_, arch_;
  bool tuning_ = false, loaded_ = false;
  std::map<std::string, Entry> entries_;
  std::mutex mu_;
};

// The process-wide cache (every tree of the process shares it).
Tuner& global_tuner();

// Key helper: ceil(log2(max(v, 1))).
int log2_bucket(long v);

}  // namespace spq
