#include "parallel.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace spq {

int host_threads() {
  static const int n = [] {
    int t = 8;
    if (const char* e = std::getenv("SPAQR_HOST_THREADS")) t = std::atoi(e);
    const int hw = (int)std::thread::hardware_concurrency();
    if (hw > 0) t = std::min(t, hw);
    return std::max(1, std::min(t, 64));
  }();
  return n;
}

int plan_chunks(int n, int grain) {
  if (n <= 0) return 0;
  const int by_grain = std::max(1, n / std::max(grain, 1));
  return std::min(by_grain, kMaxChunks);
}

namespace {

// One job at a time.  The workers sleep on `cv_` until a job is posted and
// copy its parameters under the lock; chunks are claimed from `ticket_`,
// whose high half is the job's generation, so a worker that wakes late can
// never claim a chunk of a later job with the parameters of an earlier one.
// The last chunk to finish wakes the caller.  The pool is created on first
// use and never destroyed (its threads sleep through process exit).
class Pool {
 public:
  explicit Pool(int workers) {
    for (int i = 0; i < workers; ++i) th_.emplace_back([this] { work(); });
  }
  std::mutex run_mu;  // held by the thread whose job the pool is running

  void run(int nchunks, int n, const std::function<void(int, int, int)>& f) {
    Job j;
    {
      std::lock_guard<std::mutex> lk(mu_);
      j = Job{&f, n, nchunks, ++gen_};
      job_ = j;
      left_ = nchunks;
      err_ = nullptr;
      ticket_.store((j.gen & 0xffffffffull) << 32);
    }
    cv_.notify_all();
    take_chunks(j);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return left_ == 0; });
    if (err_) std::rethrow_exception(err_);
  }

 private:
  struct Job {
    const std::function<void(int, int, int)>* f = nullptr;
    int n = 0, nchunks = 0;
    unsigned long long gen = 0;
  };
  void take_chunks(const Job& j) {
    for (;;) {
      unsigned long long t = ticket_.load();
      int c;
      do {
        if ((t >> 32) != (j.gen & 0xffffffffull)) return;  // that job is over
        c = (int)(t & 0xffffffffull);
        if (c >= j.nchunks) return;
      } while (!ticket_.compare_exchange_weak(t, t + 1));
      const int b = (int)((long long)c * j.n / j.nchunks), e = (int)((long long)(c + 1) * j.n / j.nchunks);
      try {
        (*j.f)(c, b, e);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!err_) err_ = std::current_exception();
      }
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_cv_.notify_all();
    }
  }
  void work() {
    unsigned long long seen = 0;
    for (;;) {
      Job j;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        j = job_;
      }
      take_chunks(j);
    }
  }

  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  Job job_;
  int left_ = 0;
  std::atomic<unsigned long long> ticket_{0};
  unsigned long long gen_ = 0;
  std::exception_ptr err_;
};

Pool* pool() {
  static Pool* p = host_threads() > 1 ? new Pool(host_threads() - 1) : nullptr;
  return p;
}

}  // namespace

void parallel_for(int n, int grain, const std::function<void(int, int, int)>& f) {
  const int k = plan_chunks(n, grain);
  if (k <= 0) return;
  Pool* p = k > 1 ? pool() : nullptr;
  if (p) {
    std::unique_lock<std::mutex> lk(p->run_mu, std::try_to_lock);
    if (lk.owns_lock()) {
      p->run(k, n, f);
      return;
    }
  }
  for (int c = 0; c < k; ++c) f(c, (int)((long long)c * n / k), (int)((long long)(c + 1) * n / k));
}

}  // namespace spq
