// Host-time sampler for the library's planning code (diagnostics, off by
// default).  With SPAQR_HOST_PROF=1 the thread that creates the first tree is
// sampled on its own CPU-time clock every 100 us; each sample keeps a short
// call stack.  At process exit the samples are attributed to the innermost
// frame inside libspaqr.so (and to the leaf frame's library) and the top
// entries are printed on stderr.
#include <cxxabi.h>
#include <dlfcn.h>
#include <execinfo.h>
#include <pthread.h>
#include <signal.h>
#include <sys/syscall.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace spq {

namespace {
constexpr int kDepth = 12;
constexpr int kMaxSamples = 1 << 18;
struct Sample {
  void* pc[kDepth];
  int n;
};
Sample* g_samples = nullptr;
std::atomic<int> g_count{0};
timer_t g_timer;
bool g_on = false;

void on_prof(int, siginfo_t*, void*) {
  const int i = g_count.fetch_add(1, std::memory_order_relaxed);
  if (i >= kMaxSamples) return;
  g_samples[i].n = backtrace(g_samples[i].pc, kDepth);
}

std::string name_of(void* pc, std::string* lib) {
  Dl_info info{};
  if (!dladdr(pc, &info)) { *lib = "?"; return "?"; }
  *lib = info.dli_fname ? info.dli_fname : "?";
  const size_t s = lib->rfind('/');
  if (s != std::string::npos) *lib = lib->substr(s + 1);
  if (!info.dli_sname) return "?";
  int st = 0;
  char* dm = abi::__cxa_demangle(info.dli_sname, nullptr, nullptr, &st);
  std::string r = st == 0 && dm ? dm : info.dli_sname;
  free(dm);
  if (r.size() > 110) r = r.substr(0, 110);
  return r;
}

void report() {
  if (!g_on) return;
  timer_delete(g_timer);
  const int n = std::min(g_count.load(), kMaxSamples);
  std::map<std::string, int> inner, leaf;
  for (int i = 0; i < n; ++i) {
    const Sample& s = g_samples[i];
    std::string lib, in = "(outside libspaqr)";
    // frames 0-1 are the handler and the signal trampoline
    for (int f = 2; f < s.n; ++f) {
      std::string l;
      const std::string nm = name_of(s.pc[f], &l);
      if (f == 2) leaf[l + ": " + nm]++;
      if (l.find("libspaqr") != std::string::npos) { in = nm; break; }
    }
    inner[in]++;
  }
  auto top = [&](const std::map<std::string, int>& m, const char* title) {
    std::vector<std::pair<int, std::string>> v;
    for (auto& kv : m) v.push_back({kv.second, kv.first});
    std::sort(v.rbegin(), v.rend());
    fprintf(stderr, "host profile (%d samples of 100 us CPU time): %s\n", n, title);
    for (size_t i = 0; i < v.size() && i < 30; ++i)
      fprintf(stderr, "  %6.1f%%  %s\n", 100.0 * v[i].first / std::max(n, 1), v[i].second.c_str());
  };
  top(inner, "innermost libspaqr frame");
  top(leaf, "leaf frame");
}
}  // namespace

// Starts the sampler on the calling thread (once per process).
void host_prof_start() {
  static bool tried = false;
  if (tried) return;
  tried = true;
  const char* e = std::getenv("SPAQR_HOST_PROF");
  if (!e || std::atoi(e) == 0) return;
  g_samples = (Sample*)calloc(kMaxSamples, sizeof(Sample));
  void* warm[4];
  backtrace(warm, 4);  // loads the unwinder outside the handler
  struct sigaction sa;
  memset(&sa, 0, sizeof sa);
  sa.sa_sigaction = on_prof;
  sa.sa_flags = SA_SIGINFO | SA_RESTART;
  sigaction(SIGPROF, &sa, nullptr);
  clockid_t cid;
  pthread_getcpuclockid(pthread_self(), &cid);
  sigevent sev;
  memset(&sev, 0, sizeof sev);
  sev.sigev_notify = SIGEV_THREAD_ID;
  sev.sigev_signo = SIGPROF;
  sev._sigev_un._tid = (pid_t)syscall(SYS_gettid);
  if (timer_create(cid, &sev, &g_timer) != 0) return;
  itimerspec its;
  memset(&its, 0, sizeof its);
  its.it_interval.tv_nsec = 100000;
  its.it_value.tv_nsec = 100000;
  timer_settime(g_timer, 0, &its, nullptr);
  g_on = true;
  atexit(report);
}

}  // namespace spq
