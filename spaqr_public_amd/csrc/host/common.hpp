// Host-side infrastructure of the spaQR engine: separator / cluster IDs,
// device memory arenas, the pinned descriptor stager and error handling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace spq {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
enum : int { kOk = 0, kErrHip = -1, kErrArg = -2, kErrNumeric = -3, kErrOom = -4, kErrState = -5 };

#define SPQ_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::spq::Error(::spq::kErrHip, std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                                             __FILE__ + ":" + std::to_string(__LINE__));            \
  } while (0)

// Stream capture and device-memory management are serialised process-wide:
// a free (or pinned allocation) synchronises the device, which invalidates a
// stream capture running in another thread (factorize captures the solve
// graph), so trees that run in parallel threads take this lock around both.
std::recursive_mutex& capture_lock();
inline void dev_free(void* p) {
  if (!p) return;
  std::lock_guard<std::recursive_mutex> lk(capture_lock());
  (void)hipFree(p);
}
inline void host_free(void* p) {
  if (!p) return;
  std::lock_guard<std::recursive_mutex> lk(capture_lock());
  (void)hipHostFree(p);
}
inline hipError_t dev_malloc(void** p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(capture_lock());
  return hipMalloc(p, bytes);
}
inline hipError_t host_malloc(void** p, size_t bytes) {
  std::lock_guard<std::recursive_mutex> lk(capture_lock());
  return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

inline double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Separator ID (level, index) and cluster ID (self, left, right) with the
// reference's lexicographic order (cluster.h:22-111).
struct SepID {
  int lvl = -1, sep = 0;
  SepID() {}
  SepID(int l, int s) : lvl(l), sep(s) {}
  bool operator==(const SepID& o) const { return lvl == o.lvl && sep == o.sep; }
  bool operator!=(const SepID& o) const { return !(*this == o); }
  bool operator<(const SepID& o) const { return lvl < o.lvl || (lvl == o.lvl && sep < o.sep); }
};
struct CID {
  SepID self, l, r;
  int part = 2, section = 0;
  bool operator==(const CID& o) const {
    return self == o.self && l == o.l && r == o.r && part == o.part && section == o.section;
  }
  bool operator!=(const CID& o) const { return !(*this == o); }
  bool operator<(const CID& o) const {
    if (self != o.self) return self < o.self;
    if (l != o.l) return l < o.l;
    if (r != o.r) return r < o.r;
    if (section != o.section) return section < o.section;
    return part < o.part;
  }
};
// cluster.cpp:19-32
inline SepID sep_merge(const SepID& s) { return SepID(s.lvl + 1, s.sep / 2); }
inline CID cid_merge_if(const CID& c, int lvl) {
  CID o;
  o.self = c.self;
  o.l = c.l.lvl < lvl ? sep_merge(c.l) : c.l;
  o.r = c.r.lvl < lvl ? sep_merge(c.r) : c.r;
  o.section = lvl > 0 ? c.section / 2 : c.section;
  o.part = c.self.lvl == lvl ? 2 : c.part;
  return o;
}

// Bump allocator over large hipMalloc'd chunks.  Permanent storage (factors,
// blocks) lives until the arena is destroyed; a scratch arena is rewound
// between phases (stream order guarantees earlier kernels are done with it).
class Arena {
 public:
  explicit Arena(size_t chunk = size_t(256) << 20) : chunk_(chunk) {}
  ~Arena() { release(); }
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  void* alloc(size_t bytes, size_t align = 256);
  void rewind();  // keep chunks, reset offsets
  void release();
  size_t used() const { return used_; }
  // chunk allocations of all arenas: count and host seconds (diagnostics)
  static long n_chunks;
  static double t_chunks;
  size_t reserved() const { return reserved_; }

 private:
  struct Chunk { char* p; size_t size, off; };
  size_t chunk_;
  std::vector<Chunk> chunks_;
  size_t cur_ = 0, used_ = 0, reserved_ = 0;
};

// Process-wide pool of staging buffers (pinned host or device memory of the
// current device), so trees created one after another reuse them instead of
// allocating and pinning again.  pool_take returns a buffer of at least
// `bytes` (the smallest pooled one that fits, or the largest with `largest`,
// else a new one) and sets `bytes` to its size; pool_give returns it to the pool of device `dev` (the device
// current when it was taken; the caller's stream work on it must be
// complete).  At most 2 GiB per kind and device stay pooled.
char* pool_take(bool host, size_t& bytes, bool largest = false);
void pool_give(bool host, char* p, size_t bytes, int dev);

// Pinned host staging + device ring for descriptor tables.  push() copies a
// table into the pinned ring, enqueues the H2D copy on the stream and returns
// the device address.  The ring is reset at host synchronisation points.
// A table that does not fit behind the last one moves to a ring of twice the
// size (the old ring is kept until the next synchronisation, then pooled),
// up to kMaxRing; past that the ring wraps, which waits for the stream.
class Stager {
 public:
  static constexpr size_t kMaxRing = size_t(1) << 30;
  Stager() {}
  ~Stager();
  void init(hipStream_t s, size_t bytes);
  long n_grow = 0, n_wrap = 0;  // diagnostics
  template <typename X>
  const X* push(const std::vector<X>& v) {
    return (const X*)push_raw(v.data(), sizeof(X) * v.size());
  }
  const void* push_raw(const void* src, size_t bytes);
  // Copies everything pushed since the last flush with one transfer; every
  // launch that reads pushed tables flushes first.
  void flush();
  // only after the stream has been synchronised (uploads still pending are
  // issued and waited for first)
  void reset() {
    if (up_pend_bytes_) {
      flush_uploads();
      SPQ_HIP(hipStreamSynchronize(stream_));
    }
    off_ = 0; pend_ = 0; up_off_ = 0;
    release_retired();
  }
  // Persistent copy (for solve programs): allocated from `perm` and kept.
  // The table goes through a pinned buffer with an asynchronous copy; the
  // buffer is reused after the next synchronisation (or a forced one when
  // it is full).
  template <typename X>
  X* upload(Arena& perm, const std::vector<X>& v) {
    if (v.empty()) return nullptr;
    X* d = (X*)perm.alloc(sizeof(X) * v.size());
    upload_raw(d, v.data(), sizeof(X) * v.size());
    return d;
  }
  // Persistent uploads whose device destinations follow one another (the
  // arena hands them out in order) are merged into one copy, issued here;
  // flush() and every synchronisation issue it first.
  void upload_raw(void* dst, const void* src, size_t bytes);
  void flush_uploads();

 private:
  struct Ring { char* host; char* dev; size_t host_bytes, dev_bytes; };
  void release_retired();
  hipStream_t stream_ = nullptr;
  int device_ = 0;
  char* host_ = nullptr;
  char* dev_ = nullptr;
  size_t host_bytes_ = 0, dev_bytes_ = 0;
  std::vector<Ring> retired_;              // outgrown rings still read by queued work
  size_t cap_ = 0, off_ = 0, pend_ = 0;  // [pend_, off_) not yet copied
  char* up_host_ = nullptr;               // pinned source of persistent uploads
  size_t up_cap_ = 0, up_off_ = 0;
  size_t up_pend_host_ = 0, up_pend_bytes_ = 0;  // merged upload not yet issued:
  char* up_pend_dev_ = nullptr;                   // [up_pend_host_, +bytes) -> up_pend_dev_
};

// Prefix sums of per-descriptor tile counts (tstart table of tiled kernels).
inline int prefix_tiles(const std::vector<int>& counts, std::vector<int>& start) {
  start.resize(counts.size());
  int s = 0;
  for (size_t i = 0; i < counts.size(); ++i) { start[i] = s; s += counts[i]; }
  return s;
}

}  // namespace spq
