#include "common.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace spq {

std::recursive_mutex& capture_lock() {
  static std::recursive_mutex m;
  return m;
}

long Arena::n_chunks = 0;
double Arena::t_chunks = 0;

void* Arena::alloc(size_t bytes, size_t align) {
  if (bytes == 0) bytes = align;
  while (cur_ < chunks_.size()) {
    Chunk& c = chunks_[cur_];
    size_t off = (c.off + align - 1) / align * align;
    if (off + bytes <= c.size) {
      c.off = off + bytes;
      used_ += bytes;
      return c.p + off;
    }
    ++cur_;
  }
  size_t sz = bytes > chunk_ ? (bytes + align) : chunk_;
  void* p = nullptr;
  const double t0 = wall();
  hipError_t e = dev_malloc(&p, sz);
  t_chunks += wall() - t0;
  ++n_chunks;
  if (e != hipSuccess)
    throw Error(kErrOom, "device allocation of " + std::to_string(sz) + " bytes failed: " + hipGetErrorString(e));
  chunks_.push_back({(char*)p, sz, 0});
  reserved_ += sz;
  cur_ = chunks_.size() - 1;
  Chunk& c = chunks_.back();
  c.off = bytes;
  used_ += bytes;
  return c.p;
}

void Arena::rewind() {
  for (auto& c : chunks_) c.off = 0;
  cur_ = 0;
  used_ = 0;
}

void Arena::release() {
  for (auto& c : chunks_) dev_free(c.p);
  chunks_.clear();
  cur_ = 0;
  used_ = 0;
  reserved_ = 0;
}

namespace {
struct PoolEntry { int dev; bool host; size_t bytes; char* p; };
constexpr size_t kPoolBytes = size_t(2) << 30;
std::mutex& pool_mu() {
  static std::mutex m;
  return m;
}
// never destroyed: buffers still pooled at exit are left to the process end
// (freeing them would race the HIP runtime's own teardown)
std::vector<PoolEntry>& pool() {
  static std::vector<PoolEntry>* v = new std::vector<PoolEntry>();
  return *v;
}
}  // namespace

char* pool_take(bool host, size_t& bytes, bool largest) {
  int dev = 0;
  SPQ_HIP(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(pool_mu());
    auto& v = pool();
    int best = -1;
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i].dev == dev && v[i].host == host && v[i].bytes >= bytes &&
          (best < 0 || (largest ? v[i].bytes > v[best].bytes : v[i].bytes < v[best].bytes)))
        best = (int)i;
    if (best >= 0) {
      char* p = v[best].p;
      bytes = v[best].bytes;
      v.erase(v.begin() + best);
      return p;
    }
  }
  void* p = nullptr;
  if (host) SPQ_HIP(host_malloc(&p, bytes));
  else SPQ_HIP(dev_malloc(&p, bytes));
  return (char*)p;
}

void pool_give(bool host, char* p, size_t bytes, int dev) {
  if (!p) return;
  std::vector<PoolEntry> drop;
  {
    std::lock_guard<std::mutex> lk(pool_mu());
    auto& v = pool();
    v.push_back({dev, host, bytes, p});
    // over the limit: free the smallest buffers of this kind and device
    for (;;) {
      size_t tot = 0;
      int small = -1;
      for (size_t i = 0; i < v.size(); ++i) {
        if (v[i].dev != dev || v[i].host != host) continue;
        tot += v[i].bytes;
        if (small < 0 || v[i].bytes < v[small].bytes) small = (int)i;
      }
      if (tot <= kPoolBytes || small < 0) break;
      drop.push_back(v[small]);
      v.erase(v.begin() + small);
    }
  }
  for (auto& e : drop) {
    if (e.host) host_free(e.p);
    else dev_free(e.p);
  }
}

Stager::~Stager() {
  release_retired();
  pool_give(true, host_, host_bytes_, device_);
  pool_give(false, dev_, dev_bytes_, device_);
  if (up_host_) pool_give(true, up_host_, up_cap_, device_);
}

void Stager::release_retired() {
  for (auto& r : retired_) {
    pool_give(true, r.host, r.host_bytes, device_);
    pool_give(false, r.dev, r.dev_bytes, device_);
  }
  retired_.clear();
}

void Stager::upload_raw(void* dst, const void* src, size_t bytes) {
  constexpr size_t kUpCap = size_t(32) << 20;
  if (!up_host_) {
    up_cap_ = kUpCap;
    up_host_ = pool_take(true, up_cap_);
    up_off_ = 0;
  }
  if (bytes > up_cap_) {  // larger than the buffer: a synchronous copy
    flush_uploads();
    SPQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
    SPQ_HIP(hipStreamSynchronize(stream_));
    return;
  }
  if (up_off_ + bytes > up_cap_) {  // every earlier upload must have landed
    flush_uploads();
    SPQ_HIP(hipStreamSynchronize(stream_));
    up_off_ = 0;
  }
  std::memcpy(up_host_ + up_off_, src, bytes);
  char* d = (char*)dst;
  // the pending copy grows when this table sits at the same distance from
  // its start on the device as in the staging buffer (the bytes in between
  // are alignment padding of the earlier destination)
  if (up_pend_bytes_ && d == up_pend_dev_ + (up_off_ - up_pend_host_)) {
    up_pend_bytes_ = up_off_ - up_pend_host_ + bytes;
  } else {
    flush_uploads();
    up_pend_host_ = up_off_;
    up_pend_dev_ = d;
    up_pend_bytes_ = bytes;
  }
  up_off_ += (bytes + 255) / 256 * 256;
}

void Stager::flush_uploads() {
  if (!up_pend_bytes_) return;
  SPQ_HIP(hipMemcpyAsync(up_pend_dev_, up_host_ + up_pend_host_, up_pend_bytes_, hipMemcpyHostToDevice, stream_));
  up_pend_bytes_ = 0;
}

void Stager::init(hipStream_t s, size_t bytes) {
  stream_ = s;
  SPQ_HIP(hipGetDevice(&device_));
  host_bytes_ = bytes;
  host_ = pool_take(true, host_bytes_, true);  // a ring an earlier tree grew
  dev_bytes_ = host_bytes_;
  dev_ = pool_take(false, dev_bytes_);
  cap_ = host_bytes_;
  off_ = 0;
}

const void* Stager::push_raw(const void* src, size_t bytes) {
  if (bytes == 0) return nullptr;
  size_t need = (bytes + 255) / 256 * 256;
  if (off_ + need > cap_ && cap_ < kMaxRing) {
    // grow: what is pending is copied from the old ring, which stays with
    // its readers until the next synchronisation.  Without memory for a
    // larger ring the old one wraps below.
    size_t want = cap_;
    while (want < kMaxRing && want < need) want *= 2;
    if (want < kMaxRing) want *= 2;
    want = std::max(std::min(want, kMaxRing), need);
    size_t hb = want, db = 0;
    char *nh = nullptr, *nd = nullptr;
    try {
      nh = pool_take(true, hb);
      db = hb;
      nd = pool_take(false, db);
    } catch (const Error&) {
      (void)hipGetLastError();
      if (nh) pool_give(true, nh, hb, device_);
      nh = nd = nullptr;
    }
    if (nh && nd) {
      flush();
      retired_.push_back({host_, dev_, host_bytes_, dev_bytes_});
      host_ = nh; host_bytes_ = hb;
      dev_ = nd; dev_bytes_ = db;
      cap_ = hb;
      off_ = 0;
      pend_ = 0;
      ++n_grow;
    }
  }
  if (need > cap_) throw Error(kErrOom, "descriptor table larger than the staging ring");
  if (off_ + need > cap_) {
    // wrap: copy what is pending, let every reader of the ring finish
    flush();
    SPQ_HIP(hipStreamSynchronize(stream_));
    off_ = 0;
    pend_ = 0;
    release_retired();
    ++n_wrap;
  }
  std::memcpy(host_ + off_, src, bytes);
  const void* d = dev_ + off_;
  off_ += need;
  return d;
}

void Stager::flush() {
  flush_uploads();
  if (off_ > pend_) SPQ_HIP(hipMemcpyAsync(dev_ + pend_, host_ + pend_, off_ - pend_, hipMemcpyHostToDevice, stream_));
  pend_ = off_;
}

}  // namespace spq
