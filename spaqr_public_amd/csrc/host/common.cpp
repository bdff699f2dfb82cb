#include "common.hpp"

#include <cstring>

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

Stager::~Stager() {
  host_free(host_);
  host_free(up_host_);
  dev_free(dev_);
}

void Stager::upload_raw(void* dst, const void* src, size_t bytes) {
  constexpr size_t kUpCap = size_t(32) << 20;
  if (!up_host_) {
    SPQ_HIP(host_malloc((void**)&up_host_, kUpCap));
    up_cap_ = kUpCap;
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
  cap_ = bytes;
  SPQ_HIP(host_malloc((void**)&host_, bytes));
  SPQ_HIP(dev_malloc((void**)&dev_, bytes));
  off_ = 0;
}

const void* Stager::push_raw(const void* src, size_t bytes) {
  if (bytes == 0) return nullptr;
  size_t need = (bytes + 255) / 256 * 256;
  if (need > cap_) throw Error(kErrOom, "descriptor table larger than the staging ring");
  if (off_ + need > cap_) {
    // wrap: copy what is pending, let every reader of the ring finish
    flush();
    SPQ_HIP(hipStreamSynchronize(stream_));
    off_ = 0;
    pend_ = 0;
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
