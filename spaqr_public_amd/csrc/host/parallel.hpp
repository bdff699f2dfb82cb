// Host worker pool for the planner's independent loops (descriptor tables,
// per-cluster scans).  Work is cut into a fixed number of contiguous chunks
// that depends only on the loop length and grain, so callers that write
// per-chunk results and join them in chunk order get the same output whether
// the chunks ran on the pool or one after another on the calling thread.
#pragma once
#include <functional>

namespace spq {

// Host threads the planner may use: SPAQR_HOST_THREADS (default 8, at most
// the machine's hardware threads; 1 = serial).
int host_threads();

// Number of chunks parallel_for splits [0, n) into with at least `grain`
// items per chunk, at most kMaxChunks (0 when n <= 0); it does not depend on
// the thread count.  Chunk c covers [c*n/k, (c+1)*n/k).
constexpr int kMaxChunks = 32;
int plan_chunks(int n, int grain);

// Calls f(chunk, begin, end) for every chunk of [0, n), on the pool's
// threads and the calling thread; returns when all chunks are done.  When
// another thread is using the pool, the caller runs every chunk itself.  The
// first exception thrown by f is rethrown here after all chunks have ended.
void parallel_for(int n, int grain, const std::function<void(int, int, int)>& f);

}  // namespace spq
