// GMRES that allocates nothing: spaqr_gmres_device_noalloc (include/spaqr.h),
// for callers that cannot call hipMalloc (graph capture, serving loops).
//
// The work vectors of gmres_dev (FP64 trees) or of gmres_ir_dev (FP32 trees)
// live in `scratch`, a fixed device array of kNoAllocScratchDoubles doubles
// (64 MiB) that every GPU the library runs on holds until the process ends;
// the dot products are read back into a fixed host array per device.  The
// device address of each GPU's copy is looked up once per device, with that
// device current, and only that address is handed to kernels and copies.
// The kernels write every scratch element before reading it (partial sums,
// basis columns, residuals), so nothing needs clearing between calls.  Every
// call on one GPU shares its array: calls that overlap on one GPU race.  The
// solve graph is instantiated by factorize, so the preconditioner applies
// inside allocate nothing either.
#include <mutex>
#include <string>

#include "capi_impl.hpp"
#include "kernels/launch.hpp"

namespace {
constexpr size_t kNoAllocScratchDoubles = size_t(1) << 23;
constexpr int kNoAllocMaxRestart = 1024;
constexpr int kNoAllocMaxDevices = 16;
constexpr int kNoAllocHostDoubles = 2 * (kNoAllocMaxRestart + 2) + 4;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace spq {
namespace {

double* g_scratch_ptr[kNoAllocMaxDevices] = {};
double g_host_read[kNoAllocMaxDevices][kNoAllocHostDoubles];
std::mutex g_scratch_mu;

// Device address of this GPU's copy of `scratch` (device current).
double* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    SPQ_HIP(hipGetDevice(&cur));
    if (cur != device) throw Error(kErrState, "gmres_noalloc: the tree's device is not current");
    void* p = nullptr;
    SPQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(kErrHip, "gmres_noalloc: no device address for the scratch array");
    g_scratch_ptr[device] = (double*)p;
  }
  return g_scratch_ptr[device];
}

size_t up32(size_t e) { return (e + 31) / 32 * 32; }  // sub-buffers on 256-byte boundaries

// Scratch doubles one run needs for n unknowns (restart already clamped).
size_t noalloc_doubles(size_t n, int restart, bool refine) {
  size_t e = up32(n * (restart + 1)) + 2 * up32(n) + up32((size_t)2 * (restart + 2)) +
             up32(dots_part_elems(restart + 1));
  if (refine) e += 2 * up32(n) + up32(2) + up32(dots_part_elems(1));
  return e;
}

KrylovResult gmres_noalloc(EngineBase& M, const DevMatrix& A, const double* b, double* x, int iters, int restart,
                           double tol, int verbose, int device) {
  if (device < 0 || device >= kNoAllocMaxDevices)
    throw Error(kErrArg, "gmres_noalloc: device ordinal beyond " + std::to_string(kNoAllocMaxDevices - 1));
  const bool refine = M.precision_bytes() == 4;
  const size_t n = (size_t)std::max(A.m, 1);
  restart = std::max(1, std::min(restart, (int)n));
  if (restart > kNoAllocMaxRestart)
    throw Error(kErrArg, "gmres_noalloc: restart " + std::to_string(restart) + " exceeds the maximum " +
                             std::to_string(kNoAllocMaxRestart));
  const size_t need = noalloc_doubles(n, restart, refine);
  if (need > kNoAllocScratchDoubles)
    throw Error(kErrArg, "gmres_noalloc: n = " + std::to_string(n) + " with restart " + std::to_string(restart) +
                             " needs " + std::to_string(need) + " doubles of scratch; the fixed array holds " +
                             std::to_string(kNoAllocScratchDoubles) + " (lower restart, or use spaqr_gmres_device)");
  double* p = scratch_on(device);
  auto take = [&](size_t e) {
    double* q = p;
    p += up32(e);
    return q;
  };
  GmresWork w{};
  w.V = take(n * (restart + 1));
  w.w = take(n);
  w.p0 = take(n);
  w.h = take((size_t)2 * (restart + 2));
  w.dpart = take(dots_part_elems(restart + 1));
  w.hh = g_host_read[device];
  if (refine) {
    w.r = take(n);
    w.d = take(n);
    w.h_ir = take(2);
    w.dpart_ir = take(dots_part_elems(1));
    w.hh_ir = g_host_read[device] + 2 * (kNoAllocMaxRestart + 2);
  }
  return refine ? gmres_ir_dev(M, A, b, x, iters, restart, tol, verbose, false, &w)
                : gmres_dev(M, A, b, x, iters, restart, tol, verbose, false, &w);
}

}  // namespace
}  // namespace spq

using namespace spq;

extern "C" int spaqr_gmres_device_noalloc(spaqr_tree* t, const spaqr_matrix* A, const double* b, double* x,
                                          int iters, int restart, double tol, int verbose, int* iters_out,
                                          double* est_out) {
  return guard([&] {
    EngineBase& e = eng(t);
    if (!A || !A->d) throw Error(kErrArg, "null matrix");
    if (A->d->m != A->d->n || A->d->m != e.rows()) throw Error(kErrArg, "gmres needs the square matrix of the tree");
    SPQ_HIP(hipSetDevice(t->device));
    KrylovResult r = gmres_noalloc(e, *A->d, b, x, iters, restart, tol, verbose, t->device);
    e.check_solve();
    if (iters_out) *iters_out = r.iters;
    if (est_out) *est_out = r.est;
  });
}
