// GPU Krylov drivers preconditioned by the spaQR factors (is.cpp:95-346).
#pragma once
#include "engine.hpp"

namespace spq {

// Device CSR copy of A (and of A^T for the normal equations), FP64.
struct DevMatrix {
  int m = 0, n = 0, nnz = 0;
  int *rp = nullptr, *ci = nullptr;
  double* v = nullptr;
  int *rpt = nullptr, *cit = nullptr;  // CSR of A^T
  double* vt = nullptr;
  ~DevMatrix();
};
DevMatrix* upload_matrix(const CscMat& A, bool with_transpose);

struct KrylovResult {
  int iters = 0;
  double est = 0;       // the solver's own stopping quantity
  double t_total = 0;   // seconds, device synchronised
  double t_precond = 0; // seconds spent in preconditioner applies (host clock)
  int n_precond = 0;
  int outer = 0;        // refinement steps (gmres_ir_dev)
};

// Device and host work vectors of one GMRES run: the Krylov basis and
// partial sums of gmres_dev, and the residual / correction of the refinement
// loop of gmres_ir_dev around it.  hh / hh_ir are host buffers the dot
// products are read back into.
struct GmresWork {
  double *V, *w, *p0, *h, *dpart, *hh;      // gmres_dev: n (restart+1), n, n, 2 (restart+2), dots, 2 (restart+2)
  double *r, *d, *h_ir, *dpart_ir, *hh_ir;  // gmres_ir_dev: n, n, 2, dots, 2
};
// The tree's own work vectors for n unknowns and `restart` (grown on
// demand, kept across calls); refine: also those of the refinement loop.
GmresWork gmres_work_tree(EngineBase& M, int n, int restart, bool refine);

// Left-preconditioned restarted GMRES: r0 = M^{-1}(b - A x), stop when
// |g_k| / ||r0|| < tol (is.cpp:217-218), Arnoldi by classical Gram-Schmidt
// with one reorthogonalisation, FP64 throughout.  b, x are device vectors.
// trans: solve A^T x = b instead, with A^T from the uploaded transpose and
// M^{-T} as the preconditioner (the adjoint solve).
// wk: work vectors to use (nullptr: the tree's own).
KrylovResult gmres_dev(EngineBase& M, const DevMatrix& A, const double* b, double* x, int iters, int restart,
                       double tol, int verbose, bool trans = false, const GmresWork* wk = nullptr);

// Mixed form: FP64 iterative refinement around gmres_dev for a tree factored
// in FP32.  Each step forms the true residual r = b - A x with the FP64 SpMV,
// solves A d = r by gmres_dev (FP64 Arnoldi, FP32 factors as the
// preconditioner) to a relative estimate of max(kIrInnerTol, tol / (2 rel)),
// rel the current true relative residual, and adds d.
// Stops when ||b - A x|| / ||b|| <= tol, when a step fails to halve the true
// residual, or when `iters` inner iterations are spent.  est = final true
// relative residual; iters = total inner iterations.
constexpr double kIrInnerTol = 1e-6;
constexpr int kIrMaxOuter = 20;
KrylovResult gmres_ir_dev(EngineBase& M, const DevMatrix& A, const double* b, double* x, int iters, int restart,
                          double tol, int verbose, bool trans = false, const GmresWork* wk = nullptr);

// Preconditioned CGLS on the normal equations (is.cpp:275-346): stop when
// ||A^T r||^2 < tol^2 ||A^T b||^2; est = ||A^T r|| / ||A^T b||.  b has length
// m, x length n (device vectors); A must carry its transpose.
KrylovResult cgls_dev(EngineBase& M, const DevMatrix& A, const double* b, double* x, int iters, double tol,
                      int verbose);

}  // namespace spq
