// Internal declarations shared by the HIP translation units of liboa_ortho.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/oa_ortho.h"

#define OA_HIP_CHECK(expr)                                                    \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      oa_note_hip_error(_e, #expr, __FILE__, __LINE__);                       \
      return OA_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

void oa_note_hip_error(hipError_t e, const char *what, const char *file, int line);

// Device workspace owned by a handle.  Grows on demand outside graph
// capture; oa_reserve() pre-sizes it.
struct oa_workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct oa_context {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  oa_workspace ws;          // split-K partials and digit slices
  oa_workspace ns;          // S->T ping-pong buffers (kept apart so digits survive)
  oa_ns_report *report = nullptr;  // device report for synchronous S->T
  oa_ns_report *host_report = nullptr;  // pinned mirror
};

int oa_ws_require(oa_context *c, size_t bytes);
int oa_ws_require_into(oa_context *c, oa_workspace &w, size_t bytes);

// ---------------------------------------------------------------- FP64 GEMM
// Row-major C[M][N] = alpha * op(X)[M][K] * op(Y)[K][N] + beta * C
//                     + diag_add * I.
//   op(X) = X ([M][K], ldx) or X^T (X stored [K][M]) when tx.
//   op(Y) = Y ([K][N], ldy) or Y^T (Y stored [N][K]) when ty.
struct oa_gemm_problem {
  int M, N, K;
  const double *X;
  long ldx;
  const double *Y;
  long ldy;
  double *C;
  long ldc;
  double alpha, beta, diag_add;
};

// Launch one or two independent problems of the same shape class in one grid
// (grid.y selects the problem).  `skip` (device int, may be null) turns the
// launch into a no-op when nonzero (used by guarded S->T iterations).
int oa_launch_gemm_f64(hipStream_t s, bool tx, bool ty, const oa_gemm_problem *p,
                       int nprob, const int32_t *skip);

// Symmetric overlap, FP64 MFMA, upper-triangle tiles, split-K.
// S (row-major n x n, full) = alpha * sum_k w_k psi[x][k] psi[y][k] with
// w_k = 1/2 for k < kx0_end and 1 otherwise.
size_t oa_overlap_f64_ws_bytes(int n, int K);
int oa_launch_overlap_f64(hipStream_t s, int n, int K, const double *psi,
                          long ld, int kx0_end, double alpha, double *S,
                          double diag, double *maxdev, void *ws, size_t ws_bytes);

// Rectangular split-K product C[M][N] = alpha * X[M][K] * Y[N][K]^T (Lambda).
size_t oa_ntsplit_f64_ws_bytes(int M, int N, int K);
int oa_launch_ntsplit_f64(hipStream_t s, int M, int N, int K, const double *X,
                          long ldx, const double *Y, long ldy, double alpha,
                          double *C, long ldc, void *ws, size_t ws_bytes);

// S->T on the device.  Workspace: oa_ns_ws_bytes(n).  begin / iterate /
// finish let a host loop stop early; oa_launch_s_to_t enqueues `maxit`
// guarded iterations (no host decision, capturable).
size_t oa_ns_ws_bytes(int n);
int oa_ns_begin(hipStream_t s, int n, const double *S, int nspin,
                oa_ns_report *rep, void *ws);
int oa_ns_iterate(hipStream_t s, int n, int it, double tol, int maxit,
                  oa_ns_report *rep, void *ws);
int oa_ns_finish(hipStream_t s, int n, double *T, const oa_ns_report *rep,
                 void *ws);
int oa_launch_s_to_t(hipStream_t s, int n, const double *S, int nspin,
                     double tol, int maxit, double *T, oa_ns_report *report,
                     void *ws);

// ---------------------------------------------------------------- split
size_t oa_split_ws_bytes(int prec, int n, int K);
int oa_launch_overlap_split(hipStream_t s, int prec, int n, int K,
                            const double *psi, long ld, int kx0_end,
                            double alpha, double *S, double diag,
                            double *maxdev, void *ws, size_t ws_bytes);
int oa_launch_backmul_split(hipStream_t s, int prec, int n, int K,
                            const double *M, int transM, const double *psi,
                            long ldpsi, int kx0_end, double alpha, double beta,
                            double *out, long ldout, void *ws, size_t ws_bytes);
