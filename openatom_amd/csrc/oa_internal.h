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

// What the digit slices in a handle's workspace were made from: the last
// split slicing on the handle.  Any other use of the workspace clears it, so a
// back-multiply asked to reuse digits re-slices unless they are still there.
struct oa_digit_tag {
  bool valid = false;
  const double *psi = nullptr;
  long ld = 0;
  int n = 0, K = 0, prec = 0;
};

struct oa_context {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  oa_workspace ws;          // split-K partials and digit slices
  oa_workspace ns;          // S->T ping-pong buffers (kept apart so digits survive)
  oa_ns_report *report = nullptr;  // device report for synchronous S->T
  oa_ns_report *host_report = nullptr;  // pinned mirror
  oa_digit_tag digits;      // contents of ws's digit area
  oa_workspace pinned;      // pinned host staging of the transport calls
};

int oa_ws_require(oa_context *c, size_t bytes);

// Looks up device `device`'s copy of the no-alloc step's scratch array
// (once per device; the device is current).
int oa_noalloc_init(int device);

// Synchronous S->T on the handle's stream with the buffers at ws
// (oa_ns_ws_bytes(n)) and the device report rep; the host reads rep through
// the handle's pinned mirror every two iterations.
int oa_s_to_t_sync(oa_context *c, int n, const double *S, int nspin, double tol, int maxit,
                   double *T, int *iters, double *err, void *ws, oa_ns_report *rep);

// Makes the handle's device current for the scope of an API call and
// restores the caller's device afterwards.
struct DeviceGuard {
  int prev = -1;
  bool ok = true;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};
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
  // columns n < cw of C use alpha_w instead of alpha (the kx = 0 weight of
  // the GSpace force update); cw = 0 leaves every column at alpha
  int cw = 0;
  double alpha_w = 0.0;
};

// Launch one or two independent problems of the same shape class in one grid
// (grid.y selects the problem).  `skip` (device int, may be null) turns the
// launch into a no-op when nonzero (used by guarded S->T iterations).
int oa_launch_gemm_f64(hipStream_t s, bool tx, bool ty, const oa_gemm_problem *p,
                       int nprob, const int32_t *skip);

// Symmetric overlap, FP64 MFMA, upper-triangle tiles, split-K.
// S (row-major n x n, full) = alpha * sum_k w_k psi[x][k] psi[y][k] with
// w_k = 1/2 for k < kx0_end and 1 otherwise.
// X[r][c] *= s for r < rows, c < ncols (row-major, leading dimension ld).
int oa_launch_scale_cols(hipStream_t s, int rows, int ncols, double *X, long ld, double scale);

size_t oa_overlap_f64_ws_bytes(int n, int K);
int oa_launch_overlap_f64(hipStream_t s, int n, int K, const double *psi,
                          long ld, int kx0_end, double alpha, double *S,
                          double diag, double *maxdev, void *ws, size_t ws_bytes);

// The same product with the split depth capped so that its partial tiles fit
// ws_bytes (OA_ERR_CAPACITY when one chunk does not); the sums differ from
// oa_launch_overlap_f64's only by the chunking of the k sum.
int oa_launch_overlap_f64_fit(hipStream_t s, int n, int K, const double *psi, long ld,
                              int kx0_end, double alpha, double *S, double diag,
                              double *maxdev, void *ws, size_t ws_bytes);

// Rectangular split-K product C[M][N] = alpha * X[M][K] * Y[N][K]^T (Lambda).
size_t oa_ntsplit_f64_ws_bytes(int M, int N, int K);
int oa_launch_ntsplit_f64(hipStream_t s, int M, int N, int K, const double *X,
                          long ldx, const double *Y, long ldy, double alpha,
                          double *C, long ldc, void *ws, size_t ws_bytes);
// Upper tiles of C = alpha X Y^T (n x n) for two distinct [n][K] operands,
// the rest of C the mirror (the strict upper triangle is the product); same
// workspace as oa_launch_overlap_f64.
int oa_launch_overlap_xy_f64(hipStream_t s, int n, int K, const double *X, long ldx,
                             const double *Y, long ldy, double alpha, double *C, void *ws,
                             size_t ws_bytes);

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

// ---------------------------------------------------------------- complex
// Column-major ZGEMM on device buffers (COMPLEX*16 interleaved); alpha and
// beta are (re, im) pairs in host memory.  Workspace: oa_zgemm_ws_bytes.
size_t oa_zgemm_ws_bytes(char ta, char tb, int m, int n, int k);
int oa_launch_zgemm(hipStream_t s, char ta, char tb, int m, int n, int k, const double *alpha,
                    const double *a, int lda, const double *b, int ldb, const double *beta,
                    double *c, int ldc, void *ws, size_t ws_bytes);

// ---------------------------------------------------------------- split
size_t oa_split_ws_bytes(int prec, int n, int K);
int oa_launch_overlap_split(hipStream_t s, int prec, int n, int K,
                            const double *psi, long ld, int kx0_end,
                            double alpha, double *S, double diag,
                            double *maxdev, void *ws, size_t ws_bytes);
// Asymmetric split product L[x][y] = alpha * sum_k psi[x][k] F[y][k] (the
// Lambda forward) on digit slices of psi and F, all tiles, split-K.
size_t oa_split_lambda_ws_bytes(int prec, int n, int K);
int oa_launch_lambda_split(hipStream_t s, int prec, int n, int K, const double *psi, long ld,
                           const double *F, long ldF, double alpha, double *L, void *ws,
                           size_t ws_bytes);
int oa_launch_backmul_split(hipStream_t s, int prec, int n, int K,
                            const double *M, int transM, const double *psi,
                            long ldpsi, int kx0_end, double alpha, double beta,
                            double *out, long ldout, void *ws, size_t ws_bytes,
                            int cw = 0, double alpha_w = 0.0);
// In place on the back-multiply digits (DB) a split slicing of an [n][K]
// block left in ws: digits of psi become those of J psi, J (re, im) =
// (-im, re), exactly (the chunk maxima do not change).
int oa_launch_jswap_db(hipStream_t s, int prec, int n, int K, void *ws, size_t ws_bytes);
// Im S = alpha psi (J psi)^T from the digits a split overlap of the same
// [n][K] block without kx = 0 columns left in ws (no slicing); needs
// oa_split_zimag_ws_bytes, which includes the overlap's own workspace.
size_t oa_split_zimag_ws_bytes(int prec, int n, int K);
int oa_launch_zimag_split(hipStream_t s, int prec, int n, int K, double alpha, double *Li,
                          void *ws, size_t ws_bytes);

// ---------------------------------------------------------------- grouped
// A grouped launch walks `total` work items spread over `count` instances:
// item t belongs to instance i with prefix[i] <= t < prefix[i + 1], and
// args[i] holds that instance's arguments.  Both tables live in device
// memory; the launch is persistent (grid sized to the device, each workgroup
// loops over items) and maps items to XCDs in contiguous ranges.
struct oa_group {
  const void *args;   // device array of per-instance argument structs
  const int *prefix;  // device int[count + 1]
  int count;
  int total;          // prefix[count]
};

// Per-instance arguments of the grouped kernels (plain data, uploaded as
// arrays).  Field meanings follow the single-instance kernels.
struct oa_chunkmax_args {
  const double *psi;
  long ld;
  int n, K, kx0_end, kchunk, nz, pad;
  unsigned long long *maxS, *maxB;
};
struct oa_slice_args {
  const double *psi;
  long ld;
  int n, K, kx0_end, kchunk;
  const unsigned long long *maxS, *maxB;
  void *DS, *DB;
  int np, Kp, want_ds, vec;
  int want_db, pad;
  const int *slices;  // per chunk digit count (adaptive), null: all S
};
// Per chunk slice counts of the adaptive split (one workgroup per instance).
struct oa_count_args {
  const unsigned long long *maxB;
  int *slices;
  int nz, K, kchunk, smax;
};
struct oa_digit_args {
  const void *X;  // k-blocked digits, rows m
  long ldx, sx;   // rows of the layout and slice stride (elements)
  const void *Y;  // k-blocked digits, rows n
  long ldy, sy;
  int M, N;       // output extent
  int kc;         // contraction length per split (elements) or total
  int kfull;      // contraction extent of the digit arrays (padded)
  int ntiles, nz; // MODE 0: symmetric tiles, splits
  double *part;   // MODE 0 output
  const unsigned long long *chunk_max;  // MODE 0: per chunk scale source
  double *out;    // MODE 1 output
  long ldo;
  double alpha, beta;
  const int *fexp;                   // per output row exponent (T rows)
  const double *diagv;               // FP64 diagonal of T
  const void *DB;                    // psi digits (rows Kp, contraction np)
  long dbs;                          // DB slice stride (elements)
  int np, Kp;
  const unsigned long long *colmax;  // per chunk of output columns
  int kchunk;
  int cw;                            // MODE 1: columns < cw use alpha_w
  double alpha_w;
  const int *slices;                 // adaptive: digits per chunk (null: S)
};
struct oa_sreduce_args {
  const double *part;
  int n, nz;
  double alpha;
  const double *psi;
  long ld;
  int kx0_end, pad;
  double *S;
  double diag;
  unsigned long long *maxdev;
};
struct oa_rows_args {
  const double *M;
  int n, trans;
  void *D;
  int npk, pad;
  int *fexp;
  double *diagv;
};

// Split-precision layout of one instance (digit workspace offsets).
struct oa_split_plan {
  int dt, S, es, tile;
  int adaptive;     // per chunk slice counts (OA_PREC_ADAPTIVE)
  int np, Kp;       // padded rows / columns (multiples of the tile)
  int kchunk, nz;   // psi chunks: scale blocks and split-K units
  int ntiles;       // upper-triangle tiles of S
  size_t off_ds, off_db, off_dt, off_maxS, off_maxB, off_fexp, off_diag, off_part, off_md,
      off_sl;
  size_t total;
};
oa_split_plan oa_split_plan_for(int prec, int n, int K, int tile);
// OA_SUCCESS when a split back-multiply of n states fits the digit kernels'
// exact int32 sums (OA_ERR_CAPACITY otherwise), OA_ERR_UNSUPPORTED for an
// unknown split precision.
int oa_split_check(int prec, int n, int tile);
// Fill the per-phase argument structs of one instance whose digit workspace
// starts at `ws` (layout `pl`).  maxS/maxB may point outside ws (batched
// counters); null selects the plan's own slots.
struct oa_split_inst_args {
  oa_chunkmax_args cm;
  oa_count_args cnt;
  oa_slice_args sl;
  oa_digit_args d0;
  oa_sreduce_args rd;
  oa_rows_args rows;
  oa_digit_args d1;
  int items_cm, items_sl, items_d0, items_rd, items_rows, items_d1;
};
void oa_split_inst_fill(const oa_split_plan &pl, int prec, int n, int K, const double *psi,
                        long ld, int kx0_end, double alpha, double *S, const double *T,
                        int transT, double beta, double *out, long ldout, char *ws,
                        unsigned long long *maxS, unsigned long long *maxB,
                        oa_split_inst_args *a);
enum oa_split_phase {
  OA_SP_CHUNKMAX = 0, OA_SP_SLICE, OA_SP_DIGIT0, OA_SP_REDUCE, OA_SP_ROWS, OA_SP_DIGIT1,
  OA_SP_COUNT
};
// One grouped launch of a split phase for instances sharing `prec` and `tile`.
int oa_split_grouped(hipStream_t s, int prec, int tile, int phase, const oa_group &g);
int oa_fill_u64(hipStream_t s, unsigned long long *p, long n);
// Chunk maxima and adaptive slice counts of psi (a split plan of `prec` in
// ws); *counts receives the device array of nz counts.
int oa_launch_slice_counts(hipStream_t s, int prec, int n, int K, const double *psi, long ld,
                           int kx0_end, void *ws, size_t ws_bytes, const int **counts,
                           int *nz, int *kchunk);

// FP64 grouped kernels: GEMM (row-major NN, per-problem skip flag) and the
// S->T bookkeeping kernels.
struct oa_gemm_item {
  oa_gemm_problem p;
  const int32_t *skip;  // may be null
};
struct oa_ns_args {
  const double *S;
  int n, nspin;
  double *A[2], *B[2], *C;
  double *partial;
  int nwg, maxit;
  double tol;
  double *T;
  oa_ns_report *rep;
};
enum oa_ns_phase { OA_NS_INIT = 0, OA_NS_RESID0, OA_NS_RESID1, OA_NS_UPDATE, OA_NS_FINISH };
int oa_gemm_grouped(hipStream_t s, int wm, const oa_group &g);
int oa_ns_grouped(hipStream_t s, int phase, const oa_group &g);
int oa_ns_items(int phase, int n, int nwg);
