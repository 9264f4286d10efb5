// C ABI of liboa_ortho.so: handles, the dgemm_ drop-in and the path-level
// entry points.  Every arithmetic operation runs in a HIP kernel on the
// device; there is no host compute path.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "oa_internal.h"

void oa_note_hip_error(hipError_t e, const char *what, const char *file, int line) {
  fprintf(stderr, "[oa_ortho] HIP error '%s' in %s (%s:%d)\n", hipGetErrorString(e), what,
          file, line);
}

namespace {

bool is_capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) return false;
  return cs != hipStreamCaptureStatusNone;
}

bool lsame(char a, char b) {
  if (a >= 'a' && a <= 'z') a = (char)(a - 'a' + 'A');
  return a == b;
}

// dgemm.f:161-194 argument checks; returns INFO.
int dgemm_info(char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc) {
  const bool nota = lsame(ta, 'N'), notb = lsame(tb, 'N');
  const int nrowa = nota ? m : k, nrowb = notb ? k : n;
  if (!nota && !lsame(ta, 'C') && !lsame(ta, 'T')) return 1;
  if (!notb && !lsame(tb, 'C') && !lsame(tb, 'T')) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (k < 0) return 5;
  if (lda < (nrowa > 1 ? nrowa : 1)) return 8;
  if (ldb < (nrowb > 1 ? nrowb : 1)) return 10;
  if (ldc < (m > 1 ? m : 1)) return 13;
  return 0;
}

// Column-major DGEMM on device pointers, expressed as the row-major problem
// C^T[n][m] = alpha * op(B)^T * op(A)^T + beta * C^T.
int enqueue_dgemm(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha,
                  const double *a, int lda, const double *b, int ldb, double beta, double *c,
                  int ldc) {
  const bool nota = lsame(ta, 'N'), notb = lsame(tb, 'N');
  oa_gemm_problem p;
  p.M = n;
  p.N = m;
  p.K = (alpha == 0.0) ? 0 : k;  // alpha == 0: C = beta*C without touching A, B
  p.X = b;
  p.ldx = ldb;
  p.Y = a;
  p.ldy = lda;
  p.C = c;
  p.ldc = ldc;
  p.alpha = alpha;
  p.beta = beta;
  p.diag_add = 0.0;
  return oa_launch_gemm_f64(s, !notb, !nota, &p, 1, nullptr);
}

bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

thread_local oa_context *tl_default = nullptr;

oa_context *default_handle() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  if (tl_default && tl_default->device == dev) return tl_default;
  oa_handle_t h = nullptr;
  if (oa_create(&h, dev) != OA_SUCCESS) return nullptr;
  tl_default = h;  // the previous default (other device) stays alive for reuse-free exit
  return h;
}

size_t span_elems(int ld, int rows, int cols) {
  if (rows <= 0 || cols <= 0) return 0;
  return (size_t)ld * (size_t)(cols - 1) + (size_t)rows;
}

}  // namespace

int oa_ws_require_into(oa_context *c, oa_workspace &w, size_t bytes) {
  // every user of the main workspace may overwrite the digit area
  if (&w == &c->ws) c->digits.valid = false;
  if (bytes <= w.bytes) return OA_SUCCESS;
  if (is_capturing(c->stream)) return OA_ERR_CAPACITY;  // no allocation under capture
  OA_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (w.ptr) OA_HIP_CHECK(hipFree(w.ptr));
  w.ptr = nullptr;
  w.bytes = 0;
  const size_t gran = (size_t)1 << 24;
  const size_t sz = (bytes + gran - 1) / gran * gran;
  OA_HIP_CHECK(hipMalloc(&w.ptr, sz));
  w.bytes = sz;
  return OA_SUCCESS;
}

int oa_ws_require(oa_context *c, size_t bytes) { return oa_ws_require_into(c, c->ws, bytes); }

extern "C" {

const char *oa_status_string(int st) {
  switch (st) {
    case OA_SUCCESS: return "success";
    case OA_ERR_HIP: return "HIP runtime error";
    case OA_ERR_NOT_CONVERGED: return "S->T iteration limit reached (invsqr_max_iter)";
    case OA_ERR_NO_DEVICE: return "no usable gfx950 device";
    case OA_ERR_UNSUPPORTED: return "unsupported precision or variant";
    case OA_ERR_CAPACITY: return "request exceeds a limit (workspace under capture, size)";
    case OA_ERR_IO: return "state file cannot be opened, parsed or written";
    default: return st > 0 ? "invalid argument (BLAS INFO index)" : "unknown status";
  }
}

int oa_create(oa_handle_t *out, int device) {
  if (!out) return 1;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return OA_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return OA_ERR_NO_DEVICE;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    fprintf(stderr, "[oa_ortho] device %d is %s; this library is built for gfx950 only\n",
            device, prop.gcnArchName);
    return OA_ERR_NO_DEVICE;
  }
  DeviceGuard g(device);
  if (!g.ok) return OA_ERR_NO_DEVICE;
  oa_context *c = new oa_context();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->report, sizeof(oa_ns_report)) != hipSuccess ||
      hipHostMalloc(&c->host_report, sizeof(oa_ns_report), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    oa_destroy(c);
    return OA_ERR_HIP;
  }
  c->stream = c->own_stream;
  const int st = oa_noalloc_init(device);
  if (st) {
    oa_destroy(c);
    return st;
  }
  *out = c;
  return OA_SUCCESS;
}

int oa_destroy(oa_handle_t c) {
  if (!c) return OA_SUCCESS;
  DeviceGuard g(c->device);
  if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
  if (c->stream && c->stream != c->own_stream) (void)hipStreamSynchronize(c->stream);
  if (c->ws.ptr) (void)hipFree(c->ws.ptr);
  if (c->ns.ptr) (void)hipFree(c->ns.ptr);
  if (c->report) (void)hipFree(c->report);
  if (c->host_report) (void)hipHostFree(c->host_report);
  if (c->pinned.ptr) (void)hipHostFree(c->pinned.ptr);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  if (tl_default == c) tl_default = nullptr;
  delete c;
  return OA_SUCCESS;
}

int oa_set_stream(oa_handle_t c, void *stream) {
  if (!c) return 1;
  c->stream = (hipStream_t)stream;  // NULL is the device's default stream
  return OA_SUCCESS;
}

void *oa_get_stream(oa_handle_t c) { return c ? (void *)c->stream : nullptr; }

int oa_synchronize(oa_handle_t c) {
  if (!c) return 1;
  DeviceGuard g(c->device);
  OA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return OA_SUCCESS;
}

// Record that the workspace holds the digits of psi (both layouts after an
// overlap, the back-multiply layout after a back-multiply; a reusing
// back-multiply reads only the latter).
static void tag_digits(oa_context *c, const double *psi, long ld, int n, int K, int prec) {
  c->digits.valid = true;
  c->digits.psi = psi;
  c->digits.ld = ld;
  c->digits.n = n;
  c->digits.K = K;
  c->digits.prec = prec & ~OA_PREC_REUSE_DIGITS;
}

static bool digits_match(const oa_context *c, const double *psi, long ld, int n, int K,
                         int prec) {
  const oa_digit_tag &t = c->digits;
  return t.valid && t.psi == psi && t.ld == ld && t.n == n && t.K == K &&
         t.prec == (prec & ~OA_PREC_REUSE_DIGITS);
}

static size_t ws_for(int n, int npw) {
  const int K = 2 * npw;
  size_t b = oa_overlap_f64_ws_bytes(n, K);
  const size_t l = oa_ntsplit_f64_ws_bytes(n, n, K);
  if (l > b) b = l;
  return b;
}

int oa_reserve(oa_handle_t c, int n, int npw) {
  if (!c) return 1;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  DeviceGuard g(c->device);
  const int st = oa_ws_require(c, ws_for(n, npw));
  if (st) return st;
  return oa_ws_require_into(c, c->ns, oa_ns_ws_bytes(n));
}

int oa_reserve_prec(oa_handle_t c, int n, int npw, int prec) {
  if (!c) return 1;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  int st = oa_reserve(c, n, npw);
  if (st) return st;
  DeviceGuard g(c->device);
  return oa_ws_require(c, oa_split_ws_bytes(prec, n, 2 * npw));
}

// ------------------------------------------------------------- BLAS drop-ins

int oa_dgemm(oa_handle_t c, char ta, char tb, int m, int n, int k, double alpha,
             const double *a, int lda, const double *b, int ldb, double beta, double *cc,
             int ldc) {
  if (!c) return OA_ERR_NO_DEVICE;
  const int info = dgemm_info(ta, tb, m, n, k, lda, ldb, ldc);
  if (info) return info;
  if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return OA_SUCCESS;
  DeviceGuard g(c->device);
  return enqueue_dgemm(c->stream, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, cc, ldc);
}

// Run a column-major GEMM of `esz`-byte elements on host or device buffers:
// device buffers go straight to `run`; host buffers (Charm++ message
// payloads) are staged through the workspace, after `extra` bytes the GEMM
// itself needs.  Synchronous.
extern "C++" {
template <class Run>
int staged_gemm(oa_context *h, size_t esz, char ta, char tb, int m, int n, int k,
                bool need_ab, bool beta0, const void *a, int lda, const void *b, int ldb,
                void *c, int ldc, size_t extra, Run run) {
  const bool nota = lsame(ta, 'N'), notb = lsame(tb, 'N');
  const size_t na = need_ab ? span_elems(lda, nota ? m : k, nota ? k : m) : 0;
  const size_t nb = need_ab ? span_elems(ldb, notb ? k : n, notb ? n : k) : 0;
  const size_t nc = span_elems(ldc, m, n);
  const bool dev = (!need_ab || (is_device_ptr(a) && is_device_ptr(b))) && is_device_ptr(c);
  hipStream_t s = h->stream;
  const size_t al = 256;
  auto upb = [&](size_t e) { return (e * esz + al - 1) / al * al; };
  int st;
  if (dev) {
    st = oa_ws_require(h, extra);
    if (st == OA_SUCCESS) st = run(a, b, c, (char *)h->ws.ptr, h->ws.bytes);
    if (st == OA_SUCCESS && hipStreamSynchronize(s) != hipSuccess) st = OA_ERR_HIP;
    return st;
  }
  st = oa_ws_require(h, upb(na) + upb(nb) + upb(nc) + extra);
  if (st) return st;
  char *da = (char *)h->ws.ptr;
  char *db = da + upb(na);
  char *dc = db + upb(nb);
  char *dx = dc + upb(nc);
  hipError_t e = hipSuccess;
  if (na) e = hipMemcpyAsync(da, a, na * esz, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && nb) e = hipMemcpyAsync(db, b, nb * esz, hipMemcpyHostToDevice, s);
  // C is read when beta != 0; with beta == 0 and ldc > m the rows between m
  // and ldc are not part of C and are carried through unchanged
  if (e == hipSuccess && (!beta0 || ldc != m))
    e = hipMemcpyAsync(dc, c, nc * esz, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) st = run(da, db, dc, dx, h->ws.bytes - (size_t)(dx - da));
  else st = OA_ERR_HIP;
  if (st == OA_SUCCESS && hipMemcpyAsync(c, dc, nc * esz, hipMemcpyDeviceToHost, s) != hipSuccess)
    st = OA_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) st = OA_ERR_HIP;
  return st;
}
}  // extern "C++"

void dgemm_(const char *transa, const char *transb, const int *pm, const int *pn, const int *pk,
            const double *palpha, const double *a, const int *plda, const double *b,
            const int *pldb, const double *pbeta, double *c, const int *pldc) {
  const char ta = *transa, tb = *transb;
  const int m = *pm, n = *pn, k = *pk, lda = *plda, ldb = *pldb, ldc = *pldc;
  const double alpha = *palpha, beta = *pbeta;
  if (dgemm_info(ta, tb, m, n, k, lda, ldb, ldc)) return;  // XERBLA is silent (xerbla.f)
  if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return;
  oa_context *h = default_handle();
  if (!h) {
    fprintf(stderr, "[oa_ortho] dgemm_: no usable gfx950 device; C left unchanged\n");
    return;
  }
  const int st = staged_gemm(
      h, sizeof(double), ta, tb, m, n, k, alpha != 0.0 && k != 0, beta == 0.0, a, lda, b, ldb,
      c, ldc, 0,
      [&](const void *da, const void *db, void *dc, char *, size_t) {
        return enqueue_dgemm(h->stream, ta, tb, m, n, k, alpha, (const double *)da, lda,
                             (const double *)db, ldb, beta, (double *)dc, ldc);
      });
  if (st != OA_SUCCESS) fprintf(stderr, "[oa_ortho] dgemm_ failed: %s\n", oa_status_string(st));
}

int oa_zgemm(oa_handle_t c, char ta, char tb, int m, int n, int k, const double *alpha,
             const double *a, int lda, const double *b, int ldb, const double *beta,
             double *cc, int ldc) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (!alpha || !beta) return !alpha ? 6 : 11;
  const int info = dgemm_info(ta, tb, m, n, k, lda, ldb, ldc);  // zgemm.f:175-196, same codes
  if (info) return info;
  const bool a0 = alpha[0] == 0.0 && alpha[1] == 0.0, b1 = beta[0] == 1.0 && beta[1] == 0.0;
  if (m == 0 || n == 0 || ((a0 || k == 0) && b1)) return OA_SUCCESS;
  DeviceGuard g(c->device);
  const int st = oa_ws_require(c, oa_zgemm_ws_bytes(ta, tb, m, n, k));
  if (st) return st;
  return oa_launch_zgemm(c->stream, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, cc, ldc,
                         c->ws.ptr, c->ws.bytes);
}

void zgemm_(const char *transa, const char *transb, const int *pm, const int *pn, const int *pk,
            const double *alpha, const double *a, const int *plda, const double *b,
            const int *pldb, const double *beta, double *c, const int *pldc) {
  const char ta = *transa, tb = *transb;
  const int m = *pm, n = *pn, k = *pk, lda = *plda, ldb = *pldb, ldc = *pldc;
  if (dgemm_info(ta, tb, m, n, k, lda, ldb, ldc)) return;  // XERBLA is silent (xerbla.f)
  const bool a0 = alpha[0] == 0.0 && alpha[1] == 0.0;
  const bool b0 = beta[0] == 0.0 && beta[1] == 0.0, b1 = beta[0] == 1.0 && beta[1] == 0.0;
  if (m == 0 || n == 0 || ((a0 || k == 0) && b1)) return;
  oa_context *h = default_handle();
  if (!h) {
    fprintf(stderr, "[oa_ortho] zgemm_: no usable gfx950 device; C left unchanged\n");
    return;
  }
  const int st = staged_gemm(
      h, 2 * sizeof(double), ta, tb, m, n, k, !a0 && k != 0, b0, a, lda, b, ldb, c, ldc,
      oa_zgemm_ws_bytes(ta, tb, m, n, k),
      [&](const void *da, const void *db, void *dc, char *wx, size_t wbytes) {
        return oa_launch_zgemm(h->stream, ta, tb, m, n, k, alpha, (const double *)da, lda,
                               (const double *)db, ldb, beta, (double *)dc, ldc, wx, wbytes);
      });
  if (st != OA_SUCCESS) fprintf(stderr, "[oa_ortho] zgemm_ failed: %s\n", oa_status_string(st));
}

// ------------------------------------------------------------- path level

int oa_overlap_sym(oa_handle_t c, int n, int npw, const double *psi, long ldpsi, int nkx0,
                   double alpha, int prec, double *S, double diag, double *maxdev) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && npw > 0 && !psi) return 4;
  if (ldpsi < 2L * npw) return 5;
  if (nkx0 < 0 || nkx0 > npw) return 6;
  if (n > 0 && !S) return 9;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  const int K = 2 * npw;
  if ((prec & 0xff) == OA_PREC_FP64) {
    int st = oa_ws_require(c, oa_overlap_f64_ws_bytes(n, K));
    if (st) return st;
    return oa_launch_overlap_f64(c->stream, n, K, psi, ldpsi, 2 * nkx0, alpha, S, diag, maxdev,
                                 c->ws.ptr, c->ws.bytes);
  }
  const size_t need = oa_split_ws_bytes(prec, n, K);
  if (need == 0) return OA_ERR_UNSUPPORTED;
  int st = oa_ws_require(c, need);
  if (st) return st;
  st = oa_launch_overlap_split(c->stream, prec, n, K, psi, ldpsi, 2 * nkx0, alpha, S, diag,
                               maxdev, c->ws.ptr, c->ws.bytes);
  if (st == OA_SUCCESS) tag_digits(c, psi, ldpsi, n, K, prec);
  return st;
}

int oa_slice_counts(oa_handle_t c, int n, int npw, const double *psi, long ldpsi, int nkx0,
                    int prec, int *counts, int capacity, int *nchunks, int *kchunk) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 1) return 2;
  if (npw < 1) return 3;
  if (!psi) return 4;
  if (ldpsi < 2L * npw) return 5;
  if (nkx0 < 0 || nkx0 > npw) return 6;
  if (capacity < 0 || (capacity > 0 && !counts)) return 8;
  DeviceGuard g(c->device);
  const int K = 2 * npw;
  const size_t need = oa_split_ws_bytes(prec, n, K);
  if (need == 0 || !(prec & OA_PREC_ADAPTIVE)) return OA_ERR_UNSUPPORTED;
  int st = oa_ws_require(c, need);
  if (st) return st;
  const int *dcounts = nullptr;
  int nz = 0, kc = 0;
  st = oa_launch_slice_counts(c->stream, prec, n, K, psi, ldpsi, 2 * nkx0, c->ws.ptr,
                              c->ws.bytes, &dcounts, &nz, &kc);
  if (st) return st;
  if (nchunks) *nchunks = nz;
  if (kchunk) *kchunk = kc;
  if (nz > capacity) {
    OA_HIP_CHECK(hipStreamSynchronize(c->stream));
    return OA_ERR_CAPACITY;
  }
  OA_HIP_CHECK(hipMemcpyAsync(counts, dcounts, (size_t)nz * sizeof(int), hipMemcpyDeviceToHost,
                              c->stream));
  OA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return OA_SUCCESS;
}

int oa_lambda(oa_handle_t c, int n, int npw, const double *psi, long ldpsi, const double *F,
              long ldF, double alpha, int prec, double *L) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (ldpsi < 2L * npw) return 5;
  if (ldF < 2L * npw) return 7;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  const int K = 2 * npw;
  if ((prec & 0xff) != OA_PREC_FP64) {
    const size_t need = oa_split_lambda_ws_bytes(prec, n, K);
    if (need == 0) return OA_ERR_UNSUPPORTED;
    const int st = oa_ws_require(c, need);
    if (st) return st;
    return oa_launch_lambda_split(c->stream, prec, n, K, psi, ldpsi, F, ldF, alpha, L,
                                  c->ws.ptr, c->ws.bytes);
  }
  int st = oa_ws_require(c, oa_ntsplit_f64_ws_bytes(n, n, K));
  if (st) return st;
  return oa_launch_ntsplit_f64(c->stream, n, n, K, psi, ldpsi, F, ldF, alpha, L, n, c->ws.ptr,
                               c->ws.bytes);
}

}  // extern "C"

int oa_s_to_t_sync(oa_context *c, int n, const double *S, int nspin, double tol, int maxit,
                   double *T, int *iters, double *err, void *ws, oa_ns_report *rep) {
  int st = oa_ns_begin(c->stream, n, S, nspin, rep, ws);
  if (st) return st;
  // Host loop with early exit: two guarded iterations per report read.
  int it = 0;
  for (;;) {
    const int upto = (it + 2 < maxit) ? it + 2 : maxit;
    for (; it < upto; it++) {
      st = oa_ns_iterate(c->stream, n, it, tol, maxit, rep, ws);
      if (st) return st;
    }
    OA_HIP_CHECK(hipMemcpyAsync(c->host_report, rep, sizeof(oa_ns_report),
                                hipMemcpyDeviceToHost, c->stream));
    OA_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (c->host_report->done || it >= maxit) break;
  }
  st = oa_ns_finish(c->stream, n, T, rep, ws);
  if (st) return st;
  OA_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (iters) *iters = c->host_report->iterations;
  if (err) *err = c->host_report->error;
  return c->host_report->done ? c->host_report->status : OA_ERR_NOT_CONVERGED;
}

extern "C" {

int oa_s_to_t(oa_handle_t c, int n, const double *S, int nspin, double tol, int maxit,
              double *T, int *iters, double *err) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (n > 0 && !S) return 3;
  if (nspin != 1 && nspin != 2) return 4;
  if (maxit < 1) return 6;
  if (n > 0 && !T) return 7;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  int st = oa_ws_require_into(c, c->ns, oa_ns_ws_bytes(n));
  if (st) return st;
  return oa_s_to_t_sync(c, n, S, nspin, tol, maxit, T, iters, err, c->ns.ptr, c->report);
}

int oa_s_to_t_async(oa_handle_t c, int n, const double *S, int nspin, double tol, int maxit,
                    double *T, oa_ns_report *report) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (n > 0 && !S) return 3;
  if (nspin != 1 && nspin != 2) return 4;
  if (maxit < 1) return 6;
  if (n > 0 && !T) return 7;
  if (!report) return 8;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  int st = oa_ws_require_into(c, c->ns, oa_ns_ws_bytes(n));
  if (st) return st;
  return oa_launch_s_to_t(c->stream, n, S, nspin, tol, maxit, T, report, c->ns.ptr);
}

}  // extern "C"

// out = beta out + alpha w .* op(M) psi, w = 1 except columns < cw (alpha_w).
static int backmul_cols(oa_context *c, int n, int npw, const double *M, int transM,
                        const double *psi, long ldpsi, double alpha, double beta, double *out,
                        long ldout, int prec, int cw, double alpha_w) {
  DeviceGuard g(c->device);
  const int K = 2 * npw;
  if ((prec & 0xff) != OA_PREC_FP64) {
    // reuse is a hint: honoured only while the workspace still holds this
    // psi's digits at this precision
    const int chk = oa_split_check(prec, n, 128);
    if (chk) return chk;
    if ((prec & OA_PREC_REUSE_DIGITS) && !digits_match(c, psi, ldpsi, n, K, prec))
      prec &= ~OA_PREC_REUSE_DIGITS;
    const size_t need = oa_split_ws_bytes(prec, n, K);
    if (need == 0) return OA_ERR_UNSUPPORTED;
    int st = oa_ws_require(c, need);
    if (st) return st;
    st = oa_launch_backmul_split(c->stream, prec, n, K, M, transM, psi, ldpsi, 0, alpha, beta,
                                 out, ldout, c->ws.ptr, c->ws.bytes, cw, alpha_w);
    if (st == OA_SUCCESS) tag_digits(c, psi, ldpsi, n, K, prec);
    return st;
  }
  oa_gemm_problem p;
  p.M = n;
  p.N = K;
  p.K = n;
  p.X = M;
  p.ldx = n;
  p.Y = psi;
  p.ldy = ldpsi;
  p.C = out;
  p.ldc = ldout;
  p.alpha = alpha;
  p.beta = beta;
  p.diag_add = 0.0;
  p.cw = cw;
  p.alpha_w = alpha_w;
  return oa_launch_gemm_f64(c->stream, transM != 0, false, &p, 1, nullptr);
}

extern "C" {

int oa_backmul(oa_handle_t c, int n, int npw, const double *M, int transM, const double *psi,
               long ldpsi, double alpha, double beta, double *out, long ldout, int prec) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && !M) return 4;
  if (ldpsi < 2L * npw) return 7;
  if (ldout < 2L * npw) return 11;
  if (n == 0 || npw == 0) return OA_SUCCESS;
  return backmul_cols(c, n, npw, M, transM, psi, ldpsi, alpha, beta, out, ldout, prec, 0, 0.0);
}

int oa_lambda_force(oa_handle_t c, int n, int npw, const double *M, int transM,
                    const double *psi, long ldpsi, int nkx0, double wkx0, double alpha,
                    double beta, double *force, long ldf, int prec) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && !M) return 4;
  if (n > 0 && npw > 0 && !psi) return 6;
  if (ldpsi < 2L * npw) return 7;
  if (nkx0 < 0 || nkx0 > npw) return 8;
  if (n > 0 && npw > 0 && !force) return 12;
  if (ldf < 2L * npw) return 13;
  if (n == 0 || npw == 0) return OA_SUCCESS;
  return backmul_cols(c, n, npw, M, transM, psi, ldpsi, alpha, beta, force, ldf, prec,
                      2 * nkx0, alpha * wkx0);
}

int oa_scale_kx0(oa_handle_t c, int n, int npw, double *X, long ld, int nkx0, double s) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && nkx0 > 0 && !X) return 4;
  if (ld < 2L * npw) return 5;
  if (nkx0 < 0 || nkx0 > npw) return 6;
  if (n == 0 || nkx0 == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  return oa_launch_scale_cols(c->stream, n, 2 * nkx0, X, ld, s);
}

int oa_ortho_step(oa_handle_t c, int n, int npw, const double *psi, long ldpsi, int nkx0,
                  int nspin, double tol, int maxit, int prec, double *S, double *T,
                  double *psi_out, long ldout, int *iters, double *err) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && !psi) return 4;
  if (ldpsi < 2L * npw) return 5;
  if (nkx0 < 0 || nkx0 > npw) return 6;
  if (nspin != 1 && nspin != 2) return 7;
  if (maxit < 1) return 9;
  if (n > 0 && (!S || !T || !psi_out)) return 11;
  if (ldout < 2L * npw) return 14;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  if ((prec & 0xff) != OA_PREC_FP64) {
    const int chk = oa_split_check(prec, n, 128);
    if (chk) return chk;
  }
  int st = oa_reserve_prec(c, n, npw, prec);
  if (st) return st;
  st = oa_overlap_sym(c, n, npw, psi, ldpsi, nkx0, 2.0, prec, S, 0.0, nullptr);
  if (st) return st;
  // the split back-multiply reuses the digits the overlap just produced
  const int bprec = ((prec & 0xff) == OA_PREC_FP64) ? prec : (prec | OA_PREC_REUSE_DIGITS);
  if (iters || err) {
    const int st_ns = oa_s_to_t_sync(c, n, S, nspin, tol, maxit, T, iters, err, c->ns.ptr,
                                     c->report);
    if (st_ns != OA_SUCCESS && st_ns != OA_ERR_NOT_CONVERGED) return st_ns;
    st = oa_backmul(c, n, npw, T, 0, psi, ldpsi, 1.0, 0.0, psi_out, ldout, bprec);
    return st ? st : st_ns;
  }
  st = oa_launch_s_to_t(c->stream, n, S, nspin, tol, maxit, T, c->report, c->ns.ptr);
  if (st) return st;
  return oa_backmul(c, n, npw, T, 0, psi, ldpsi, 1.0, 0.0, psi_out, ldout, bprec);
}

}  // extern "C"
