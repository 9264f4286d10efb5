// The FP64 orthonormalization step for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Every buffer beyond the caller's inputs and outputs
// lives in one file-scope device array of OA_NOALLOC_SCRATCH_BYTES: the S->T
// report, the Newton-Schulz ping-pong buffers and the overlap's split-K
// partial tiles.  Each device holds its own copy of the array until the
// process ends; its address is looked up once per device when the first
// handle on that device is created.  Calls that overlap on one device share
// the array and race.
//
// The step is the one oa_ortho_step runs in FP64 (overlap, S->T, T psi), with
// the overlap's split depth capped to what fits beside the S->T buffers; the
// sums then differ from oa_ortho_step's only in how the k sum is chunked.
#include <mutex>

#include "oa_internal.h"

namespace {

constexpr size_t OA_NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int OA_NOALLOC_MAX_DEVICES = 64;

// Only ever addressed through the per-device pointer taken below: in host
// code the name is a host-side stand-in, not a device address.
static __device__ __align__(256) unsigned char scratch[OA_NOALLOC_SCRATCH_BYTES];

std::mutex scratch_mu;
void *scratch_dev[OA_NOALLOC_MAX_DEVICES] = {};

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// Looks up this device's copy of the scratch array (the device is current).
int oa_noalloc_init(int device) {
  if (device < 0 || device >= OA_NOALLOC_MAX_DEVICES) return OA_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return OA_SUCCESS;
  void *p = nullptr;
  OA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return OA_ERR_HIP;
  scratch_dev[device] = p;
  return OA_SUCCESS;
}

extern "C" {

size_t oa_noalloc_scratch_bytes(void) { return OA_NOALLOC_SCRATCH_BYTES; }

int oa_ortho_step_noalloc(oa_handle_t c, int n, int npw, const double *psi, long ldpsi,
                          int nkx0, int nspin, double tol, int maxit, int prec, double *S,
                          double *T, double *psi_out, long ldout, int *iters, double *err) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (npw < 1) return 3;
  if (n > 0 && !psi) return 4;
  if (ldpsi < 2L * npw) return 5;
  if (nkx0 < 0 || nkx0 > npw) return 6;
  if (nspin != 1 && nspin != 2) return 7;
  if (maxit < 1) return 9;
  if (prec != OA_PREC_FP64) return OA_ERR_UNSUPPORTED;
  if (n > 0 && (!S || !T || !psi_out)) return 11;
  if (ldout < 2L * npw) return 14;
  if (n == 0) return OA_SUCCESS;
  if (c->device < 0 || c->device >= OA_NOALLOC_MAX_DEVICES) return OA_ERR_NO_DEVICE;
  char *base;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    base = static_cast<char *>(scratch_dev[c->device]);
  }
  if (!base) return OA_ERR_HIP;
  // layout: report | S->T buffers | split-K partial tiles
  const size_t off_ns = al256(sizeof(oa_ns_report));
  const size_t off_part = off_ns + al256(oa_ns_ws_bytes(n));
  if (off_part >= OA_NOALLOC_SCRATCH_BYTES) return OA_ERR_CAPACITY;
  oa_ns_report *rep = reinterpret_cast<oa_ns_report *>(base);
  void *ns = base + off_ns;
  void *part = base + off_part;
  const size_t part_bytes = OA_NOALLOC_SCRATCH_BYTES - off_part;
  DeviceGuard g(c->device);
  const int K = 2 * npw;
  int st = oa_launch_overlap_f64_fit(c->stream, n, K, psi, ldpsi, 2 * nkx0, 2.0, S, 0.0,
                                     nullptr, part, part_bytes);
  if (st) return st;
  oa_gemm_problem p{};
  p.M = n;
  p.N = K;
  p.K = n;
  p.X = T;
  p.ldx = n;
  p.Y = psi;
  p.ldy = ldpsi;
  p.C = psi_out;
  p.ldc = ldout;
  p.alpha = 1.0;
  p.beta = 0.0;
  p.diag_add = 0.0;
  if (iters || err) {
    const int st_ns = oa_s_to_t_sync(c, n, S, nspin, tol, maxit, T, iters, err, ns, rep);
    if (st_ns != OA_SUCCESS && st_ns != OA_ERR_NOT_CONVERGED) return st_ns;
    st = oa_launch_gemm_f64(c->stream, false, false, &p, 1, nullptr);
    return st ? st : st_ns;
  }
  st = oa_launch_s_to_t(c->stream, n, S, nspin, tol, maxit, T, rep, ns);
  if (st) return st;
  return oa_launch_gemm_f64(c->stream, false, false, &p, 1, nullptr);
}

}  // extern "C"
