// Complex-math (k-point) form of the orthonormalization path.
//
// CP_PAIRCALC_USES_COMPLEX_MATH (pcFwdDeclarations.h:17-24) makes the
// PairCalculator's internal type complex: the forward is ZGEMM('C','N')
// (ckPairCalculator.C:840-889), Ortho iterates on complex matrices (ortho.C
// with internalType complex, CLA_Matrix products through ZGEMM('T','T')) and
// the backward is ZGEMM('N','N') / ('N','C') (:1792-1842).  The states carry
// the full G set of a k-point: no double packing, no kx = 0 weighting.
//
// Every complex product runs as real products on the FP64-MFMA kernels:
//   * a block psi [n][npw complex] is read as the real [n][2npw] matrix of
//     interleaved (re, im) doubles, and J psi = i psi = (-im, re) per point
//     is formed once, so that
//       Re sum_k psi_x conj(phi_y) = psi_x . phi_y          (row dot product)
//       Im sum_k psi_x conj(phi_y) = psi_x . (J phi)_y
//       M psi = Re(M) psi + Im(M) (J psi);
//   * n x n matrices are kept as separate real and imaginary planes, and a
//     complex product is four real products in two launches of two problems.
// With a split precision the forward and backward products run on the digit
// kernels instead (Re S on the symmetric overlap, Im S and Lambda on the
// asymmetric product, T psi as two back-multiplies); S -> T stays FP64.
#include <math.h>

#include <vector>

#include "oa_internal.h"

namespace {

constexpr int NT = 256;
constexpr int ZDIFF_BLOCKS = 128;

unsigned blocks_for(long elems) {
  const long b = (elems + NT - 1) / NT;
  return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// out[r][2k] = -psi[r][2k+1], out[r][2k+1] = psi[r][2k]  (J psi = i psi)
__global__ void jswap_kernel(int n, int npw, const double *__restrict__ psi, long ld,
                             double *__restrict__ out, long ldo) {
  const long total = (long)n * npw;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long r = i / npw, k = i % npw;
    const double re = psi[r * ld + 2 * k], im = psi[r * ld + 2 * k + 1];
    out[r * ldo + 2 * k] = -im;
    out[r * ldo + 2 * k + 1] = re;
  }
}

// Interleaved complex Z[nn] -> planes scale*re, scale*im.
__global__ void zplanes_kernel(long nn, const double *__restrict__ Z, double scale,
                               double *__restrict__ re, double *__restrict__ im) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
       i += (long)gridDim.x * blockDim.x) {
    re[i] = scale * Z[2 * i];
    im[i] = scale * Z[2 * i + 1];
  }
}

// Planes -> interleaved complex Z[nn].
__global__ void zmerge_kernel(long nn, const double *__restrict__ re,
                              const double *__restrict__ im, double *__restrict__ Z) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
       i += (long)gridDim.x * blockDim.x) {
    Z[2 * i] = re[i];
    Z[2 * i + 1] = im[i];
  }
}

// Hermitian S from the symmetric real part Sr and the imaginary part Si of
// the upper triangle: S[x][y] = Sr[x][y] + i Si[x][y] (x < y), its conjugate
// mirror below, a real diagonal.
// With md != null it also takes Ortho::array_diag_max of the merged S
// (ortho.h:282-313 with myabs = std::abs): |S_ij| off the diagonal and
// ||S_ii| - diag| on it, maximised through atomicMax on the IEEE bits of
// non-negative doubles (md cleared beforehand).
__global__ void zherm_kernel(int n, const double *__restrict__ Sr, const double *__restrict__ Si,
                             double *__restrict__ S, double diag, unsigned long long *md) {
  const long nn = (long)n * n;
  double mx = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
       i += (long)gridDim.x * blockDim.x) {
    const long x = i / n, y = i % n;
    const double re = Sr[i], im = (x < y) ? Si[i] : ((x > y) ? -Si[y * n + x] : 0.0);
    S[2 * i] = re;
    S[2 * i + 1] = im;
    const double a = hypot(re, im);
    mx = fmax(mx, (x == y) ? fabs(a - diag) : a);
  }
  if (md) {
    // wave maximum, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) atomicMax(md, (unsigned long long)__double_as_longlong(mx));
  }
}

// re = d I, im = 0 (n x n planes).
__global__ void zeye_kernel(int n, double d, double *__restrict__ re, double *__restrict__ im) {
  const long nn = (long)n * n;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
       i += (long)gridDim.x * blockDim.x) {
    re[i] = (i / n == i % n) ? d : 0.0;
    im[i] = 0.0;
  }
}

// Per-block sums of (ar - br)^2 and (ai - bi)^2 over nn elements, in a fixed
// order for a fixed grid: part[2b] and part[2b + 1].
__global__ __launch_bounds__(NT) void zdiff_kernel(long nn, const double *__restrict__ ar,
                                                   const double *__restrict__ ai,
                                                   const double *__restrict__ br,
                                                   const double *__restrict__ bi,
                                                   double *__restrict__ part) {
  __shared__ double sr[NT], si[NT];
  double r = 0.0, m = 0.0;
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < nn; i += (long)gridDim.x * NT) {
    const double dr = ar[i] - br[i], di = ai[i] - bi[i];
    r += dr * dr;
    m += di * di;
  }
  sr[threadIdx.x] = r;
  si[threadIdx.x] = m;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr[0];
    part[2 * blockIdx.x + 1] = si[0];
  }
}

#define OA_KLAUNCH(kern, elems, s, ...)                                                     \
  do {                                                                                      \
    hipLaunchKernelGGL(kern, dim3(blocks_for(elems)), dim3(NT), 0, s, __VA_ARGS__);        \
    if (hipPeekAtLastError() != hipSuccess) return OA_ERR_HIP;                              \
  } while (0)

// Two n x n x n problems of one launch: C0 = a0 X0 Y0 + b0 C0 + d0 I and
// C1 = a1 X1 Y1 + b1 C1 (row-major, leading dimension n).
int gemm2(hipStream_t s, int n, const double *X0, const double *Y0, double *C0, double a0,
          double b0, double d0, const double *X1, const double *Y1, double *C1, double a1,
          double b1) {
  const oa_gemm_problem p[2] = {{n, n, n, X0, n, Y0, n, C0, n, a0, b0, d0},
                                {n, n, n, X1, n, Y1, n, C1, n, a1, b1, 0.0}};
  return oa_launch_gemm_f64(s, false, false, p, 2, nullptr);
}

// C = alpha (A B) + beta C + d I on complex planes (re, im):
//   Cr = alpha (Ar Br - Ai Bi) + beta Cr + d I,  Ci = alpha (Ar Bi + Ai Br) + beta Ci.
int zmul_planes(hipStream_t s, int n, double alpha, const double *Ar, const double *Ai,
                const double *Br, const double *Bi, double beta, double d, double *Cr,
                double *Ci) {
  int st = gemm2(s, n, Ar, Br, Cr, alpha, beta, d, Ar, Bi, Ci, alpha, beta);
  if (st) return st;
  return gemm2(s, n, Ai, Bi, Cr, -alpha, 1.0, 0.0, Ai, Br, Ci, alpha, 1.0);
}

// Workspace of the complex forward and backward: J psi ([n][K]) where the
// products read it, two n x n planes, then the kernels' own workspace; every
// part starts on a 256-byte boundary.  J psi is formed in FP64 by the FP64
// products and by the general split forward (which slices it); the split
// Hermitian forward and the split back-multiply make J psi's digits from
// psi's and need no J region, so a split step lays out no J.
bool is_split(int prec) { return (prec & 0xff) != OA_PREC_FP64; }
size_t al256(size_t b) { return (b + 255) / 256 * 256; }
bool needs_j(bool sym_or_back, int prec) { return !(sym_or_back && is_split(prec)); }
size_t off_planes(int n, int K, bool jn) { return jn ? al256((size_t)n * K * sizeof(double)) : 0; }
size_t plane_bytes(int n) { return al256((size_t)n * n * sizeof(double)); }
size_t off_kws(int n, int K, bool jn) { return off_planes(n, K, jn) + 2 * plane_bytes(n); }

size_t zoverlap_ws_bytes(int n, int K, int prec, bool sym) {
  size_t kws = oa_overlap_f64_ws_bytes(n, K);
  if (oa_ntsplit_f64_ws_bytes(n, n, K) > kws) kws = oa_ntsplit_f64_ws_bytes(n, n, K);
  if (is_split(prec)) {
    if (oa_split_ws_bytes(prec, n, K) > kws) kws = oa_split_ws_bytes(prec, n, K);
    if (oa_split_lambda_ws_bytes(prec, n, K) > kws) kws = oa_split_lambda_ws_bytes(prec, n, K);
    if (oa_split_zimag_ws_bytes(prec, n, K) > kws) kws = oa_split_zimag_ws_bytes(prec, n, K);
  }
  return off_kws(n, K, needs_j(sym, prec)) + kws;
}

size_t zbackmul_ws_bytes(int n, int K, int prec) {
  return off_kws(n, K, needs_j(true, prec)) + (is_split(prec) ? oa_split_ws_bytes(prec, n, K) : 0);
}

// Hermitian (sym) or general (!sym) product L[x][y] = alpha sum_k psi_x[k]
// conj(phi_y[k]); `J` receives J phi ([n][K], leading dimension K), the
// real and imaginary planes and the kernel workspace follow it in ws.
int zforward(hipStream_t s, bool sym, int n, int npw, const double *psi, long ld,
             const double *phi, long ldphi, double alpha, double *L, char *ws, size_t ws_bytes,
             int prec, double diag = 0.0, double *maxdev = nullptr) {
  const int K = 2 * npw;
  const bool jn = needs_j(sym, prec);
  double *J = jn ? reinterpret_cast<double *>(ws) : nullptr;
  double *Lr = reinterpret_cast<double *>(ws + off_planes(n, K, jn));
  double *Li = reinterpret_cast<double *>(ws + off_planes(n, K, jn) + plane_bytes(n));
  char *kws = ws + off_kws(n, K, jn);
  const size_t kbytes = ws_bytes - off_kws(n, K, jn);
  // the split Hermitian forward takes J psi's digits from psi's; every other
  // form multiplies J phi in FP64 or slices it
  if (!(sym && is_split(prec)))
    OA_KLAUNCH(jswap_kernel, (long)n * npw, s, n, npw, phi, ldphi, J, (long)K);
  int st;
  if (is_split(prec)) {
    if (sym) {
      // psi is sliced once: Re S on its digits, Im S on them and on J psi's
      // made from them; psi's digits (both layouts) stay in the kernel
      // workspace, where a following back-multiply can reuse them
      st = oa_launch_overlap_split(s, prec, n, K, psi, ld, 0, alpha, Lr, 0.0, nullptr, kws,
                                   kbytes);
      if (st) return st;
      st = oa_launch_zimag_split(s, prec, n, K, alpha, Li, kws, kbytes);
    } else {
      st = oa_launch_lambda_split(s, prec, n, K, psi, ld, J, K, alpha, Li, kws, kbytes);
      if (st) return st;
      st = oa_launch_lambda_split(s, prec, n, K, psi, ld, phi, ldphi, alpha, Lr, kws, kbytes);
    }
  } else {
    if (sym) {
      st = oa_launch_overlap_f64(s, n, K, psi, ld, 0, alpha, Lr, 0.0, nullptr, kws, kbytes);
      if (st) return st;
      // the Hermitian merge reads only Im S's strict upper triangle
      st = oa_launch_overlap_xy_f64(s, n, K, psi, ld, J, K, alpha, Li, kws, kbytes);
    } else {
      st = oa_launch_ntsplit_f64(s, n, n, K, psi, ld, phi, ldphi, alpha, Lr, n, kws, kbytes);
      if (st) return st;
      st = oa_launch_ntsplit_f64(s, n, n, K, psi, ld, J, K, alpha, Li, n, kws, kbytes);
    }
  }
  if (st) return st;
  if (sym) {
    auto *md = reinterpret_cast<unsigned long long *>(maxdev);
    if (md && oa_fill_u64(s, md, 1)) return OA_ERR_HIP;
    OA_KLAUNCH(zherm_kernel, (long)n * n, s, n, Lr, Li, L, diag, md);
  } else
    OA_KLAUNCH(zmerge_kernel, (long)n * n, s, (long)n * n, Lr, Li, L);
  return OA_SUCCESS;
}

// out = alpha op(M) psi + beta out, op(M) = M or M^H; `J` holds J psi
// already when have_j; with reuse_psi the kernel workspace holds psi's
// digits from a symmetric split forward on the same psi (no re-slicing).
int zbackward(hipStream_t s, int n, int npw, const double *M, int transM, const double *psi,
              long ld, double alpha, double beta, double *out, long ldo, char *ws,
              size_t ws_bytes, bool have_j, int prec, bool reuse_psi = false) {
  const int K = 2 * npw;
  const bool jn = needs_j(true, prec);
  double *J = jn ? reinterpret_cast<double *>(ws) : nullptr;
  double *Mr = reinterpret_cast<double *>(ws + off_planes(n, K, jn));
  double *Mi = reinterpret_cast<double *>(ws + off_planes(n, K, jn) + plane_bytes(n));
  // split modes make J psi's digits from psi's and never read J psi
  if (!have_j && !is_split(prec))
    OA_KLAUNCH(jswap_kernel, (long)n * npw, s, n, npw, psi, ld, J, (long)K);
  OA_KLAUNCH(zplanes_kernel, (long)n * n, s, (long)n * n, M, 1.0, Mr, Mi);
  // Re(M^H) = Mr^T and Im(M^H) = -Mi^T
  if (is_split(prec)) {
    char *kws = ws + off_kws(n, K, jn);
    const size_t kbytes = ws_bytes - off_kws(n, K, jn);
    const int p = prec & ~OA_PREC_REUSE_DIGITS;
    int st = oa_launch_backmul_split(s, reuse_psi ? (p | OA_PREC_REUSE_DIGITS) : p, n, K, Mr,
                                     transM, psi, ld, 0, alpha, beta, out, ldo, kws, kbytes);
    if (st) return st;
    // J psi's digits from psi's (left by the first product), in place,
    // instead of slicing J psi; the second product reuses them
    st = oa_launch_jswap_db(s, p, n, K, kws, kbytes);
    if (st) return st;
    // psi's pointer and pitch stand in for J psi's: with the digits reused
    // the product reads neither
    return oa_launch_backmul_split(s, p | OA_PREC_REUSE_DIGITS, n, K, Mi, transM, psi, ld, 0,
                                   transM ? -alpha : alpha, 1.0, out, ldo, kws, kbytes);
  }
  oa_gemm_problem p{n, K, n, Mr, n, psi, ld, out, ldo, alpha, beta, 0.0};
  int st = oa_launch_gemm_f64(s, transM != 0, false, &p, 1, nullptr);
  if (st) return st;
  oa_gemm_problem q{n, K, n, Mi, n, J, K, out, ldo, transM ? -alpha : alpha, 1.0, 0.0};
  return oa_launch_gemm_f64(s, transM != 0, false, &q, 1, nullptr);
}

// Ortho's loop on complex planes, with the stop rule of the reference's
// complex tolerance_check (ortho.C:950-953 accumulates tmp * tmp as a complex
// square; the sum_double reduction and collect_error, :113, keep its real
// part): err = sqrt(Re sum (dA)^2 / n^2), NaN when that sum is negative, which
// ends the loop because `err > tol` is then false.  modulus != 0 uses
// sum |dA|^2 instead.
int zs_to_t_run(oa_context *c, int n, const double *S, int nspin, double tol, int maxit,
                int modulus, double *T, int *iters, double *err) {
  hipStream_t s = c->stream;
  const size_t nn = (size_t)n * n;
  const size_t np = plane_bytes(n) / sizeof(double);  // plane stride, 256-byte aligned
  double *P = reinterpret_cast<double *>(c->ns.ptr);
  double *Br = P, *Bi = P + np, *Nr = P + 2 * np, *Ni = P + 3 * np, *Cr = P + 4 * np,
         *Ci = P + 5 * np, *Ar = P + 6 * np, *Ai = P + 7 * np, *Or = P + 8 * np,
         *Oi = P + 9 * np, *part = P + 10 * np;
  OA_KLAUNCH(zplanes_kernel, (long)nn, s, (long)nn, S, nspin == 1 ? 0.5 : 1.0, Br, Bi);
  OA_KLAUNCH(zeye_kernel, (long)nn, s, n, 1.0, Ar, Ai);
  std::vector<double> hp(2 * ZDIFF_BLOCKS);
  int it = 0;
  double error = 0.0;
  for (;;) {
    // do_iteration: C = 3I - A B
    int st = zmul_planes(s, n, -1.0, Ar, Ai, Br, Bi, 0.0, 3.0, Cr, Ci);
    // step_2: Bn = 0.5 B C
    if (!st) st = zmul_planes(s, n, 0.5, Br, Bi, Cr, Ci, 0.0, 0.0, Nr, Ni);
    if (st) return st;
    // step_3: Aold = A, A = 0.5 C Aold
    double *t = Or; Or = Ar; Ar = t;
    t = Oi; Oi = Ai; Ai = t;
    st = zmul_planes(s, n, 0.5, Cr, Ci, Or, Oi, 0.0, 0.0, Ar, Ai);
    if (st) return st;
    hipLaunchKernelGGL(zdiff_kernel, dim3(ZDIFF_BLOCKS), dim3(NT), 0, s, (long)nn, Or, Oi, Ar,
                       Ai, part);
    if (hipPeekAtLastError() != hipSuccess) return OA_ERR_HIP;
    OA_HIP_CHECK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    OA_HIP_CHECK(hipStreamSynchronize(s));
    t = Br; Br = Nr; Nr = t;
    t = Bi; Bi = Ni; Ni = t;
    it++;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < ZDIFF_BLOCKS; b++) {
      sr += hp[2 * b];
      si += hp[2 * b + 1];
    }
    error = sqrt((modulus ? sr + si : sr - si) / ((double)n * (double)n));
    if (error > tol && it < maxit) continue;
    break;
  }
  OA_KLAUNCH(zmerge_kernel, (long)nn, s, (long)nn, Ar, Ai, T);
  OA_HIP_CHECK(hipStreamSynchronize(s));
  if (iters) *iters = it;
  if (err) *err = error;
  return it >= maxit ? OA_ERR_NOT_CONVERGED : OA_SUCCESS;
}

size_t zs_to_t_ws_bytes(int n) {
  return 10 * plane_bytes(n) + 2 * ZDIFF_BLOCKS * sizeof(double);
}

}  // namespace

extern "C" {

int oa_zoverlap(oa_handle_t c, int n, int npw, const double *psi, long ldpsi, double alpha,
                double *S, int prec, double diag, double *maxdev) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (prec & OA_PREC_ADAPTIVE) return OA_ERR_UNSUPPORTED;  // uniform slices only
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && npw > 0 && !psi) return 4;
  if (ldpsi < 2L * npw) return 5;
  if (n > 0 && !S) return 7;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  if (npw == 0) {
    OA_HIP_CHECK(hipMemsetAsync(S, 0, 2 * (size_t)n * n * sizeof(double), c->stream));
    if (maxdev) {
      // S = 0: array_diag_max is |0 - diag| on the diagonal
      const double d = fabs(diag);
      OA_HIP_CHECK(hipMemcpyAsync(maxdev, &d, sizeof(double), hipMemcpyHostToDevice, c->stream));
      OA_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    return OA_SUCCESS;
  }
  const int st = oa_ws_require(c, zoverlap_ws_bytes(n, 2 * npw, prec, true));
  if (st) return st;
  return zforward(c->stream, true, n, npw, psi, ldpsi, psi, ldpsi, alpha, S, (char *)c->ws.ptr,
                  c->ws.bytes, prec, diag, maxdev);
}

int oa_zlambda(oa_handle_t c, int n, int npw, const double *psi, long ldpsi, const double *F,
               long ldF, double alpha, double *L, int prec) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (prec & OA_PREC_ADAPTIVE) return OA_ERR_UNSUPPORTED;  // uniform slices only
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && npw > 0 && !psi) return 4;
  if (ldpsi < 2L * npw) return 5;
  if (n > 0 && npw > 0 && !F) return 6;
  if (ldF < 2L * npw) return 7;
  if (n > 0 && !L) return 9;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  if (npw == 0) {
    OA_HIP_CHECK(hipMemsetAsync(L, 0, 2 * (size_t)n * n * sizeof(double), c->stream));
    return OA_SUCCESS;
  }
  const int st = oa_ws_require(c, zoverlap_ws_bytes(n, 2 * npw, prec, false));
  if (st) return st;
  return zforward(c->stream, false, n, npw, psi, ldpsi, F, ldF, alpha, L, (char *)c->ws.ptr,
                  c->ws.bytes, prec);
}

int oa_zs_to_t(oa_handle_t c, int n, const double *S, int nspin, double tol, int maxit,
               int modulus, double *T, int *iters, double *err) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (n > 0 && !S) return 3;
  if (nspin != 1 && nspin != 2) return 4;
  if (maxit < 1) return 6;
  if (n > 0 && !T) return 8;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  const int st = oa_ws_require_into(c, c->ns, zs_to_t_ws_bytes(n));
  if (st) return st;
  return zs_to_t_run(c, n, S, nspin, tol, maxit, modulus, T, iters, err);
}

int oa_zbackmul(oa_handle_t c, int n, int npw, const double *M, int transM, const double *psi,
                long ldpsi, double alpha, double beta, double *out, long ldout, int prec) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (prec & OA_PREC_ADAPTIVE) return OA_ERR_UNSUPPORTED;  // uniform slices only
  if (n < 0) return 2;
  if (npw < 0) return 3;
  if (n > 0 && !M) return 4;
  if (n > 0 && npw > 0 && !psi) return 6;
  if (ldpsi < 2L * npw) return 7;
  if (n > 0 && npw > 0 && !out) return 10;
  if (ldout < 2L * npw) return 11;
  if (n == 0 || npw == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  const int st = oa_ws_require(c, zbackmul_ws_bytes(n, 2 * npw, prec));
  if (st) return st;
  return zbackward(c->stream, n, npw, M, transM, psi, ldpsi, alpha, beta, out, ldout,
                   (char *)c->ws.ptr, c->ws.bytes, false, prec);
}

int oa_zortho_step(oa_handle_t c, int n, int npw, const double *psi, long ldpsi, int nspin,
                   double tol, int maxit, int modulus, double *S, double *T, double *psi_out,
                   long ldout, int *iters, double *err, int prec) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (prec & OA_PREC_ADAPTIVE) return OA_ERR_UNSUPPORTED;  // uniform slices only
  if (n < 0) return 2;
  if (npw < 1) return 3;
  if (n > 0 && !psi) return 4;
  if (ldpsi < 2L * npw) return 5;
  if (nspin != 1 && nspin != 2) return 6;
  if (maxit < 1) return 8;
  if (n > 0 && (!S || !T || !psi_out)) return 10;
  if (ldout < 2L * npw) return 13;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  const int K = 2 * npw;
  size_t need = zoverlap_ws_bytes(n, K, prec, true);
  if (zbackmul_ws_bytes(n, K, prec) > need) need = zbackmul_ws_bytes(n, K, prec);
  int st = oa_ws_require(c, need);
  if (!st) st = oa_ws_require_into(c, c->ns, zs_to_t_ws_bytes(n));
  if (st) return st;
  char *ws = (char *)c->ws.ptr;
  st = zforward(c->stream, true, n, npw, psi, ldpsi, psi, ldpsi, 1.0, S, ws, c->ws.bytes, prec);
  if (st) return st;
  const int st_ns = zs_to_t_run(c, n, S, nspin, tol, maxit, modulus, T, iters, err);
  if (st_ns != OA_SUCCESS && st_ns != OA_ERR_NOT_CONVERGED) return st_ns;
  // the forward left J psi at the start of the workspace (FP64) or psi's
  // digits in the kernel workspace (split modes; S->T has buffers of its
  // own), from which the back-multiply makes J psi's: J psi itself is not
  // read there
  st = zbackward(c->stream, n, npw, T, 0, psi, ldpsi, 1.0, 0.0, psi_out, ldout, ws, c->ws.bytes,
                 true, prec, is_split(prec));
  return st ? st : st_ns;
}

}  // extern "C"
