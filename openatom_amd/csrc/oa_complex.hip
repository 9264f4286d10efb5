// Complex GEMM (ZGEMM semantics, column-major COMPLEX*16) on the FP64-MFMA
// kernel: the complex-math PairCalculator's GEMM (ckPairCalculator.h:26-48,
// zgemm.f).
//
//   C := alpha * op(A) * op(B) + beta * C,   op in {N, T, C}.
//
// The operands are split once into real and imaginary planes (column-major,
// conjugation folded into the sign of the imaginary plane), the four real
// products Ar.Br, Ai.Bi, Ar.Bi, Ai.Br run as two launches of two problems on
// gemm_f64_kernel, and one pass combines them with the complex alpha and
// beta: C = alpha (P1 - P2 + i (P3 + P4)) + beta C.
#include "oa_internal.h"

namespace {

typedef double2 z_t;

__global__ void zsplit_kernel(const z_t *__restrict__ X, int rows, int cols, long ld, int conj,
                              double *__restrict__ re, double *__restrict__ im) {
  const long total = (long)rows * cols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const int r = (int)(i % rows);
    const long c = i / rows;
    const z_t v = X[r + c * ld];
    re[i] = v.x;
    im[i] = conj ? -v.y : v.y;
  }
}

// P holds P1..P4 (m x n each, column-major, ld m).
__global__ void zcombine_kernel(int m, int n, const double *__restrict__ P, double ar, double ai,
                                double br, double bi, z_t *C, long ldc) {
  const long mn = (long)m * n;
  const bool beta0 = (br == 0.0 && bi == 0.0);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < mn;
       i += (long)gridDim.x * blockDim.x) {
    const int r = (int)(i % m);
    const long c = i / m;
    const double pr = P[i] - P[mn + i];
    const double pi = P[2 * mn + i] + P[3 * mn + i];
    z_t out;
    out.x = ar * pr - ai * pi;
    out.y = ar * pi + ai * pr;
    z_t *dst = C + r + c * ldc;
    if (!beta0) {
      const z_t cv = *dst;
      out.x += br * cv.x - bi * cv.y;
      out.y += br * cv.y + bi * cv.x;
    }
    *dst = out;
  }
}

// alpha == 0: C = beta C (zero when beta == 0), as zgemm.f:208-223.
__global__ void zscale_kernel(int m, int n, double br, double bi, z_t *C, long ldc) {
  const long mn = (long)m * n;
  const bool beta0 = (br == 0.0 && bi == 0.0);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < mn;
       i += (long)gridDim.x * blockDim.x) {
    z_t *dst = C + (i % m) + (i / m) * ldc;
    z_t out = {0.0, 0.0};
    if (!beta0) {
      const z_t cv = *dst;
      out.x = br * cv.x - bi * cv.y;
      out.y = br * cv.y + bi * cv.x;
    }
    *dst = out;
  }
}

bool up(char c) { return c == 'N' || c == 'n'; }
bool isconj(char c) { return c == 'C' || c == 'c'; }

unsigned blocks_for(long n) {
  const long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

}  // namespace

size_t oa_zgemm_ws_bytes(char ta, char tb, int m, int n, int k) {
  const bool nota = up(ta), notb = up(tb);
  const long ra = nota ? m : k, ca = nota ? k : m;
  const long rb = notb ? k : n, cb = notb ? n : k;
  const size_t d = 2 * (size_t)ra * ca + 2 * (size_t)rb * cb + 4 * (size_t)m * n;
  return d * sizeof(double) + 1024;
}

int oa_launch_zgemm(hipStream_t s, char ta, char tb, int m, int n, int k, const double *alpha,
                    const double *a, int lda, const double *b, int ldb, const double *beta,
                    double *c, int ldc, void *ws, size_t ws_bytes) {
  z_t *C = reinterpret_cast<z_t *>(c);
  const double ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
  if (m == 0 || n == 0) return OA_SUCCESS;
  if ((ar == 0.0 && ai == 0.0) || k == 0) {
    if (ar == 0.0 && ai == 0.0) {
      hipLaunchKernelGGL(zscale_kernel, dim3(blocks_for((long)m * n)), dim3(256), 0, s, m, n, br,
                         bi, C, (long)ldc);
      return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
    }
  }
  if (oa_zgemm_ws_bytes(ta, tb, m, n, k) > ws_bytes) return OA_ERR_CAPACITY;
  const bool nota = up(ta), notb = up(tb);
  const int ra = nota ? m : k, ca = nota ? k : m;
  const int rb = notb ? k : n, cb = notb ? n : k;
  double *Ar = reinterpret_cast<double *>(ws);
  double *Ai = Ar + (size_t)ra * ca;
  double *Br = Ai + (size_t)ra * ca;
  double *Bi = Br + (size_t)rb * cb;
  double *P = Bi + (size_t)rb * cb;
  const size_t mn = (size_t)m * n;
  if (k > 0) {
    hipLaunchKernelGGL(zsplit_kernel, dim3(blocks_for((long)ra * ca)), dim3(256), 0, s,
                       reinterpret_cast<const z_t *>(a), ra, ca, (long)lda, isconj(ta) ? 1 : 0,
                       Ar, Ai);
    hipLaunchKernelGGL(zsplit_kernel, dim3(blocks_for((long)rb * cb)), dim3(256), 0, s,
                       reinterpret_cast<const z_t *>(b), rb, cb, (long)ldb, isconj(tb) ? 1 : 0,
                       Br, Bi);
  }
  // column-major P = op(Xa) op(Xb)  <=>  row-major P^T[n][m] = op(Xb)^T op(Xa)^T
  auto prob = [&](const double *xa, const double *xb, double *out) {
    return oa_gemm_problem{n, m, k, xb, rb, xa, ra, out, m, 1.0, 0.0, 0.0};
  };
  const oa_gemm_problem p1[2] = {prob(Ar, Br, P), prob(Ar, Bi, P + 2 * mn)};
  const oa_gemm_problem p2[2] = {prob(Ai, Bi, P + mn), prob(Ai, Br, P + 3 * mn)};
  int st = oa_launch_gemm_f64(s, !notb, !nota, p1, 2, nullptr);
  if (st) return st;
  st = oa_launch_gemm_f64(s, !notb, !nota, p2, 2, nullptr);
  if (st) return st;
  hipLaunchKernelGGL(zcombine_kernel, dim3(blocks_for((long)mn)), dim3(256), 0, s, m, n,
                     (const double *)P, ar, ai, br, bi, C, (long)ldc);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}
