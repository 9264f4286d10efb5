/* Host AddressSanitizer / UBSan driver for oracle/oa_oracle.c.
 *
 * Built by tests/test_host_asan.py with gcc -fsanitize=address,undefined and
 * run on the CPU.  Every entry point of the oracle is called on small odd
 * sizes (ragged planes, all transpose forms, complex and real, nspin 1 and 2,
 * non-convergence) so that an out-of-bounds access in a loop nest, a plane
 * split or a scratch buffer is reported by the sanitizer.  The values are
 * not checked here (the parity tests do that); the exit status is 0 unless a
 * call misbehaves visibly. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef void (*gemm_fn)(const char *, const char *, const int *, const int *, const int *,
                        const double *, const double *, const int *, const double *,
                        const int *, const double *, double *, const int *);

int oao_dgemm_info(char ta, char tb, int m, int n, int k, double alpha, const double *A,
                   int lda, const double *B, int ldb, double beta, double *C, int ldc);
int oao_zgemm_info(char ta, char tb, int m, int n, int k, const double *alpha, const double *A,
                   int lda, const double *B, int ldb, const double *beta, double *C, int ldc);
void oao_overlap(int n, int npw, const double *psi, int nkx0, double *S, int nplanes,
                 gemm_fn gemm);
void oao_lambda(int n, int npw, const double *psi, const double *F, double *L, gemm_fn gemm);
void oao_cla_multiply(int m, int n, int K, double alpha, const double *A, const double *B,
                      double beta, double *dest, gemm_fn gemm);
int oao_s_to_t(int n, const double *S, int nspin, double tol, int maxit, double *T,
               int *iters, double *err, gemm_fn gemm);
int oao_piny_iter_tmat(const double *S, double *Titer, int nstate, int *iters, double *err,
                       gemm_fn gemm);
void oao_backmul(int n, int K, const double *T, int transT, const double *psi, double alpha,
                 double beta, double *out, int nplanes, gemm_fn gemm);
int oao_ortho_step(int n, int npw, const double *psi, int nkx0, int nspin, double tol,
                   int maxit, double *S, double *T, double *psi_out, int *iters, double *err,
                   int nplanes, gemm_fn gemm);
void oao_gsp_unpack(int n, int npw, const double *data, int nkx0, int double_pack, int mode,
                    int lsda, double *force);
void oao_zoverlap(int n, int npw, const double *psi, double *S, gemm_fn zg);
void oao_zlambda(int n, int npw, const double *psi, const double *F, double *L, gemm_fn zg);
int oao_zs_to_t(int n, const double *S, int nspin, double tol, int maxit, int modulus,
                double *T, int *iters, double *err, gemm_fn zg);
void oao_zbackmul(int n, int npw, const double *M, int transM, const double *psi, double alpha,
                  double beta, double *out, gemm_fn zg);
int oao_zortho_step(int n, int npw, const double *psi, int nspin, double tol, int maxit,
                    int modulus, double *S, double *T, double *psi_out, int *iters, double *err,
                    gemm_fn zg);
double oao_array_diag_max(int n, const double *S, int og, int lsda, int cplx);
double oao_orthonormality_residual(int n, int npw, const double *psi, int nkx0, double occ,
                                   int nplanes, gemm_fn gemm);

static unsigned long long rs = 88172645463325252ull;
static double rnd(void) {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return (double)(rs >> 11) / 9007199254740992.0 - 0.5;
}

/* exact-size heap arrays, so a read one past the end is reported */
static double *arr(size_t n) {
  double *p = (double *)malloc((n ? n : 1) * sizeof(double));
  for (size_t i = 0; i < n; i++) p[i] = rnd();
  return p;
}

/* rows of a near-orthonormal block: e_i scaled plus a small perturbation, in
 * [n][K] doubles */
static double *near_ortho(int n, int K, double eps) {
  double *p = arr((size_t)n * K);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < K; k++) p[(size_t)i * K + k] *= eps;
  for (int i = 0; i < n && 2 * i + 2 < K; i++) p[(size_t)i * K + 2 + 2 * i] += sqrt(0.5);
  return p;
}

int main(void) {
  int bad = 0;
  /* every transpose form of both GEMMs on odd shapes with padded leading
   * dimensions, plus the quick-return and argument-error cases */
  const char tr[3] = {'N', 'T', 'C'};
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const int m = 7, n = 5, k = 9;
      const int nota = tr[a] == 'N', notb = tr[b] == 'N';
      const int lda = (nota ? m : k) + 2, ldb = (notb ? k : n) + 1, ldc = m + 3;
      const int acols = nota ? k : m, bcols = notb ? n : k;
      double *A = arr((size_t)lda * acols * 2), *B = arr((size_t)ldb * bcols * 2);
      double *C = arr((size_t)ldc * n * 2);
      const double al[2] = {1.5, -0.5}, be[2] = {0.25, 0.5}, z0[2] = {0.0, 0.0};
      if (a < 2 && b < 2) {
        bad |= oao_dgemm_info(tr[a], tr[b], m, n, k, 1.5, A, lda, B, ldb, 0.25, C, ldc) != 0;
        bad |= oao_dgemm_info(tr[a], tr[b], m, n, k, 0.0, A, lda, B, ldb, 0.0, C, ldc) != 0;
        bad |= oao_dgemm_info(tr[a], tr[b], m, n, 0, 1.0, A, lda, B, ldb, 1.0, C, ldc) != 0;
      }
      bad |= oao_zgemm_info(tr[a], tr[b], m, n, k, al, A, lda, B, ldb, be, C, ldc) != 0;
      bad |= oao_zgemm_info(tr[a], tr[b], m, n, k, al, A, lda, B, ldb, z0, C, ldc) != 0;
      free(A), free(B), free(C);
    }
  {
    double x = 0.0;
    bad |= oao_dgemm_info('X', 'N', 1, 1, 1, 1.0, &x, 1, &x, 1, 0.0, &x, 1) != 1;
    bad |= oao_dgemm_info('N', 'N', 2, 1, 1, 1.0, &x, 1, &x, 1, 0.0, &x, 2) != 8;
  }

  /* the gamma-point path: ragged plane splits, kx0 weights, both spins */
  for (int nspin = 1; nspin <= 2; nspin++) {
    const int n = 13, npw = 37, K = 2 * npw, nkx0 = 3;
    double *psi = near_ortho(n, K, 1e-3);
    double *S = arr((size_t)n * n), *T = arr((size_t)n * n), *out = arr((size_t)n * K);
    double *F = arr((size_t)n * K), *L = arr((size_t)n * n);
    int it = 0;
    double err = 0.0;
    for (int np = 1; np <= 5; np += 2) oao_overlap(n, npw, psi, nkx0, S, np, NULL);
    oao_lambda(n, npw, psi, F, L, NULL);
    oao_cla_multiply(n, n, n, 1.0, S, L, 0.5, T, NULL);
    oao_cla_multiply(n, n - 2, n, 1.0, S, L, 0.0, T, NULL);
    const double occ = nspin == 1 ? 2.0 : 1.0;
    double *Sn = arr((size_t)n * n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        Sn[(size_t)i * n + j] = (i == j ? occ : 0.0) + 1e-3 * (S[(size_t)i * n + j] +
                                                               S[(size_t)j * n + i]);
    bad |= oao_s_to_t(n, Sn, nspin, 1e-13, 50, T, &it, &err, NULL) != 0;
    bad |= oao_s_to_t(n, Sn, nspin, 0.0, 2, T, &it, &err, NULL) != 1; /* abort at maxit */
    (void)oao_piny_iter_tmat(Sn, T, n, &it, &err, NULL);
    for (int tt = 0; tt < 2; tt++)
      for (int np = 1; np <= 4; np += 3) oao_backmul(n, K, T, tt, psi, 1.0, 0.5, out, np, NULL);
    (void)oao_ortho_step(n, npw, psi, nkx0, nspin, 1e-12, 50, S, T, out, &it, &err, 3, NULL);
    for (int mode = 0; mode < 3; mode++)
      for (int dp = 0; dp < 2; dp++) oao_gsp_unpack(n, npw, out, nkx0, dp, mode, nspin == 2, F);
    (void)oao_orthonormality_residual(n, npw, out, nkx0, 1.0, 2, NULL);
    for (int og = 0; og <= 5; og += 5) (void)oao_array_diag_max(n, Sn, og, nspin == 2, 0);
    free(psi), free(S), free(T), free(out), free(F), free(L), free(Sn);
  }

  /* the complex-math (k-point) path */
  {
    const int n = 11, npw = 29;
    double *psi = near_ortho(n, 2 * npw, 1e-3);
    double *S = arr((size_t)2 * n * n), *T = arr((size_t)2 * n * n);
    double *out = arr((size_t)2 * n * npw), *F = arr((size_t)2 * n * npw);
    double *L = arr((size_t)2 * n * n);
    int it = 0;
    double err = 0.0;
    oao_zoverlap(n, npw, psi, S, NULL);
    oao_zlambda(n, npw, psi, F, L, NULL);
    for (int i = 0; i < n; i++) S[2 * ((size_t)i * n + i)] += 0.5;
    for (int mod = 0; mod < 2; mod++) (void)oao_zs_to_t(n, S, 2, 1e-12, 60, mod, T, &it, &err, NULL);
    for (int tm = 0; tm < 2; tm++) oao_zbackmul(n, npw, T, tm, psi, 1.0, 0.5, out, NULL);
    (void)oao_zortho_step(n, npw, psi, 2, 1e-12, 60, 0, S, T, out, &it, &err, NULL);
    for (int og = 0; og <= 4; og += 4) (void)oao_array_diag_max(n, S, og, 1, 1);
    free(psi), free(S), free(T), free(out), free(F), free(L);
  }
  if (bad) fprintf(stderr, "oracle_driver: an entry point returned an unexpected status\n");
  else printf("oracle_driver: ok\n");
  return bad;
}
