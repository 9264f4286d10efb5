/*
 * oa_oracle.c -- CPU oracle for OpenAtom's orthonormalization path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
 * oracle/liboa_oracle.so and loaded only by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg.  The product library never links it and
 * has no CPU fallback that routes through it.
 *
 * It restates, in plain C, the arithmetic the reference performs on the path:
 *
 *   oao_dgemm            src_mathlib/dgemm.f:161-313 (netlib reference DGEMM:
 *                        argument checks -> XERBLA (empty body, silent return),
 *                        quick returns, and the four loop nests with their exact
 *                        summation order).  Pinned bitwise against the compiled
 *                        reference dgemm_ in oracle/_ref.
 *   oao_pc_forward       src_charm_driver/paircalc/ckPairCalculator.C:846-889
 *                        (DGEMM('T','N', nY, nX, K, alpha, right, K, left, K, 0,
 *                        out, nY); diagonal tiles use right == left, :856-864).
 *   oao_overlap          CP_State_GSpacePlane.C:3025-3028 (psi[kx0] *= 1/sqrt2)
 *                        followed by oao_pc_forward with alpha = 2
 *                        (ckPairCalculator.C:874-875, flag_dp), K split into
 *                        "planes" whose partial tiles are summed in plane order
 *                        (sumMatrixDouble, ckPairCalculator.C:20-53).
 *   oao_cla_multiply     CLA_Matrix.C:528-569 (row-major C = alpha*A*B + beta*C
 *                        realised as transpose(C) ; DGEMM('T','T') ; transpose).
 *   oao_s_to_t           orthog_ctrl/ortho.C:233-253 (B = S/2 or S, A = I),
 *                        :848-929 (C = 3I - A*B ; Bn = 0.5*B*C ; A = 0.5*C*Aold),
 *                        :936-971 (sum of squared T change), :109-146
 *                        (err = sqrt(sum/n^2); continue while err > tol and
 *                        it < maxit; abort status when it >= maxit at exit).
 *   oao_piny_iter_tmat   src_piny_physics_v1.0/abinito_physics/cp_orthog/
 *                        cp_lowdin.C:180-265 (serial Schulz iteration, tol
 *                        1e-15, at most 10 iterations, EXIT on failure).
 *   oao_backmul          ckPairCalculator.C:1835 (symmetric: DGEMM('N','N',
 *                        2np, nY, nX, alpha, psi, 2np, T, nX, beta, out, 2np))
 *                        and :1842 (asymmetric: DGEMM('N','T', ...)).
 *   oao_ortho_step       the whole symmetric step: kx0 scale, forward, S->T,
 *                        backward, kx0 rescale (CP_State_GSpacePlane.C:3278-3281).
 *   oao_lambda           asymmetric forward, alpha = 1 (cpaimd.C:2510-2516,
 *                        ckPairCalculator.C:889): L[x][y] = sum_k psi_x[k] F_y[k].
 *   oao_zgemm            src_mathlib/zgemm.f:160-411 (netlib reference ZGEMM,
 *                        the complex-math PairCalculator's GEMM,
 *                        ckPairCalculator.h:26-48): argument checks, quick
 *                        returns and the nine loop nests of op(A) x op(B) in
 *                        {N, T, C}^2 with their summation order.  Pinned
 *                        bitwise against the compiled reference zgemm_.
 *   oao_zoverlap, oao_zlambda, oao_zs_to_t, oao_zbackmul, oao_zortho_step
 *                        the complex-math (k-point) build
 *                        (CP_PAIRCALC_USES_COMPLEX_MATH): ZGEMM('C','N')
 *                        forward (ckPairCalculator.C:840-889), the Ortho loop
 *                        on complex matrices with CLA products as
 *                        ZGEMM('T','T') and the reference's complex stop rule
 *                        (ortho.C:109-146, 936-971: Re of the complex sum of
 *                        squares), ZGEMM('N','N') / ('N','C') backward
 *                        (:1792-1842).  Every GEMM goes through a zgemm
 *                        pointer (oao_zgemm or the compiled zgemm_).
 *   oao_array_diag_max   Ortho::array_diag_max (ortho.h:282-313) per Ortho
 *                        tile and the max_double reduction into maxCheck
 *                        (ortho.C:255-262, 455-475), real or complex S.
 *   oao_gsp_unpack       the GSpace accumulation of the backward result into
 *                        the force: unpackLambda(partialResultMsg)
 *                        CP_State_GSpacePlane.C:2168-2195, unpackLambda(
 *                        CkReductionMsg) :2068-2106, and for dynamics
 *                        force = -data (:2180-2183) with doLambda's kx = 0
 *                        rescaling (:2240-2245).  Pinned by the constraint it
 *                        enforces: for states with S = 2I (lda) or I (lsda)
 *                        the updated force has zero Lambda with them.
 *
 * Every matrix of states is row-major [n][K] doubles with K = 2*npw (complex
 * coefficients interleaved re,im), exactly the collated layout the reference
 * hands to DGEMM (MessageDataCollator.h:222, ckPairCalculator.C:871-872).
 * n x n matrices are row-major [x][y].
 *
 * A dgemm function pointer may be supplied to every routine that issues GEMMs,
 * so the same call sequence can run on the oracle's own oao_dgemm or on the
 * reference's compiled dgemm_ (oracle/_ref/librefblas.so).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef void (*oao_dgemm_fn)(const char *ta, const char *tb, const int *m,
                             const int *n, const int *k, const double *alpha,
                             const double *A, const int *lda, const double *B,
                             const int *ldb, const double *beta, double *C,
                             const int *ldc);

static int lsame_c(char a, char b) {
  if (a >= 'a' && a <= 'z') a = (char)(a - 'a' + 'A');
  if (b >= 'a' && b <= 'z') b = (char)(b - 'a' + 'A');
  return a == b;
}

#define IDX(M, i, j, ld) (M)[(size_t)(i) + (size_t)(j) * (size_t)(ld)]

/* Restatement of dgemm.f:161-313.  Returns INFO (0 on success); on a bad
 * argument the reference calls XERBLA, whose body is empty, and returns with
 * C untouched -- the return value here only makes that observable to tests. */
int oao_dgemm_info(char ta, char tb, int m, int n, int k, double alpha,
                   const double *A, int lda, const double *B, int ldb,
                   double beta, double *C, int ldc) {
  int nota = lsame_c(ta, 'N');
  int notb = lsame_c(tb, 'N');
  int nrowa = nota ? m : k;
  int nrowb = notb ? k : n;
  int info = 0;
  if (!nota && !lsame_c(ta, 'C') && !lsame_c(ta, 'T')) info = 1;
  else if (!notb && !lsame_c(tb, 'C') && !lsame_c(tb, 'T')) info = 2;
  else if (m < 0) info = 3;
  else if (n < 0) info = 4;
  else if (k < 0) info = 5;
  else if (lda < (nrowa > 1 ? nrowa : 1)) info = 8;
  else if (ldb < (nrowb > 1 ? nrowb : 1)) info = 10;
  else if (ldc < (m > 1 ? m : 1)) info = 13;
  if (info != 0) return info;

  if (m == 0 || n == 0 || ((alpha == 0.0 || k == 0) && beta == 1.0)) return 0;

  if (alpha == 0.0) {
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++)
        IDX(C, i, j, ldc) = (beta == 0.0) ? 0.0 : beta * IDX(C, i, j, ldc);
    return 0;
  }

  if (notb) {
    if (nota) { /* C := alpha*A*B + beta*C  (dgemm.f:232-249) */
      for (int j = 0; j < n; j++) {
        if (beta == 0.0) {
          for (int i = 0; i < m; i++) IDX(C, i, j, ldc) = 0.0;
        } else if (beta != 1.0) {
          for (int i = 0; i < m; i++) IDX(C, i, j, ldc) = beta * IDX(C, i, j, ldc);
        }
        for (int l = 0; l < k; l++) {
          if (IDX(B, l, j, ldb) != 0.0) {
            double temp = alpha * IDX(B, l, j, ldb);
            for (int i = 0; i < m; i++)
              IDX(C, i, j, ldc) = IDX(C, i, j, ldc) + temp * IDX(A, i, l, lda);
          }
        }
      }
    } else { /* C := alpha*A'*B + beta*C  (dgemm.f:254-266) */
      for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
          double temp = 0.0;
          for (int l = 0; l < k; l++) temp = temp + IDX(A, l, i, lda) * IDX(B, l, j, ldb);
          if (beta == 0.0) IDX(C, i, j, ldc) = alpha * temp;
          else IDX(C, i, j, ldc) = alpha * temp + beta * IDX(C, i, j, ldc);
        }
    }
  } else {
    if (nota) { /* C := alpha*A*B' + beta*C  (dgemm.f:271-288) */
      for (int j = 0; j < n; j++) {
        if (beta == 0.0) {
          for (int i = 0; i < m; i++) IDX(C, i, j, ldc) = 0.0;
        } else if (beta != 1.0) {
          for (int i = 0; i < m; i++) IDX(C, i, j, ldc) = beta * IDX(C, i, j, ldc);
        }
        for (int l = 0; l < k; l++) {
          if (IDX(B, j, l, ldb) != 0.0) {
            double temp = alpha * IDX(B, j, l, ldb);
            for (int i = 0; i < m; i++)
              IDX(C, i, j, ldc) = IDX(C, i, j, ldc) + temp * IDX(A, i, l, lda);
          }
        }
      }
    } else { /* C := alpha*A'*B' + beta*C  (dgemm.f:293-305) */
      for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
          double temp = 0.0;
          for (int l = 0; l < k; l++) temp = temp + IDX(A, l, i, lda) * IDX(B, j, l, ldb);
          if (beta == 0.0) IDX(C, i, j, ldc) = alpha * temp;
          else IDX(C, i, j, ldc) = alpha * temp + beta * IDX(C, i, j, ldc);
        }
    }
  }
  return 0;
}

/* Fortran-ABI shaped wrapper, so oao_dgemm and the reference dgemm_ are
 * interchangeable through oao_dgemm_fn. */
void oao_dgemm(const char *ta, const char *tb, const int *m, const int *n,
               const int *k, const double *alpha, const double *A,
               const int *lda, const double *B, const int *ldb,
               const double *beta, double *C, const int *ldc) {
  (void)oao_dgemm_info(*ta, *tb, *m, *n, *k, *alpha, A, *lda, B, *ldb, *beta, C, *ldc);
}

static oao_dgemm_fn pick(oao_dgemm_fn f) { return f ? f : oao_dgemm; }

/* ---- complex (COMPLEX*16 = two doubles, re then im) ------------------- */
typedef struct { double re, im; } oz;
static oz zmul(oz a, oz b) {
  oz r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static oz zadd(oz a, oz b) {
  oz r = {a.re + b.re, a.im + b.im};
  return r;
}
static oz zconj(oz a) {
  oz r = {a.re, -a.im};
  return r;
}
static int zis0(oz a) { return a.re == 0.0 && a.im == 0.0; }
static int zis1(oz a) { return a.re == 1.0 && a.im == 0.0; }
#define ZX(M, i, j, ld) (M)[(size_t)(i) + (size_t)(j) * (size_t)(ld)]

/* Restatement of zgemm.f:160-411; A, B, C are column-major COMPLEX*16. */
int oao_zgemm_info(char ta, char tb, int m, int n, int k, const double *palpha,
                   const double *pA, int lda, const double *pB, int ldb,
                   const double *pbeta, double *pC, int ldc) {
  const oz *A = (const oz *)pA, *B = (const oz *)pB;
  oz *C = (oz *)pC;
  const oz alpha = {palpha[0], palpha[1]}, beta = {pbeta[0], pbeta[1]};
  const oz zero = {0.0, 0.0};
  int nota = lsame_c(ta, 'N'), notb = lsame_c(tb, 'N');
  int conja = lsame_c(ta, 'C'), conjb = lsame_c(tb, 'C');
  int nrowa = nota ? m : k;
  int nrowb = notb ? k : n;
  int info = 0;
  if (!nota && !conja && !lsame_c(ta, 'T')) info = 1;
  else if (!notb && !conjb && !lsame_c(tb, 'T')) info = 2;
  else if (m < 0) info = 3;
  else if (n < 0) info = 4;
  else if (k < 0) info = 5;
  else if (lda < (nrowa > 1 ? nrowa : 1)) info = 8;
  else if (ldb < (nrowb > 1 ? nrowb : 1)) info = 10;
  else if (ldc < (m > 1 ? m : 1)) info = 13;
  if (info != 0) return info;
  if (m == 0 || n == 0 || ((zis0(alpha) || k == 0) && zis1(beta))) return 0;
  if (zis0(alpha)) {
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++)
        ZX(C, i, j, ldc) = zis0(beta) ? zero : zmul(beta, ZX(C, i, j, ldc));
    return 0;
  }
  if (notb && nota) { /* C := alpha*A*B + beta*C  (zgemm.f:232-249) */
    for (int j = 0; j < n; j++) {
      if (zis0(beta)) {
        for (int i = 0; i < m; i++) ZX(C, i, j, ldc) = zero;
      } else if (!zis1(beta)) {
        for (int i = 0; i < m; i++) ZX(C, i, j, ldc) = zmul(beta, ZX(C, i, j, ldc));
      }
      for (int l = 0; l < k; l++) {
        if (!zis0(ZX(B, l, j, ldb))) {
          oz temp = zmul(alpha, ZX(B, l, j, ldb));
          for (int i = 0; i < m; i++)
            ZX(C, i, j, ldc) = zadd(ZX(C, i, j, ldc), zmul(temp, ZX(A, i, l, lda)));
        }
      }
    }
    return 0;
  }
  if (nota) { /* C := alpha*A*op(B) + beta*C, op = conjg(B') or B' (zgemm.f:288-335) */
    for (int j = 0; j < n; j++) {
      if (zis0(beta)) {
        for (int i = 0; i < m; i++) ZX(C, i, j, ldc) = zero;
      } else if (!zis1(beta)) {
        for (int i = 0; i < m; i++) ZX(C, i, j, ldc) = zmul(beta, ZX(C, i, j, ldc));
      }
      for (int l = 0; l < k; l++) {
        if (!zis0(ZX(B, j, l, ldb))) {
          oz bjl = conjb ? zconj(ZX(B, j, l, ldb)) : ZX(B, j, l, ldb);
          oz temp = zmul(alpha, bjl);
          for (int i = 0; i < m; i++)
            ZX(C, i, j, ldc) = zadd(ZX(C, i, j, ldc), zmul(temp, ZX(A, i, l, lda)));
        }
      }
    }
    return 0;
  }
  /* op(A) = A' or conjg(A'): dot products (zgemm.f:250-285, 336-405) */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      oz temp = zero;
      for (int l = 0; l < k; l++) {
        oz a = conja ? zconj(ZX(A, l, i, lda)) : ZX(A, l, i, lda);
        oz bv;
        if (notb) bv = ZX(B, l, j, ldb);
        else bv = conjb ? zconj(ZX(B, j, l, ldb)) : ZX(B, j, l, ldb);
        temp = zadd(temp, zmul(a, bv));
      }
      if (zis0(beta)) ZX(C, i, j, ldc) = zmul(alpha, temp);
      else ZX(C, i, j, ldc) = zadd(zmul(alpha, temp), zmul(beta, ZX(C, i, j, ldc)));
    }
  return 0;
}

/* ckPairCalculator.C:846-889: out (row-major [nX][nY]) = alpha * left * right^T
 * over K doubles per row. */
void oao_pc_forward(int nX, int nY, int K, const double *left,
                    const double *right, double alpha, double *out,
                    oao_dgemm_fn gemm) {
  char tT = 'T', tN = 'N';
  double beta = 0.0;
  int m_in = nY, n_in = nX, k_in = K;
  pick(gemm)(&tT, &tN, &m_in, &n_in, &k_in, &alpha, right, &k_in, left, &k_in,
             &beta, out, &m_in);
}

/* Number of "planes" the K range is split into: plane p covers doubles
 * [p*K/np, (p+1)*K/np), rounded to whole complex coefficients. */
static void plane_range(int K, int nplanes, int p, int *k0, int *k1) {
  int npw = K / 2;
  *k0 = 2 * (int)(((long)npw * p) / nplanes);
  *k1 = 2 * (int)(((long)npw * (p + 1)) / nplanes);
}

static void fwd_planes(int n, int K, const double *psi, double alpha,
                       double *S, int nplanes, oao_dgemm_fn gemm) {
  size_t nn = (size_t)n * n;
  double *part = (double *)calloc(nn * (size_t)nplanes, sizeof(double));
  double *tmp = (double *)malloc((size_t)n * K * sizeof(double));
  /* gather each plane's points into a contiguous [n][Kp] message, as the
   * collator does for one PairCalculator (MessageDataCollator.h:222) */
#pragma omp parallel for schedule(static)
  for (int p = 0; p < nplanes; p++) {
    int k0, k1;
    plane_range(K, nplanes, p, &k0, &k1);
    int Kp = k1 - k0;
    if (Kp == 0) continue;
    double *buf = tmp + (size_t)n * k0;
    for (int i = 0; i < n; i++)
      memcpy(buf + (size_t)i * Kp, psi + (size_t)i * K + k0, (size_t)Kp * sizeof(double));
    oao_pc_forward(n, n, Kp, buf, buf, alpha, part + nn * p, gemm);
  }
  /* sumMatrixDouble (ckPairCalculator.C:20-53), fixed plane order */
  memcpy(S, part, nn * sizeof(double));
  for (int p = 1; p < nplanes; p++)
    for (size_t d = 0; d < nn; d++) S[d] += part[nn * p + d];
  free(tmp);
  free(part);
}

/* Symmetric overlap as the reference computes it: the leading nkx0 complex
 * coefficients (the kx=0 points, stored first by GStateSlab::setKRange,
 * stateSlab.C:285-380) are scaled by 1/sqrt(2) (CP_State_GSpacePlane.C:
 * 3025-3028), then the symmetric PairCalculator multiplies with alpha = 2.
 * For a unique-half layout holding G=0 once (nkx0 = 1) this is
 * S = 2 Re(Psi^H Psi) - Re(psi(0)^* psi(0)). */
void oao_overlap(int n, int npw, const double *psi, int nkx0, double *S,
                 int nplanes, oao_dgemm_fn gemm) {
  int K = 2 * npw;
  double *sc = (double *)malloc((size_t)n * K * sizeof(double));
  memcpy(sc, psi, (size_t)n * K * sizeof(double));
  double rad2i = 1.0 / sqrt(2.0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 2 * nkx0; k++) sc[(size_t)i * K + k] *= rad2i;
  fwd_planes(n, K, sc, 2.0, S, nplanes < 1 ? 1 : nplanes, gemm);
  free(sc);
}

/* Asymmetric (Lambda) forward: alpha = 1, no kx0 scaling in minimisation
 * (CP_State_GSpacePlane.C:2006-2007 sends psi and force unscaled when
 * cp_min_opt == 1).  L[x][y] = sum_k psi_x[k] * F_y[k]. */
void oao_lambda(int n, int npw, const double *psi, const double *F, double *L,
                oao_dgemm_fn gemm) {
  oao_pc_forward(n, n, 2 * npw, psi, F, 1.0, L, gemm);
}

static void transpose_sq(double *a, int m, int n) {
  /* CLA_Matrix.h:233-253 semantics: in-place transpose of an m x n row-major
   * block into n x m */
  double *t = (double *)malloc((size_t)m * n * sizeof(double));
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) t[(size_t)j * m + i] = a[(size_t)i * n + j];
  memcpy(a, t, (size_t)m * n * sizeof(double));
  free(t);
}

/* CLA_Matrix.C:528-569: row-major dest = alpha*A*B + beta*dest, A m x K,
 * B K x n, realised with DGEMM('T','T') on the row-major buffers. */
void oao_cla_multiply(int m, int n, int K, double alpha, const double *A,
                      const double *B, double beta, double *dest,
                      oao_dgemm_fn gemm) {
  char t = 'T';
  if (beta != 0.0) transpose_sq(dest, m, n);
  pick(gemm)(&t, &t, &m, &n, &K, &alpha, A, &K, B, &n, &beta, dest, &m);
  transpose_sq(dest, n, m);
}

/* Ortho S -> T (ortho.C).  Returns 0 on convergence, 1 when the iteration
 * count reached maxit at exit (the reference aborts there, ortho.C:129-133,
 * even when the last iterate met the tolerance).  T receives the last A. */
int oao_s_to_t(int n, const double *S, int nspin, double tol, int maxit,
               double *T, int *iters_out, double *err_out, oao_dgemm_fn gemm) {
  size_t nn = (size_t)n * n;
  double *A = T;
  double *B = (double *)malloc(nn * sizeof(double));
  double *C = (double *)malloc(nn * sizeof(double));
  double *Bn = (double *)malloc(nn * sizeof(double));
  double *Aold = (double *)malloc(nn * sizeof(double));
  for (size_t i = 0; i < nn; i++) B[i] = (nspin == 1) ? S[i] / 2.0 : S[i];
  memset(A, 0, nn * sizeof(double));
  for (int i = 0; i < n; i++) A[(size_t)i * n + i] = 1;
  int iterations = 0;
  double error = 0.0;
  for (;;) {
    /* do_iteration: C = 3I ; C = -1*A*B + 1*C */
    memset(C, 0, nn * sizeof(double));
    for (int i = 0; i < n; i++) C[(size_t)i * n + i] = 3;
    oao_cla_multiply(n, n, n, -1.0, A, B, 1.0, C, gemm);
    /* step_2: Bn = 0.5*B*C */
    oao_cla_multiply(n, n, n, 0.5, B, C, 0.0, Bn, gemm);
    /* step_3: Aold = A ; A = 0.5*C*Aold */
    memcpy(Aold, A, nn * sizeof(double));
    oao_cla_multiply(n, n, n, 0.5, C, Aold, 0.0, A, gemm);
    /* tolerance_check */
    double ret = 0;
    for (size_t i = 0; i < nn; i++) {
      double d = Aold[i] - A[i];
      ret += d * d;
    }
    double *sw = B; B = Bn; Bn = sw;
    iterations++;
    /* collect_error */
    error = sqrt(ret / ((double)n * (double)n));
    if (error > tol && iterations < maxit) continue;
    break;
  }
  free(B); free(C); free(Bn); free(Aold);
  if (iters_out) *iters_out = iterations;
  if (err_out) *err_out = error;
  return iterations >= maxit ? 1 : 0;
}

/* PINY serial Schulz iteration, cp_lowdin.C:180-265 (column-major GENMATMUL
 * calls: itransp = 0 -> 'T', inorm = 1 -> 'N').  Returns 0 on success, 1 when
 * the tolerance 1e-15 is not met within 10 iterations (the reference EXITs). */
int oao_piny_iter_tmat(const double *S, double *Titer, int nstate,
                       int *iters_out, double *tol_out, oao_dgemm_fn gemm) {
  size_t nsq = (size_t)nstate * nstate;
  double *s1 = (double *)malloc(nsq * sizeof(double));
  double *s2 = (double *)malloc(nsq * sizeof(double));
  double *s3 = (double *)malloc(nsq * sizeof(double));
  for (size_t i = 0; i < nsq; i++) s1[i] = S[i] / 2.0;
  memset(Titer, 0, nsq * sizeof(double));
  for (int i = 0; i < nstate; i++) Titer[(size_t)i + (size_t)i * nstate] = 1.0;
  int iter = 0;
  double tol_now = 1.0;
  char N = 'N', T = 'T';
  while (tol_now > 1.0e-15 && iter < 10) {
    iter++;
    double alpha = -1.0, beta = 1.0;
    memset(s2, 0, nsq * sizeof(double));
    for (int i = 0; i < nstate; i++) s2[(size_t)i + (size_t)i * nstate] = 3.0;
    pick(gemm)(&N, &T, &nstate, &nstate, &nstate, &alpha, Titer, &nstate, s1,
               &nstate, &beta, s2, &nstate);
    alpha = 0.5; beta = 0.0;
    memcpy(s3, s1, nsq * sizeof(double));
    pick(gemm)(&N, &T, &nstate, &nstate, &nstate, &alpha, s3, &nstate, s2,
               &nstate, &beta, s1, &nstate);
    memcpy(s3, Titer, nsq * sizeof(double));
    pick(gemm)(&N, &T, &nstate, &nstate, &nstate, &alpha, s2, &nstate, s3,
               &nstate, &beta, Titer, &nstate);
    tol_now = 0.0;
    for (size_t i = 0; i < nsq; i++) {
      double tmp = s3[i] - Titer[i];
      tol_now += tmp * tmp;
    }
    tol_now /= (double)nsq;
    tol_now = sqrt(tol_now);
  }
  free(s1); free(s2); free(s3);
  if (iters_out) *iters_out = iter;
  if (tol_out) *tol_out = tol_now;
  return tol_now > 1.0e-15 ? 1 : 0;
}

/* Backward multiply (ckPairCalculator.C:1835 / :1842) over K doubles:
 * transT == 0: out[y] = alpha * sum_x T[y][x] psi[x] + beta * out[y]
 *   (DGEMM('N','N', K, n, n, alpha, psi, K, T, n, beta, out, K) with T
 *   row-major, i.e. column-major T^T)
 * transT != 0: out[y] = alpha * sum_x T[x][y] psi[x] + beta * out[y]
 *   (DGEMM('N','T', ...), the asymmetric Lambda back-multiply). */
void oao_backmul(int n, int K, const double *T, int transT, const double *psi,
                 double alpha, double beta, double *out, int nplanes,
                 oao_dgemm_fn gemm) {
  char tN = 'N', tT = 'T';
  if (nplanes < 1) nplanes = 1;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < nplanes; p++) {
    int k0, k1;
    plane_range(K, nplanes, p, &k0, &k1);
    int Kp = k1 - k0;
    if (Kp == 0) continue;
    double *pin = (double *)malloc((size_t)n * Kp * sizeof(double));
    double *pout = (double *)malloc((size_t)n * Kp * sizeof(double));
    for (int i = 0; i < n; i++) {
      memcpy(pin + (size_t)i * Kp, psi + (size_t)i * K + k0, (size_t)Kp * sizeof(double));
      memcpy(pout + (size_t)i * Kp, out + (size_t)i * K + k0, (size_t)Kp * sizeof(double));
    }
    pick(gemm)(&tN, transT ? &tT : &tN, &Kp, &n, &n, &alpha, pin, &Kp, T, &n,
               &beta, pout, &Kp);
    for (int i = 0; i < n; i++)
      memcpy(out + (size_t)i * K + k0, pout + (size_t)i * Kp, (size_t)Kp * sizeof(double));
    free(pin);
    free(pout);
  }
}

/* One symmetric orthonormalisation step as the reference runs it at its
 * default decomposition (one state tile, K split across planes):
 *   GSP: psi[kx0] *= 1/sqrt2              CP_State_GSpacePlane.C:3025-3028
 *   PC:  S = 2 * psi psi^T (per plane)    ckPairCalculator.C:889
 *   Ortho: S -> T                         ortho.C:153-971
 *   PC:  out = T psi (per plane)          ckPairCalculator.C:1835
 *   GSP: out[kx0] *= sqrt2                CP_State_GSpacePlane.C:3278-3281
 * S, T and psi_out may be NULL-free outputs of n*n, n*n and n*2npw doubles.
 * Returns the S->T status (0 ok, 1 iteration limit). */
int oao_ortho_step(int n, int npw, const double *psi, int nkx0, int nspin,
                   double tol, int maxit, double *S, double *T,
                   double *psi_out, int *iters, double *err, int nplanes,
                   oao_dgemm_fn gemm) {
  int K = 2 * npw;
  double *sc = (double *)malloc((size_t)n * K * sizeof(double));
  memcpy(sc, psi, (size_t)n * K * sizeof(double));
  double rad2i = 1.0 / sqrt(2.0), rad2 = sqrt(2.0);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 2 * nkx0; k++) sc[(size_t)i * K + k] *= rad2i;
  fwd_planes(n, K, sc, 2.0, S, nplanes < 1 ? 1 : nplanes, gemm);
  int st = oao_s_to_t(n, S, nspin, tol, maxit, T, iters, err, gemm);
  memset(psi_out, 0, (size_t)n * K * sizeof(double));
  oao_backmul(n, K, T, 0, sc, 1.0, 0.0, psi_out, nplanes, gemm);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 2 * nkx0; k++) psi_out[(size_t)i * K + k] *= rad2;
  free(sc);
  return st;
}

/* GSpace accumulation of a PairCalculator backward result `data` ([n][2npw])
 * into `force` (same layout).  The kx = 0 points are the leading nkx0
 * complex points of every row.
 *   mode 0, minimisation, partialResultMsg (CP_State_GSpacePlane.C:
 *     2168-2195): double pack: force -= wght*data, wght = 0.5 on kx = 0
 *     points, 1 elsewhere; single pack: force -= 0.5*data.
 *   mode 1, minimisation, CkReductionMsg (:2068-2106): overlap = 2 (lda) or
 *     1 (lsda); double pack: wght = 1/overlap on kx = 0 points, 2/overlap
 *     elsewhere; single pack: wght = 1/overlap.
 *   mode 2, dynamics: force = -data (:2180-2183, one contribution), then
 *     doLambda multiplies the kx = 0 points by 1/sqrt(2) (:2240-2245;
 *     double pack only: the single-pack layout has no kx = 0 block). */
void oao_gsp_unpack(int n, int npw, const double *data, int nkx0, int double_pack,
                    int mode, int lsda, double *force) {
  const size_t K = 2 * (size_t)npw;
  const double overlap = lsda ? 1.0 : 2.0;
  const double rad2i = 1.0 / sqrt(2.0);
  for (int r = 0; r < n; r++) {
    const double *d = data + r * K;
    double *f = force + r * K;
    for (int k = 0; k < npw; k++) {
      const int kx0 = double_pack && k < nkx0;
      if (mode == 2) {
        for (int c = 0; c < 2; c++) {
          f[2 * k + c] = d[2 * k + c] * (-1.0);
          if (kx0) f[2 * k + c] *= rad2i;
        }
        continue;
      }
      double wght;
      if (mode == 0)
        wght = double_pack ? (kx0 ? 0.5 : 1.0) : 0.5;
      else
        wght = double_pack ? (kx0 ? 1.0 / overlap : 2.0 / overlap) : 1.0 / overlap;
      f[2 * k] -= wght * d[2 * k];
      f[2 * k + 1] -= wght * d[2 * k + 1];
    }
  }
}

/* ---- complex-math (k-point) build ------------------------------------- */

typedef void (*oao_zgemm_fn)(const char *ta, const char *tb, const int *m, const int *n,
                             const int *k, const double *alpha, const double *A,
                             const int *lda, const double *B, const int *ldb,
                             const double *beta, double *C, const int *ldc);

/* Fortran-ABI form of the ZGEMM restatement (INFO discarded, as XERBLA is
 * silent). */
void oao_zgemm(const char *ta, const char *tb, const int *m, const int *n, const int *k,
               const double *alpha, const double *A, const int *lda, const double *B,
               const int *ldb, const double *beta, double *C, const int *ldc) {
  (void)oao_zgemm_info(*ta, *tb, *m, *n, *k, alpha, A, *lda, B, *ldb, beta, C, *ldc);
}

static oao_zgemm_fn zpick(oao_zgemm_fn f) { return f ? f : oao_zgemm; }

/* ckPairCalculator.C:840-889 with internalType complex: myGEMM promotes alpha
 * to (alpha, 0) (:56-60) and calls ZGEMM('C','N', nY, nX, npw, alpha, right,
 * npw, left, npw, 0, out, nY); out row-major [x][y] = alpha * sum_k
 * left_x[k] conj(right_y[k]). */
void oao_zpc_forward(int nX, int nY, int npw, const double *left, const double *right,
                     double alpha, double *out, oao_zgemm_fn zg) {
  char tC = 'C', tN = 'N';
  double al[2] = {alpha, 0.0}, be[2] = {0.0, 0.0};
  int m_in = nY, n_in = nX, k_in = npw;
  zpick(zg)(&tC, &tN, &m_in, &n_in, &k_in, al, right, &k_in, left, &k_in, be, out, &m_in);
}

/* Symmetric complex forward (right = left, alpha = 1: no double pack). */
void oao_zoverlap(int n, int npw, const double *psi, double *S, oao_zgemm_fn zg) {
  oao_zpc_forward(n, n, npw, psi, psi, 1.0, S, zg);
}

/* Asymmetric complex forward, L = psi F^H. */
void oao_zlambda(int n, int npw, const double *psi, const double *F, double *L,
                 oao_zgemm_fn zg) {
  oao_zpc_forward(n, n, npw, psi, F, 1.0, L, zg);
}

/* Plain (unconjugated) transpose of a row-major complex m x n block. */
static void ztranspose(double *a, int m, int n) {
  double *t = (double *)malloc((size_t)2 * m * n * sizeof(double));
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) {
      t[2 * ((size_t)j * m + i)] = a[2 * ((size_t)i * n + j)];
      t[2 * ((size_t)j * m + i) + 1] = a[2 * ((size_t)i * n + j) + 1];
    }
  memcpy(a, t, (size_t)2 * m * n * sizeof(double));
  free(t);
}

/* CLA_Matrix.C:528-569 with complex data: row-major dest = alpha*A*B +
 * beta*dest through ZGEMM('T','T') and transposes, alpha and beta real. */
static void zcla_multiply(int m, int n, int K, double alpha, const double *A, const double *B,
                          double beta, double *dest, oao_zgemm_fn zg) {
  char t = 'T';
  double al[2] = {alpha, 0.0}, be[2] = {beta, 0.0};
  if (beta != 0.0) ztranspose(dest, m, n);
  zpick(zg)(&t, &t, &m, &n, &K, al, A, &K, B, &n, be, dest, &m);
  ztranspose(dest, n, m);
}

/* Ortho S -> T on complex matrices: the loop of oao_s_to_t with complex
 * arithmetic.  tolerance_check (ortho.C:950-953) accumulates tmp * tmp in
 * complex arithmetic, the sum_double reduction adds both doubles and
 * collect_error (:113) reads the first: err = sqrt(Re(sum) / n^2), NaN when
 * Re(sum) < 0, which stops the loop (`error > tol` is false).  modulus != 0
 * accumulates |tmp|^2 instead.  Returns 0, or 1 when the count reached maxit. */
int oao_zs_to_t(int n, const double *S, int nspin, double tol, int maxit, int modulus,
                double *T, int *iters_out, double *err_out, oao_zgemm_fn zg) {
  const size_t nn2 = (size_t)2 * n * n;
  double *A = T;
  double *B = (double *)malloc(nn2 * sizeof(double));
  double *C = (double *)malloc(nn2 * sizeof(double));
  double *Bn = (double *)malloc(nn2 * sizeof(double));
  double *Aold = (double *)malloc(nn2 * sizeof(double));
  for (size_t i = 0; i < nn2; i++) B[i] = (nspin == 1) ? S[i] / 2.0 : S[i];
  memset(A, 0, nn2 * sizeof(double));
  for (int i = 0; i < n; i++) A[2 * ((size_t)i * n + i)] = 1;
  int iterations = 0;
  double error = 0.0;
  for (;;) {
    memset(C, 0, nn2 * sizeof(double));
    for (int i = 0; i < n; i++) C[2 * ((size_t)i * n + i)] = 3;
    zcla_multiply(n, n, n, -1.0, A, B, 1.0, C, zg);
    zcla_multiply(n, n, n, 0.5, B, C, 0.0, Bn, zg);
    memcpy(Aold, A, nn2 * sizeof(double));
    zcla_multiply(n, n, n, 0.5, C, Aold, 0.0, A, zg);
    double ret_re = 0.0, ret_im = 0.0;
    for (size_t i = 0; i < nn2 / 2; i++) {
      const double dr = Aold[2 * i] - A[2 * i], di = Aold[2 * i + 1] - A[2 * i + 1];
      if (modulus) {
        ret_re += dr * dr + di * di;
      } else {
        /* ret += tmp * tmp, complex */
        const double sq_re = dr * dr - di * di, sq_im = dr * di + di * dr;
        ret_re += sq_re;
        ret_im += sq_im;
      }
    }
    (void)ret_im;
    double *sw = B; B = Bn; Bn = sw;
    iterations++;
    error = sqrt(ret_re / ((double)n * (double)n));
    if (error > tol && iterations < maxit) continue;
    break;
  }
  free(B); free(C); free(Bn); free(Aold);
  if (iters_out) *iters_out = iterations;
  if (err_out) *err_out = error;
  return iterations >= maxit ? 1 : 0;
}

/* Complex backward (ckPairCalculator.C:1792-1842): symmetric ZGEMM('N','N',
 * npw, n, n, alpha, psi, npw, M, n, beta, out, npw), out_y = alpha * sum_x
 * M[y][x] psi_x + beta out_y; asymmetric ZGEMM('N','C'), out_y = alpha * sum_x
 * conj(M[x][y]) psi_x + beta out_y.  M row-major complex n x n. */
void oao_zbackmul(int n, int npw, const double *M, int transM, const double *psi, double alpha,
                  double beta, double *out, oao_zgemm_fn zg) {
  char tN = 'N', tC = 'C';
  double al[2] = {alpha, 0.0}, be[2] = {beta, 0.0};
  zpick(zg)(&tN, transM ? &tC : &tN, &npw, &n, &n, al, psi, &npw, M, &n, be, out, &npw);
}

/* One complex step: S = psi psi^H, S -> T (reference stop rule unless
 * modulus), out = T psi. */
int oao_zortho_step(int n, int npw, const double *psi, int nspin, double tol, int maxit,
                    int modulus, double *S, double *T, double *psi_out, int *iters,
                    double *err, oao_zgemm_fn zg) {
  oao_zoverlap(n, npw, psi, S, zg);
  int st = oao_zs_to_t(n, S, nspin, tol, maxit, modulus, T, iters, err, zg);
  memset(psi_out, 0, (size_t)2 * n * npw * sizeof(double));
  oao_zbackmul(n, npw, T, 0, psi, 1.0, 0.0, psi_out, zg);
  return st;
}

/* Ortho::array_diag_max (orthog_ctrl/ortho.h:282-313) over the og x og
 * Ortho tiles of a row-major n x n S, combined by the max_double reduction
 * that feeds Ortho::maxCheck (ortho.C:255-262, 455-475).  Off-diagonal tiles:
 * max |S_ij| from the tile's first element on.  Diagonal tiles: the first
 * value is |S_00 - d| and every later one is |S_ij|, with ||S_ii| - d| on the
 * diagonal; d = 1 under lsda, else 2.  complex != 0 reads S as row-major
 * complex (re, im) pairs and takes myabs = std::abs, the modulus. */
static double dmodulus(const double *S, size_t idx, int cplx) {
  return cplx ? hypot(S[2 * idx], S[2 * idx + 1]) : fabs(S[idx]);
}

double oao_array_diag_max(int n, const double *S, int og, int lsda, int cplx) {
  const double diag = lsda ? 1.0 : 2.0;
  double result = 0.0;
  int first = 1;
  if (og < 1) og = n;
  for (int x0 = 0; x0 < n; x0 += og)
    for (int y0 = 0; y0 < n; y0 += og) {
      const int m = (x0 + og <= n) ? og : n - x0, w = (y0 + og <= n) ? og : n - y0;
      const size_t i00 = (size_t)x0 * n + y0;
      double mx;
      if (x0 != y0) {
        mx = dmodulus(S, i00, cplx);
        for (int i = 0; i < m; i++)
          for (int j = 0; j < w; j++) {
            const double a = dmodulus(S, (size_t)(x0 + i) * n + y0 + j, cplx);
            mx = (mx > a) ? mx : a;
          }
      } else {
        mx = cplx ? hypot(S[2 * i00] - diag, S[2 * i00 + 1]) : fabs(S[i00] - diag);
        for (int i = 0; i < m; i++)
          for (int j = 0; j < w; j++) {
            double a = dmodulus(S, (size_t)(x0 + i) * n + y0 + j, cplx);
            if (i == j) a = fabs(a - diag);
            mx = (mx > a) ? mx : a;
          }
      }
      if (first || mx > result) result = mx;
      first = 0;
    }
  return result;
}

/* max |S(psi) - occ*I| with S in the 2Re - G0 metric of oao_overlap. */
double oao_orthonormality_residual(int n, int npw, const double *psi, int nkx0,
                                   double occ, int nplanes, oao_dgemm_fn gemm) {
  double *S = (double *)malloc((size_t)n * n * sizeof(double));
  oao_overlap(n, npw, psi, nkx0, S, nplanes, gemm);
  double r = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double v = fabs(S[(size_t)i * n + j] - (i == j ? occ : 0.0));
      if (v > r) r = v;
    }
  free(S);
  return r;
}

int oao_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oao_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}
