// Split-precision (Ozaki digit slice) products of the orthonormalization path.
//
// Representation.  The K plane-wave doubles are cut into chunks of kchunk
// columns (a multiple of 32).  Every value x of chunk z is written as
//     x = 2^(e_z - 7) * sum_{p=1..S} a_p R^(-(p-1))  + r,   R = 254,
// |a_p| <= 127, |r| <= 2^(e_z - 8) R^(-(S-1)), with one power-of-two exponent
// e_z per chunk, shared by all rows.  Each digit is the rounded remainder:
// |u - rint(u)| <= 1/2, so the next u = R (u - rint(u)) stays within 127 and
// every digit uses the full signed 8-bit range (log2 254 = 7.99 bits per
// slice).  The digits are exact integers; the recurrence rounds only in the
// product by R, 2^-53 relative to the remainder, far below 2^-40.
//
// Products.  sum_k a_p[i,k] b_q[j,k] over one chunk is accumulated exactly:
// int8 digits on v_mfma_i32_32x32x32_i8 into int32 (bounded by the chunk
// length, which the host caps), fp16 / bf16 digits on
// v_mfma_f32_32x32x16_{f16,bf16} into fp32 promoted to FP64 every 128 k
// (sums stay below 2^24 and are therefore exact).  Products are grouped by
// level L = p + q <= S + 1 (the Ozaki triangular truncation); level sums are
// combined in FP64 with weights R^(-(L-2)) and the chunk scales.
//
// Layout.  Every digit matrix (rows R, contraction C) is stored k-blocked:
// element (p, r, c) at p*R*C + ((c / B)*R + r)*B + c % B with B = 32 bytes of
// digits, so one pipeline stage of a workgroup (128 rows x 32 bytes) is one
// contiguous 4 KB block per slice.
//
// Kernels:
//   chunk_max_kernel   max |psi| per chunk (two variants: without / with the
//                      leading kx0 columns), order-free via atomicMax on the
//                      IEEE bits of non-negative doubles.
//   slice_psi_kernel   psi -> digits in two layouts: DS[p][row][k] (rows
//                      contiguous in k, for S = psi psi^T; kx0 columns zeroed,
//                      their term is added in FP64) and DB[p][k][row] (rows
//                      contiguous in the state index, for T psi), the second
//                      through an LDS transpose.
//   slice_rows_kernel  an n x n matrix -> digits with one exponent per row.
//   digit_nt_kernel    C[m][n] = sum_k X[m][k] Y[n][k] over digit slices:
//                      MODE 0 = split-K partial tiles of S (upper triangle),
//                      MODE 1 = back-multiply epilogue out = a*T*psi + b*out.
//   split_reduce_kernel  fixed-order sum of the partial tiles, FP64 G=0 term,
//                      mirror, max|S - d I|.
#include <math.h>

#include "oa_internal.h"

namespace {

typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

constexpr int DNT = 512;  // threads per digit-GEMM workgroup (8 waves, 4 x 2)
constexpr int DBM = 128;  // workgroup tile
constexpr int DBK = 32;   // chunk granularity (elements); a stage holds 32 bytes per row
constexpr int SMAX = 6;
constexpr long KCHUNK_TARGET = 8192;  // columns per scale block / split-K chunk

__host__ __device__ inline int cdv(int a, int b) { return (a + b - 1) / b; }

// Digit element type: int8 or 16-bit float.  ES = bytes per element.
enum DType { DT_I8 = 0, DT_F16 = 1, DT_BF16 = 2 };

template <int DT>
struct DTraits;
template <>
struct DTraits<DT_I8> {
  typedef int8_t elem;
  static constexpr int ES = 1;
  static constexpr int KI = 32;  // contraction per MFMA
  typedef i16v acc_t;
};
template <>
struct DTraits<DT_F16> {
  typedef _Float16 elem;
  static constexpr int ES = 2;
  static constexpr int KI = 16;
  typedef f16v acc_t;
};
template <>
struct DTraits<DT_BF16> {
  typedef __bf16 elem;
  static constexpr int ES = 2;
  static constexpr int KI = 16;
  typedef f16v acc_t;
};

template <int DT>
__device__ __forceinline__ typename DTraits<DT>::elem to_digit(double a) {
  if constexpr (DT == DT_I8) return (int8_t)(int)a;
  else if constexpr (DT == DT_F16) return (_Float16)(float)a;
  else return (__bf16)(float)a;
}

template <int DT>
__device__ __forceinline__ typename DTraits<DT>::acc_t mfma_digit(i4v a, i4v b,
                                                                  typename DTraits<DT>::acc_t c) {
  if constexpr (DT == DT_I8) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
  } else if constexpr (DT == DT_F16) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, a),
                                                  __builtin_bit_cast(h8v, b), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8v, a),
                                                   __builtin_bit_cast(b8v, b), c, 0, 0, 0);
  }
}

// Chunk exponent e from the maximum magnitude m: |x| * 2^(7-e) <= 127 for
// every |x| <= m.  max == 0 gives e = 0 (all digits zero).
__device__ __forceinline__ int chunk_exp(double m) {
  if (!(m > 0.0)) return 0;
  int E;
  const double f = frexp(m, &E);  // m = f 2^E, f in [0.5, 1)
  return (f <= 127.0 / 128.0) ? E : E + 1;
}

// Pack 16 bytes of digits (16 int8 or 8 16-bit floats starting at d[0]).
template <int DT>
__device__ __forceinline__ i4v pack_digits(const double *d) {
  i4v r;
  if constexpr (DT == DT_I8) {
#pragma unroll
    for (int w = 0; w < 4; w++) {
      unsigned u = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) u |= ((unsigned)(int)d[4 * w + b] & 0xffu) << (8 * b);
      r[w] = (int)u;
    }
  } else {
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const typename DTraits<DT>::elem lo = to_digit<DT>(d[2 * w]), hi = to_digit<DT>(d[2 * w + 1]);
      const unsigned u = (unsigned)__builtin_bit_cast(unsigned short, lo) |
                         ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
      r[w] = (int)u;
    }
  }
  return r;
}

// Digit base R and the level weights R^-l (R^l is exact in FP64 for l <= 12,
// so each weight is the correctly rounded reciprocal).
constexpr double RADIX = 254.0;
__host__ __device__ constexpr double rinv_c(int l) {
  double p = 1.0;
  for (int i = 0; i < l; i++) p *= RADIX;
  return 1.0 / p;
}
// Called with indices that are constants after unrolling: the table folds.
__host__ __device__ __forceinline__ double rinv(int l) {
  constexpr double t[2 * SMAX] = {rinv_c(0), rinv_c(1), rinv_c(2), rinv_c(3),
                                  rinv_c(4), rinv_c(5), rinv_c(6), rinv_c(7),
                                  rinv_c(8), rinv_c(9), rinv_c(10), rinv_c(11)};
  return t[l];
}

// Split one value into S digits given its exponent.
template <int S>
__device__ __forceinline__ void split_value(double x, int e, double (&d)[S]) {
  double u = (e == 0) ? x * 128.0 : ldexp(x, 7 - e);
#pragma unroll
  for (int p = 0; p < S; p++) {
    const double a = rint(u);
    d[p] = a;
    u = (u - a) * RADIX;
  }
}

__global__ void fill_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// grid (nz, row groups of 8): max over rows x chunk columns; maxS skips the
// leading kx0 columns, maxB includes them.
__global__ __launch_bounds__(256) void chunk_max_kernel(const double *__restrict__ psi, long ld,
                                                        int n, int K, int kx0_end, int kchunk,
                                                        unsigned long long *maxS,
                                                        unsigned long long *maxB) {
  const int z = blockIdx.x;
  const int k0 = z * kchunk, k1 = min(K, k0 + kchunk);
  double ms = 0.0, mb = 0.0;
  const int r0 = blockIdx.y * 8;
  for (int r = r0; r < min(n, r0 + 8); r++) {
    const double *row = psi + (long)r * ld;
    for (int k = k0 + threadIdx.x; k < k1; k += 256) {
      const double v = fabs(row[k]);
      mb = fmax(mb, v);
      if (k >= kx0_end) ms = fmax(ms, v);
    }
  }
  __shared__ double rs[256], rb[256];
  rs[threadIdx.x] = ms;
  rb[threadIdx.x] = mb;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      rs[threadIdx.x] = fmax(rs[threadIdx.x], rs[threadIdx.x + o]);
      rb[threadIdx.x] = fmax(rb[threadIdx.x], rb[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMax(maxS + z, (unsigned long long)__double_as_longlong(rs[0]));
    atomicMax(maxB + z, (unsigned long long)__double_as_longlong(rb[0]));
  }
}

// Per-chunk scale factors 2^(7 - e) for both digit layouts.
__global__ void chunk_scale_kernel(const unsigned long long *__restrict__ maxS,
                                   const unsigned long long *__restrict__ maxB, int nz,
                                   double *scS, double *scB) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  scS[z] = ldexp(1.0, 7 - chunk_exp(__longlong_as_double((long long)maxS[z])));
  scB[z] = ldexp(1.0, 7 - chunk_exp(__longlong_as_double((long long)maxB[z])));
}

// k-blocked element offset (elements) of (row r, contraction c) in one slice.
template <int DT>
__host__ __device__ __forceinline__ size_t kb_off(int r, int c, size_t rows) {
  constexpr int B = 32 / DTraits<DT>::ES;
  return ((size_t)(c / B) * rows + r) * B + (c % B);
}

// grid (Kp/64, np/64), 256 threads; tile 64 rows x 64 columns.  Thread t owns
// columns 2(t%32), 2(t%32)+1 of rows t/32 + 8q (q < 8): coalesced 16-byte
// reads.  DS[p] (rows np, contraction Kp) gets the S-layout digits (kx0
// columns zeroed, their term is exact FP64); DB[p] (rows Kp, contraction np)
// gets all columns through an LDS transpose.  want_ds == 0 skips DS.
template <int DT, int S>
__global__ __launch_bounds__(256) void slice_psi_kernel(
    const double *__restrict__ psi, long ld, int n, int K, int kx0_end, int kchunk,
    const double *__restrict__ scS, const double *__restrict__ scB,
    typename DTraits<DT>::elem *DS, typename DTraits<DT>::elem *DB, int np, int Kp,
    int want_ds, int vec) {
  typedef typename DTraits<DT>::elem E;
  constexpr int ES = DTraits<DT>::ES;
  constexpr int PITCH = 64 + 16 / ES;  // LDS row pitch in elements (16-byte aligned rows)
  constexpr int PER = 16 / ES;
  __shared__ __attribute__((aligned(16))) E lt[S][64 * PITCH];
  const int t = threadIdx.x;
  const int i0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  const int cp = t % 32, rg = t / 32;
  const int kc = k0 + 2 * cp;  // first of this thread's two columns
  const int z = min(kc, K - 1 > 0 ? K - 1 : 0) / kchunk;
  const double sS = scS[z], sB = scB[z];
  const size_t sds = (size_t)np * Kp, sdb = (size_t)Kp * np;
#pragma unroll 2
  for (int q = 0; q < 8; q++) {
    const int ii = rg + 8 * q, i = i0 + ii;
    double x[2];
    if (vec && i < n && kc + 1 < K) {
      const double2 v = *reinterpret_cast<const double2 *>(psi + (long)i * ld + kc);
      x[0] = v.x;
      x[1] = v.y;
    } else {
#pragma unroll
      for (int e = 0; e < 2; e++) x[e] = (i < n && kc + e < K) ? psi[(long)i * ld + kc + e] : 0.0;
    }
    double dB[2][S], dS[2][S];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      split_value<S>(x[e] * sB / 128.0, 0, dB[e]);  // x * 2^(7-e) with e folded in sB
      if (want_ds) {
        if (kc + e < kx0_end) {
#pragma unroll
          for (int p = 0; p < S; p++) dS[e][p] = 0.0;
        } else if (sS == sB) {
#pragma unroll
          for (int p = 0; p < S; p++) dS[e][p] = dB[e][p];
        } else {
          split_value<S>(x[e] * sS / 128.0, 0, dS[e]);
        }
      }
#pragma unroll
      for (int p = 0; p < S; p++) lt[p][(2 * cp + e) * PITCH + ii] = to_digit<DT>(dB[e][p]);
    }
    if (want_ds) {
#pragma unroll
      for (int p = 0; p < S; p++) {
        E *dst = DS + p * sds + kb_off<DT>(i, kc, np);
        if constexpr (ES == 1) {
          const unsigned short u = (unsigned short)(((unsigned)(int)dS[0][p] & 0xffu) |
                                                    (((unsigned)(int)dS[1][p] & 0xffu) << 8));
          *reinterpret_cast<unsigned short *>(dst) = u;
        } else {
          const E lo = to_digit<DT>(dS[0][p]), hi = to_digit<DT>(dS[1][p]);
          *reinterpret_cast<unsigned *>(dst) =
              (unsigned)__builtin_bit_cast(unsigned short, lo) |
              ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
        }
      }
    }
  }
  __syncthreads();
  // DB: thread t writes 16 consecutive contraction entries (rows i) of row k0 + t/4
  const int kk = t / 4, ib = (t % 4) * 16;
#pragma unroll
  for (int p = 0; p < S; p++) {
    E *dst = DB + p * sdb + kb_off<DT>(k0 + kk, i0 + ib, Kp);
#pragma unroll
    for (int c = 0; c < 16; c += PER)
      *reinterpret_cast<i4v *>(dst + c) =
          *reinterpret_cast<const i4v *>(&lt[p][kk * PITCH + ib + c]);
  }
}

// One workgroup (256 threads) per row r of an n x n matrix M (row-major, or
// its transpose when trans): the off-diagonal part X = M - diag(M) goes to
// digits D[p][r][0..npk) (zero padded) with exponent f_r from the row's
// off-diagonal maximum; the diagonal M[r][r] is kept in FP64 in diagv[r].
// Rows r >= n are zero.
template <int DT, int S>
__global__ __launch_bounds__(256) void slice_rows_kernel(const double *__restrict__ M, int n,
                                                         int trans, typename DTraits<DT>::elem *D,
                                                         int npk, int *fexp, double *diagv) {
  const int r = blockIdx.x;
  __shared__ double red[256];
  double m = 0.0;
  if (r < n)
    for (int c = threadIdx.x; c < n; c += 256)
      if (c != r) m = fmax(m, fabs(trans ? M[(long)c * n + r] : M[(long)r * n + c]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const int e = chunk_exp(red[0]);
  if (threadIdx.x == 0) {
    fexp[r] = e;
    diagv[r] = (r < n) ? M[(long)r * n + r] : 0.0;
  }
  for (int c = threadIdx.x; c < npk; c += 256) {
    double d[S];
    const bool in = (r < n && c < n && c != r);
    const double x = in ? (trans ? M[(long)c * n + r] : M[(long)r * n + c]) : 0.0;
    split_value<S>(x, e, d);
#pragma unroll
    for (int p = 0; p < S; p++)
      D[(size_t)p * npk * npk + kb_off<DT>(r, c, npk)] = to_digit<DT>(d[p]);
  }
}

struct DigitArgs {
  const void *X;  // k-blocked digits, rows m
  long ldx, sx;   // rows of the layout and slice stride (elements)
  const void *Y;  // k-blocked digits, rows n
  long ldy, sy;
  int M, N;       // output extent
  int kc;         // contraction length per split (elements, multiple of DBK) or total
  int kfull;      // contraction extent of the digit arrays (padded)
  int ntiles, nz; // MODE 0: symmetric tiles, splits
  double *part;   // MODE 0 output
  const unsigned long long *chunk_max;  // MODE 0: per chunk scale source
  // MODE 1
  double *out;
  long ldo;
  double alpha, beta;
  const int *fexp;                       // per output row exponent (T rows)
  const double *diagv;                   // FP64 diagonal of T
  const void *DB;                        // psi digits (rows Kp, contraction np), k-blocked
  long dbs;                              // DB slice stride (elements)
  int np;                                // DB contraction extent
  int Kp;                                // DB rows
  const unsigned long long *colmax;      // per chunk of output columns (psi chunks)
  int kchunk;                            // psi chunk length for the column exponent
};

__device__ __forceinline__ void upper_tile_d(int tile, int T, int &bi, int &bj) {
  int b = 0, rem = tile;
  while (rem >= T - b) {
    rem -= T - b;
    b++;
  }
  bi = b;
  bj = b + rem;
}

// Level-grouped digit products of one 32-byte stage for a 32 x 64 wave tile:
// acc[l] += sum over (p, q) with p < SA, q < SB, p + q = l < NL of
// X_p Y_q^T.  Panels hold 128 rows x 32 bytes per slice with the two 16-byte
// halves of row r swapped when bit 3 of r is set, which spreads any 16
// consecutive lanes' 16-byte fragment reads over all 64 LDS banks.
template <int DT, int SA, int SB, int NL>
__device__ __forceinline__ void digit_stage(const char *Xs, const char *Ys, int xrow, int yrow,
                                            int lane,
                                            typename DTraits<DT>::acc_t (&acc)[NL][2]) {
  constexpr int PANEL = DBM * 32;
  const int r = lane & 31;
  const int boff = r * 32 + 16 * ((lane >> 5) ^ ((r >> 3) & 1));
  i4v a[SA];
#pragma unroll
  for (int p = 0; p < SA; p++) a[p] = *reinterpret_cast<const i4v *>(Xs + p * PANEL + xrow * 32 + boff);
#pragma unroll
  for (int q = 0; q < SB && q < NL; q++) {
    i4v b[2];
#pragma unroll
    for (int j = 0; j < 2; j++)
      b[j] = *reinterpret_cast<const i4v *>(Ys + q * PANEL + (yrow + j * 32) * 32 + boff);
#pragma unroll
    for (int p = 0; p < SA && p + q < NL; p++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[p + q][j] = mfma_digit<DT>(a[p], b[j], acc[p + q][j]);
  }
}

// s_waitcnt with vmcnt = N and the other counters left at their maximum.
template <int N>
__device__ __forceinline__ void wait_vm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Digit of DB at element offset idx of a row of 32 (4 consecutive rows per
// dword group for int8, per 8-byte group for 16-bit digits).
template <int DT>
__device__ __forceinline__ void load_digit_quad(const void *base, size_t off, double (&q)[4]) {
  if constexpr (DT == DT_I8) {
    const int w = *reinterpret_cast<const int *>(reinterpret_cast<const int8_t *>(base) + off);
#pragma unroll
    for (int r = 0; r < 4; r++) q[r] = (double)(int8_t)((w >> (8 * r)) & 0xff);
  } else {
    typedef typename DTraits<DT>::elem E;
    const E *e = reinterpret_cast<const E *>(base) + off;
    const int2 w = *reinterpret_cast<const int2 *>(e);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const unsigned u = (unsigned)((r < 2) ? w.x : w.y);
      const unsigned short h = (unsigned short)((r & 1) ? (u >> 16) : (u & 0xffff));
      q[r] = (double)(float)__builtin_bit_cast(E, h);
    }
  }
}

// C[m][n] = sum over the contraction of level-grouped digit products.
//   MODE 0: symmetric split-K partial tiles of S; S slices, levels < S.
//   MODE 1: back-multiply with T = D + X: out[m][n] = alpha (D_m psi~[m][n]
//           + 2^(f_m + e(n) - 14) C) + beta out, where C uses S-1 slices of
//           X and of psi (levels < S-1) and psi~ is psi rebuilt from all S
//           digits of DB.
// 8 waves as 4 (rows) x 2 (cols); wave tile 32 x 64; double-buffered LDS.
template <int DT, int S, int MODE, int NLB = S - 1>
__global__ __launch_bounds__(DNT) void digit_nt_kernel(DigitArgs a) {
  typedef typename DTraits<DT>::acc_t ACC;
  constexpr int NL = (MODE == 0) ? S : NLB;  // slices in the product and levels kept
  constexpr int ES = DTraits<DT>::ES;
  constexpr int SK = 32 / ES;                 // contraction elements per stage
  constexpr int PANEL = DBM * 32;             // bytes of one slice panel per stage
  constexpr int STAGE = 2 * NL * PANEL;       // X and Y slices of one stage
  constexpr int NBUF = 3;                     // LDS ring depth
  constexpr int NI = 2 * NL * 4;              // 1 KB wave-loads per stage
  constexpr int PERW = NI / 8;                // per wave (8 waves)
  // fp32 accumulation is exact only below 2^24 (level sums < 6 * 127^2 =
  // 96774 per k for S <= 6): promote every PROMOTE stages (128 k) into one
  // FP64 sum per output
  constexpr int PROMOTE = (DT == DT_I8) ? (1 << 30) : 128 / SK;
  __shared__ __attribute__((aligned(1024))) char smem[NBUF * STAGE];

  int bi, bj, z = 0, tile;
  if (MODE == 0) {
    const int T = cdv(a.M, DBM);
    const int b = blockIdx.x, nb = gridDim.x, per = nb >> 3;
    const int logical = (b & 7) * per + (b >> 3);
    if (logical >= a.ntiles * a.nz) return;
    tile = logical % a.ntiles;
    z = logical / a.ntiles;
    upper_tile_d(tile, T, bi, bj);
  } else {
    const int tm = cdv(a.M, DBM);
    const int b = blockIdx.x;
    const int nb = gridDim.x;
    const int per = nb >> 3;
    const int logical = (b & 7) * per + (b >> 3);
    if (logical >= tm * cdv(a.N, DBM)) return;
    bi = logical % tm;
    bj = logical / tm;
    tile = logical;
  }
  const int m0 = bi * DBM, n0 = bj * DBM;
  const int ka = (MODE == 0) ? z * a.kc : 0;
  const int kb = (MODE == 0) ? min(a.kfull, ka + a.kc) : a.kfull;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xrow = (w >> 1) * 32, yrow = (w & 1) * 64;

  ACC acc[NL][2];
  double facc[2][16];
#pragma unroll
  for (int l = 0; l < NL; l++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[l][j] = ACC{};
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int v = 0; v < 16; v++) facc[j][v] = 0.0;

  const char *Xg = reinterpret_cast<const char *>(a.X);
  const char *Yg = reinterpret_cast<const char *>(a.Y);
  // Wave-load i of a stage: operand o, slice p, quarter q (rows 32q..32q+31).
  // Lane L fills LDS slot q*64 + L of the panel, i.e. row r = (q*64+L)/2,
  // stored half hs = L%2, which holds global half hs ^ ((r >> 3) & 1).
  auto issue = [&](int st) {
    const int kblk = (ka + st * SK) / SK;
    char *buf = smem + (st % NBUF) * STAGE;
#pragma unroll
    for (int j = 0; j < PERW; j++) {
      const int i = w + 8 * j;
      const int o = i / (NL * 4), p = (i / 4) % NL, q = i % 4;
      const int pos = q * 64 + lane, r = pos >> 1, h = (pos & 1) ^ ((r >> 3) & 1);
      const char *g = (o == 0)
          ? Xg + ((size_t)p * a.sx + ((size_t)kblk * a.ldx + m0 + r) * SK) * ES + h * 16
          : Yg + ((size_t)p * a.sy + ((size_t)kblk * a.ldy + n0 + r) * SK) * ES + h * 16;
      char *l = buf + o * NL * PANEL + p * PANEL + q * 1024;
      __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void *)l, 16, 0, 0);
    }
  };
  const int nst = (kb - ka) / SK;
  if (nst > 0) issue(0);
  if (nst > 1) issue(1);
  for (int s = 0; s < nst; s++) {
    if (s + 1 < nst) wait_vm<PERW>();
    else wait_vm<0>();
    __syncthreads();
    if (s + 2 < nst) issue(s + 2);
    const char *cur = smem + (s % NBUF) * STAGE;
    digit_stage<DT, NL, NL, NL>(cur, cur + NL * PANEL, xrow, yrow, lane, acc);
    if (DT != DT_I8 && ((s + 1) % PROMOTE == 0 || s + 1 == nst)) {
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
          double t = 0.0;
#pragma unroll
          for (int l = NL - 1; l >= 0; l--) t += (double)acc[l][j][v] * rinv(l);
          facc[j][v] += t;
        }
#pragma unroll
      for (int l = 0; l < NL; l++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[l][j] = ACC{};
    }
  }

  // epilogue: combine levels in FP64
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int ln = yrow + j * 32 + (lane & 31);
    const int nn = n0 + ln;
    // MODE 1: psi~[m][nn] for this lane's 16 rows, rebuilt from all S digits
    double pd[16];
    int ez = 0;
    if (MODE == 1) {
      ez = chunk_exp(__longlong_as_double((long long)a.colmax[min(nn, a.N - 1) / a.kchunk]));
#pragma unroll
      for (int v = 0; v < 16; v++) pd[v] = 0.0;
#pragma unroll
      for (int p = S - 1; p >= 0; p--) {
#pragma unroll
        for (int g = 0; g < 4; g++) {
          double q4[4];
          const size_t off = (size_t)p * a.dbs +
                             kb_off<DT>(min(nn, a.N - 1), m0 + xrow + 8 * g + 4 * (lane >> 5),
                                        (size_t)a.Kp);
          load_digit_quad<DT>(a.DB, off, q4);
#pragma unroll
          for (int r = 0; r < 4; r++) pd[4 * g + r] += q4[r] * rinv(p);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
      double val = 0.0;
      if (DT == DT_I8) {
#pragma unroll
        for (int l = NL - 1; l >= 0; l--) val += (double)acc[l][j][v] * rinv(l);
      } else {
        val = facc[j][v];
      }
      const int lm = xrow + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
      if (MODE == 0) {
        const int e = chunk_exp(__longlong_as_double((long long)a.chunk_max[z]));
        a.part[((size_t)z * a.ntiles + tile) * (DBM * DBM) + lm * DBM + ln] =
            ldexp(val, 2 * (e - 7));
      } else {
        const int m = m0 + lm;
        if (m < a.M && nn < a.N) {
          const double r = a.alpha * (a.diagv[m] * ldexp(pd[v], ez - 7) +
                                      ldexp(val, a.fexp[m] + ez - 14));
          double *o = a.out + (size_t)m * a.ldo + nn;
          *o = (a.beta != 0.0) ? r + a.beta * (*o) : r;
        }
      }
    }
  }
}

// Fixed-order sum of the split-K partial tiles + FP64 kx0 term.
__global__ __launch_bounds__(256) void split_reduce_kernel(
    const double *__restrict__ part, int n, int nz, double alpha, const double *__restrict__ psi,
    long ld, int kx0_end, double *S, double diag, unsigned long long *maxdev) {
  const int T = cdv(n, DBM);
  const int ntiles = T * (T + 1) / 2;
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const int tile = (int)(idx / (DBM * DBM));
  if (tile >= ntiles) return;
  const int e = (int)(idx % (DBM * DBM));
  const int lm = e / DBM, ln = e % DBM;
  int bi, bj;
  upper_tile_d(tile, T, bi, bj);
  if (bi == bj && ln < lm) return;
  const int i = bi * DBM + lm, j = bj * DBM + ln;
  if (i >= n || j >= n) return;
  double s = 0.0;
  for (int z = 0; z < nz; z++) s += part[((long)z * ntiles + tile) * (DBM * DBM) + e];
  double g = 0.0;
  for (int k = 0; k < kx0_end; k++) g += psi[(long)i * ld + k] * psi[(long)j * ld + k];
  s = alpha * s + 0.5 * alpha * g;
  S[(long)i * n + j] = s;
  S[(long)j * n + i] = s;
  if (maxdev) {
    const double dv = fabs(s - (i == j ? diag : 0.0));
    atomicMax(maxdev, (unsigned long long)__double_as_longlong(dv));
  }
}

// ---------------------------------------------------------------- host plan
struct SplitPlan {
  int dt, S, es;
  int np, Kp;      // padded rows (multiple of DBM) and columns (multiple of 64 and DBK)
  int kchunk, nz;  // psi chunks (exponents and split-K)
  int ntiles;
  size_t off_ds, off_db, off_dt, off_maxS, off_maxB, off_scS, off_scB, off_fexp, off_diag,
      off_part, off_md, total;
};

bool decode_prec(int prec, int &dt, int &S) {
  const int base = prec & 0xff;
  S = (prec >> 8) & 0xf;
  if (S == 0) S = 3;
  if (S < 2 || S > SMAX) return false;
  if (base == OA_PREC_INT8) dt = DT_I8;
  else if (base == OA_PREC_FP16) dt = DT_F16;
  else if (base == OA_PREC_BF16) dt = DT_BF16;
  else return false;
  return true;
}

// Largest level sum of |a_p a_q| per contraction element (|a_p| <= 127):
// the int32 chunk length must keep chunk * bound < 2^31.
long level_bound(int S) {
  long best = 0;
  for (int L = 2; L <= S + 1; L++) {
    long b = 0;
    for (int p = 1; p < L; p++) {
      const int q = L - p;
      if (q < 1 || p > S || q > S) continue;
      b += 127L * 127L;
    }
    best = b > best ? b : best;
  }
  return best;
}

SplitPlan make_plan(int prec, int n, int K) {
  SplitPlan pl{};
  decode_prec(prec, pl.dt, pl.S);
  pl.es = (pl.dt == DT_I8) ? 1 : 2;
  pl.np = cdv(n, DBM) * DBM;
  pl.Kp = cdv(K, DBM) * DBM;
  const int T = cdv(n, DBM);
  pl.ntiles = T * (T + 1) / 2;
  // The chunk is both the split-K unit and the scale block: short chunks
  // keep each chunk's exponent close to its own coefficients (the envelope
  // falls steeply with |G|), and the int32 sums stay exact below the cap.
  long cap = (pl.dt == DT_I8) ? (2147483647L / level_bound(pl.S)) : (1L << 30);
  cap = cap / DBK * DBK;
  long target = KCHUNK_TARGET;
  // enough workgroups for small problems: at least ~256 of them
  while (target > 512 && (long)pl.ntiles * cdv(pl.Kp, (int)target) < 256) target /= 2;
  int kchunk = (int)(target < cap ? target : cap);
  kchunk = max(DBK, kchunk / DBK * DBK);
  pl.kchunk = kchunk;
  pl.nz = cdv(pl.Kp, kchunk);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  pl.off_ds = take((size_t)pl.S * pl.np * pl.Kp * pl.es);
  pl.off_db = take((size_t)pl.S * pl.Kp * pl.np * pl.es);
  pl.off_dt = take((size_t)pl.S * pl.np * pl.np * pl.es);
  pl.off_maxS = take((size_t)pl.nz * 8);
  pl.off_maxB = take((size_t)pl.nz * 8);
  pl.off_scS = take((size_t)pl.nz * 8);
  pl.off_scB = take((size_t)pl.nz * 8);
  pl.off_fexp = take((size_t)pl.np * 4);
  pl.off_diag = take((size_t)pl.np * 8);
  pl.off_md = take(8);
  pl.off_part = take((size_t)pl.nz * pl.ntiles * DBM * DBM * 8);
  pl.total = off;
  return pl;
}

template <int DT, int S>
void launch_slice_psi(hipStream_t s, const SplitPlan &pl, const double *psi, long ld, int n, int K,
                      int kx0_end, bool want_ds, char *ws) {
  typedef typename DTraits<DT>::elem E;
  auto *maxS = reinterpret_cast<unsigned long long *>(ws + pl.off_maxS);
  auto *maxB = reinterpret_cast<unsigned long long *>(ws + pl.off_maxB);
  double *scS = reinterpret_cast<double *>(ws + pl.off_scS);
  double *scB = reinterpret_cast<double *>(ws + pl.off_scB);
  hipLaunchKernelGGL(fill_u64, dim3(cdv(pl.nz, 256)), dim3(256), 0, s, maxS, pl.nz);
  hipLaunchKernelGGL(fill_u64, dim3(cdv(pl.nz, 256)), dim3(256), 0, s, maxB, pl.nz);
  hipLaunchKernelGGL(chunk_max_kernel, dim3(pl.nz, cdv(n, 8)), dim3(256), 0, s, psi, ld, n, K,
                     kx0_end, pl.kchunk, maxS, maxB);
  hipLaunchKernelGGL(chunk_scale_kernel, dim3(cdv(pl.nz, 256)), dim3(256), 0, s,
                     (const unsigned long long *)maxS, (const unsigned long long *)maxB, pl.nz,
                     scS, scB);
  const int vec = (((uintptr_t)psi & 15) == 0 && (ld % 2) == 0) ? 1 : 0;
  hipLaunchKernelGGL((slice_psi_kernel<DT, S>), dim3(pl.Kp / 64, pl.np / 64), dim3(256), 0, s,
                     psi, ld, n, K, kx0_end, pl.kchunk, (const double *)scS, (const double *)scB,
                     reinterpret_cast<E *>(ws + pl.off_ds), reinterpret_cast<E *>(ws + pl.off_db),
                     pl.np, pl.Kp, want_ds ? 1 : 0, vec);
}

template <int DT, int S>
int run_overlap(hipStream_t s, const SplitPlan &pl, int n, int K, const double *psi, long ld,
                int kx0_end, double alpha, double *Sout, double diag, double *maxdev, char *ws) {
  launch_slice_psi<DT, S>(s, pl, psi, ld, n, K, kx0_end, true, ws);
  DigitArgs a{};
  a.X = ws + pl.off_ds;
  a.ldx = pl.np;
  a.sx = (long)pl.np * pl.Kp;
  a.Y = a.X;
  a.ldy = a.ldx;
  a.sy = a.sx;
  a.M = n;
  a.N = n;
  a.kc = pl.kchunk;
  a.kfull = pl.Kp;
  a.ntiles = pl.ntiles;
  a.nz = pl.nz;
  a.part = reinterpret_cast<double *>(ws + pl.off_part);
  a.chunk_max = reinterpret_cast<const unsigned long long *>(ws + pl.off_maxS);
  hipLaunchKernelGGL((digit_nt_kernel<DT, S, 0>), dim3(cdv(pl.ntiles * pl.nz, 8) * 8), dim3(DNT),
                     0, s, a);
  unsigned long long *md = reinterpret_cast<unsigned long long *>(maxdev);
  if (md) hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, s, md, 1);
  const long total = (long)pl.ntiles * DBM * DBM;
  hipLaunchKernelGGL(split_reduce_kernel, dim3((unsigned)cdv((int)total, 256)), dim3(256), 0, s,
                     (const double *)a.part, n, pl.nz, alpha, psi, ld, kx0_end, Sout, diag, md);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

template <int DT, int S, int NLB>
int run_backmul(hipStream_t s, const SplitPlan &pl, int n, int K, const double *M, int transM,
                const double *psi, long ldpsi, int kx0_end, bool reuse, double alpha,
                double beta, double *out, long ldout, char *ws) {
  typedef typename DTraits<DT>::elem E;
  if (!reuse) launch_slice_psi<DT, S>(s, pl, psi, ldpsi, n, K, kx0_end, false, ws);
  hipLaunchKernelGGL((slice_rows_kernel<DT, NLB>), dim3(pl.np), dim3(256), 0, s,
                     M, n, transM, reinterpret_cast<E *>(ws + pl.off_dt), pl.np,
                     reinterpret_cast<int *>(ws + pl.off_fexp),
                     reinterpret_cast<double *>(ws + pl.off_diag));
  DigitArgs a{};
  a.X = ws + pl.off_dt;
  a.ldx = pl.np;
  a.sx = (long)pl.np * pl.np;
  a.Y = ws + pl.off_db;
  a.ldy = pl.Kp;
  a.sy = (long)pl.Kp * pl.np;
  a.M = n;
  a.N = K;
  a.kc = pl.np;
  a.kfull = pl.np;
  a.out = out;
  a.ldo = ldout;
  a.alpha = alpha;
  a.beta = beta;
  a.fexp = reinterpret_cast<const int *>(ws + pl.off_fexp);
  a.diagv = reinterpret_cast<const double *>(ws + pl.off_diag);
  a.DB = ws + pl.off_db;
  a.dbs = (long)pl.Kp * pl.np;
  a.np = pl.np;
  a.Kp = pl.Kp;
  a.colmax = reinterpret_cast<const unsigned long long *>(ws + pl.off_maxB);
  a.kchunk = pl.kchunk;
  const int tiles = cdv(n, DBM) * cdv(K, DBM);
  hipLaunchKernelGGL((digit_nt_kernel<DT, S, 1, NLB>), dim3(cdv(tiles, 8) * 8), dim3(DNT), 0, s,
                     a);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

template <int DT>
int dispatch_overlap(hipStream_t st, const SplitPlan &pl, int n, int K, const double *psi,
                     long ld, int kx0_end, double alpha, double *S, double diag, double *md,
                     char *ws) {
  switch (pl.S) {
    case 2: return run_overlap<DT, 2>(st, pl, n, K, psi, ld, kx0_end, alpha, S, diag, md, ws);
    case 3: return run_overlap<DT, 3>(st, pl, n, K, psi, ld, kx0_end, alpha, S, diag, md, ws);
    case 4: return run_overlap<DT, 4>(st, pl, n, K, psi, ld, kx0_end, alpha, S, diag, md, ws);
    case 5: return run_overlap<DT, 5>(st, pl, n, K, psi, ld, kx0_end, alpha, S, diag, md, ws);
    case 6: return run_overlap<DT, 6>(st, pl, n, K, psi, ld, kx0_end, alpha, S, diag, md, ws);
  }
  return OA_ERR_UNSUPPORTED;
}

template <int DT>
int dispatch_backmul(hipStream_t st, const SplitPlan &pl, bool lean, int n, int K,
                     const double *M, int transM, const double *psi, long ldpsi, int kx0_end,
                     bool reuse, double alpha, double beta, double *out, long ldout, char *ws) {
#define OA_BM(SS, NL)                                                                        \
  return run_backmul<DT, SS, NL>(st, pl, n, K, M, transM, psi, ldpsi, kx0_end, reuse, alpha, \
                                 beta, out, ldout, ws);
  switch (pl.S) {
    case 2: OA_BM(2, 1)
    case 3: if (lean) { OA_BM(3, 1) } else { OA_BM(3, 2) }
    case 4: if (lean) { OA_BM(4, 2) } else { OA_BM(4, 3) }
    case 5: if (lean) { OA_BM(5, 3) } else { OA_BM(5, 4) }
    case 6: if (lean) { OA_BM(6, 4) } else { OA_BM(6, 5) }
  }
#undef OA_BM
  return OA_ERR_UNSUPPORTED;
}

}  // namespace

size_t oa_split_ws_bytes(int prec, int n, int K) {
  int dt, S;
  if ((prec & 0xff) == OA_PREC_FP64 || !decode_prec(prec & 0xffff, dt, S)) return 0;
  return make_plan(prec & 0xffff, n, K).total;
}

int oa_launch_overlap_split(hipStream_t s, int prec, int n, int K, const double *psi, long ld,
                            int kx0_end, double alpha, double *S, double diag, double *maxdev,
                            void *ws, size_t ws_bytes) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return OA_ERR_UNSUPPORTED;
  const SplitPlan pl = make_plan(prec & 0xffff, n, K);
  if (pl.total > ws_bytes) return OA_ERR_CAPACITY;
  char *w = (char *)ws;
  if (dt == DT_I8) return dispatch_overlap<DT_I8>(s, pl, n, K, psi, ld, kx0_end, alpha, S, diag, maxdev, w);
  if (dt == DT_F16) return dispatch_overlap<DT_F16>(s, pl, n, K, psi, ld, kx0_end, alpha, S, diag, maxdev, w);
  return dispatch_overlap<DT_BF16>(s, pl, n, K, psi, ld, kx0_end, alpha, S, diag, maxdev, w);
}

int oa_launch_backmul_split(hipStream_t s, int prec, int n, int K, const double *M, int transM,
                            const double *psi, long ldpsi, int kx0_end, double alpha, double beta,
                            double *out, long ldout, void *ws, size_t ws_bytes) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return OA_ERR_UNSUPPORTED;
  const SplitPlan pl = make_plan(prec & 0xffff, n, K);
  if (pl.total > ws_bytes) return OA_ERR_CAPACITY;
  const bool reuse = (prec & OA_PREC_REUSE_DIGITS) != 0;
  const bool lean = (prec & OA_PREC_LEAN_BACKMUL) != 0;
  char *w = (char *)ws;
  if (dt == DT_I8)
    return dispatch_backmul<DT_I8>(s, pl, lean, n, K, M, transM, psi, ldpsi, kx0_end, reuse,
                                   alpha, beta, out, ldout, w);
  if (dt == DT_F16)
    return dispatch_backmul<DT_F16>(s, pl, lean, n, K, M, transM, psi, ldpsi, kx0_end, reuse,
                                    alpha, beta, out, ldout, w);
  return dispatch_backmul<DT_BF16>(s, pl, lean, n, K, M, transM, psi, ldpsi, kx0_end, reuse,
                                   alpha, beta, out, ldout, w);
}
