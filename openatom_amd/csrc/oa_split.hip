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
// v_mfma_f32_32x32x16_{f16,bf16} and fp32 digits on v_mfma_f32_32x32x2_f32
// into fp32 promoted to FP64 every 128 k (sums stay below 2^24 and are
// therefore exact).  Products are grouped by
// level L = p + q <= S + 1 (the Ozaki triangular truncation); level sums are
// combined in FP64 with weights R^(-(L-2)) and the chunk scales.  Because the
// level sums are exact, every result is independent of the tile shape, the
// split-K grouping and the launch (single or grouped): only the chunk length
// enters the numerics, and it depends on (n, K) alone.
//
// Layout.  Every digit matrix (rows R, contraction C) is stored k-blocked:
// element (p, r, c) at p*R*C + ((c / B)*R + r)*B + c % B with B = 32 bytes of
// digits, so one pipeline stage of a workgroup (TILE rows x 32 bytes) is one
// contiguous block per slice.
//
// Kernels (each a body over work items, launched either for one instance
// with arguments by value or grouped over many instances from device tables):
//   chunk_max     max |psi| per chunk (two variants: for the S layout, the
//                 leading kx0 columns left out or at 1/sqrt(2); plain),
//                 order-free via atomicMax on the IEEE bits of non-negative
//                 doubles.
//   slice_psi     psi -> digits in two layouts: DS[p][row][k] (rows
//                 contiguous in k, for S = psi psi^T; kx0 columns zeroed
//                 and their term added in FP64, or for a wide plane kept
//                 times 1/sqrt(2)) and DB[p][k][row] (rows
//                 contiguous in the state index, for T psi), the second
//                 through an LDS transpose.
//   slice_rows    an n x n matrix -> digits with one exponent per row.
//   digit_nt      C[m][n] = sum_k X[m][k] Y[n][k] over digit slices:
//                 MODE 0 = split-K partial tiles of S (upper triangle),
//                 MODE 1 = back-multiply epilogue out = a*T*psi + b*out,
//                 MODE 2 = split-K partial tiles of psi F^T (all tiles, the
//                 asymmetric Lambda product).
//                 TILE 128: 8 waves as 4 x 2 with 32 x 64 wave tiles, two
//                 waves per SIMD;
//                 TILE 64: 256 threads, 4 waves as 2 x 2, wave tile 32 x 32.
//   split_reduce  fixed-order sum of the partial tiles, FP64 kx0 term,
//                 mirror, max|S - d I|.
#include <math.h>

#include <type_traits>

#include "oa_internal.h"

namespace {

typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

constexpr int DBK = 32;   // chunk granularity (elements); a stage holds 32 bytes per row
constexpr int SMAX = 6;
constexpr int SLT = 64;   // slicing tile (rows x columns)
// The leading kx0 columns (the kx = 0 plane, stored once for +G and -G)
// count half in S.  Up to KX0_FP64_MAX of them are left out of the S-layout
// digits and their term is added in FP64 by the reduce (exact, and cheap for
// a few columns); a wider plane is kept in the digits times 1/sqrt(2), so the
// digit products carry the weight 1/2 at no extra cost per column.
constexpr double KX0W = 0.70710678118654752440;
constexpr int KX0_FP64_MAX = 32;
constexpr int ADA_SMIN = 2;      // fewest slices an adaptive chunk keeps
constexpr double ADA_THETA = 254.0 / 4.0;  // budget of the dropped chunks, in units of B_S
constexpr int ADA_SMIN_MAX = 4;  // adaptive mode needs at least this many slices at most
constexpr long KCHUNK_TARGET = 8192;  // columns per scale block / split-K chunk

__host__ __device__ inline int cdv(int a, int b) { return (a + b - 1) / b; }

// A workgroup tile of TILE x TILE outputs is split over NW waves, each
// owning WI x WJ blocks of 32 x 32 (one MFMA accumulator per block and
// level); WPE waves per SIMD are allowed, so a wave may hold 512 / WPE
// registers.  WIDE (64 x 64 wave tiles, one wave per SIMD) reads a third
// fewer LDS bytes per MFMA; it is used where the accumulators fit.
template <int TILE, bool WIDE>
struct TileCfg;
template <>
struct TileCfg<128, true> {
  static constexpr int NW = 4;  // waves: 2 rows x 2 columns, 64 x 64 each
  static constexpr int WI = 2;  // 32-row MFMA blocks per wave
  static constexpr int WJ = 2;  // 32-column MFMA blocks per wave
  static constexpr int WPE = 1;
};
template <>
struct TileCfg<128, false> {
  static constexpr int NW = 8;  // waves: 4 rows x 2 columns, 32 x 64 each
  static constexpr int WI = 1;
  static constexpr int WJ = 2;
  static constexpr int WPE = 2;
};
template <bool WIDE>
struct TileCfg<64, WIDE> {
  static constexpr int NW = 4;  // waves: 2 rows x 2 columns, 32 x 32 each
  static constexpr int WI = 1;
  static constexpr int WJ = 1;
  static constexpr int WPE = 2;
};

// Digit element type: int8, 16-bit float or fp32.  ES = bytes per element.
enum DType { DT_I8 = 0, DT_F16 = 1, DT_BF16 = 2, DT_F32 = 3 };

template <int DT>
struct DTraits;
template <>
struct DTraits<DT_I8> {
  typedef int8_t elem;
  static constexpr int ES = 1;
  typedef i16v acc_t;
};
template <>
struct DTraits<DT_F16> {
  typedef _Float16 elem;
  static constexpr int ES = 2;
  typedef f16v acc_t;
};
template <>
struct DTraits<DT_BF16> {
  typedef __bf16 elem;
  static constexpr int ES = 2;
  typedef f16v acc_t;
};
template <>
struct DTraits<DT_F32> {
  typedef float elem;
  static constexpr int ES = 4;
  typedef f16v acc_t;
};

template <int DT>
__device__ __forceinline__ typename DTraits<DT>::elem to_digit(double a) {
  if constexpr (DT == DT_I8) return (int8_t)(int)a;
  else if constexpr (DT == DT_F16) return (_Float16)(float)a;
  else if constexpr (DT == DT_BF16) return (__bf16)(float)a;
  else return (float)a;
}

template <int DT>
__device__ __forceinline__ typename DTraits<DT>::acc_t mfma_digit(i4v a, i4v b,
                                                                  typename DTraits<DT>::acc_t c) {
  if constexpr (DT == DT_I8) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
  } else if constexpr (DT == DT_F16) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(h8v, a),
                                                  __builtin_bit_cast(h8v, b), c, 0, 0, 0);
  } else if constexpr (DT == DT_BF16) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(b8v, a),
                                                   __builtin_bit_cast(b8v, b), c, 0, 0, 0);
  } else {
    // four k = 2 steps: lanes < 32 hold contraction 4h+e of their row, lanes
    // >= 32 the other half, the same pairing for both operands
    const float fa[4] = {__int_as_float(a.x), __int_as_float(a.y), __int_as_float(a.z),
                         __int_as_float(a.w)};
    const float fb[4] = {__int_as_float(b.x), __int_as_float(b.y), __int_as_float(b.z),
                         __int_as_float(b.w)};
#pragma unroll
    for (int e = 0; e < 4; e++) c = __builtin_amdgcn_mfma_f32_32x32x2f32(fa[e], fb[e], c, 0, 0, 0);
    return c;
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

__device__ __forceinline__ double chunk_scale(const unsigned long long *mx, int z) {
  return ldexp(1.0, 7 - chunk_exp(__longlong_as_double((long long)mx[z])));
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
// Only the leading nd digits are formed (the rest are left at 0): an
// adaptive split keeps s_z <= S digits of chunk z.
template <int S>
__device__ __forceinline__ void split_value(double x, int e, double (&d)[S], int nd = S) {
  double u = (e == 0) ? x * 128.0 : ldexp(x, 7 - e);
#pragma unroll
  for (int p = 0; p < S; p++) {
    if (p >= nd) {
      d[p] = 0.0;
      continue;
    }
    const double a = rint(u);
    d[p] = a;
    u = (u - a) * RADIX;
  }
}

// ------------------------------------------------------------ work items
// Instance owning item t: the largest i with prefix[i] <= t.
__device__ __forceinline__ int group_find(const int *__restrict__ prefix, int count, int t) {
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The dispatcher places workgroup b on XCD b % 8.  Workgroups of XCD x walk
// the contiguous item range [x*per, (x+1)*per), so items that share operand
// panels (neighbouring tiles of one instance) share one L2.  The grid is a
// multiple of 8; with grid = 8*per every workgroup takes one item.
template <class F>
__device__ __forceinline__ void for_items(int total, F &&f) {
  const int per = cdv(total, 8), slots = (int)gridDim.x >> 3, x = blockIdx.x & 7;
  for (int pos = blockIdx.x >> 3; pos < per; pos += slots) {
    const int item = x * per + pos;
    if (item < total) f(item);
  }
}

// Single-instance launches of the streaming kernels (chunk max, slicing,
// reduce) use one workgroup per item (grid = 8*per), so no loop is needed:
// workgroup b takes item (b % 8)*per + b/8.
__device__ __forceinline__ int single_item(int total) {
  const int per = cdv(total, 8);
  return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

__global__ void fill_u64(unsigned long long *p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// ------------------------------------------------------------ chunk max
// Item (z, row group of 8): max over rows x chunk columns.  maxB is the
// plain maximum; maxS skips the leading kx0 columns, or counts them at
// |x| / sqrt(2) when the S-layout digits hold them (kx0_end > KX0_FP64_MAX).
__device__ __forceinline__ void chunk_max_body(const oa_chunkmax_args &a, int item) {
  const int z = item % a.nz, r0 = (item / a.nz) * 8;
  const int k0 = z * a.kchunk, k1 = min(a.K, k0 + a.kchunk);
  double ms = 0.0, mb = 0.0;
  const bool vec = ((a.ld & 1) == 0) && (((uintptr_t)a.psi & 15) == 0) && ((k0 & 1) == 0);
  const double w0 = (a.kx0_end > KX0_FP64_MAX) ? KX0W : 0.0;  // weight of a kx0 column in maxS
  for (int r = r0; r < min(a.n, r0 + 8); r++) {
    const double *row = a.psi + (long)r * a.ld;
    int k = k0 + 2 * (int)threadIdx.x;
    if (vec) {
      // 16-byte loads, four in flight per thread
      for (; k + 3 * 512 + 1 < k1; k += 4 * 512) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = *reinterpret_cast<const double2 *>(row + k + u * 512);
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const double x0 = fabs(v[u].x), x1 = fabs(v[u].y);
          const int kk = k + u * 512;
          mb = fmax(mb, fmax(x0, x1));
          ms = fmax(ms, (kk >= a.kx0_end) ? x0 : x0 * w0);
          ms = fmax(ms, (kk + 1 >= a.kx0_end) ? x1 : x1 * w0);
        }
      }
    }
    // remainder (and the unaligned case): this thread's pairs (k, k+1)
    for (; k < k1; k += 512) {
      const double x0 = fabs(row[k]);
      mb = fmax(mb, x0);
      ms = fmax(ms, (k >= a.kx0_end) ? x0 : x0 * w0);
      if (k + 1 < k1) {
        const double x1 = fabs(row[k + 1]);
        mb = fmax(mb, x1);
        ms = fmax(ms, (k + 1 >= a.kx0_end) ? x1 : x1 * w0);
      }
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
    atomicMax(a.maxS + z, (unsigned long long)__double_as_longlong(rs[0]));
    atomicMax(a.maxB + z, (unsigned long long)__double_as_longlong(rb[0]));
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void chunk_max_kernel(oa_chunkmax_args a, int total) {
  const int item = single_item(total);
  if (item < total) chunk_max_body(a, item);
}

__global__ __launch_bounds__(256) void chunk_max_grouped(const oa_chunkmax_args *__restrict__ t,
                                                         const int *__restrict__ prefix,
                                                         int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    chunk_max_body(t[i], item - prefix[i]);
  });
}

// ------------------------------------------------------------ slice counts
// Adaptive split: the digit count s_z of every chunk from the chunk maxima
// alone (one workgroup per instance).  The truncation of a chunk's digits
// contributes at most c_s K_z 4^(e_z) 254^-(s_z - 1) to every element of S
// (kept levels p + q < s_z plus the remainders of s_z digits; c_s grows
// slowly with s), and the same weight W_z = K_z 4^(e_z) prices the error the
// chunk's columns of T psi add to the new states' overlap.  With
// W = sum_z W_z, chunk z drops d slices while W_z 254^d <= ADA_THETA W / nz.
// The kept chunks add at most the bound B_S of S uniform slices, the dropped
// ones at most ADA_THETA B_S, so the total stays within (1 + ADA_THETA) B_S;
// with ADA_THETA = 254 / 4 that is a quarter of the bound of S - 1 uniform
// slices (plus B_S).  Weights are taken relative to the largest exponent,
// and both sums run in a fixed order, so every launch form gives the same
// counts.
__device__ __forceinline__ double chunk_weight(const oa_count_args &a, int z, int eref) {
  const double m = __longlong_as_double((long long)a.maxB[z]);
  if (!(m > 0.0)) return 0.0;
  const long k0 = (long)z * a.kchunk, k1 = min((long)a.K, k0 + a.kchunk);
  const double kz = (double)max(0L, k1 - k0);
  return ldexp(kz, 2 * (chunk_exp(m) - eref));
}

__device__ __forceinline__ void slice_count_body(const oa_count_args &a) {
  __shared__ double red[256];
  __shared__ int ered[256];
  const int t = threadIdx.x;
  int em = -2147483647;
  for (int z = t; z < a.nz; z += 256) {
    const double m = __longlong_as_double((long long)a.maxB[z]);
    if (m > 0.0) em = max(em, chunk_exp(m));
  }
  ered[t] = em;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) ered[t] = max(ered[t], ered[t + o]);
    __syncthreads();
  }
  const int eref = ered[0];
  double w = 0.0;
  for (int z = t; z < a.nz; z += 256) w += chunk_weight(a, z, eref);
  red[t] = w;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  const double budget = ADA_THETA * red[0] / (double)a.nz;
  for (int z = t; z < a.nz; z += 256) {
    double wz = chunk_weight(a, z, eref);
    int s = a.smax;
    while (s > ADA_SMIN && wz * RADIX <= budget) {
      wz *= RADIX;
      s--;
    }
    a.slices[z] = s;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void slice_count_kernel(oa_count_args a) {
  slice_count_body(a);
}

__global__ __launch_bounds__(256) void slice_count_grouped(const oa_count_args *__restrict__ t,
                                                           const int *__restrict__ prefix,
                                                           int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    slice_count_body(t[i]);
  });
}

// Calls f(integral_constant<int, V>) with V = v clamped into [LO, HI].
template <int LO, int HI, class F>
__device__ __forceinline__ void level_switch(int v, F &&f) {
  if constexpr (LO < HI) {
    if (v <= LO) f(std::integral_constant<int, LO>{});
    else level_switch<LO + 1, HI>(v, f);
  } else {
    f(std::integral_constant<int, HI>{});
  }
}

// ------------------------------------------------------------ slicing
// k-blocked element offset (elements) of (row r, contraction c) in one slice.
template <int DT>
__host__ __device__ __forceinline__ size_t kb_off(int r, int c, size_t rows) {
  constexpr int B = 32 / DTraits<DT>::ES;
  return ((size_t)(c / B) * rows + r) * B + (c % B);
}

// Item = 64 rows x 64 columns (item % (Kp/64) selects the columns).  Thread t
// owns columns 2(t%32), 2(t%32)+1 of rows t/32 + 8q (q < 8): coalesced
// 16-byte reads.  DS[p] (rows np, contraction Kp) gets the S-layout digits
// (kx0 columns zeroed, or times 1/sqrt(2) for a wide plane); DB[p] (rows Kp,
// contraction
// np) gets all columns through an LDS transpose.  want_ds == 0 skips DS,
// want_db == 0 skips DB.  (Staging DS through LDS as well, for 16-byte stores,
// measured slower: the second buffer halves the resident workgroups.)
constexpr int SLPITCH = SLT + 16;  // LDS row pitch in bytes (16-byte aligned rows)

// SZ: digits kept for the item's chunk (the item's 64 columns lie in one
// chunk), a compile-time count so that the digit loops carry no per-digit
// branch.
template <int DT, int S, int SZ>
__device__ __forceinline__ void slice_psi_sz(const oa_slice_args &a, int item,
                                             int8_t (*lt)[SLT * SLPITCH]) {
  typedef typename DTraits<DT>::elem E;
  constexpr int ES = DTraits<DT>::ES;
  constexpr int PITCH = SLPITCH;
  constexpr int sz = SZ;
  E *DS = reinterpret_cast<E *>(a.DS);
  E *DB = reinterpret_cast<E *>(a.DB);
  const int nkt = a.Kp / SLT;
  const int t = threadIdx.x;
  const int i0 = (item / nkt) * SLT, k0 = (item % nkt) * SLT;
  const int cp = t % 32, rg = t / 32;
  const int kc = k0 + 2 * cp;  // first of this thread's two columns
  const int z = min(kc, a.K - 1 > 0 ? a.K - 1 : 0) / a.kchunk;
  const double sS = chunk_scale(a.maxS, z), sB = chunk_scale(a.maxB, z);
  const size_t sds = (size_t)a.np * a.Kp, sdb = (size_t)a.Kp * a.np;
  // all eight 16-byte loads first (eight in flight per thread)
  double xq[8][2];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int i = i0 + rg + 8 * q;
    if (a.vec && i < a.n && kc + 1 < a.K) {
      const double2 v = *reinterpret_cast<const double2 *>(a.psi + (long)i * a.ld + kc);
      xq[q][0] = v.x;
      xq[q][1] = v.y;
    } else {
#pragma unroll
      for (int e = 0; e < 2; e++)
        xq[q][e] = (i < a.n && kc + e < a.K) ? a.psi[(long)i * a.ld + kc + e] : 0.0;
    }
  }
#pragma unroll 2
  for (int q = 0; q < 8; q++) {
    const int ii = rg + 8 * q, i = i0 + ii;
    const double x[2] = {xq[q][0], xq[q][1]};
    double dB[2][S], dS[2][S];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      split_value<S>(x[e] * sB / 128.0, 0, dB[e], sz);  // x * 2^(7-e) with e folded in sB
      if (a.want_ds) {
        if (kc + e < a.kx0_end) {
          if (a.kx0_end > KX0_FP64_MAX) {
            split_value<S>(x[e] * KX0W * sS / 128.0, 0, dS[e], sz);
          } else {
#pragma unroll
            for (int p = 0; p < S; p++) dS[e][p] = 0.0;
          }
        } else if (sS == sB) {
#pragma unroll
          for (int p = 0; p < S; p++) dS[e][p] = dB[e][p];
        } else {
          split_value<S>(x[e] * sS / 128.0, 0, dS[e], sz);
        }
      }
#pragma unroll
      for (int p = 0; p < S; p++) {
        if (p >= sz) break;
        lt[p][(2 * cp + e) * PITCH + ii] = (int8_t)(int)dB[e][p];
      }
    }
    if (a.want_ds) {
#pragma unroll
      for (int p = 0; p < S; p++) {
        if (p >= sz) break;
        E *dst = DS + p * sds + kb_off<DT>(i, kc, a.np);
        if constexpr (ES == 1) {
          const unsigned short u = (unsigned short)(((unsigned)(int)dS[0][p] & 0xffu) |
                                                    (((unsigned)(int)dS[1][p] & 0xffu) << 8));
          *reinterpret_cast<unsigned short *>(dst) = u;
        } else if constexpr (ES == 2) {
          const E lo = to_digit<DT>(dS[0][p]), hi = to_digit<DT>(dS[1][p]);
          *reinterpret_cast<unsigned *>(dst) =
              (unsigned)__builtin_bit_cast(unsigned short, lo) |
              ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
        } else {
          *reinterpret_cast<float2 *>(dst) = float2{(float)dS[0][p], (float)dS[1][p]};
        }
      }
    }
  }
  __syncthreads();
  // DB: thread t writes 16 consecutive contraction entries (rows i) of row k0 + t/4
  const int kk = t / 4, ib = (t % 4) * 16;
#pragma unroll
  for (int p = 0; p < S; p++) {
    if (!a.want_db || p >= sz) break;
    E *dst = DB + p * sdb;
    const i4v src = *reinterpret_cast<const i4v *>(&lt[p][kk * PITCH + ib]);
    if constexpr (ES == 1) {
      *reinterpret_cast<i4v *>(dst + kb_off<DT>(k0 + kk, i0 + ib, a.Kp)) = src;
    } else {
      E v[16];
#pragma unroll
      for (int c = 0; c < 16; c++)
        v[c] = to_digit<DT>((double)(int8_t)((src[c / 4] >> (8 * (c % 4))) & 0xff));
      // 16 bytes per store; a store never crosses a 32-byte k-block
#pragma unroll
      for (int c = 0; c < 16; c += 16 / ES)
        *reinterpret_cast<i4v *>(dst + kb_off<DT>(k0 + kk, i0 + ib + c, a.Kp)) =
            *reinterpret_cast<const i4v *>(&v[c]);
    }
  }
  __syncthreads();
}

template <int DT, int S>
__device__ __forceinline__ void slice_psi_body(const oa_slice_args &a, int item) {
  // digits are integers |a| <= 127 in every type: the transpose holds int8
  __shared__ __attribute__((aligned(16))) int8_t lt[S][SLT * SLPITCH];
  const int nkt = a.Kp / SLT;
  const int k0 = (item % nkt) * SLT;
  const int z = min(k0, a.K - 1 > 0 ? a.K - 1 : 0) / a.kchunk;
  const int sz = a.slices ? a.slices[z] : S;
  level_switch<1, S>(sz, [&](auto c) { slice_psi_sz<DT, S, decltype(c)::value>(a, item, lt); });
}

template <int DT, int S>
__global__ __launch_bounds__(256) void slice_psi_kernel(oa_slice_args a, int total) {
  const int item = single_item(total);
  if (item < total) slice_psi_body<DT, S>(a, item);
}

template <int DT, int S>
__global__ __launch_bounds__(256) void slice_psi_grouped(const oa_slice_args *__restrict__ t,
                                                         const int *__restrict__ prefix,
                                                         int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    slice_psi_body<DT, S>(t[i], item - prefix[i]);
  });
}

// Item = row r of an n x n matrix M (row-major, or its transpose when trans):
// the off-diagonal part X = M - diag(M) goes to digits D[p][r][0..npk) (zero
// padded) with exponent f_r from the row's off-diagonal maximum; the diagonal
// M[r][r] is kept in FP64 in diagv[r].  Rows r >= n are zero.
template <int DT, int S>
__device__ __forceinline__ void slice_rows_body(const oa_rows_args &a, int r) {
  typedef typename DTraits<DT>::elem E;
  E *D = reinterpret_cast<E *>(a.D);
  const double *M = a.M;
  const int n = a.n, npk = a.npk;
  __shared__ double red[256];
  double m = 0.0;
  if (r < n)
    for (int c = threadIdx.x; c < n; c += 256)
      if (c != r) m = fmax(m, fabs(a.trans ? M[(long)c * n + r] : M[(long)r * n + c]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const int e = chunk_exp(red[0]);
  if (threadIdx.x == 0) {
    a.fexp[r] = e;
    a.diagv[r] = (r < n) ? M[(long)r * n + r] : 0.0;
  }
  for (int c = threadIdx.x; c < npk; c += 256) {
    double d[S];
    const bool in = (r < n && c < n && c != r);
    const double x = in ? (a.trans ? M[(long)c * n + r] : M[(long)r * n + c]) : 0.0;
    split_value<S>(x, e, d);
#pragma unroll
    for (int p = 0; p < S; p++)
      D[(size_t)p * npk * npk + kb_off<DT>(r, c, npk)] = to_digit<DT>(d[p]);
  }
  __syncthreads();
}

template <int DT, int S>
__global__ __launch_bounds__(256) void slice_rows_kernel(oa_rows_args a, int total) {
  const int item = single_item(total);
  if (item < total) slice_rows_body<DT, S>(a, item);
}

template <int DT, int S>
__global__ __launch_bounds__(256) void slice_rows_grouped(const oa_rows_args *__restrict__ t,
                                                          const int *__restrict__ prefix,
                                                          int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    slice_rows_body<DT, S>(t[i], item - prefix[i]);
  });
}

// ------------------------------------------------------------ digit GEMM
__device__ __forceinline__ void upper_tile_d(int tile, int T, int &bi, int &bj) {
  int b = 0, rem = tile;
  while (rem >= T - b) {
    rem -= T - b;
    b++;
  }
  bi = b;
  bj = b + rem;
}

// Level-grouped digit products of one 32-byte stage for a (32*WI) x (32*WJ)
// wave tile: acc[l] += sum over (p, q) with p + q = l < NL of X_p Y_q^T.
// Panels hold TILE rows x 32 bytes per slice with the two 16-byte halves of
// row r swapped when bit 3 of r is set, which spreads any 16 consecutive
// lanes' 16-byte fragment reads over all 64 LDS banks.
template <int DT, int NL, int WI, int WJ, int TILE, int TN = TILE>
__device__ __forceinline__ void digit_stage(const char *Xs, const char *Ys, int xrow, int yrow,
                                            int lane,
                                            typename DTraits<DT>::acc_t (&acc)[NL][WI][WJ]) {
  constexpr int PANEL = TILE * 32;   // X slice panel (TILE rows of 32 bytes)
  constexpr int YPANEL = TN * 32;    // Y slice panel (TN rows)
  const int r = lane & 31;
  const int boff = r * 32 + 16 * ((lane >> 5) ^ ((r >> 3) & 1));
  i4v a[NL][WI];
#pragma unroll
  for (int p = 0; p < NL; p++)
#pragma unroll
    for (int i = 0; i < WI; i++)
      a[p][i] = *reinterpret_cast<const i4v *>(Xs + p * PANEL + (xrow + i * 32) * 32 + boff);
#pragma unroll
  for (int q = 0; q < NL; q++) {
    i4v b[WJ];
#pragma unroll
    for (int j = 0; j < WJ; j++)
      b[j] = *reinterpret_cast<const i4v *>(Ys + q * YPANEL + (yrow + j * 32) * 32 + boff);
#pragma unroll
    for (int p = 0; p + q < NL; p++)
#pragma unroll
      for (int i = 0; i < WI; i++)
#pragma unroll
        for (int j = 0; j < WJ; j++)
          acc[p + q][i][j] = mfma_digit<DT>(a[p][i], b[j], acc[p + q][i][j]);
  }
}

// s_waitcnt with vmcnt = N and the other counters left at their maximum.
template <int N>
__device__ __forceinline__ void wait_vm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup barrier of the LDS ring.  __syncthreads() carries a workgroup
// fence, for which the compiler drains every vector memory operation of the
// wave (vmcnt(0)) before s_barrier: the direct-to-LDS loads of the stages
// meant to stay in flight would be waited for at every stage.  The ring only
// needs the wave's own LDS reads finished (lgkmcnt(0)) before the barrier;
// the loads of a stage are awaited with wait_vm before it is read.  The
// memory clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void ring_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Digits of DB for 4 consecutive contraction rows (one dword group for int8,
// one 8-byte group for 16-bit digits).
template <int DT>
__device__ __forceinline__ void load_digit_quad(const void *base, size_t off, double (&q)[4]) {
  if constexpr (DT == DT_I8) {
    const int w = *reinterpret_cast<const int *>(reinterpret_cast<const int8_t *>(base) + off);
#pragma unroll
    for (int r = 0; r < 4; r++) q[r] = (double)(int8_t)((w >> (8 * r)) & 0xff);
  } else if constexpr (DT == DT_F32) {
    const float4 w = *reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(base) + off);
    q[0] = w.x;
    q[1] = w.y;
    q[2] = w.z;
    q[3] = w.w;
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

template <int DT, int S, int MODE, int NLB, int TILE>
struct DigitCfg {
  static constexpr int NL = (MODE != 1) ? S : NLB;  // slices in the product and levels kept
  // 64 x 64 wave tiles (one wave per SIMD) fit int32 accumulators of up to
  // 5 levels but measured slower than two 32 x 64 waves per SIMD (latency is
  // hidden by the second wave, not by the wider tile): kept off
  static constexpr bool WIDE = false;
  typedef TileCfg<TILE, WIDE> T;
  static constexpr int WI = T::WI, WJ = T::WJ, WPE = T::WPE;
  static constexpr int PANEL = TILE * 32;           // bytes of one slice panel per stage
  static constexpr int STAGE = 2 * NL * PANEL;      // X and Y slices of one stage
  // LDS ring depth: as many stages (up to five, NBUF - 1 in flight) as fit
  // the 160 KB of a workgroup
  static constexpr int NBUF =
      (5 * STAGE <= 160 * 1024) ? 5 : ((4 * STAGE <= 160 * 1024) ? 4 : 3);
  static constexpr int SMEM = NBUF * STAGE;
  static constexpr int NW = T::NW;
  static constexpr int NTHR = 64 * NW;
};

// C[m][n] = sum over the contraction of level-grouped digit products, for one
// output tile (bi, bj) and, in MODES 0 and 2, one chunk z.
//   MODE 0: symmetric split-K partial tiles of S; NLR slices, levels < NLR.
//   MODE 1: back-multiply with T = D + X: out[m][n] = alpha (D_m psi~[m][n]
//           + 2^(f_m + e(n) - 14) C) + beta out, where C uses NLR slices of
//           X and of psi (levels < NLR) and psi~ is psi rebuilt from the
//           first SD digits of DB.
//   MODE 2: all-tiles split-K partial tiles (Lambda).
// NLR is NL (S or NLB) for a uniform split; an adaptive split selects it per
// item from the chunk's slice count.  The X and Y panels of NBUF stages
// (slots sized for NL slices) rotate through LDS, filled by direct-to-LDS
// loads (global_load_lds_dwordx4) NBUF-1 stages ahead.
//   TN: output columns of the tile (TILE, or 2 TILE for the wide back-multiply
//       tile of 64 x 64 wave tiles, used at two levels or fewer).
template <int DT, int S, int MODE, int NLB, int TILE, int NLR, int TN = TILE>
__device__ __forceinline__ void digit_core(const oa_digit_args &a, int bi, int bj, int z,
                                           int tile, int SD, char *smem) {
  typedef typename DTraits<DT>::acc_t ACC;
  typedef DigitCfg<DT, S, MODE, NLB, TILE> C;
  constexpr bool WIDE = TN != TILE;
  static_assert(!WIDE || (MODE == 1 && DT == DT_I8 && TN == 2 * TILE && NLR <= 3),
                "the wide tile is the int8 back-multiply's, at three levels or fewer");
  constexpr int NL = NLR;
  static_assert(NLR >= 1 && NLR <= C::NL, "levels within the ring's slots");
  constexpr int ES = DTraits<DT>::ES;
  constexpr int SK = 32 / ES;             // contraction elements per stage
  constexpr int PANEL = C::PANEL;         // X slice panel: TILE rows x 32 bytes
  constexpr int YPANEL = TN * 32;         // Y slice panel: TN rows x 32 bytes
  // ring slots: the configuration's (sized for its largest level count), or
  // for the wide tile as many of its own stages as the same LDS holds (<= 5)
  constexpr int STAGE = WIDE ? NL * (PANEL + YPANEL) : C::STAGE;
  constexpr int NBUF = WIDE ? (C::SMEM / STAGE >= 5 ? 5 : C::SMEM / STAGE) : C::NBUF;
  static_assert(NBUF >= 3, "ring of at least three stages");
  constexpr int NW = C::NW;
  constexpr int WI = WIDE ? 2 : C::WI;
  constexpr int WJ = WIDE ? 2 : C::WJ;
  constexpr int WC = TN / (32 * WJ);      // waves per tile row
  static_assert(NW == (TILE / (32 * WI)) * WC, "waves must cover the tile");
  constexpr int NQ = TILE / 32;           // 1 KB wave-loads per X slice panel
  constexpr int NQY = TN / 32;            // per Y slice panel
  constexpr int NI = NL * (NQ + NQY);     // wave-loads per stage
  constexpr int PERW = (NI + NW - 1) / NW;  // per wave (waves w >= NHI issue PERW - 1)
  constexpr int NHI = NI % NW;
  // fp32 accumulation is exact only below 2^24 (level sums < 6 * 127^2 =
  // 96774 per k for S <= 6): promote every PROMOTE stages (128 k) into one
  // FP64 sum per output
  constexpr int PROMOTE = (DT == DT_I8) ? (1 << 30) : 128 / SK;

  const int m0 = bi * TILE, n0 = bj * TN;
  const int ka = (MODE != 1) ? z * a.kc : 0;
  const int kb = (MODE != 1) ? min(a.kfull, ka + a.kc) : a.kfull;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xrow = (w / WC) * 32 * WI, yrow = (w % WC) * 32 * WJ;

  ACC acc[NL][WI][WJ];
  double facc[WI][WJ][16];
#pragma unroll
  for (int l = 0; l < NL; l++)
#pragma unroll
    for (int i = 0; i < WI; i++)
#pragma unroll
      for (int j = 0; j < WJ; j++) acc[l][i][j] = ACC{};
#pragma unroll
  for (int i = 0; i < WI; i++)
#pragma unroll
    for (int j = 0; j < WJ; j++)
#pragma unroll
      for (int v = 0; v < 16; v++) facc[i][j][v] = 0.0;

  const char *Xg = reinterpret_cast<const char *>(a.X);
  const char *Yg = reinterpret_cast<const char *>(a.Y);
  const size_t sx = a.sx, sy = a.sy, ldx = a.ldx, ldy = a.ldy;
  const int nst = (kb - ka) / SK;
  // MODE 1 with int8 digits contracts the tile's own ND row blocks last, so
  // their NL slices are still in the ring when the epilogue rebuilds psi~
  // (those blocks of DB are the psi digits of the tile's output rows).  The
  // int32 level sums do not depend on the order.
  constexpr int ND = TILE / SK;
  constexpr bool DIAG_LDS = (MODE == 1) && (DT == DT_I8) && (ND <= NBUF);
  const int rot = DIAG_LDS ? m0 / SK + ND : 0;
  // Wave-load i of a stage: operand o, slice p, quarter q (rows 32q..32q+31).
  // Lane L fills LDS slot q*64 + L of the panel, i.e. row r = (q*64+L)/2,
  // stored half hs = L%2, which holds global half hs ^ ((r >> 3) & 1).
  // Load j of the wave's share of a stage: operand o, slice p and quarter q
  // are wave-uniform, so the global address splits into a scalar part (the
  // operand, slice and tile rows, plus k-block times its pitch) and a
  // per-lane offset (row r of the quarter, 16-byte half h) fixed for the
  // tile; each stage then adds one scalar product per load.
  const int wu = __builtin_amdgcn_readfirstlane(w);
  const bool whi = NHI == 0 || wu < NHI;  // this wave issues PERW loads per stage
  const char *sbase[PERW];
  size_t spitch[PERW];
  unsigned voff[PERW];
  int loff[PERW];
#pragma unroll
  for (int j = 0; j < PERW; j++) {
    const int i = min(wu + NW * j, NI - 1);   // loads past NI are not issued
    const bool x = i < NL * NQ;               // loads 0 .. NL*NQ-1: X, then Y
    const int iy = i - NL * NQ;
    const int p = x ? i / NQ : iy / NQY, q = x ? i % NQ : iy % NQY;
    const int pos = q * 64 + lane, r = pos >> 1, h = (pos & 1) ^ ((r >> 3) & 1);
    sbase[j] = x ? Xg + ((size_t)p * sx + (size_t)m0 * SK) * ES
                 : Yg + ((size_t)p * sy + (size_t)n0 * SK) * ES;
    spitch[j] = (size_t)(x ? ldx : ldy) * SK * ES;
    voff[j] = (unsigned)(r * SK * ES + h * 16);
    loff[j] = x ? p * PANEL + q * 1024 : NL * PANEL + p * YPANEL + q * 1024;
  }
  auto issue = [&](int st) {
    int kblk = DIAG_LDS ? st + rot : (ka + st * SK) / SK;
    if (DIAG_LDS && kblk >= nst) kblk -= nst;
    char *buf = smem + (st % NBUF) * STAGE;
#pragma unroll
    for (int j = 0; j < PERW; j++) {
      if (NHI != 0 && j == PERW - 1 && !whi) break;
      const char *g = sbase[j] + (size_t)kblk * spitch[j] + voff[j];
      __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void *)(buf + loff[j]),
                                       16, 0, 0);
    }
  };
  // waits until at most c stages of this wave's direct-to-LDS loads are in
  // flight (the loop issues no other vector memory operation)
  auto wait_stages = [&](int c) {
    if (NHI == 0 || whi) {
      if (NBUF >= 5 && c >= 3) wait_vm<3 * PERW>();
      else if (NBUF >= 4 && c >= 2) wait_vm<2 * PERW>();
      else if (c >= 1) wait_vm<PERW>();
      else wait_vm<0>();
    } else {
      if (NBUF >= 5 && c >= 3) wait_vm<3 * (PERW - 1)>();
      else if (NBUF >= 4 && c >= 2) wait_vm<2 * (PERW - 1)>();
      else if (c >= 1) wait_vm<PERW - 1>();
      else wait_vm<0>();
    }
  };
#pragma unroll
  for (int st = 0; st < NBUF - 1; st++)
    if (st < nst) issue(st);
  for (int s = 0; s < nst; s++) {
    // stages s+1 .. min(nst-1, s+NBUF-2) may stay in flight
    wait_stages(min(nst - 1 - s, NBUF - 2));
    ring_barrier();
    if (s + NBUF - 1 < nst) issue(s + NBUF - 1);
    const char *cur = smem + (s % NBUF) * STAGE;
    digit_stage<DT, NL, WI, WJ, TILE, TN>(cur, cur + NL * PANEL, xrow, yrow, lane, acc);
    if (DT != DT_I8 && ((s + 1) % PROMOTE == 0 || s + 1 == nst)) {
#pragma unroll
      for (int i = 0; i < WI; i++)
#pragma unroll
        for (int j = 0; j < WJ; j++)
#pragma unroll
          for (int v = 0; v < 16; v++) {
            double t = 0.0;
#pragma unroll
            for (int l = NL - 1; l >= 0; l--) t = fma((double)acc[l][i][j][v], rinv(l), t);
            facc[i][j][v] += t;
          }
#pragma unroll
      for (int l = 0; l < NL; l++)
#pragma unroll
        for (int i = 0; i < WI; i++)
#pragma unroll
          for (int j = 0; j < WJ; j++) acc[l][i][j] = ACC{};
    }
  }

  // int8: the exact level sums fold into one FP64 value per output right
  // after the loop (smallest level first, explicit fma so every
  // instantiation rounds alike), which frees the NL int32 accumulators before
  // the epilogue's temporaries are live
  if constexpr (DT == DT_I8) {
#pragma unroll
    for (int i = 0; i < WI; i++)
#pragma unroll
      for (int j = 0; j < WJ; j++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
          double val = 0.0;
#pragma unroll
          for (int l = NL - 1; l >= 0; l--) val = fma((double)acc[l][i][j][v], rinv(l), val);
          facc[i][j][v] = val;
        }
  }

  // epilogue: combine levels in FP64.  MODE 1 works on groups of 4 rows per
  // lane: it loads T's per-row exponent and FP64 diagonal together with the
  // psi~ digits of those rows, and tests beta and the tile edge outside the
  // element loop, so an interior tile pays one memory round trip per group
  // (a per-element bounds test kept the compiler from hoisting the loads: one
  // round trip per output).
  const bool edge = (m0 + TILE > a.M) || (n0 + TN > a.N);
#pragma unroll
  for (int i = 0; i < WI; i++)
#pragma unroll
  for (int j = 0; j < WJ; j++) {
    const int ln = yrow + j * 32 + (lane & 31);
    const int nn = n0 + ln;
    const int xr = xrow + i * 32;
    auto level_sum = [&](int v) { return facc[i][j][v]; };
    if (MODE != 1) {
      // the chunk's scale, read once (the partial-tile stores could alias it)
      const int ex = chunk_exp(__longlong_as_double((long long)a.chunk_max[z]));
      const int sc = (MODE == 0) ? 2 * (ex - 7)
                                 : ex + chunk_exp(__longlong_as_double((long long)a.colmax[z])) - 14;
      double *pt = a.part + ((size_t)z * a.ntiles + tile) * (TILE * TILE) + ln;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int lm = xr + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        pt[lm * TILE] = ldexp(level_sum(v), sc);
      }
      continue;
    }
    // a tile's columns lie in one chunk (kchunk is a multiple of 128 and a
    // wide tile pairs two tiles of one chunk): one exponent per tile
    const int ez = chunk_exp(__longlong_as_double((long long)a.colmax[n0 / a.kchunk]));
    const double al = (nn < a.cw) ? a.alpha_w : a.alpha;
#pragma unroll
    for (int g = 0; g < 4; g++) {
      // elements v = 4g + r sit in rows mg + r, mg = m0 + xr + 8g + 4 (lane / 32)
      const int mg = m0 + xr + 8 * g + 4 * (lane >> 5);
      double pd[4] = {0.0, 0.0, 0.0, 0.0}, dv[4];
      int fe[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = min(mg + r, a.M - 1);
        dv[r] = a.diagv[m];
        fe[r] = a.fexp[m];
      }
      // psi~[m][nn] for these rows, rebuilt from all S digits; with DIAG_LDS
      // the leading NL slices come from the ring: row block m0/SK + xr/SK was
      // stage nst - ND + xr/SK, and the 4 rows are bytes b0..b0+3 of LDS row
      // ln (16-byte halves swapped on odd groups of 8 rows, as stored)
      const int b0 = 8 * g + 4 * (lane >> 5);
      if constexpr (DT == DT_I8) {
        // the four rows' digits of slice p sit in one dword (byte r = row
        // mg + r); the leading min(S, 4) digits combine exactly in int32
        // (|sum| <= 127 (254^3 + 254^2 + 254 + 1) < 2^31), absent digits
        // (p >= SD) as zeros, and the rest (S > 4) add in FP64
        constexpr int PI = S < 4 ? S : 4;
        int wq[S];
#pragma unroll
        for (int p = 0; p < S; p++) {
          wq[p] = 0;
          if (p >= SD) continue;
          if (DIAG_LDS && p < NL) {
            const char *yb = smem + ((nst - ND + xr / SK) % NBUF) * STAGE + NL * PANEL + p * YPANEL;
            wq[p] = *reinterpret_cast<const int *>(
                yb + ln * 32 + (((b0 >> 4) ^ ((ln >> 3) & 1)) << 4) + (b0 & 15));
          } else {
            wq[p] = *reinterpret_cast<const int *>(
                reinterpret_cast<const int8_t *>(a.DB) + (size_t)p * a.dbs +
                kb_off<DT>(min(nn, a.N - 1), mg, (size_t)a.Kp));
          }
        }
        // explicit fma pins the rounding, so every instantiation (uniform or
        // per chunk slice counts) gives the same bits
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int iv = 0;
#pragma unroll
          for (int p = 0; p < PI; p++) iv = iv * 254 + (int)(int8_t)((wq[p] >> (8 * r)) & 0xff);
          double t = 0.0;
#pragma unroll
          for (int p = S - 1; p >= PI; p--)
            t = fma((double)(int8_t)((wq[p] >> (8 * r)) & 0xff), rinv(p), t);
          pd[r] = fma((double)iv, rinv(PI - 1), t);
        }
      } else {
#pragma unroll
        for (int p = S - 1; p >= 0; p--) {
          if (p >= SD) continue;
          double q4[4];
          load_digit_quad<DT>(a.DB, (size_t)p * a.dbs + kb_off<DT>(min(nn, a.N - 1), mg,
                                                                     (size_t)a.Kp), q4);
#pragma unroll
          for (int r = 0; r < 4; r++) pd[r] += q4[r] * rinv(p);
        }
      }
      // explicit fma / __dmul_rn: the same rounding in every kernel that
      // inlines this body (single and grouped launches) and every build
      auto value = [&](int r) {
        return __dmul_rn(al, fma(dv[r], ldexp(pd[r], ez - 7),
                                 ldexp(level_sum(4 * g + r), fe[r] + ez - 14)));
      };
      double *o = a.out + (size_t)mg * a.ldo + nn;
      if (!edge) {
        if (a.beta == 0.0) {
#pragma unroll
          for (int r = 0; r < 4; r++) o[(size_t)r * a.ldo] = value(r);
        } else {
          double old[4];
#pragma unroll
          for (int r = 0; r < 4; r++) old[r] = o[(size_t)r * a.ldo];
#pragma unroll
          for (int r = 0; r < 4; r++) o[(size_t)r * a.ldo] = fma(a.beta, old[r], value(r));
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (mg + r < a.M && nn < a.N) {
            double *ov = o + (size_t)r * a.ldo;
            *ov = (a.beta != 0.0) ? fma(a.beta, *ov, value(r)) : value(r);
          }
      }
    }
  }
  // the next item's first loads overwrite the ring (stores may stay in flight)
  ring_barrier();
}


// Item decoding.  An adaptive split spreads the chunks over the XCDs: the
// dispatcher gives XCD x the contiguous item range x*per .. (x+1)*per - 1,
// which covers consecutive chunks; with nz = 8c the chunk order is permuted
// so that XCD x takes chunks x, x + 8, ..., and every XCD gets a like mix of
// heavy (leading, more slices) and light chunks.  Tiles of one chunk stay on
// one XCD (their operand panels share its L2).
template <int DT, int S, int MODE, int NLB, int TILE>
__device__ __forceinline__ void digit_body(const oa_digit_args &a, int logical, char *smem) {
  constexpr int NL = DigitCfg<DT, S, MODE, NLB, TILE>::NL;
  constexpr bool ADA_OK = (DT == DT_I8) && (MODE != 2) && (S >= ADA_SMIN_MAX);
  const bool ada = ADA_OK && a.slices != nullptr;
  int bi, bj, z = 0, tile = logical;
  if (MODE == 0) {
    tile = logical % a.ntiles;
    z = logical / a.ntiles;
    if (ada && (a.nz & 7) == 0) {
      const int c = a.nz >> 3;
      z = (z % c) * 8 + z / c;
    }
    upper_tile_d(tile, cdv(a.M, TILE), bi, bj);
  } else if (MODE == 2) {
    const int T = cdv(a.M, TILE);
    tile = logical % a.ntiles;
    z = logical / a.ntiles;
    bi = tile % T;
    bj = tile / T;
  } else {
    const int tm = cdv(a.M, TILE);
    bi = logical % tm;
    bj = logical / tm;
    if (ada) {
      // chunk of the output columns (kchunk is a multiple of TILE)
      const int cpt = a.kchunk / TILE;
      if ((a.nz & 7) == 0 && cdv(a.N, TILE) == a.nz * cpt) {
        const int c = a.nz >> 3, zl = bj / cpt;
        bj = ((zl % c) * 8 + zl / c) * cpt + bj % cpt;
      }
      z = bj / cpt;
    }
  }
  if constexpr (ADA_OK) {
    if (ada) {
      const int sz = a.slices[z];
      if constexpr (MODE == 0) {
        level_switch<ADA_SMIN, NL>(sz, [&](auto c) {
          digit_core<DT, S, MODE, NLB, TILE, decltype(c)::value>(a, bi, bj, z, tile, S, smem);
        });
      } else {
        constexpr int LO = (NLB - (S - ADA_SMIN) > 1) ? NLB - (S - ADA_SMIN) : 1;
        const int nlr = NLB - (S - sz);
        if constexpr (TILE == 128 && LO <= 2) {
          // chunks of at most two levels run 128 x 256 tiles: the even
          // column tile of each pair takes both, the odd one returns (an
          // even cpt keeps a pair inside one chunk).  At three levels the
          // 64 x 64 wave tiles' accumulators leave the epilogue too few
          // registers and the wide tile measured slower (4.54 -> 4.81 ms).
          const int cpt = a.kchunk / TILE;
          if (nlr <= 2 && (cpt & 1) == 0 && (bj | 1) < cdv(a.N, TILE)) {
            if (bj & 1) return;
            level_switch<LO, (NLB < 2 ? NLB : 2)>(nlr, [&](auto c) {
              digit_core<DT, S, MODE, NLB, TILE, decltype(c)::value, 2 * TILE>(
                  a, bi, bj >> 1, z, tile, sz, smem);
            });
            return;
          }
        }
        level_switch<LO, NLB>(nlr, [&](auto c) {
          digit_core<DT, S, MODE, NLB, TILE, decltype(c)::value>(a, bi, bj, z, tile, sz, smem);
        });
      }
      return;
    }
  }
  digit_core<DT, S, MODE, NLB, TILE, NL>(a, bi, bj, z, tile, S, smem);
}

template <int DT, int S, int MODE, int NLB, int TILE>
__global__ __attribute__((amdgpu_flat_work_group_size(1, DigitCfg<DT, S, MODE, NLB, TILE>::NTHR)))
__attribute__((amdgpu_waves_per_eu(DigitCfg<DT, S, MODE, NLB, TILE>::WPE,
                                   DigitCfg<DT, S, MODE, NLB, TILE>::WPE))) void digit_nt_kernel(
    oa_digit_args a, int total) {
  __shared__ __attribute__((aligned(1024))) char smem[DigitCfg<DT, S, MODE, NLB, TILE>::SMEM];
  // persistent: one workgroup per CU walks its XCD's items (a workgroup per
  // item paid the 160 KB LDS allocation and wave launch at every tile:
  // int8x6a back-multiply 4.62 -> 4.54 ms, int8x5 5.97 -> 5.79 ms)
  for_items(total, [&](int item) { digit_body<DT, S, MODE, NLB, TILE>(a, item, smem); });
}

template <int DT, int S, int MODE, int NLB, int TILE>
__global__ __attribute__((amdgpu_flat_work_group_size(1, DigitCfg<DT, S, MODE, NLB, TILE>::NTHR)))
__attribute__((amdgpu_waves_per_eu(DigitCfg<DT, S, MODE, NLB, TILE>::WPE,
                                   DigitCfg<DT, S, MODE, NLB, TILE>::WPE))) void digit_nt_grouped(
    const oa_digit_args *__restrict__ t, const int *__restrict__ prefix, int count) {
  __shared__ __attribute__((aligned(1024))) char smem[DigitCfg<DT, S, MODE, NLB, TILE>::SMEM];
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    digit_body<DT, S, MODE, NLB, TILE>(t[i], item - prefix[i], smem);
  });
}

// ------------------------------------------------------------ reduce
// Item = 256 consecutive elements of the upper-triangle partial tiles: fixed
// order sum over the splits, the FP64 kx0 term (weight 1/2) of a plane of
// at most KX0_FP64_MAX columns (a wider one is in the digits), mirror,
// max|S - d I|.
template <int TILE>
__device__ __forceinline__ void split_reduce_body(const oa_sreduce_args &a, int item) {
  const int n = a.n;
  const int T = cdv(n, TILE);
  const int ntiles = T * (T + 1) / 2;
  const long idx = (long)item * 256 + threadIdx.x;
  const int tile = (int)(idx / (TILE * TILE));
  if (tile >= ntiles) return;
  const int e = (int)(idx % (TILE * TILE));
  const int lm = e / TILE, ln = e % TILE;
  int bi, bj;
  upper_tile_d(tile, T, bi, bj);
  if (bi == bj && ln < lm) return;
  const int i = bi * TILE + lm, j = bj * TILE + ln;
  if (i >= n || j >= n) return;
  double s = 0.0;
  for (int z = 0; z < a.nz; z++) s += a.part[((long)z * ntiles + tile) * (TILE * TILE) + e];
  double g = 0.0;
  if (a.kx0_end <= KX0_FP64_MAX)
    for (int k = 0; k < a.kx0_end; k++)
      g = fma(a.psi[(long)i * a.ld + k], a.psi[(long)j * a.ld + k], g);
  s = fma(a.alpha, s, __dmul_rn(0.5 * a.alpha, g));
  a.S[(long)i * n + j] = s;
  a.S[(long)j * n + i] = s;
  if (a.maxdev) {
    const double dv = fabs(s - (i == j ? a.diag : 0.0));
    atomicMax(a.maxdev, (unsigned long long)__double_as_longlong(dv));
  }
}

template <int TILE>
__global__ __launch_bounds__(256) void split_reduce_kernel(oa_sreduce_args a, int total) {
  const int item = single_item(total);
  if (item < total) split_reduce_body<TILE>(a, item);
}

template <int TILE>
__global__ __launch_bounds__(256) void split_reduce_grouped(const oa_sreduce_args *__restrict__ t,
                                                            const int *__restrict__ prefix,
                                                            int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    split_reduce_body<TILE>(t[i], item - prefix[i]);
  });
}

// Asymmetric reduce: L[i][j] = alpha * sum_z part (fixed order), all tiles.
template <int TILE>
__global__ __launch_bounds__(256) void lambda_reduce_kernel(const double *__restrict__ part,
                                                            int n, int nz, double alpha,
                                                            double *L) {
  const int T = cdv(n, TILE);
  const int ntiles = T * T;
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  const int tile = (int)(idx / (TILE * TILE));
  if (tile >= ntiles) return;
  const int e = (int)(idx % (TILE * TILE));
  const int i = (tile % T) * TILE + e / TILE, j = (tile / T) * TILE + e % TILE;
  if (i >= n || j >= n) return;
  double s = 0.0;
  for (int z = 0; z < nz; z++) s += part[((long)z * ntiles + tile) * (TILE * TILE) + e];
  L[(long)i * n + j] = alpha * s;
}

// ---------------------------------------------------------------- host plan
bool decode_prec(int prec, int &dt, int &S) {
  const int base = prec & 0xff;
  S = (prec >> 8) & 0xf;
  if (S == 0) S = 3;
  if (S < 2 || S > SMAX) return false;
  if (base == OA_PREC_INT8) dt = DT_I8;
  else if (base == OA_PREC_FP16) dt = DT_F16;
  else if (base == OA_PREC_BF16) dt = DT_BF16;
  else if (base == OA_PREC_FP32) dt = DT_F32;
  else return false;
  // adaptive slice counts: int8 digits, 4..6 slices (the kernels carry the
  // per chunk level switch only there)
  if ((prec & OA_PREC_ADAPTIVE) && (dt != DT_I8 || S < ADA_SMIN_MAX)) return false;
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

// Chunk length from (n, K) alone, so that every tile shape and launch form
// sees the same digits.
int chunk_len(int dt, int S, int n, int K) {
  const int T128 = cdv(n, 128);
  const long t128 = (long)T128 * (T128 + 1) / 2;
  const int K128 = cdv(K, 128) * 128;
  // The chunk is both the split-K unit and the scale block: short chunks
  // keep each chunk's exponent close to its own coefficients (the envelope
  // falls steeply with |G|), and the int32 sums stay exact below the cap.
  long cap = (dt == DT_I8) ? (2147483647L / level_bound(S)) : (1L << 30);
  cap = cap / DBK * DBK;
  long target = KCHUNK_TARGET;
  // enough workgroups for one instance alone: at least ~256 of them
  while (target > 512 && t128 * cdv(K128, (int)target) < 256) target /= 2;
  int kchunk = (int)(target < cap ? target : cap);
  // a multiple of 128, so back-multiply output tiles never straddle chunks
  return max(128, kchunk / 128 * 128);
}

// Levels the back-multiply's X psi products keep (MODE 1).
int backmul_levels(int S, bool lean) {
  const int nl = lean ? S - 2 : S - 1;
  return nl < 1 ? 1 : nl;
}

// The int8 back-multiply contracts over all np padded states in one int32
// sum per level: level l holds l + 1 products of |digit| <= 127, so the sum
// stays exact while np * NLB * 127^2 < 2^31 (np < 33,288 at NLB = 4).
bool backmul_fits(int dt, int S, bool lean, int np) {
  if (dt != DT_I8) return true;  // 16-bit and fp32 digits promote to FP64 every 128 states
  return (long)np * backmul_levels(S, lean) * 127L * 127L < 2147483648L;
}

oa_split_plan make_plan(int prec, int n, int K, int tile) {
  oa_split_plan pl{};
  decode_prec(prec, pl.dt, pl.S);
  pl.tile = tile;
  pl.es = (pl.dt == DT_I8) ? 1 : (pl.dt == DT_F32 ? 4 : 2);
  pl.np = cdv(n, tile) * tile;
  const int kq = tile > SLT ? tile : SLT;
  pl.Kp = cdv(K, kq) * kq;
  const int T = pl.np / tile;
  pl.ntiles = T * (T + 1) / 2;
  pl.kchunk = chunk_len(pl.dt, pl.S, n, K);
  pl.nz = cdv(pl.Kp, pl.kchunk);
  pl.adaptive = (prec & OA_PREC_ADAPTIVE) ? 1 : 0;
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
  pl.off_fexp = take((size_t)pl.np * 4);
  pl.off_diag = take((size_t)pl.np * 8);
  pl.off_md = take(8);
  pl.off_part = take((size_t)pl.nz * pl.ntiles * tile * tile * 8);
  pl.off_sl = take((size_t)pl.nz * 4);
  pl.total = off;
  return pl;
}

// ---------------------------------------------------------------- launches
int device_cus() {
  static thread_local int cached_dev = -1, cached_cus = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev != cached_dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached_dev = dev;
    cached_cus = cus;
  }
  return cached_cus;
}

// Grid of a launch over `total` items: one workgroup per item (a multiple of
// 8) for single-instance launches; for grouped launches at most the number
// of co-resident workgroups, so the grid is persistent.
template <class K>
int grid_for(K kernel, int threads, int total, bool persistent) {
  int g = cdv(max(total, 1), 8) * 8;
  if (persistent) {
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    const int cap = cdv(device_cus() * per_cu, 8) * 8;
    g = min(g, cap);
  }
  return g;
}

#define OA_LAUNCH_ITEMS(KERN, THREADS, TOTAL, STREAM, ARGS)                                    \
  do {                                                                                       \
    const int _t = (TOTAL);                                                                  \
    const int _g = grid_for(KERN, THREADS, _t, false);                                       \
    hipLaunchKernelGGL(KERN, dim3(_g), dim3(THREADS), 0, STREAM, ARGS, _t);                  \
  } while (0)

// Persistent single-instance launch: as many workgroups as fit the device
// (a multiple of 8), each walking items with for_items.
#define OA_LAUNCH_PERSISTENT(KERN, THREADS, TOTAL, STREAM, ARGS)                               \
  do {                                                                                       \
    const int _t = (TOTAL);                                                                  \
    const int _g = grid_for(KERN, THREADS, _t, true);                                        \
    hipLaunchKernelGGL(KERN, dim3(_g), dim3(THREADS), 0, STREAM, ARGS, _t);                  \
  } while (0)

#define OA_LAUNCH_GROUPED(KERN, THREADS, STREAM, G, T)                                       \
  do {                                                                                       \
    const int _g = grid_for(KERN, THREADS, (G).total, true);                                 \
    hipLaunchKernelGGL(KERN, dim3(_g), dim3(THREADS), 0, STREAM,                             \
                       reinterpret_cast<const T *>((G).args), (G).prefix, (G).count);        \
  } while (0)

// Argument structs of one instance for every phase.
void fill_inst(const oa_split_plan &pl, int n, int K, const double *psi, long ld, int kx0_end,
               double alpha, double *Sout, const double *T, int transT, double beta,
               double *out, long ldout, char *ws, unsigned long long *maxS,
               unsigned long long *maxB, oa_split_inst_args *r) {
  const int tile = pl.tile;
  if (!maxS) maxS = reinterpret_cast<unsigned long long *>(ws + pl.off_maxS);
  if (!maxB) maxB = reinterpret_cast<unsigned long long *>(ws + pl.off_maxB);
  const int vec = (((uintptr_t)psi & 15) == 0 && (ld % 2) == 0) ? 1 : 0;
  r->cm = oa_chunkmax_args{psi, ld, n, K, kx0_end, pl.kchunk, pl.nz, 0, maxS, maxB};
  r->items_cm = pl.nz * cdv(n, 8);
  int *slices = pl.adaptive ? reinterpret_cast<int *>(ws + pl.off_sl) : nullptr;
  r->cnt = oa_count_args{maxB, slices, pl.nz, K, pl.kchunk, pl.S};
  r->sl = oa_slice_args{psi, ld, n, K, kx0_end, pl.kchunk, maxS, maxB, ws + pl.off_ds,
                        ws + pl.off_db, pl.np, pl.Kp, 1, vec, 1, 0, slices};
  r->items_sl = (pl.Kp / SLT) * (pl.np / SLT);
  oa_digit_args &a = r->d0;
  a = oa_digit_args{};
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
  a.chunk_max = maxS;
  a.slices = slices;
  r->items_d0 = pl.ntiles * pl.nz;
  r->rd = oa_sreduce_args{a.part, n, pl.nz, alpha, psi, ld, kx0_end, 0, Sout, 0.0, nullptr};
  r->items_rd = cdv(pl.ntiles * tile * tile, 256);
  r->rows = oa_rows_args{T, n, transT, ws + pl.off_dt, pl.np, 0,
                         reinterpret_cast<int *>(ws + pl.off_fexp),
                         reinterpret_cast<double *>(ws + pl.off_diag)};
  r->items_rows = pl.np;
  oa_digit_args &b = r->d1;
  b = oa_digit_args{};
  b.X = ws + pl.off_dt;
  b.ldx = pl.np;
  b.sx = (long)pl.np * pl.np;
  b.Y = ws + pl.off_db;
  b.ldy = pl.Kp;
  b.sy = (long)pl.Kp * pl.np;
  b.M = n;
  b.N = K;
  b.kc = pl.np;
  b.kfull = pl.np;
  b.out = out;
  b.ldo = ldout;
  b.alpha = 1.0;
  b.beta = beta;
  b.fexp = reinterpret_cast<const int *>(ws + pl.off_fexp);
  b.diagv = reinterpret_cast<const double *>(ws + pl.off_diag);
  b.DB = ws + pl.off_db;
  b.dbs = (long)pl.Kp * pl.np;
  b.np = pl.np;
  b.Kp = pl.Kp;
  b.colmax = maxB;
  b.kchunk = pl.kchunk;
  b.nz = pl.nz;
  b.slices = slices;
  r->items_d1 = cdv(n, tile) * cdv(K, tile);
}

template <int DT, int S>
void launch_slice_psi(hipStream_t s, const oa_split_plan &pl, const oa_split_inst_args &r,
                      bool want_ds) {
  hipLaunchKernelGGL(fill_u64, dim3(cdv(pl.nz, 256)), dim3(256), 0, s, r.cm.maxS, (long)pl.nz);
  hipLaunchKernelGGL(fill_u64, dim3(cdv(pl.nz, 256)), dim3(256), 0, s, r.cm.maxB, (long)pl.nz);
  OA_LAUNCH_ITEMS(chunk_max_kernel, 256, r.items_cm, s, r.cm);
  if (pl.adaptive) hipLaunchKernelGGL(slice_count_kernel, dim3(1), dim3(256), 0, s, r.cnt);
  oa_slice_args sl = r.sl;
  sl.want_ds = want_ds ? 1 : 0;
  OA_LAUNCH_ITEMS((slice_psi_kernel<DT, S>), 256, r.items_sl, s, sl);
}

template <int DT, int S>
int run_overlap(hipStream_t s, const oa_split_plan &pl, int n, int K, const double *psi,
                long ld, int kx0_end, double alpha, double *Sout, double diag, double *maxdev,
                char *ws) {
  oa_split_inst_args r;
  fill_inst(pl, n, K, psi, ld, kx0_end, alpha, Sout, nullptr, 0, 0.0, nullptr, 0, ws, nullptr,
            nullptr, &r);
  launch_slice_psi<DT, S>(s, pl, r, true);
  OA_LAUNCH_PERSISTENT((digit_nt_kernel<DT, S, 0, S - 1, 128>),
                  (DigitCfg<DT, S, 0, S - 1, 128>::NTHR), r.items_d0, s, r.d0);
  unsigned long long *md = reinterpret_cast<unsigned long long *>(maxdev);
  if (md) hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, s, md, 1L);
  r.rd.diag = diag;
  r.rd.maxdev = md;
  OA_LAUNCH_ITEMS(split_reduce_kernel<128>, 256, r.items_rd, s, r.rd);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

template <int DT, int S, int NLB>
int run_backmul(hipStream_t s, const oa_split_plan &pl, int n, int K, const double *M, int transM,
                const double *psi, long ldpsi, int kx0_end, bool reuse, double alpha,
                double beta, double *out, long ldout, char *ws, int cw, double alpha_w) {
  oa_split_inst_args r;
  fill_inst(pl, n, K, psi, ldpsi, kx0_end, 1.0, nullptr, M, transM, beta, out, ldout, ws,
            nullptr, nullptr, &r);
  r.d1.alpha = alpha;
  r.d1.cw = cw;
  r.d1.alpha_w = alpha_w;
  if (!reuse) launch_slice_psi<DT, S>(s, pl, r, false);
  OA_LAUNCH_ITEMS((slice_rows_kernel<DT, NLB>), 256, r.items_rows, s, r.rows);
  OA_LAUNCH_PERSISTENT((digit_nt_kernel<DT, S, 1, NLB, 128>), (DigitCfg<DT, S, 1, NLB, 128>::NTHR),
                  r.items_d1, s, r.d1);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

// Workspace of the asymmetric (Lambda) product: digits of psi and of F in
// the S layout (all columns), chunk maxima of both, partial tiles.
struct LambdaPlan {
  oa_split_plan pl;
  int ntiles;
  size_t off_x, off_y, off_mx, off_my, off_part, total;
};

LambdaPlan make_lambda_plan(int prec, int n, int K) {
  LambdaPlan lp;
  lp.pl = make_plan(prec, n, K, 128);
  const oa_split_plan &pl = lp.pl;
  const int T = pl.np / 128;
  lp.ntiles = T * T;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  lp.off_x = take((size_t)pl.S * pl.np * pl.Kp * pl.es);
  lp.off_y = take((size_t)pl.S * pl.np * pl.Kp * pl.es);
  lp.off_mx = take((size_t)2 * pl.nz * 8);
  lp.off_my = take((size_t)2 * pl.nz * 8);
  lp.off_part = take((size_t)pl.nz * lp.ntiles * 128 * 128 * 8);
  lp.total = off;
  return lp;
}

template <int DT, int S>
int run_lambda(hipStream_t s, const LambdaPlan &lp, int n, int K, const double *psi, long ld,
               const double *F, long ldF, double alpha, double *L, char *ws) {
  const oa_split_plan &pl = lp.pl;
  auto *mx = reinterpret_cast<unsigned long long *>(ws + lp.off_mx);
  auto *my = reinterpret_cast<unsigned long long *>(ws + lp.off_my);
  hipLaunchKernelGGL(fill_u64, dim3(cdv(4 * pl.nz, 256)), dim3(256), 0, s, mx, 2L * pl.nz);
  hipLaunchKernelGGL(fill_u64, dim3(cdv(4 * pl.nz, 256)), dim3(256), 0, s, my, 2L * pl.nz);
  const int items_cm = pl.nz * cdv(n, 8), items_sl = (pl.Kp / SLT) * (pl.np / SLT);
  const struct { const double *x; long ld; unsigned long long *m; size_t off; } ops[2] = {
      {psi, ld, mx, lp.off_x}, {F, ldF, my, lp.off_y}};
  for (const auto &o : ops) {
    const int vec = (((uintptr_t)o.x & 15) == 0 && (o.ld % 2) == 0) ? 1 : 0;
    oa_chunkmax_args cm{o.x, o.ld, n, K, 0, pl.kchunk, pl.nz, 0, o.m, o.m + pl.nz};
    OA_LAUNCH_ITEMS(chunk_max_kernel, 256, items_cm, s, cm);
    oa_slice_args sl{o.x, o.ld, n, K, 0, pl.kchunk, o.m, o.m + pl.nz, ws + o.off, nullptr,
                     pl.np, pl.Kp, 1, vec, 0, 0};
    OA_LAUNCH_ITEMS((slice_psi_kernel<DT, S>), 256, items_sl, s, sl);
  }
  oa_digit_args a{};
  a.X = ws + lp.off_x;
  a.Y = ws + lp.off_y;
  a.ldx = a.ldy = pl.np;
  a.sx = a.sy = (long)pl.np * pl.Kp;
  a.M = n;
  a.N = n;
  a.kc = pl.kchunk;
  a.kfull = pl.Kp;
  a.ntiles = lp.ntiles;
  a.nz = pl.nz;
  a.part = reinterpret_cast<double *>(ws + lp.off_part);
  a.chunk_max = mx;
  a.colmax = my;
  OA_LAUNCH_PERSISTENT((digit_nt_kernel<DT, S, 2, S - 1, 128>),
                  (DigitCfg<DT, S, 2, S - 1, 128>::NTHR), lp.ntiles * pl.nz, s, a);
  const long total = (long)lp.ntiles * 128 * 128;
  hipLaunchKernelGGL(lambda_reduce_kernel<128>, dim3((unsigned)cdv((int)total, 256)), dim3(256),
                     0, s, (const double *)a.part, n, pl.nz, alpha, L);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

template <int DT>
int dispatch_lambda(hipStream_t st, const LambdaPlan &lp, int n, int K, const double *psi,
                    long ld, const double *F, long ldF, double alpha, double *L, char *ws) {
  switch (lp.pl.S) {
    case 2: return run_lambda<DT, 2>(st, lp, n, K, psi, ld, F, ldF, alpha, L, ws);
    case 3: return run_lambda<DT, 3>(st, lp, n, K, psi, ld, F, ldF, alpha, L, ws);
    case 4: return run_lambda<DT, 4>(st, lp, n, K, psi, ld, F, ldF, alpha, L, ws);
    case 5: return run_lambda<DT, 5>(st, lp, n, K, psi, ld, F, ldF, alpha, L, ws);
    case 6: return run_lambda<DT, 6>(st, lp, n, K, psi, ld, F, ldF, alpha, L, ws);
  }
  return OA_ERR_UNSUPPORTED;
}

template <int DT>
int dispatch_overlap(hipStream_t st, const oa_split_plan &pl, int n, int K, const double *psi,
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
int dispatch_backmul(hipStream_t st, const oa_split_plan &pl, bool lean, int n, int K,
                     const double *M, int transM, const double *psi, long ldpsi, int kx0_end,
                     bool reuse, double alpha, double beta, double *out, long ldout, char *ws,
                     int cw, double alpha_w) {
#define OA_BM(SS, NL)                                                                        \
  return run_backmul<DT, SS, NL>(st, pl, n, K, M, transM, psi, ldpsi, kx0_end, reuse, alpha, \
                                 beta, out, ldout, ws, cw, alpha_w);
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

// In place: psi's digits in the back-multiply layout become J psi's, J (re,
// im) = (-im, re).  Columns 2c and 2c+1 of one 32-byte contraction block are
// adjacent 32-byte rows of DB, so every 64-byte unit has its halves swapped
// and its new first half negated.  The digits of -x are the negated digits
// of x (rint is odd), and a chunk holds whole (re, im) pairs, so the chunk
// maxima stay valid.  Thread t takes 16 bytes of each half of unit t / 2.
template <int DT>
__device__ __forceinline__ i4v neg_digits(i4v v) {
  if constexpr (DT == DT_I8) {
    typedef signed char c16 __attribute__((ext_vector_type(16)));
    return __builtin_bit_cast(i4v, -__builtin_bit_cast(c16, v));
  } else if constexpr (DT == DT_F32) {
    return v ^ (int)0x80000000;
  } else {
    return v ^ (int)0x80008000;
  }
}

// The same map on the S layout, out of place: there columns 2c and 2c+1 are
// adjacent elements of a 32-byte row, so every 16 bytes hold whole pairs.
template <int DT>
__device__ __forceinline__ i4v jpairs(i4v v) {
  if constexpr (DT == DT_I8) {
    typedef signed char c16 __attribute__((ext_vector_type(16)));
    const c16 x = __builtin_bit_cast(c16, v), m = -x;
    return __builtin_bit_cast(i4v, __builtin_shufflevector(x, m, 17, 0, 19, 2, 21, 4, 23, 6, 25,
                                                           8, 27, 10, 29, 12, 31, 14));
  } else if constexpr (DT == DT_F32) {
    const i4v m = v ^ (int)0x80000000;
    return __builtin_shufflevector(v, m, 5, 0, 7, 2);
  } else {
    typedef short s8v __attribute__((ext_vector_type(8)));
    const s8v x = __builtin_bit_cast(s8v, v), m = x ^ (short)0x8000;
    return __builtin_bit_cast(i4v, __builtin_shufflevector(x, m, 9, 0, 11, 2, 13, 4, 15, 6));
  }
}

template <int DT>
__global__ __launch_bounds__(256) void jswap_ds_kernel(const char *src, char *dst, long chunks) {
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < chunks; t += (long)gridDim.x * 256)
    *reinterpret_cast<i4v *>(dst + t * 16) = jpairs<DT>(*reinterpret_cast<const i4v *>(src + t * 16));
}

template <int DT>
__global__ __launch_bounds__(256) void jswap_db_kernel(char *DB, long units) {
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < 2 * units;
       t += (long)gridDim.x * 256) {
    char *u = DB + (t >> 1) * 64 + (t & 1) * 16;
    const i4v a = *reinterpret_cast<const i4v *>(u);
    const i4v b = *reinterpret_cast<const i4v *>(u + 32);
    *reinterpret_cast<i4v *>(u) = neg_digits<DT>(b);
    *reinterpret_cast<i4v *>(u + 32) = a;
  }
}

// ------------------------------------------------------------ grouped launches
template <int DT, int S, int NLB, int TILE>
int grouped_phase(hipStream_t s, int phase, const oa_group &g) {
  switch (phase) {
    case OA_SP_CHUNKMAX:
      OA_LAUNCH_GROUPED(chunk_max_grouped, 256, s, g, oa_chunkmax_args);
      break;
    case OA_SP_COUNT:
      OA_LAUNCH_GROUPED(slice_count_grouped, 256, s, g, oa_count_args);
      break;
    case OA_SP_SLICE:
      OA_LAUNCH_GROUPED((slice_psi_grouped<DT, S>), 256, s, g, oa_slice_args);
      break;
    case OA_SP_DIGIT0:
      OA_LAUNCH_GROUPED((digit_nt_grouped<DT, S, 0, S - 1, TILE>),
                        (DigitCfg<DT, S, 0, S - 1, TILE>::NTHR), s, g, oa_digit_args);
      break;
    case OA_SP_REDUCE:
      OA_LAUNCH_GROUPED(split_reduce_grouped<TILE>, 256, s, g, oa_sreduce_args);
      break;
    case OA_SP_ROWS:
      OA_LAUNCH_GROUPED((slice_rows_grouped<DT, NLB>), 256, s, g, oa_rows_args);
      break;
    case OA_SP_DIGIT1:
      OA_LAUNCH_GROUPED((digit_nt_grouped<DT, S, 1, NLB, TILE>),
                        (DigitCfg<DT, S, 1, NLB, TILE>::NTHR), s, g, oa_digit_args);
      break;
    default:
      return OA_ERR_UNSUPPORTED;
  }
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

template <int DT, int TILE>
int grouped_s(hipStream_t s, int S, bool lean, int phase, const oa_group &g) {
#define OA_GP(SS, NL) return grouped_phase<DT, SS, NL, TILE>(s, phase, g);
  switch (S) {
    case 2: OA_GP(2, 1)
    case 3: if (lean) { OA_GP(3, 1) } else { OA_GP(3, 2) }
    case 4: if (lean) { OA_GP(4, 2) } else { OA_GP(4, 3) }
    case 5: if (lean) { OA_GP(5, 3) } else { OA_GP(5, 4) }
    case 6: if (lean) { OA_GP(6, 4) } else { OA_GP(6, 5) }
  }
#undef OA_GP
  return OA_ERR_UNSUPPORTED;
}

template <int DT>
int grouped_dt(hipStream_t s, int S, bool lean, int tile, int phase, const oa_group &g) {
  if (tile == 128) return grouped_s<DT, 128>(s, S, lean, phase, g);
  if (tile == 64) return grouped_s<DT, 64>(s, S, lean, phase, g);
  return OA_ERR_UNSUPPORTED;
}

}  // namespace

oa_split_plan oa_split_plan_for(int prec, int n, int K, int tile) {
  return make_plan(prec & 0xffff, n, K, tile);
}

int oa_split_check(int prec, int n, int tile) {
  int dt, S;
  if (!decode_prec(prec & 0xffff, dt, S)) return OA_ERR_UNSUPPORTED;
  const int np = cdv(n, tile) * tile;
  return backmul_fits(dt, S, (prec & OA_PREC_LEAN_BACKMUL) != 0, np) ? OA_SUCCESS
                                                                     : OA_ERR_CAPACITY;
}

void oa_split_inst_fill(const oa_split_plan &pl, int prec, int n, int K, const double *psi,
                        long ld, int kx0_end, double alpha, double *S, const double *T,
                        int transT, double beta, double *out, long ldout, char *ws,
                        unsigned long long *maxS, unsigned long long *maxB,
                        oa_split_inst_args *a) {
  (void)prec;
  fill_inst(pl, n, K, psi, ld, kx0_end, alpha, S, T, transT, beta, out, ldout, ws, maxS, maxB,
            a);
}

int oa_split_grouped(hipStream_t s, int prec, int tile, int phase, const oa_group &g) {
  int dt, S;
  if (!decode_prec(prec & 0xffff, dt, S)) return OA_ERR_UNSUPPORTED;
  if (g.count <= 0 || g.total <= 0) return OA_SUCCESS;
  const bool lean = (prec & OA_PREC_LEAN_BACKMUL) != 0;
  if (dt == DT_I8) return grouped_dt<DT_I8>(s, S, lean, tile, phase, g);
  if (dt == DT_F16) return grouped_dt<DT_F16>(s, S, lean, tile, phase, g);
  if (dt == DT_F32) return grouped_dt<DT_F32>(s, S, lean, tile, phase, g);
  return grouped_dt<DT_BF16>(s, S, lean, tile, phase, g);
}

int oa_launch_slice_counts(hipStream_t s, int prec, int n, int K, const double *psi, long ld,
                           int kx0_end, void *ws, size_t ws_bytes, const int **counts,
                           int *nz, int *kchunk) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl) || !(prec & OA_PREC_ADAPTIVE))
    return OA_ERR_UNSUPPORTED;
  const oa_split_plan pl = make_plan(prec & 0xffff, n, K, 128);
  if (pl.total > ws_bytes) return OA_ERR_CAPACITY;
  oa_split_inst_args r;
  char *w = (char *)ws;
  fill_inst(pl, n, K, psi, ld, kx0_end, 1.0, nullptr, nullptr, 0, 0.0, nullptr, 0, w, nullptr,
            nullptr, &r);
  hipLaunchKernelGGL(fill_u64, dim3(cdv(pl.nz, 256)), dim3(256), 0, s, r.cm.maxS, (long)pl.nz);
  hipLaunchKernelGGL(fill_u64, dim3(cdv(pl.nz, 256)), dim3(256), 0, s, r.cm.maxB, (long)pl.nz);
  OA_LAUNCH_ITEMS(chunk_max_kernel, 256, r.items_cm, s, r.cm);
  hipLaunchKernelGGL(slice_count_kernel, dim3(1), dim3(256), 0, s, r.cnt);
  *counts = r.cnt.slices;
  *nz = pl.nz;
  *kchunk = pl.kchunk;
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

int oa_fill_u64(hipStream_t s, unsigned long long *p, long n) {
  if (n <= 0) return OA_SUCCESS;
  const long nb = (n + 255) / 256;
  hipLaunchKernelGGL(fill_u64, dim3((unsigned)(nb > 1024 ? 1024 : nb)), dim3(256), 0, s, p, n);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

size_t oa_split_ws_bytes(int prec, int n, int K) {
  int dt, S;
  if ((prec & 0xff) == OA_PREC_FP64 || !decode_prec(prec & 0xffff, dt, S)) return 0;
  return make_plan(prec & 0xffff, n, K, 128).total;
}

int oa_launch_overlap_split(hipStream_t s, int prec, int n, int K, const double *psi, long ld,
                            int kx0_end, double alpha, double *S, double diag, double *maxdev,
                            void *ws, size_t ws_bytes) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return OA_ERR_UNSUPPORTED;
  const oa_split_plan pl = make_plan(prec & 0xffff, n, K, 128);
  if (pl.total > ws_bytes) return OA_ERR_CAPACITY;
  char *w = (char *)ws;
  if (dt == DT_I8) return dispatch_overlap<DT_I8>(s, pl, n, K, psi, ld, kx0_end, alpha, S, diag, maxdev, w);
  if (dt == DT_F16) return dispatch_overlap<DT_F16>(s, pl, n, K, psi, ld, kx0_end, alpha, S, diag, maxdev, w);
  if (dt == DT_F32) return dispatch_overlap<DT_F32>(s, pl, n, K, psi, ld, kx0_end, alpha, S, diag, maxdev, w);
  return dispatch_overlap<DT_BF16>(s, pl, n, K, psi, ld, kx0_end, alpha, S, diag, maxdev, w);
}

size_t oa_split_lambda_ws_bytes(int prec, int n, int K) {
  int dt, S;
  if ((prec & 0xff) == OA_PREC_FP64 || !decode_prec(prec & 0xffff, dt, S)) return 0;
  return make_lambda_plan(prec & 0xffff, n, K).total;
}

int oa_launch_lambda_split(hipStream_t s, int prec, int n, int K, const double *psi, long ld,
                           const double *F, long ldF, double alpha, double *L, void *ws,
                           size_t ws_bytes) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return OA_ERR_UNSUPPORTED;
  if (prec & OA_PREC_ADAPTIVE) return OA_ERR_UNSUPPORTED;  // uniform slices only
  const LambdaPlan lp = make_lambda_plan(prec & 0xffff, n, K);
  if (lp.total > ws_bytes) return OA_ERR_CAPACITY;
  char *w = (char *)ws;
  if (dt == DT_I8) return dispatch_lambda<DT_I8>(s, lp, n, K, psi, ld, F, ldF, alpha, L, w);
  if (dt == DT_F16) return dispatch_lambda<DT_F16>(s, lp, n, K, psi, ld, F, ldF, alpha, L, w);
  if (dt == DT_F32) return dispatch_lambda<DT_F32>(s, lp, n, K, psi, ld, F, ldF, alpha, L, w);
  return dispatch_lambda<DT_BF16>(s, lp, n, K, psi, ld, F, ldF, alpha, L, w);
}

int oa_launch_jswap_db(hipStream_t s, int prec, int n, int K, void *ws, size_t ws_bytes) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return OA_ERR_UNSUPPORTED;
  if (prec & OA_PREC_ADAPTIVE) return OA_ERR_UNSUPPORTED;  // uniform slices only
  const oa_split_plan pl = make_plan(prec & 0xffff, n, K, 128);
  if (pl.total > ws_bytes) return OA_ERR_CAPACITY;
  // DB: S slices of Kp x np digits; Kp and np are multiples of 128, so whole
  // 64-byte units
  const long units = (long)pl.S * pl.Kp * pl.np * pl.es / 64;
  if (units <= 0) return OA_SUCCESS;
  const unsigned g = (unsigned)std::min<long>((2 * units + 255) / 256, 8192);
  char *db = (char *)ws + pl.off_db;
  if (dt == DT_I8) hipLaunchKernelGGL(jswap_db_kernel<DT_I8>, dim3(g), dim3(256), 0, s, db, units);
  else if (dt == DT_F32)
    hipLaunchKernelGGL(jswap_db_kernel<DT_F32>, dim3(g), dim3(256), 0, s, db, units);
  else hipLaunchKernelGGL(jswap_db_kernel<DT_F16>, dim3(g), dim3(256), 0, s, db, units);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

// Im S = alpha psi (J psi)^T on the digits a split overlap of the same
// [n][K] block (no kx = 0 columns) left in ws: J psi's S-layout digits are
// made from psi's (jswap_ds_kernel) in the area after the overlap plan; the
// split Lambda's slicing would give the same digits and chunk maxima.  Im S
// is exactly antisymmetric in digit arithmetic and a Hermitian S needs only
// its strict upper triangle, so the product runs on the upper tiles (the
// symmetric kernel's tile map, X = psi, Y = J psi, both exponents psi's)
// and its fixed-order reduce writes the upper triangle of Li as the
// all-tiles product would (the mirrored lower triangle is not Im S).
struct ZimagPlan {
  oa_split_plan pl;
  int ntiles;
  size_t off_yj, off_part, total;
};

ZimagPlan make_zimag_plan(int prec, int n, int K) {
  ZimagPlan zp;
  zp.pl = make_plan(prec, n, K, 128);
  const oa_split_plan &pl = zp.pl;
  const int T = pl.np / 128;
  zp.ntiles = T * (T + 1) / 2;
  zp.off_yj = (pl.total + 255) / 256 * 256;
  zp.off_part = zp.off_yj + ((size_t)pl.S * pl.np * pl.Kp * pl.es + 255) / 256 * 256;
  zp.total = zp.off_part + (size_t)pl.nz * zp.ntiles * 128 * 128 * 8;
  return zp;
}

template <int DT, int S>
int run_zimag(hipStream_t s, const ZimagPlan &zp, int n, double alpha, double *Li, char *ws) {
  const oa_split_plan &pl = zp.pl;
  const long chunks = (long)pl.S * pl.np * pl.Kp * pl.es / 16;
  const unsigned g = (unsigned)std::min<long>((chunks + 255) / 256, 8192);
  hipLaunchKernelGGL(jswap_ds_kernel<DT>, dim3(g), dim3(256), 0, s, ws + pl.off_ds,
                     ws + zp.off_yj, chunks);
  auto *mS = reinterpret_cast<unsigned long long *>(ws + pl.off_maxS);
  oa_digit_args a{};
  a.X = ws + pl.off_ds;
  a.Y = ws + zp.off_yj;
  a.ldx = a.ldy = pl.np;
  a.sx = a.sy = (long)pl.np * pl.Kp;
  a.M = n;
  a.N = n;
  a.kc = pl.kchunk;
  a.kfull = pl.Kp;
  a.ntiles = zp.ntiles;
  a.nz = pl.nz;
  a.part = reinterpret_cast<double *>(ws + zp.off_part);
  a.chunk_max = mS;
  OA_LAUNCH_PERSISTENT((digit_nt_kernel<DT, S, 0, S - 1, 128>),
                  (DigitCfg<DT, S, 0, S - 1, 128>::NTHR), zp.ntiles * pl.nz, s, a);
  const oa_sreduce_args rd{a.part, n, pl.nz, alpha, nullptr, 0, 0, 0, Li, 0.0, nullptr};
  OA_LAUNCH_ITEMS(split_reduce_kernel<128>, 256, cdv(zp.ntiles * 128 * 128, 256), s, rd);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

template <int DT>
int dispatch_zimag(hipStream_t st, const ZimagPlan &zp, int n, double alpha, double *Li,
                   char *ws) {
  switch (zp.pl.S) {
    case 2: return run_zimag<DT, 2>(st, zp, n, alpha, Li, ws);
    case 3: return run_zimag<DT, 3>(st, zp, n, alpha, Li, ws);
    case 4: return run_zimag<DT, 4>(st, zp, n, alpha, Li, ws);
    case 5: return run_zimag<DT, 5>(st, zp, n, alpha, Li, ws);
    case 6: return run_zimag<DT, 6>(st, zp, n, alpha, Li, ws);
  }
  return OA_ERR_UNSUPPORTED;
}

size_t oa_split_zimag_ws_bytes(int prec, int n, int K) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return 0;
  return make_zimag_plan(prec & 0xffff, n, K).total;
}

int oa_launch_zimag_split(hipStream_t s, int prec, int n, int K, double alpha, double *Li,
                          void *ws, size_t ws_bytes) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return OA_ERR_UNSUPPORTED;
  if (prec & OA_PREC_ADAPTIVE) return OA_ERR_UNSUPPORTED;  // uniform slices only
  const ZimagPlan zp = make_zimag_plan(prec & 0xffff, n, K);
  if (zp.total > ws_bytes) return OA_ERR_CAPACITY;
  char *w = (char *)ws;
  if (dt == DT_I8) return dispatch_zimag<DT_I8>(s, zp, n, alpha, Li, w);
  if (dt == DT_F16) return dispatch_zimag<DT_F16>(s, zp, n, alpha, Li, w);
  if (dt == DT_F32) return dispatch_zimag<DT_F32>(s, zp, n, alpha, Li, w);
  return dispatch_zimag<DT_BF16>(s, zp, n, alpha, Li, w);
}

int oa_launch_backmul_split(hipStream_t s, int prec, int n, int K, const double *M, int transM,
                            const double *psi, long ldpsi, int kx0_end, double alpha, double beta,
                            double *out, long ldout, void *ws, size_t ws_bytes, int cw,
                            double alpha_w) {
  int dt, Sl;
  if (!decode_prec(prec & 0xffff, dt, Sl)) return OA_ERR_UNSUPPORTED;
  const oa_split_plan pl = make_plan(prec & 0xffff, n, K, 128);
  if (pl.total > ws_bytes) return OA_ERR_CAPACITY;
  const bool reuse = (prec & OA_PREC_REUSE_DIGITS) != 0;
  const bool lean = (prec & OA_PREC_LEAN_BACKMUL) != 0;
  if (!backmul_fits(pl.dt, pl.S, lean, pl.np)) return OA_ERR_CAPACITY;
  char *w = (char *)ws;
  if (dt == DT_I8)
    return dispatch_backmul<DT_I8>(s, pl, lean, n, K, M, transM, psi, ldpsi, kx0_end, reuse,
                                   alpha, beta, out, ldout, w, cw, alpha_w);
  if (dt == DT_F16)
    return dispatch_backmul<DT_F16>(s, pl, lean, n, K, M, transM, psi, ldpsi, kx0_end, reuse,
                                    alpha, beta, out, ldout, w, cw, alpha_w);
  if (dt == DT_F32)
    return dispatch_backmul<DT_F32>(s, pl, lean, n, K, M, transM, psi, ldpsi, kx0_end, reuse,
                                    alpha, beta, out, ldout, w, cw, alpha_w);
  return dispatch_backmul<DT_BF16>(s, pl, lean, n, K, M, transM, psi, ldpsi, kx0_end, reuse,
                                   alpha, beta, out, ldout, w, cw, alpha_w);
}
