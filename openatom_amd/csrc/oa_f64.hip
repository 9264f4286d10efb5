// FP64-MFMA kernels of the orthonormalization path (gfx950).
//
//   gemm_f64_kernel     row-major C = alpha*op(X)*op(Y) + beta*C + d*I, the
//                       back-multiply T*psi / Lambda*psi and the S->T GEMMs.
//   ovl_f64_kernel      split-K product over the plane-wave index, upper
//                       triangle tiles only for the symmetric overlap (and
//                       all tiles for the asymmetric Lambda product).
//   ovl_reduce_kernel   fixed-order sum of the split-K partials with the kx=0
//                       weight, mirror to the lower triangle, max|S - d I|.
//   ns_* kernels        guarded Newton-Schulz S->T iteration (device flags).
//
// Tiling: 256-thread workgroups (4 wave64s in a 2x2 grid), each wave owns a
// (16*WM) x (16*WM) block of v_mfma_f64_16x16x4f64 accumulators.  Operand
// lanes: lane l supplies A[l%16][l/16] and B[l/16][l%16]; result register v
// of lane l holds D[l/16 + 4v][l%16] (measured on gfx950).  Operand
// panels of BK=16 doubles are staged through LDS with a row stride of 18
// doubles, which makes the 16-byte fragment reads of 16 lanes cover all 64
// banks once.  Each lane reads two consecutive k per LDS access and feeds
// them to two MFMAs; the k order inside a panel is therefore permuted the
// same way for both operands, which leaves every product sum unchanged.
#include "oa_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

constexpr int NT = 256;
constexpr int BK = 16;
constexpr int LS = BK + 2;

__host__ __device__ __forceinline__ int cdiv(int a, int b) { return (a + b - 1) / b; }

// The dispatcher places workgroup b on XCD b % 8.  Give each XCD a contiguous
// range of logical ids so workgroups sharing operand panels share one L2.
__device__ __forceinline__ int xcd_logical(int b, int nb) {
  const int per = nb >> 3;
  return (b & 7) * per + (b >> 3);
}

// Per-thread loader of one BR x BK operand panel, 16 bytes per load.
//   KC  (element (r, k) at p[r*ld + k], k contiguous): a row holds 8 chunks;
//       thread t owns chunk (t % 8) of rows t/8 + 32j.
//   !KC (element (r, k) at p[k*ld + r], r contiguous): a k-row holds BR/2
//       chunks; thread t owns chunk (t % (BR/2)) of k-rows t/(BR/2) + NT/(BR/2)*j.
// load_fast() needs a 16-byte aligned, fully in-range panel; load_masked()
// reads element by element and zero-fills outside [rlim) x [klim).
template <int BR, bool KC>
struct PanelLoader {
  static constexpr int CH = BR * BK / 2 / NT;
  static constexpr int KSTEP = KC ? 0 : NT / (BR / 2);
  const double *base;  // matrix origin
  long ld;
  int rb, kb;          // this thread's first (row, k) inside the panel
  d2v v[CH];

  __device__ __forceinline__ void init(const double *p, long ld_, int tid) {
    base = p;
    ld = ld_;
    if (KC) { rb = tid / 8; kb = (tid % 8) * 2; }
    else { rb = (tid % (BR / 2)) * 2; kb = tid / (BR / 2); }
  }
  __device__ __forceinline__ void load_fast(int r0, int k0) {
    if (KC) {
      const double *q = base + (long)(r0 + rb) * ld + k0 + kb;
#pragma unroll
      for (int j = 0; j < CH; j++) v[j] = *reinterpret_cast<const d2v *>(q + (long)(32 * j) * ld);
    } else {
      const double *q = base + (long)(k0 + kb) * ld + r0 + rb;
#pragma unroll
      for (int j = 0; j < CH; j++)
        v[j] = *reinterpret_cast<const d2v *>(q + (long)(KSTEP * j) * ld);
    }
  }
  __device__ __forceinline__ void load_masked(int r0, int rlim, int k0, int klim) {
#pragma unroll
    for (int j = 0; j < CH; j++)
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int r = KC ? r0 + rb + 32 * j : r0 + rb + e;
        const int k = KC ? k0 + kb + e : k0 + kb + KSTEP * j;
        v[j][e] = (r < rlim && k < klim) ? base[KC ? (long)r * ld + k : (long)k * ld + r] : 0.0;
      }
  }
  __device__ __forceinline__ void store(double *s) const {
    if (KC) {
#pragma unroll
      for (int j = 0; j < CH; j++) *reinterpret_cast<d2v *>(s + (rb + 32 * j) * LS + kb) = v[j];
    } else {
#pragma unroll
      for (int j = 0; j < CH; j++) {
        s[rb * LS + kb + KSTEP * j] = v[j][0];
        s[(rb + 1) * LS + kb + KSTEP * j] = v[j][1];
      }
    }
  }
};

// One BK-deep LDS stage of MFMAs for a wave's WM x WM block grid.
template <int WM>
__device__ __forceinline__ void mma_stage(const double *Xs, const double *Ys, int xrow0,
                                          int yrow0, int lane, d4 (&acc)[WM][WM]) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int kk2 = 0; kk2 < BK / 8; kk2++) {
    d2v a[WM], b[WM];
#pragma unroll
    for (int i = 0; i < WM; i++)
      a[i] = *reinterpret_cast<const d2v *>(Xs + (xrow0 + i * 16 + r) * LS + kk2 * 8 + 2 * g);
#pragma unroll
    for (int j = 0; j < WM; j++)
      b[j] = *reinterpret_cast<const d2v *>(Ys + (yrow0 + j * 16 + r) * LS + kk2 * 8 + 2 * g);
#pragma unroll
    for (int e = 0; e < 2; e++)
#pragma unroll
      for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WM; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][e], b[j][e], acc[i][j], 0, 0, 0);
  }
}

// Main K loop shared by every FP64 kernel.  Two LDS buffers per operand and
// one barrier per BK stage: stage s+1 is fetched into registers while stage s
// feeds the MFMAs, then written to the other buffer.  `fast` selects 16-byte
// unmasked loads (interior, aligned tiles); the last stage falls back to
// masked loads when K is not a multiple of BK.  When YSAME the Y panel is the
// X panel (diagonal tile of a symmetric product) and is loaded once.
template <int WM, bool XKC, bool YKC, bool YSAME>
__device__ __forceinline__ void kloop(const double *X, long ldx, int m0, int mlim,
                                      const double *Y, long ldy, int n0, int nlim, int ka,
                                      int kb, bool fast, double *smem, d4 (&acc)[WM][WM]) {
  constexpr int BM = 32 * WM;
  constexpr int PANEL = BM * LS;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xrow0 = (w >> 1) * 16 * WM, yrow0 = (w & 1) * 16 * WM;
  PanelLoader<BM, XKC> lx;
  PanelLoader<BM, YKC> ly;
  lx.init(X, ldx, tid);
  if (!YSAME) ly.init(Y, ldy, tid);
  const int nst = (kb - ka + BK - 1) / BK;
  if (nst <= 0) return;
  auto fetch = [&](int k0) {
    if (fast && k0 + BK <= kb) {
      lx.load_fast(m0, k0);
      if (!YSAME) ly.load_fast(n0, k0);
    } else {
      lx.load_masked(m0, mlim, k0, kb);
      if (!YSAME) ly.load_masked(n0, nlim, k0, kb);
    }
  };
  // buffers: X at smem + b*PANEL, Y at smem + (2 + b)*PANEL (plain pointer
  // arithmetic keeps the LDS address space visible to the compiler)
  fetch(ka);
  lx.store(smem);
  if (!YSAME) ly.store(smem + 2 * PANEL);
  __syncthreads();
  for (int s = 0; s < nst; s++) {
    const int cur = (s & 1) * PANEL, nxt = PANEL - cur;
    const bool more = s + 1 < nst;
    if (more) fetch(ka + (s + 1) * BK);
    mma_stage<WM>(smem + cur, YSAME ? smem + cur : smem + 2 * PANEL + cur, xrow0, yrow0, lane,
                  acc);
    if (more) {
      lx.store(smem + nxt);
      if (!YSAME) ly.store(smem + 2 * PANEL + nxt);
    }
    __syncthreads();
  }
}

// kloop with the loads issued two stages ahead in two register sets, for the
// 64 x 64 tiles of the small n x n products (S->T, small back-multiplies),
// where one stage of MFMA work is shorter than a global load's latency.  The
// stages and their products run in kloop's order, so the sums are
// bit-identical.  Register set s % 2 holds stage s + 1 when stage s starts.
template <int WM, bool XKC, bool YKC>
__device__ __forceinline__ void kloop_pf2(const double *X, long ldx, int m0, int mlim,
                                          const double *Y, long ldy, int n0, int nlim, int ka,
                                          int kb, bool fast, double *smem, d4 (&acc)[WM][WM]) {
  constexpr int BM = 32 * WM;
  constexpr int PANEL = BM * LS;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xrow0 = (w >> 1) * 16 * WM, yrow0 = (w & 1) * 16 * WM;
  PanelLoader<BM, XKC> lx0, lx1;
  PanelLoader<BM, YKC> ly0, ly1;
  lx0.init(X, ldx, tid);
  lx1.init(X, ldx, tid);
  ly0.init(Y, ldy, tid);
  ly1.init(Y, ldy, tid);
  const int nst = (kb - ka + BK - 1) / BK;
  if (nst <= 0) return;
  auto fetch = [&](PanelLoader<BM, XKC> &lx, PanelLoader<BM, YKC> &ly, int k0) {
    if (fast && k0 + BK <= kb) {
      lx.load_fast(m0, k0);
      ly.load_fast(n0, k0);
    } else {
      lx.load_masked(m0, mlim, k0, kb);
      ly.load_masked(n0, nlim, k0, kb);
    }
  };
  fetch(lx0, ly0, ka);
  lx0.store(smem);
  ly0.store(smem + 2 * PANEL);
  if (nst > 1) fetch(lx0, ly0, ka + BK);
  if (nst > 2) fetch(lx1, ly1, ka + 2 * BK);
  __syncthreads();
  auto step = [&](int s, PanelLoader<BM, XKC> &lx, PanelLoader<BM, YKC> &ly) {
    const int cur = (s & 1) * PANEL, nxt = PANEL - cur;
    mma_stage<WM>(smem + cur, smem + 2 * PANEL + cur, xrow0, yrow0, lane, acc);
    if (s + 1 < nst) {
      lx.store(smem + nxt);
      ly.store(smem + 2 * PANEL + nxt);
      if (s + 3 < nst) fetch(lx, ly, ka + (s + 3) * BK);
    }
    __syncthreads();
  };
  for (int s = 0; s < nst; s += 2) {
    step(s, lx0, ly0);
    if (s + 1 < nst) step(s + 1, lx1, ly1);
  }
}

__device__ __forceinline__ bool aligned16(const void *p) {
  return ((uintptr_t)p & 15) == 0;
}

// Field-by-field selection keeps the problem in scalar registers (indexing
// the kernel-argument array with blockIdx.y would copy it to scratch).
struct ProbRegs {
  int M, N, K;
  const double *X;
  long ldx;
  const double *Y;
  long ldy;
  double *C;
  long ldc;
  double alpha, beta, diag_add;
  int cw;
  double alpha_w;
};

__device__ __forceinline__ ProbRegs pick_problem(const oa_gemm_problem &p0,
                                                 const oa_gemm_problem &p1) {
  const bool y = blockIdx.y != 0;
  ProbRegs r;
  r.M = y ? p1.M : p0.M;
  r.N = y ? p1.N : p0.N;
  r.K = y ? p1.K : p0.K;
  r.X = y ? p1.X : p0.X;
  r.ldx = y ? p1.ldx : p0.ldx;
  r.Y = y ? p1.Y : p0.Y;
  r.ldy = y ? p1.ldy : p0.ldy;
  r.C = y ? p1.C : p0.C;
  r.ldc = y ? p1.ldc : p0.ldc;
  r.alpha = y ? p1.alpha : p0.alpha;
  r.beta = y ? p1.beta : p0.beta;
  r.diag_add = y ? p1.diag_add : p0.diag_add;
  r.cw = y ? p1.cw : p0.cw;
  r.alpha_w = y ? p1.alpha_w : p0.alpha_w;
  return r;
}

// One (BM x BM) output tile `logical` of C = alpha*op(X)*op(Y) + beta*C + d*I.
template <bool TX, bool TY, int WM>
__device__ __forceinline__ void gemm_body(const ProbRegs &P, int logical, double *smem) {
  constexpr int BM = 32 * WM;
  const int tmn = cdiv(P.M, BM), tnn = cdiv(P.N, BM);
  if (logical >= tmn * tnn) return;
  const int tm = logical % tmn, tn = logical / tmn;
  const int m0 = tm * BM, n0 = tn * BM;
  d4 acc[WM][WM];
#pragma unroll
  for (int i = 0; i < WM; i++)
#pragma unroll
    for (int j = 0; j < WM; j++) acc[i][j] = d4{0, 0, 0, 0};
  const bool fast = m0 + BM <= P.M && n0 + BM <= P.N && aligned16(P.X) && aligned16(P.Y) &&
                    (P.ldx % 2) == 0 && (P.ldy % 2) == 0;
  if constexpr (WM == 2)
    kloop_pf2<WM, !TX, TY>(P.X, P.ldx, m0, P.M, P.Y, P.ldy, n0, P.N, 0, P.K, fast, smem, acc);
  else
    kloop<WM, !TX, TY, false>(P.X, P.ldx, m0, P.M, P.Y, P.ldy, n0, P.N, 0, P.K, fast, smem,
                              acc);

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xrow0 = (w >> 1) * 16 * WM, yrow0 = (w & 1) * 16 * WM;
  const int r = lane & 15, g = lane >> 4;
  const double alpha = P.alpha, beta = P.beta, dadd = P.diag_add;
  if (beta == 0.0) {
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = m0 + xrow0 + i * 16 + g + 4 * v;
        if (m >= P.M) continue;
        double *crow = P.C + (long)m * P.ldc;
#pragma unroll
        for (int j = 0; j < WM; j++) {
          const int n = n0 + yrow0 + j * 16 + r;
          const double al = (n < P.cw) ? P.alpha_w : alpha;
          // explicit fma: single and grouped launches inline this body and
          // must round alike
          if (n < P.N) crow[n] = fma(al, acc[i][j][v], (m == n) ? dadd : 0.0);
        }
      }
  } else {
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = m0 + xrow0 + i * 16 + g + 4 * v;
        if (m >= P.M) continue;
        double *crow = P.C + (long)m * P.ldc;
#pragma unroll
        for (int j = 0; j < WM; j++) {
          const int n = n0 + yrow0 + j * 16 + r;
          const double al = (n < P.cw) ? P.alpha_w : alpha;
          if (n < P.N)
            crow[n] = fma(al, acc[i][j][v], fma(beta, crow[n], (m == n) ? dadd : 0.0));
        }
      }
  }
}

template <bool TX, bool TY, int WM>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void gemm_f64_kernel(oa_gemm_problem p0, oa_gemm_problem p1, const int32_t *skip) {
  if (skip && *skip) return;
  const ProbRegs P = pick_problem(p0, p1);
  constexpr int BM = 32 * WM;
  __shared__ __attribute__((aligned(16))) double smem[4 * BM * LS];
  gemm_body<TX, TY, WM>(P, xcd_logical(blockIdx.x, gridDim.x), smem);
}

// ------------------------------------------------------------------ grouped
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

// Workgroups of XCD x (= b % 8) walk the contiguous item range
// [x*per, (x+1)*per); the grid is a multiple of 8.
template <class F>
__device__ __forceinline__ void for_items(int total, F &&f) {
  const int per = cdiv(total, 8), slots = (int)gridDim.x >> 3, x = blockIdx.x & 7;
  for (int pos = blockIdx.x >> 3; pos < per; pos += slots) {
    const int item = x * per + pos;
    if (item < total) f(item);
  }
}

// Many row-major NN problems in one persistent launch; items are output
// tiles; a problem whose skip flag is set contributes no work.
template <int WM>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void gemm_f64_grouped(const oa_gemm_item *__restrict__ t, const int *__restrict__ prefix,
                      int count) {
  constexpr int BM = 32 * WM;
  __shared__ __attribute__((aligned(16))) double smem[4 * BM * LS];
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    const oa_gemm_item &it = t[i];
    if (it.skip && *it.skip) return;
    ProbRegs P;
    P.M = it.p.M; P.N = it.p.N; P.K = it.p.K;
    P.X = it.p.X; P.ldx = it.p.ldx; P.Y = it.p.Y; P.ldy = it.p.ldy;
    P.C = it.p.C; P.ldc = it.p.ldc;
    P.alpha = it.p.alpha; P.beta = it.p.beta; P.diag_add = it.p.diag_add;
    P.cw = it.p.cw; P.alpha_w = it.p.alpha_w;
    gemm_body<false, false, WM>(P, item - prefix[i], smem);
    __syncthreads();
  });
}

// Split-K partial products over the plane-wave index.
//   SYM: tiles (bi <= bj) of X X^T; else all tiles of X Y^T.
//   Chunk 0 covers [0, kx0_end) when kx0_end > 0 (weighted 1/2 later).
struct OvlArgs {
  const double *X;
  long ldx;
  const double *Y;
  long ldy;
  int M, N, K;
  int kx0_end;
  int nz_main, kchunk;
  double *part;  // [nz][ntiles][BM][BM]
};

__device__ __forceinline__ void upper_tile(int tile, int T, int &bi, int &bj) {
  int b = 0, rem = tile;
  while (rem >= T - b) {
    rem -= T - b;
    b++;
  }
  bi = b;
  bj = b + rem;
}

// XY (with SYM): the upper-tile map of the symmetric product over two
// distinct operands X and Y (the complex overlap's Im S = psi (J psi)^T,
// whose strict upper triangle is all a Hermitian S needs).
template <int WM, bool SYM, bool XY = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void ovl_f64_kernel(OvlArgs a) {
  constexpr int BM = 32 * WM;
  __shared__ __attribute__((aligned(16))) double smem[4 * BM * LS];
  const int Tm = cdiv(a.M, BM), Tn = cdiv(a.N, BM);
  const int ntiles = SYM ? Tm * (Tm + 1) / 2 : Tm * Tn;
  // The short kx = 0 slice (z = 0) takes the leading workgroups, which the
  // dispatcher spreads over all XCDs; the main chunks follow in XCD ranges.
  const int hx = a.kx0_end > 0 ? 1 : 0;
  const int nkx = hx ? cdiv(ntiles, 8) * 8 : 0;
  int tile, z;
  if ((int)blockIdx.x < nkx) {
    if ((int)blockIdx.x >= ntiles) return;
    tile = blockIdx.x;
    z = 0;
  } else {
    const int logical = xcd_logical(blockIdx.x - nkx, gridDim.x - nkx);
    if (logical >= ntiles * a.nz_main) return;
    tile = logical % ntiles;
    z = logical / ntiles + hx;
  }
  int bi, bj;
  if (SYM) upper_tile(tile, Tm, bi, bj);
  else { bi = tile % Tm; bj = tile / Tm; }
  int ka, kb;
  if (a.kx0_end > 0) {
    if (z == 0) { ka = 0; kb = a.kx0_end; }
    else { ka = a.kx0_end + (z - 1) * a.kchunk; kb = min(a.K, ka + a.kchunk); }
  } else {
    ka = z * a.kchunk;
    kb = min(a.K, ka + a.kchunk);
  }
  d4 acc[WM][WM];
#pragma unroll
  for (int i = 0; i < WM; i++)
#pragma unroll
    for (int j = 0; j < WM; j++) acc[i][j] = d4{0, 0, 0, 0};
  const int m0 = bi * BM, n0 = bj * BM;
  const double *Yp = (SYM && !XY) ? a.X : a.Y;
  const long ldy = (SYM && !XY) ? a.ldx : a.ldy;
  const bool fast = m0 + BM <= a.M && n0 + BM <= a.N && aligned16(a.X) && aligned16(Yp) &&
                    (a.ldx % 2) == 0 && (ldy % 2) == 0 && (ka % 2) == 0;
  if (ka < kb) {
    if (SYM && !XY && bi == bj)
      kloop<WM, true, true, true>(a.X, a.ldx, m0, a.M, a.X, a.ldx, n0, a.N, ka, kb, fast, smem,
                                  acc);
    else
      kloop<WM, true, true, false>(a.X, a.ldx, m0, a.M, Yp, ldy, n0, a.N, ka, kb, fast, smem,
                                   acc);
  }
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int xrow0 = (w >> 1) * 16 * WM, yrow0 = (w & 1) * 16 * WM;
  const int r = lane & 15, g = lane >> 4;
  double *pt = a.part + ((long)z * ntiles + tile) * (BM * BM);
#pragma unroll
  for (int i = 0; i < WM; i++)
#pragma unroll
    for (int j = 0; j < WM; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int lm = xrow0 + i * 16 + g + 4 * v;
        const int ln = yrow0 + j * 16 + r;
        pt[lm * BM + ln] = acc[i][j][v];
      }
}

template <int WM, bool SYM>
__global__ __launch_bounds__(NT) void ovl_reduce_kernel(const double *__restrict__ part, int M,
                                                       int N, int nz, int has_kx0,
                                                       double alpha, double *C, long ldc,
                                                       double diag,
                                                       unsigned long long *maxdev) {
  constexpr int BM = 32 * WM;
  const int Tm = cdiv(M, BM), Tn = cdiv(N, BM);
  const int ntiles = SYM ? Tm * (Tm + 1) / 2 : Tm * Tn;
  const long idx = (long)blockIdx.x * NT + threadIdx.x;
  const int tile = (int)(idx / (BM * BM));
  if (tile >= ntiles) return;
  const int e = (int)(idx % (BM * BM));
  const int lm = e / BM, ln = e % BM;
  int bi, bj;
  if (SYM) upper_tile(tile, Tm, bi, bj);
  else { bi = tile % Tm; bj = tile / Tm; }
  if (SYM && bi == bj && ln < lm) return;
  const int i = bi * BM + lm, j = bj * BM + ln;
  if (i >= M || j >= N) return;
  double s = 0.0;
  for (int z = 0; z < nz; z++) {
    const double p = part[((long)z * ntiles + tile) * (BM * BM) + e];
    s += (has_kx0 && z == 0) ? 0.5 * p : p;
  }
  s *= alpha;
  C[(long)i * ldc + j] = s;
  if (SYM) C[(long)j * ldc + i] = s;
  if (maxdev) {
    const double dv = fabs(s - (i == j ? diag : 0.0));
    atomicMax(maxdev, (unsigned long long)__double_as_longlong(dv));
  }
}

__global__ void fill_u64_kernel(unsigned long long *p, unsigned long long v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ S->T
// Ping-pong buffers A[2], B[2] and C; iteration i reads index i%2 and writes
// 1-i%2, so the final T sits in A[iterations % 2].
struct NsBufs {
  double *A[2];
  double *B[2];
  double *C;
  double *partial;  // NS_RES_WG doubles
};
constexpr int NS_RES_WG = 128;

__global__ void ns_init_kernel(const double *__restrict__ S, int n, int nspin, double *A,
                               double *B, oa_ns_report *rep) {
  const long nn = (long)n * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < nn;
       idx += (long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / n), j = (int)(idx % n);
    B[idx] = (nspin == 1) ? S[idx] / 2.0 : S[idx];
    A[idx] = (i == j) ? 1.0 : 0.0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    rep->iterations = 0;
    rep->done = 0;
    rep->status = OA_SUCCESS;
    rep->pad = 0;
    rep->error = 0.0;
  }
}

// Fixed-order partial sums of (Aold - Anew)^2 (ortho.C:949-953).
__global__ __launch_bounds__(NT) void ns_resid_kernel(const double *__restrict__ Aold,
                                                      const double *__restrict__ Anew, long nn,
                                                      double *partial, const int32_t *skip) {
  if (*skip) return;
  __shared__ double red[NT];
  double s = 0.0;
  for (long idx = (long)blockIdx.x * NT + threadIdx.x; idx < nn; idx += (long)gridDim.x * NT) {
    const double d = Aold[idx] - Anew[idx];
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Ortho::collect_error (ortho.C:109-146) on the device.
__global__ void ns_update_kernel(const double *__restrict__ partial, int nwg, int n,
                                 double tol, int maxit, oa_ns_report *rep) {
  if (threadIdx.x != 0 || rep->done) return;
  double s = 0.0;
  for (int i = 0; i < nwg; i++) s += partial[i];
  const double err = sqrt(s / ((double)n * (double)n));
  const int it = rep->iterations + 1;
  rep->iterations = it;
  rep->error = err;
  if (!(err > tol && it < maxit)) {
    rep->done = 1;
    rep->status = (it >= maxit) ? OA_ERR_NOT_CONVERGED : OA_SUCCESS;
  }
}

__global__ void ns_finish_kernel(const double *__restrict__ A0, const double *__restrict__ A1,
                                 long nn, double *T, const oa_ns_report *rep) {
  const double *src = (rep->iterations & 1) ? A1 : A0;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < nn;
       idx += (long)gridDim.x * blockDim.x)
    T[idx] = src[idx];
}

// Grouped S->T bookkeeping over many instances (oa_ns_args tables).
//   INIT   items of NT elements: B = S/2 or S, A = I, report cleared.
//   RESID  item w < nwg: partial[w] = sum of (Aold - Anew)^2 over elements
//          w*NT + t + j*nwg*NT (a fixed order for a given nwg).
//   UPDATE one item per instance: Ortho::collect_error (ortho.C:109-146).
//   FINISH items of NT elements: T = A[iterations % 2].
__global__ __launch_bounds__(NT) void ns_init_grouped(const oa_ns_args *__restrict__ t,
                                                      const int *__restrict__ prefix,
                                                      int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    const oa_ns_args &a = t[i];
    const int w = item - prefix[i];
    const long nn = (long)a.n * a.n, idx = (long)w * NT + threadIdx.x;
    if (idx < nn) {
      const int r = (int)(idx / a.n), c = (int)(idx % a.n);
      a.B[0][idx] = (a.nspin == 1) ? a.S[idx] / 2.0 : a.S[idx];
      a.A[0][idx] = (r == c) ? 1.0 : 0.0;
    }
    if (w == 0 && threadIdx.x == 0) {
      a.rep->iterations = 0;
      a.rep->done = 0;
      a.rep->status = OA_SUCCESS;
      a.rep->pad = 0;
      a.rep->error = 0.0;
    }
  });
}

template <int PAR>
__global__ __launch_bounds__(NT) void ns_resid_grouped(const oa_ns_args *__restrict__ t,
                                                       const int *__restrict__ prefix,
                                                       int count) {
  __shared__ double red[NT];
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    const oa_ns_args &a = t[i];
    if (a.rep->done) return;
    const int w = item - prefix[i];
    const long nn = (long)a.n * a.n;
    const double *Aold = a.A[PAR], *Anew = a.A[PAR ^ 1];
    double s = 0.0;
    for (long idx = (long)w * NT + threadIdx.x; idx < nn; idx += (long)a.nwg * NT) {
      const double d = Aold[idx] - Anew[idx];
      s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[w] = red[0];
    __syncthreads();
  });
}

__global__ __launch_bounds__(NT) void ns_update_grouped(const oa_ns_args *__restrict__ t,
                                                        const int *__restrict__ prefix,
                                                        int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    const oa_ns_args &a = t[i];
    if (threadIdx.x != 0 || a.rep->done) return;
    double s = 0.0;
    for (int w = 0; w < a.nwg; w++) s += a.partial[w];
    const double err = sqrt(s / ((double)a.n * (double)a.n));
    const int it = a.rep->iterations + 1;
    a.rep->iterations = it;
    a.rep->error = err;
    if (!(err > a.tol && it < a.maxit)) {
      a.rep->done = 1;
      a.rep->status = (it >= a.maxit) ? OA_ERR_NOT_CONVERGED : OA_SUCCESS;
    }
  });
}

__global__ __launch_bounds__(NT) void ns_finish_grouped(const oa_ns_args *__restrict__ t,
                                                        const int *__restrict__ prefix,
                                                        int count) {
  for_items(prefix[count], [&](int item) {
    const int i = group_find(prefix, count, item);
    const oa_ns_args &a = t[i];
    const long nn = (long)a.n * a.n, idx = (long)(item - prefix[i]) * NT + threadIdx.x;
    const double *src = (a.rep->iterations & 1) ? a.A[1] : a.A[0];
    if (idx < nn) a.T[idx] = src[idx];
  });
}

template <class K>
int grouped_grid(K kernel, int total) {
  int per_cu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, NT, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  static thread_local int cached_dev = -1, cached_cus = 256;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev != cached_dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
      cached_cus = cus;
    cached_dev = dev;
  }
  const int cap = cdiv(cached_cus * per_cu, 8) * 8;
  const int g = cdiv(max(total, 1), 8) * 8;
  return g < cap ? g : cap;
}

template <bool TX, bool TY>
int launch_gemm(hipStream_t s, const oa_gemm_problem *p, int nprob, const int32_t *skip) {
  int tiles128 = 0;
  for (int i = 0; i < nprob; i++)
    tiles128 = max(tiles128, cdiv(p[i].M, 128) * cdiv(p[i].N, 128));
  const oa_gemm_problem p0 = p[0], p1 = p[nprob > 1 ? 1 : 0];
  if (tiles128 * nprob >= 512) {
    const int nb = cdiv(tiles128, 8) * 8;
    hipLaunchKernelGGL((gemm_f64_kernel<TX, TY, 4>), dim3(nb, nprob), dim3(NT), 0, s, p0, p1,
                       skip);
  } else {
    int tiles64 = 0;
    for (int i = 0; i < nprob; i++)
      tiles64 = max(tiles64, cdiv(p[i].M, 64) * cdiv(p[i].N, 64));
    const int nb = cdiv(tiles64, 8) * 8;
    hipLaunchKernelGGL((gemm_f64_kernel<TX, TY, 2>), dim3(nb, nprob), dim3(NT), 0, s, p0, p1,
                       skip);
  }
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

// Split-K plan for the plane-wave products.
struct OvlPlan {
  int WM, BM, ntiles, nz_main, kchunk, nz;
};

// Split depth nz for ntiles output tiles over K columns.  The main chunks go
// to the 8 XCDs in equal contiguous ranges of cdiv(ntiles*nz, 8) workgroups,
// of which 32 CUs x wpc are resident at once (wpc = 2 for WM = 4, whose two
// 74 KB LDS footprints fill a CU's 160 KB; 4 for WM = 2), so a depth costs
// rounds x cdiv(K, nz) columns of one workgroup's MFMA time (2 flops per
// output element and column on 1/wpc of a CU's share of the FP64 MFMA peak:
// 6.5 ps x wpc) plus writing and re-reading the nz*ntiles partial tiles
// (16 bytes per element at about 5 TB/s: 3.2 ps); both terms are per output
// element of a tile.  The cheapest depth wins; chunks keep at least 8 BK
// stages.  The choice depends on (M, N, K) only, so workspace sizing and
// launch agree for any kx = 0 split.
int ovl_depth(int ntiles, int K, int wpc) {
  const int nzmax = max(1, min(512, cdiv(K, 8 * BK)));
  const long slots = 32L * wpc;
  int best = 1;
  double best_cost = 0.0;
  for (int nz = 1; nz <= nzmax; nz++) {
    const long rounds = (cdiv(ntiles * nz, 8) + slots - 1) / slots;
    const double cost = (double)rounds * cdiv(K, nz) * wpc * 6.5 + (double)nz * ntiles * 3.2;
    if (nz == 1 || cost < best_cost) { best = nz; best_cost = cost; }
  }
  return best;
}

// max_part_bytes > 0 caps the depth so that the partial tiles (the kx = 0
// slice included) fit that many bytes; nz_main = 0 reports that even one
// chunk does not fit.
OvlPlan plan_ovl(int M, int N, int K, int kx0_end, bool sym, size_t max_part_bytes = 0) {
  OvlPlan pl;
  const int t128 = sym ? cdiv(M, 128) * (cdiv(M, 128) + 1) / 2 : cdiv(M, 128) * cdiv(N, 128);
  pl.WM = (M >= 256 || t128 >= 16) ? 4 : 2;
  pl.BM = 32 * pl.WM;
  const int Tm = cdiv(M, pl.BM), Tn = cdiv(N, pl.BM);
  pl.ntiles = sym ? Tm * (Tm + 1) / 2 : Tm * Tn;
  const int Kmain = K - kx0_end;
  int nz = ovl_depth(pl.ntiles, max(K, 1), pl.WM == 4 ? 2 : 4);
  if (max_part_bytes > 0) {
    const size_t per = (size_t)pl.ntiles * pl.BM * pl.BM * sizeof(double);
    const long fit = (long)(max_part_bytes / per) - (kx0_end > 0 ? 1 : 0);
    if (fit < 1) {
      pl.nz_main = 0;
      pl.kchunk = 0;
      pl.nz = 0;
      return pl;
    }
    if (nz > fit) nz = (int)fit;
  }
  int kchunk = cdiv(cdiv(max(Kmain, 1), nz), BK) * BK;
  nz = max(1, cdiv(max(Kmain, 1), kchunk));
  pl.nz_main = nz;
  pl.kchunk = kchunk;
  pl.nz = nz + (kx0_end > 0 ? 1 : 0);
  return pl;
}

// X[r][c] *= s over the leading ncols columns of each row: one thread per
// element, rows strided over the grid.
__global__ void scale_cols_kernel(int rows, int ncols, double *X, long ld, double scale) {
  const long total = (long)rows * ncols;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    double *p = X + (i / ncols) * ld + (i % ncols);
    *p *= scale;
  }
}

}  // namespace

int oa_launch_scale_cols(hipStream_t s, int rows, int ncols, double *X, long ld, double scale) {
  const long total = (long)rows * ncols;
  if (total <= 0) return OA_SUCCESS;
  long blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(scale_cols_kernel, dim3((unsigned)blocks), dim3(256), 0, s, rows, ncols, X,
                     ld, scale);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

int oa_launch_gemm_f64(hipStream_t s, bool tx, bool ty, const oa_gemm_problem *p, int nprob,
                       const int32_t *skip) {
  if (nprob < 1 || nprob > 2) return OA_ERR_UNSUPPORTED;
  if (!tx && !ty) return launch_gemm<false, false>(s, p, nprob, skip);
  if (!tx && ty) return launch_gemm<false, true>(s, p, nprob, skip);
  if (tx && !ty) return launch_gemm<true, false>(s, p, nprob, skip);
  return launch_gemm<true, true>(s, p, nprob, skip);
}

size_t oa_overlap_f64_ws_bytes(int n, int K) {
  const OvlPlan pl = plan_ovl(n, n, K, 0, true);
  // kx0 chunk adds one more z slice at most; maxdev slot after the partials
  return (size_t)(pl.nz + 1) * pl.ntiles * pl.BM * pl.BM * sizeof(double) + 256;
}

size_t oa_ntsplit_f64_ws_bytes(int M, int N, int K) {
  const OvlPlan pl = plan_ovl(M, N, K, 0, false);
  return (size_t)pl.nz * pl.ntiles * pl.BM * pl.BM * sizeof(double) + 256;
}

template <int WM, bool SYM, bool XY = false>
static int launch_ovl(hipStream_t s, const OvlPlan &pl, OvlArgs a, int M, int N, int kx0_end,
                      double alpha, double *C, long ldc, double diag,
                      unsigned long long *maxdev) {
  const int nkx = kx0_end > 0 ? cdiv(pl.ntiles, 8) * 8 : 0;
  const int nb = nkx + cdiv(pl.ntiles * pl.nz_main, 8) * 8;
  hipLaunchKernelGGL((ovl_f64_kernel<WM, SYM, XY>), dim3(nb), dim3(NT), 0, s, a);
  if (maxdev) hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, maxdev, 0ull, 1);
  const long total = (long)pl.ntiles * pl.BM * pl.BM;
  hipLaunchKernelGGL((ovl_reduce_kernel<WM, SYM>), dim3((unsigned)cdiv((int)total, NT)),
                     dim3(NT), 0, s, (const double *)a.part, M, N, pl.nz,
                     kx0_end > 0 ? 1 : 0, alpha, C, ldc, diag, maxdev);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

int oa_launch_overlap_f64(hipStream_t s, int n, int K, const double *psi, long ld, int kx0_end,
                          double alpha, double *S, double diag, double *maxdev, void *ws,
                          size_t ws_bytes) {
  const OvlPlan pl = plan_ovl(n, n, K, kx0_end, true);
  const size_t need = (size_t)pl.nz * pl.ntiles * pl.BM * pl.BM * sizeof(double);
  if (need > ws_bytes) return OA_ERR_CAPACITY;
  OvlArgs a;
  a.X = psi; a.ldx = ld; a.Y = psi; a.ldy = ld;
  a.M = n; a.N = n; a.K = K;
  a.kx0_end = kx0_end; a.nz_main = pl.nz_main; a.kchunk = pl.kchunk;
  a.part = (double *)ws;
  unsigned long long *md = reinterpret_cast<unsigned long long *>(maxdev);
  if (pl.WM == 4) return launch_ovl<4, true>(s, pl, a, n, n, kx0_end, alpha, S, n, diag, md);
  return launch_ovl<2, true>(s, pl, a, n, n, kx0_end, alpha, S, n, diag, md);
}

int oa_launch_overlap_f64_fit(hipStream_t s, int n, int K, const double *psi, long ld,
                              int kx0_end, double alpha, double *S, double diag,
                              double *maxdev, void *ws, size_t ws_bytes) {
  const OvlPlan pl = plan_ovl(n, n, K, kx0_end, true, ws_bytes);
  if (pl.nz_main < 1) return OA_ERR_CAPACITY;
  OvlArgs a;
  a.X = psi; a.ldx = ld; a.Y = psi; a.ldy = ld;
  a.M = n; a.N = n; a.K = K;
  a.kx0_end = kx0_end; a.nz_main = pl.nz_main; a.kchunk = pl.kchunk;
  a.part = (double *)ws;
  unsigned long long *md = reinterpret_cast<unsigned long long *>(maxdev);
  if (pl.WM == 4) return launch_ovl<4, true>(s, pl, a, n, n, kx0_end, alpha, S, n, diag, md);
  return launch_ovl<2, true>(s, pl, a, n, n, kx0_end, alpha, S, n, diag, md);
}

int oa_launch_ntsplit_f64(hipStream_t s, int M, int N, int K, const double *X, long ldx,
                          const double *Y, long ldy, double alpha, double *C, long ldc, void *ws,
                          size_t ws_bytes) {
  const OvlPlan pl = plan_ovl(M, N, K, 0, false);
  const size_t need = (size_t)pl.nz * pl.ntiles * pl.BM * pl.BM * sizeof(double);
  if (need > ws_bytes) return OA_ERR_CAPACITY;
  OvlArgs a;
  a.X = X; a.ldx = ldx; a.Y = Y; a.ldy = ldy;
  a.M = M; a.N = N; a.K = K;
  a.kx0_end = 0; a.nz_main = pl.nz_main; a.kchunk = pl.kchunk;
  a.part = (double *)ws;
  if (pl.WM == 4) return launch_ovl<4, false>(s, pl, a, M, N, 0, alpha, C, ldc, 0.0, nullptr);
  return launch_ovl<2, false>(s, pl, a, M, N, 0, alpha, C, ldc, 0.0, nullptr);
}

int oa_launch_overlap_xy_f64(hipStream_t s, int n, int K, const double *X, long ldx,
                             const double *Y, long ldy, double alpha, double *C, void *ws,
                             size_t ws_bytes) {
  const OvlPlan pl = plan_ovl(n, n, K, 0, true);
  const size_t need = (size_t)pl.nz * pl.ntiles * pl.BM * pl.BM * sizeof(double);
  if (need > ws_bytes) return OA_ERR_CAPACITY;
  OvlArgs a;
  a.X = X; a.ldx = ldx; a.Y = Y; a.ldy = ldy;
  a.M = n; a.N = n; a.K = K;
  a.kx0_end = 0; a.nz_main = pl.nz_main; a.kchunk = pl.kchunk;
  a.part = (double *)ws;
  if (pl.WM == 4) return launch_ovl<4, true, true>(s, pl, a, n, n, 0, alpha, C, n, 0.0, nullptr);
  return launch_ovl<2, true, true>(s, pl, a, n, n, 0, alpha, C, n, 0.0, nullptr);
}

size_t oa_ns_ws_bytes(int n) {
  return (size_t)5 * n * n * sizeof(double) + NS_RES_WG * sizeof(double) + 256;
}

// S->T on the device in three parts so a host loop can stop early:
// begin (B = S/2 or S, A = I, report cleared), guarded iterations, finish.
static void ns_bufs(void *ws, int n, NsBufs &b) {
  const long nn = (long)n * n;
  double *base = (double *)ws;
  b.A[0] = base;
  b.A[1] = base + nn;
  b.B[0] = base + 2 * nn;
  b.B[1] = base + 3 * nn;
  b.C = base + 4 * nn;
  b.partial = base + 5 * nn;
}

int oa_ns_begin(hipStream_t s, int n, const double *S, int nspin, oa_ns_report *rep, void *ws) {
  NsBufs b;
  ns_bufs(ws, n, b);
  const long nn = (long)n * n;
  const long nblk = (nn + NT - 1) / NT;
  const int eb = (int)(nblk < 1 ? 1 : (nblk > 1024 ? 1024 : nblk));
  hipLaunchKernelGGL(ns_init_kernel, dim3(eb), dim3(NT), 0, s, S, n, nspin, b.A[0], b.B[0], rep);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

// Iteration `it` (0-based) of Ortho::do_iteration .. collect_error; a no-op
// once rep->done is set.
int oa_ns_iterate(hipStream_t s, int n, int it, double tol, int maxit, oa_ns_report *rep,
                  void *ws) {
  NsBufs b;
  ns_bufs(ws, n, b);
  const long nn = (long)n * n;
  const int32_t *skip = &rep->done;
  const int cur = it & 1, nxt = cur ^ 1;
  // do_iteration (ortho.C:848-870): C = 3I - A*B
  oa_gemm_problem p1{n, n, n, b.A[cur], n, b.B[cur], n, b.C, n, -1.0, 0.0, 3.0};
  int st = oa_launch_gemm_f64(s, false, false, &p1, 1, skip);
  if (st) return st;
  // step_2 / step_3 (ortho.C:880-929): B' = B*C/2 ; A' = C*A/2
  oa_gemm_problem p2[2] = {{n, n, n, b.B[cur], n, b.C, n, b.B[nxt], n, 0.5, 0.0, 0.0},
                           {n, n, n, b.C, n, b.A[cur], n, b.A[nxt], n, 0.5, 0.0, 0.0}};
  st = oa_launch_gemm_f64(s, false, false, p2, 2, skip);
  if (st) return st;
  hipLaunchKernelGGL(ns_resid_kernel, dim3(NS_RES_WG), dim3(NT), 0, s, (const double *)b.A[cur],
                     (const double *)b.A[nxt], nn, b.partial, skip);
  hipLaunchKernelGGL(ns_update_kernel, dim3(1), dim3(64), 0, s, (const double *)b.partial,
                     NS_RES_WG, n, tol, maxit, rep);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

int oa_ns_finish(hipStream_t s, int n, double *T, const oa_ns_report *rep, void *ws) {
  NsBufs b;
  ns_bufs(ws, n, b);
  const long nn = (long)n * n;
  const long nblk = (nn + NT - 1) / NT;
  const int eb = (int)(nblk < 1 ? 1 : (nblk > 1024 ? 1024 : nblk));
  hipLaunchKernelGGL(ns_finish_kernel, dim3(eb), dim3(NT), 0, s, (const double *)b.A[0],
                     (const double *)b.A[1], nn, T, rep);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

int oa_launch_s_to_t(hipStream_t s, int n, const double *S, int nspin, double tol, int maxit,
                     double *T, oa_ns_report *rep, void *ws) {
  int st = oa_ns_begin(s, n, S, nspin, rep, ws);
  for (int it = 0; it < maxit && st == OA_SUCCESS; it++)
    st = oa_ns_iterate(s, n, it, tol, maxit, rep, ws);
  if (st) return st;
  return oa_ns_finish(s, n, T, rep, ws);
}

int oa_gemm_grouped(hipStream_t s, int wm, const oa_group &g) {
  if (g.count <= 0 || g.total <= 0) return OA_SUCCESS;
  const oa_gemm_item *t = reinterpret_cast<const oa_gemm_item *>(g.args);
  if (wm == 4)
    hipLaunchKernelGGL(gemm_f64_grouped<4>, dim3(grouped_grid(gemm_f64_grouped<4>, g.total)),
                       dim3(NT), 0, s, t, g.prefix, g.count);
  else if (wm == 2)
    hipLaunchKernelGGL(gemm_f64_grouped<2>, dim3(grouped_grid(gemm_f64_grouped<2>, g.total)),
                       dim3(NT), 0, s, t, g.prefix, g.count);
  else
    return OA_ERR_UNSUPPORTED;
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

int oa_ns_items(int phase, int n, int nwg) {
  const long nn = (long)n * n;
  switch (phase) {
    case OA_NS_INIT:
    case OA_NS_FINISH: return (int)((nn + NT - 1) / NT);
    case OA_NS_RESID0:
    case OA_NS_RESID1: return nwg;
    case OA_NS_UPDATE: return 1;
  }
  return 0;
}

int oa_ns_grouped(hipStream_t s, int phase, const oa_group &g) {
  if (g.count <= 0 || g.total <= 0) return OA_SUCCESS;
  const oa_ns_args *t = reinterpret_cast<const oa_ns_args *>(g.args);
#define OA_NSG(K)                                                                      \
  hipLaunchKernelGGL(K, dim3(grouped_grid(K, g.total)), dim3(NT), 0, s, t, g.prefix, \
                     g.count)
  switch (phase) {
    case OA_NS_INIT: OA_NSG(ns_init_grouped); break;
    case OA_NS_RESID0: OA_NSG(ns_resid_grouped<0>); break;
    case OA_NS_RESID1: OA_NSG(ns_resid_grouped<1>); break;
    case OA_NS_UPDATE: OA_NSG(ns_update_grouped); break;
    case OA_NS_FINISH: OA_NSG(ns_finish_grouped); break;
    default: return OA_ERR_UNSUPPORTED;
  }
#undef OA_NSG
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}
