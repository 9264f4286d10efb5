// State files and the kx = 0 plane layout of the gamma-point half sphere.
//
// A state file (utility/util.C:1107-1330 readState, :1955-2180
// writeStateFile) holds a header `ncoef nx ny nz` and one record
// `re im kx ky kz` per plane wave, as text (fprintf "%g %g %d %d %d \n"), as
// raw binary (4 ints, then 2 doubles + 3 ints per record) or as either of
// them gzip-compressed in `<path>.gz` (formats 0, 1, 2, 3 -- the reference's
// ibinary_opt).
//
// Gamma point (double packing): the file lists the upper half of the kx = 0
// plane, then the kx > 0 points, and G = 0 last; reading stops at G = 0.  In
// memory the GSpace keeps the whole kx = 0 plane in front (para_grp_parse.C:
// 178-264 flip_data_set): [lower half, rebuilt by conjugate symmetry
// c(-G) = conj c(G) in reverse order | G = 0 | upper half | kx > 0].  The
// writer sorts by (kx, ky, kz) (sort_psi_output, util.C:2180-2215), drops
// everything before G = 0 and writes G = 0 last.
//
// The file I/O is host code (the reference's is too); oa_kx0_expand and
// oa_kx0_contract move whole device blocks of n states between the file
// order and the memory order in one gather pass.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "oa_internal.h"

namespace {

struct Rec {
  double re, im;
  int kx, ky, kz;
};

// Sequential reader over the four formats.
class Source {
 public:
  bool open(const char *path, int format) {
    fmt_ = format;
    if (format >= 2) {
      const std::string p = std::string(path) + ".gz";
      gz_ = gzopen(p.c_str(), "rb");
      return gz_ != nullptr;
    }
    fp_ = fopen(path, format == 1 ? "rb" : "r");
    return fp_ != nullptr;
  }
  ~Source() {
    if (gz_) gzclose(gz_);
    if (fp_) fclose(fp_);
  }
  bool header(int &n, int &nx, int &ny, int &nz) {
    if (fmt_ == 0) return fscanf(fp_, "%d%d%d%d", &n, &nx, &ny, &nz) == 4;
    if (fmt_ == 2) return line() && sscanf(buf_, "%d%d%d%d", &n, &nx, &ny, &nz) == 4;
    int h[4];
    if (!raw(h, sizeof(h))) return false;
    n = h[0], nx = h[1], ny = h[2], nz = h[3];
    return true;
  }
  bool record(Rec &r) {
    if (fmt_ == 0) return fscanf(fp_, "%lg%lg%d%d%d", &r.re, &r.im, &r.kx, &r.ky, &r.kz) == 5;
    if (fmt_ == 2)
      return line() && sscanf(buf_, "%lg%lg%d%d%d", &r.re, &r.im, &r.kx, &r.ky, &r.kz) == 5;
    return raw(&r.re, 8) && raw(&r.im, 8) && raw(&r.kx, 4) && raw(&r.ky, 4) && raw(&r.kz, 4);
  }

 private:
  bool line() { return gzgets(gz_, buf_, sizeof(buf_)) != nullptr; }
  bool raw(void *p, size_t b) {
    if (gz_) return gzread(gz_, p, (unsigned)b) == (int)b;
    return fread(p, 1, b, fp_) == b;
  }
  int fmt_ = 0;
  FILE *fp_ = nullptr;
  gzFile gz_ = nullptr;
  char buf_[1000];  // records are far shorter
};

class Sink {
 public:
  bool open(const char *path, int format) {
    fmt_ = format;
    if (format >= 2) {
      const std::string p = std::string(path) + ".gz";
      gz_ = gzopen(p.c_str(), "wb");
      return gz_ != nullptr;
    }
    fp_ = fopen(path, "w");
    return fp_ != nullptr;
  }
  bool close() {
    bool ok = true;
    if (gz_) ok = gzclose(gz_) == Z_OK;
    if (fp_) ok = fclose(fp_) == 0 && ok;
    gz_ = nullptr;
    fp_ = nullptr;
    return ok;
  }
  ~Sink() { close(); }
  bool header(int n, int nx, int ny, int nz) {
    if (fmt_ == 0) return fprintf(fp_, "%d %d %d %d\n", n, nx, ny, nz) > 0;
    if (fmt_ == 2) return gzprintf(gz_, "%d %d %d %d\n", n, nx, ny, nz) > 0;
    const int h[4] = {n, nx, ny, nz};
    return raw(h, sizeof(h));
  }
  bool record(double re, double im, int kx, int ky, int kz) {
    if (fmt_ == 0) return fprintf(fp_, "%g %g %d %d %d \n", re, im, kx, ky, kz) > 0;
    if (fmt_ == 2) return gzprintf(gz_, "%g %g %d %d %d \n", re, im, kx, ky, kz) > 0;
    return raw(&re, 8) && raw(&im, 8) && raw(&kx, 4) && raw(&ky, 4) && raw(&kz, 4);
  }

 private:
  bool raw(const void *p, size_t b) {
    if (gz_) return gzwrite(gz_, p, (unsigned)b) == (int)b;
    return fwrite(p, 1, b, fp_) == b;
  }
  int fmt_ = 0;
  FILE *fp_ = nullptr;
  gzFile gz_ = nullptr;
};

// memory index j -> file index, conjugated for the rebuilt lower half
__host__ __device__ inline long expand_src(long j, long nfile, long nhalf, bool &conj) {
  conj = j < nhalf;
  if (j < nhalf) return nhalf - 1 - j;
  if (j == nhalf) return nfile - 1;
  return j - nhalf - 1;
}

// file index i -> memory index
__host__ __device__ inline long contract_src(long i, long nfile, long nhalf) {
  return (i == nfile - 1) ? nhalf : i + nhalf + 1;
}

// One (state, point) pair per thread: a gather of the (re, im) pair, with
// the imaginary part negated for the rebuilt lower half.
template <bool EXPAND>
__global__ void kx0_gather_kernel(int n, long nout, long nfile, long nhalf,
                                  const double *__restrict__ in, long ldin,
                                  double *__restrict__ out, long ldout) {
  const long total = (long)n * nout;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const long r = t / nout, j = t % nout;
    bool conj = false;
    const long s = EXPAND ? expand_src(j, nfile, nhalf, conj) : contract_src(j, nfile, nhalf);
    const double re = in[r * ldin + 2 * s], im = in[r * ldin + 2 * s + 1];
    out[r * ldout + 2 * j] = re;
    out[r * ldout + 2 * j + 1] = conj ? -im : im;
  }
}

int gather(oa_context *c, bool expand, int n, int nfile, int nhalf, const double *in, long ldin,
           double *out, long ldout) {
  const long nout = expand ? (long)nfile + nhalf : nfile;
  const long total = (long)n * nout;
  long blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (expand)
    hipLaunchKernelGGL(kx0_gather_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, c->stream,
                       n, nout, (long)nfile, (long)nhalf, in, ldin, out, ldout);
  else
    hipLaunchKernelGGL(kx0_gather_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, c->stream,
                       n, nout, (long)nfile, (long)nhalf, in, ldin, out, ldout);
  return hipPeekAtLastError() == hipSuccess ? OA_SUCCESS : OA_ERR_HIP;
}

}  // namespace

extern "C" {

int oa_state_read(const char *path, int format, int double_pack, int capacity, double *coef,
                  int *kx, int *ky, int *kz, int *npts, int *nkx0, int *grid) {
  if (!path) return 1;
  if (format < 0 || format > 3) return 2;
  if (capacity < 0) return 4;
  if (capacity > 0 && !coef) return 5;
  Source src;
  if (!src.open(path, format)) return OA_ERR_IO;
  int ncoef, nx, ny, nz;
  if (!src.header(ncoef, nx, ny, nz) || ncoef < 0) return OA_ERR_IO;
  std::vector<Rec> recs;
  // the header is not trusted for the allocation: a damaged count fails at
  // the first missing record instead of reserving gigabytes
  recs.reserve((size_t)std::min(ncoef, 1 << 20));
  for (int i = 0; i < ncoef; i++) {
    Rec r;
    if (!src.record(r)) return OA_ERR_IO;
    recs.push_back(r);
    if (double_pack && r.kx == 0 && r.ky == 0 && r.kz == 0) break;
  }
  const long nfile = (long)recs.size();
  long nhalf = 0, total = nfile;
  if (double_pack) {
    // flip_data_set's checks: the upper half of the kx = 0 plane leads, G = 0 ends
    if (nfile == 0) return OA_ERR_IO;
    const Rec &g0 = recs.back();
    if (g0.kx != 0 || g0.ky != 0 || g0.kz != 0) return OA_ERR_IO;
    while (nhalf < nfile - 1 && recs[nhalf].kx == 0) nhalf++;
    for (long i = nhalf; i < nfile - 1; i++)
      if (recs[i].kx == 0) return OA_ERR_IO;
    // the upper half starts at (0, 0, 1) (flip_data_set's last check)
    if (nfile >= 2 && (recs[0].kx != 0 || recs[0].ky != 0 || recs[0].kz != 1)) return OA_ERR_IO;
    total = nfile + nhalf;
  }
  if (npts) *npts = (int)total;
  if (nkx0) {
    long c0 = 0;
    if (double_pack) c0 = 2 * nhalf + 1;
    *nkx0 = (int)c0;
  }
  if (grid) grid[0] = nx, grid[1] = ny, grid[2] = nz;
  if (total > capacity) return OA_ERR_CAPACITY;
  for (long j = 0; j < total; j++) {
    bool conj = false;
    const long s = double_pack ? expand_src(j, nfile, nhalf, conj) : j;
    const Rec &r = recs[s];
    coef[2 * j] = r.re;
    coef[2 * j + 1] = conj ? -r.im : r.im;
    if (kx) kx[j] = r.kx;
    if (ky) ky[j] = conj ? -r.ky : r.ky;
    if (kz) kz[j] = conj ? -r.kz : r.kz;
  }
  return OA_SUCCESS;
}

int oa_state_write(const char *path, int format, int double_pack, int npts, const double *coef,
                   const int *kx, const int *ky, const int *kz, const int *grid) {
  if (!path) return 1;
  if (format < 0 || format > 3) return 2;
  if (npts < 0) return 4;
  if (npts > 0 && (!coef || !kx || !ky || !kz)) return 5;
  if (!grid) return 9;
  // sort_psi_output: order by the key (kz - kzmin) + (ky - kymin) nz + (kx - kxmin) nz ny
  std::vector<int> idx((size_t)npts);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (kx[a] != kx[b]) return kx[a] < kx[b];
    if (ky[a] != ky[b]) return ky[a] < ky[b];
    return kz[a] < kz[b];
  });
  long istrt = -1;
  if (double_pack) {
    for (long i = 0; i < npts; i++)
      if (kx[idx[i]] == 0 && ky[idx[i]] == 0 && kz[idx[i]] == 0) {
        istrt = i;
        break;
      }
    if (istrt < 0) return OA_ERR_IO;  // a gamma-point set always holds G = 0
  }
  const int ncoef_true = double_pack ? (int)(npts - istrt) : npts;
  Sink out;
  if (!out.open(path, format)) return OA_ERR_IO;
  bool ok = out.header(ncoef_true, grid[0], grid[1], grid[2]);
  for (long i = istrt + 1; ok && i < npts; i++) {
    const int p = idx[i];
    ok = out.record(coef[2 * p], coef[2 * p + 1], kx[p], ky[p], kz[p]);
  }
  if (ok && double_pack) {
    const int p = idx[istrt];
    ok = out.record(coef[2 * p], coef[2 * p + 1], kx[p], ky[p], kz[p]);
  }
  ok = out.close() && ok;
  return ok ? OA_SUCCESS : OA_ERR_IO;
}

int oa_kx0_expand(oa_handle_t c, int n, int nfile, int nhalf, const double *in, long ldin,
                  double *out, long ldout) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (nfile < 1 && n > 0) return 3;
  if (nhalf < 0 || nhalf > nfile - 1) return 4;
  if (n > 0 && !in) return 5;
  if (ldin < 2L * nfile) return 6;
  if (n > 0 && !out) return 7;
  if (ldout < 2L * ((long)nfile + nhalf)) return 8;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  return gather(c, true, n, nfile, nhalf, in, ldin, out, ldout);
}

int oa_kx0_contract(oa_handle_t c, int n, int nfile, int nhalf, const double *in, long ldin,
                    double *out, long ldout) {
  if (!c) return OA_ERR_NO_DEVICE;
  if (n < 0) return 2;
  if (nfile < 1 && n > 0) return 3;
  if (nhalf < 0 || nhalf > nfile - 1) return 4;
  if (n > 0 && !in) return 5;
  if (ldin < 2L * ((long)nfile + nhalf)) return 6;
  if (n > 0 && !out) return 7;
  if (ldout < 2L * nfile) return 8;
  if (n == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  return gather(c, false, n, nfile, nhalf, in, ldin, out, ldout);
}

}  // extern "C"
