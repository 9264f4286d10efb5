// Caller-side transport of the path on one device: the per-(state, plane)
// messages a Charm++ build exchanges between GSpace and PairCalculator chares
// become copies into and out of the device-resident [n][ld] block.
//
//   oa_collate_rows   replaces MessageDataCollator::operator()
//                     (MessageDataCollator.h:163-247): each incoming message
//                     (paircalcInputMsg, pcMessages.h:117-153: nRows x nCols
//                     points of one sender) is copied to its place in the
//                     collated block (`offset * numUnitsInMsg`, :222).  Here
//                     the place is given explicitly: rows row0 .. row0 + nrows
//                     and doubles col0 .. col0 + ncols of the block, so the
//                     planes and chunks of a state land side by side in one row.
//   oa_scatter_rows   replaces the return leg: sendBWResultDirect
//                     (ckPairCalculator.C:2365-2452) ships one partialResultMsg
//                     per output state and plane, and the GSpace chare copies
//                     the first contribution and adds later ones with fastAdd
//                     (CP_State_GSpacePlane.C:3120-3126, 3197-3203;
//                     fastadd.C:34-51).  Here each destination receives its
//                     slice of the result block, copied (accumulate 0) or added
//                     elementwise (accumulate 1), in one FP64 add per element
//                     as fastAdd does.
//
// Both are HBM- (device buffers) or PCIe-bound (host buffers) copies: one
// workgroup walks the rows of one message with 16-byte accesses where the
// addresses allow.  Host buffers are staged through a pinned area of the
// handle, so a call on host buffers is one H2D (or D2H) copy of all the
// message bytes plus one kernel.  Both calls are synchronous.
#include <cstring>
#include <vector>

#include "oa_internal.h"

namespace {

struct XferMsg {
  const double *src;  // gather: message data (device address); scatter: block
  double *dst;        // gather: block; scatter: destination (device address)
  long src_ld, dst_ld;  // row pitches (doubles)
  int nrows, ncols;
  int accumulate, pad;
};

__global__ __launch_bounds__(256) void xfer_kernel(const XferMsg *__restrict__ msgs, int count) {
  for (int m = blockIdx.x; m < count; m += gridDim.x) {
    const XferMsg x = msgs[m];
    for (int r = 0; r < x.nrows; r++) {
      const double *s = x.src + (long)r * x.src_ld;
      double *d = x.dst + (long)r * x.dst_ld;
      const bool vec = ((((uintptr_t)s) | ((uintptr_t)d)) & 15) == 0;
      int c = 0;
      if (vec) {
        for (c = 2 * (int)threadIdx.x; c + 1 < x.ncols; c += 512) {
          const double2 v = *reinterpret_cast<const double2 *>(s + c);
          double2 *o = reinterpret_cast<double2 *>(d + c);
          if (x.accumulate) {
            double2 a = *o;
            a.x += v.x;
            a.y += v.y;
            *o = a;
          } else {
            *o = v;
          }
        }
        // the last odd element
        c = x.ncols & ~1;
        if (c < x.ncols && threadIdx.x == 0) {
          if (x.accumulate) d[c] += s[c];
          else d[c] = s[c];
        }
      } else {
        for (c = (int)threadIdx.x; c < x.ncols; c += 256) {
          if (x.accumulate) d[c] += s[c];
          else d[c] = s[c];
        }
      }
    }
  }
}

// Pinned host staging owned by the handle (grown on demand, freed with it).
int pinned_require(oa_workspace &p, size_t bytes) {
  if (bytes <= p.bytes) return OA_SUCCESS;
  if (p.ptr) OA_HIP_CHECK(hipHostFree(p.ptr));
  p.ptr = nullptr;
  p.bytes = 0;
  const size_t sz = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
  OA_HIP_CHECK(hipHostMalloc(&p.ptr, sz, hipHostMallocDefault));
  p.bytes = sz;
  return OA_SUCCESS;
}

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

// Common argument checks; returns the 1-based index of a bad argument (the
// block is argument ib, its pitch ib + 1, its row count ib + 2).
int check_msgs(int count, const int *row0, const int *nrows, const long *col0, const int *ncols,
               const void *const *bufs, const double *block, long ld, int nblock_rows, int ib) {
  if (count < 0) return 2;
  if (count == 0) return 0;
  if (!row0) return 3;
  if (!nrows) return 4;
  if (!col0) return 5;
  if (!ncols) return 6;
  if (!bufs) return 7;
  if (!block) return ib;
  if (ld < 1) return ib + 1;
  if (nblock_rows < 0) return ib + 2;
  for (int m = 0; m < count; m++) {
    if (nrows[m] < 0) return 4;
    if (ncols[m] < 0) return 6;
    if (row0[m] < 0 || (long)row0[m] + nrows[m] > nblock_rows) return 3;
    if (col0[m] < 0 || col0[m] + ncols[m] > ld) return 5;
    if (nrows[m] > 0 && ncols[m] > 0 && !bufs[m]) return 7;
  }
  return 0;
}

// Runs the copy of `msgs` (device addresses filled in by the caller) with the
// table in the workspace.
int run_xfer(oa_context *c, std::vector<XferMsg> &msgs, char *dev_table) {
  if (msgs.empty()) return OA_SUCCESS;
  OA_HIP_CHECK(hipMemcpyAsync(dev_table, msgs.data(), msgs.size() * sizeof(XferMsg),
                              hipMemcpyHostToDevice, c->stream));
  const int count = (int)msgs.size();
  const int grid = count < 4096 ? count : 4096;
  hipLaunchKernelGGL(xfer_kernel, dim3(grid), dim3(256), 0, c->stream,
                     reinterpret_cast<const XferMsg *>(dev_table), count);
  OA_HIP_CHECK(hipPeekAtLastError());
  return OA_SUCCESS;
}

}  // namespace

extern "C" {

int oa_collate_rows(oa_handle_t c, int count, const int *row0, const int *nrows,
                    const long *col0, const int *ncols, const double *const *src,
                    int src_on_host, double *block, long ld, int nblock_rows) {
  if (!c) return OA_ERR_NO_DEVICE;
  const int bad = check_msgs(count, row0, nrows, col0, ncols,
                             reinterpret_cast<const void *const *>(src), block, ld, nblock_rows,
                             9);
  if (bad) return bad;
  if (count == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  // staging: message bytes (host sources only) after the table
  size_t data_bytes = 0;
  std::vector<size_t> off(count, 0);
  for (int m = 0; m < count; m++) {
    off[m] = data_bytes;
    if (src_on_host) data_bytes += al256((size_t)nrows[m] * ncols[m] * sizeof(double));
  }
  const size_t table = al256((size_t)count * sizeof(XferMsg));
  int st = oa_ws_require(c, table + data_bytes);
  if (st) return st;
  char *ws = (char *)c->ws.ptr;
  std::vector<XferMsg> msgs;
  msgs.reserve(count);
  for (int m = 0; m < count; m++) {
    if (nrows[m] == 0 || ncols[m] == 0) continue;
    XferMsg x{};
    x.src = src_on_host ? reinterpret_cast<const double *>(ws + table + off[m]) : src[m];
    x.src_ld = ncols[m];
    x.dst = block + (long)row0[m] * ld + col0[m];
    x.dst_ld = ld;
    x.nrows = nrows[m];
    x.ncols = ncols[m];
    msgs.push_back(x);
  }
  if (src_on_host && data_bytes) {
    st = pinned_require(c->pinned, data_bytes);
    if (st) return st;
    char *pin = (char *)c->pinned.ptr;
    for (int m = 0; m < count; m++)
      if (nrows[m] && ncols[m])
        memcpy(pin + off[m], src[m], (size_t)nrows[m] * ncols[m] * sizeof(double));
    OA_HIP_CHECK(hipMemcpyAsync(ws + table, pin, data_bytes, hipMemcpyHostToDevice, c->stream));
  }
  st = run_xfer(c, msgs, ws);
  if (st) return st;
  OA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return OA_SUCCESS;
}

int oa_scatter_rows(oa_handle_t c, int count, const int *row0, const int *nrows,
                    const long *col0, const int *ncols, double *const *dst, int dst_on_host,
                    const int *accumulate, const double *block, long ld, int nblock_rows) {
  if (!c) return OA_ERR_NO_DEVICE;
  const int bad = check_msgs(count, row0, nrows, col0, ncols,
                             reinterpret_cast<const void *const *>(dst), block, ld, nblock_rows,
                             10);
  if (bad) return bad;
  if (count == 0) return OA_SUCCESS;
  DeviceGuard g(c->device);
  size_t data_bytes = 0;
  std::vector<size_t> off(count, 0);
  for (int m = 0; m < count; m++) {
    off[m] = data_bytes;
    if (dst_on_host) data_bytes += al256((size_t)nrows[m] * ncols[m] * sizeof(double));
  }
  const size_t table = al256((size_t)count * sizeof(XferMsg));
  int st = oa_ws_require(c, table + data_bytes);
  if (st) return st;
  char *ws = (char *)c->ws.ptr;
  std::vector<XferMsg> msgs;
  msgs.reserve(count);
  for (int m = 0; m < count; m++) {
    if (nrows[m] == 0 || ncols[m] == 0) continue;
    XferMsg x{};
    x.src = block + (long)row0[m] * ld + col0[m];
    x.src_ld = ld;
    // host destinations: the kernel packs the slices, the host adds them
    x.dst = dst_on_host ? reinterpret_cast<double *>(ws + table + off[m]) : dst[m];
    x.dst_ld = ncols[m];
    x.nrows = nrows[m];
    x.ncols = ncols[m];
    x.accumulate = (!dst_on_host && accumulate && accumulate[m]) ? 1 : 0;
    msgs.push_back(x);
  }
  st = run_xfer(c, msgs, ws);
  if (st) return st;
  if (dst_on_host && data_bytes) {
    st = pinned_require(c->pinned, data_bytes);
    if (st) return st;
    char *pin = (char *)c->pinned.ptr;
    OA_HIP_CHECK(hipMemcpyAsync(pin, ws + table, data_bytes, hipMemcpyDeviceToHost, c->stream));
    OA_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int m = 0; m < count; m++) {
      const size_t len = (size_t)nrows[m] * ncols[m];
      if (!len) continue;
      const double *s = reinterpret_cast<const double *>(pin + off[m]);
      double *d = dst[m];
      if (accumulate && accumulate[m]) {
        for (size_t i = 0; i < len; i++) d[i] += s[i];  // fastAdd (fastadd.C:34-51)
      } else {
        memcpy(d, s, len * sizeof(double));
      }
    }
    return OA_SUCCESS;
  }
  OA_HIP_CHECK(hipStreamSynchronize(c->stream));
  return OA_SUCCESS;
}

}  // extern "C"
