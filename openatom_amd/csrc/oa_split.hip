// Split-precision (Ozaki digit slice) kernels.  Not yet available: every
// entry reports OA_ERR_UNSUPPORTED so callers fail loudly.
#include "oa_internal.h"

size_t oa_split_ws_bytes(int, int, int) { return 0; }

int oa_launch_overlap_split(hipStream_t, int, int, int, const double *, long, int, double,
                            double *, double, double *, void *, size_t) {
  return OA_ERR_UNSUPPORTED;
}

int oa_launch_backmul_split(hipStream_t, int, int, int, const double *, int, const double *,
                            long, double, double, double *, long, void *, size_t) {
  return OA_ERR_UNSUPPORTED;
}
