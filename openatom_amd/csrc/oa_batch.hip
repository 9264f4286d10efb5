// Batched independent instances: one step of many orthonormalisation
// problems of uneven size, each phase ONE persistent grouped launch over all
// instances (split precision) or one grouped launch per S->T phase (both
// precisions), replayable as a hipGraph, with a tile autotuner.
//
// The reference runs one GSP/PC/Ortho instance set per (spin, temper,
// k-point, bead) index (cpaimd.C:600-664, cpaimd.C:2566-2593); the instances
// never exchange data on this path, so they batch without communication.
//
// Sequence of one step (split precision):
//   fill counters -> chunk max -> (adaptive: slice counts) -> slice ->
//   overlap digit products (one launch per digit tile in use) -> reduce ->
//   S->T init -> maxit x (GEMM C = 3I - AB, GEMM pair B' = BC/2,
//   A' = CA/2, residual, update) -> finish -> slice T -> back-multiply
//   digit products.
// FP64: the overlap and the back-multiply are per-instance FP64-MFMA launches
// inside the same sequence; S->T is grouped as above.
// Every launch reads device tables built at plan time; the sequence performs
// no allocation, no memset and no host synchronisation, so it captures into
// a graph as is.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "oa_internal.h"

namespace {

constexpr int NS_NWG = 128;  // residual partial sums per instance (same order as oa_ns_iterate)

size_t align256(size_t b) { return (b + 255) / 256 * 256; }
int cdivi(long a, long b) { return (int)((a + b - 1) / b); }

int default_tile(int n) {
  // the 128 tile wastes (np/n)^2 of its MFMA work on padding rows
  const long n128 = (long)cdivi(n, 128) * 128, n64 = (long)cdivi(n, 64) * 64;
  return (n128 * n128 <= (n64 * n64) * 5 / 4) ? 128 : 64;
}

struct DevGuard {
  int prev = -1;
  bool ok = false;
  explicit DevGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) return;
    ok = (prev == dev) || hipSetDevice(dev) == hipSuccess;
  }
  ~DevGuard() {
    if (ok && prev >= 0) (void)hipSetDevice(prev);
  }
};

enum LaunchKind { LK_FILL = 0, LK_SPLIT, LK_GEMM, LK_NS, LK_F64_OVL, LK_F64_BACK };

struct Launch {
  int kind, phase, variant;
  oa_group g;     // device tables (grouped kinds)
  int inst;       // per-instance kinds
};

struct InstPlan {
  int n, K, kx0_end, nspin, maxit, tile, wm, bucket;
  double tol;
  const double *psi;
  long ld;
  double *out;
  long ldout;
  double *S, *T;
  char *split_ws;
  oa_split_plan sp;
  double *ns_ws;
  void *ovl_ws;
  size_t ovl_bytes;
};

// Host image of the device table region: arrays appended with 256-byte
// alignment, addressed later by their offsets.
struct Blob {
  std::vector<char> bytes;
  size_t put(const void *p, size_t n) {
    const size_t off = align256(bytes.size());
    bytes.resize(off + n);
    if (n) memcpy(bytes.data() + off, p, n);
    return off;
  }
  size_t reserve(size_t n) {
    const size_t off = align256(bytes.size());
    bytes.resize(off + n, 0);
    return off;
  }
};

}  // namespace

struct oa_batch {
  oa_context *c = nullptr;
  hipStream_t gstream = nullptr;  // capture and replay stream
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  int prec = OA_PREC_FP64, count = 0, maxit = 0;
  std::vector<InstPlan> ip;
  std::map<int, std::pair<int, int>> table;  // bucket -> (digit tile, GEMM WM)
  oa_workspace ws, tab;
  oa_ns_report *reports = nullptr;
  unsigned long long *counters = nullptr;
  long ncounters = 0;
  std::vector<Launch> seq;
  std::vector<oa_instance> inst;  // last plan, for re-planning
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(oa_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Grow-only allocation that at least doubles, so a batch whose membership
// keeps changing (the serving loop) reallocates rarely; oa_batch_reserve
// sizes it once up front.
int grow(oa_workspace &w, size_t bytes) {
  if (bytes <= w.bytes) return OA_SUCCESS;
  if (w.ptr) OA_HIP_CHECK(hipFree(w.ptr));
  w.ptr = nullptr;
  const size_t want = std::max(bytes, 2 * w.bytes);
  w.bytes = 0;
  const size_t gran = (size_t)1 << 21;
  const size_t sz = (want + gran - 1) / gran * gran;
  OA_HIP_CHECK(hipMalloc(&w.ptr, sz));
  w.bytes = sz;
  return OA_SUCCESS;
}

size_t ns_bytes(int n) { return align256((size_t)5 * n * n * 8) + align256(NS_NWG * 8); }

size_t inst_bytes(int n, int npw, int prec, int tile, bool need_S, bool need_T) {
  const int K = 2 * npw;
  size_t b = ns_bytes(n);
  if (need_S) b += align256((size_t)n * n * 8);
  if (need_T) b += align256((size_t)n * n * 8);
  if ((prec & 0xff) == OA_PREC_FP64) b += align256(oa_overlap_f64_ws_bytes(n, K));
  else b += align256(oa_split_plan_for(prec, n, K, tile).total);
  return b;
}

int validate(int count, const oa_instance *in) {
  for (int i = 0; i < count; i++) {
    const oa_instance &x = in[i];
    if (x.n < 1 || x.npw < 1 || !x.psi || !x.psi_out) return 3;
    if (x.ldpsi < 2L * x.npw || x.ldout < 2L * x.npw) return 3;
    if (x.nkx0 < 0 || x.nkx0 > x.npw) return 3;
    if (x.nspin != 1 && x.nspin != 2) return 3;
    if (x.maxit < 1) return 3;
  }
  return OA_SUCCESS;
}

// Group the instances with variant v (digit tile or WM) into one table.
template <class T, class GetArgs, class GetItems, class Sel>
void add_group(Blob &blob, std::vector<Launch> &seq, int kind, int phase, int variant,
               const std::vector<InstPlan> &ip, GetArgs args, GetItems items, Sel sel,
               std::vector<std::pair<size_t, size_t>> &fix) {
  std::vector<T> a;
  std::vector<int> prefix(1, 0);
  for (size_t i = 0; i < ip.size(); i++) {
    if (!sel(ip[i])) continue;
    const int it = items(ip[i], (int)i);
    if (it <= 0) continue;
    a.push_back(args(ip[i], (int)i));
    prefix.push_back(prefix.back() + it);
  }
  if (a.empty()) return;
  Launch L{};
  L.kind = kind;
  L.phase = phase;
  L.variant = variant;
  L.g.count = (int)a.size();
  L.g.total = prefix.back();
  const size_t oa = blob.put(a.data(), a.size() * sizeof(T));
  const size_t op = blob.put(prefix.data(), prefix.size() * sizeof(int));
  // device addresses are known after the table region is sized
  L.g.args = reinterpret_cast<const void *>(oa);
  L.g.prefix = reinterpret_cast<const int *>(op);
  fix.push_back({seq.size(), 0});
  seq.push_back(L);
}

int run_seq(oa_batch *b, hipStream_t s) {
  int st = OA_SUCCESS;
  for (const Launch &L : b->seq) {
    switch (L.kind) {
      case LK_FILL: st = oa_fill_u64(s, b->counters, b->ncounters); break;
      case LK_SPLIT: st = oa_split_grouped(s, b->prec, L.variant, L.phase, L.g); break;
      case LK_GEMM: st = oa_gemm_grouped(s, L.variant, L.g); break;
      case LK_NS: st = oa_ns_grouped(s, L.phase, L.g); break;
      case LK_F64_OVL: {
        const InstPlan &p = b->ip[L.inst];
        st = oa_launch_overlap_f64(s, p.n, p.K, p.psi, p.ld, p.kx0_end, 2.0, p.S, 0.0, nullptr,
                                   p.ovl_ws, p.ovl_bytes);
        break;
      }
      case LK_F64_BACK: {
        const InstPlan &p = b->ip[L.inst];
        oa_gemm_problem g{p.n, p.K, p.n, p.T, p.n, p.psi, p.ld, p.out, p.ldout, 1.0, 0.0, 0.0};
        st = oa_launch_gemm_f64(s, false, false, &g, 1, nullptr);
        break;
      }
      default: st = OA_ERR_UNSUPPORTED;
    }
    if (st) return st;
  }
  return OA_SUCCESS;
}

int plan(oa_batch *b, int count, const oa_instance *in, int prec) {
  oa_context *c = b->c;
  DevGuard dg(c->device);
  if (!dg.ok) return OA_ERR_NO_DEVICE;
  const bool split = (prec & 0xff) != OA_PREC_FP64;
  if (split && oa_split_ws_bytes(prec, 64, 64) == 0) return OA_ERR_UNSUPPORTED;
  // every instance must fit the digit kernels' exact int32 sums (checked
  // before the batch's state changes)
  if (split)
    for (int i = 0; i < count; i++) {
      auto t = b->table.find(cdivi(in[i].n, 32));
      const int tile = (t != b->table.end()) ? t->second.first : default_tile(in[i].n);
      const int chk = oa_split_check(prec, in[i].n, tile);
      if (chk) return chk;
    }
  OA_HIP_CHECK(hipStreamSynchronize(b->gstream));
  OA_HIP_CHECK(hipStreamSynchronize(c->stream));
  drop_graph(b);
  b->prec = prec & ~OA_PREC_REUSE_DIGITS;
  b->count = count;
  b->inst.assign(in, in + count);
  b->ip.assign(count, InstPlan{});
  b->maxit = 0;
  // instance configs and workspace layout
  size_t wtot = 0;
  std::vector<size_t> woff(count);
  for (int i = 0; i < count; i++) {
    const oa_instance &x = in[i];
    InstPlan &p = b->ip[i];
    p.n = x.n;
    p.K = 2 * x.npw;
    p.kx0_end = 2 * x.nkx0;
    p.nspin = x.nspin;
    p.maxit = x.maxit;
    p.tol = x.tol;
    p.psi = x.psi;
    p.ld = x.ldpsi;
    p.out = x.psi_out;
    p.ldout = x.ldout;
    p.bucket = cdivi(x.n, 32);
    auto t = b->table.find(p.bucket);
    p.tile = (t != b->table.end()) ? t->second.first : default_tile(x.n);
    p.wm = (t != b->table.end()) ? t->second.second : 2;
    woff[i] = wtot;
    wtot += inst_bytes(x.n, x.npw, prec, p.tile, !x.S, !x.T);
    b->maxit = std::max(b->maxit, x.maxit);
  }
  int st = grow(b->ws, wtot);
  if (st) return st;
  for (int i = 0; i < count; i++) {
    const oa_instance &x = in[i];
    InstPlan &p = b->ip[i];
    char *w = (char *)b->ws.ptr + woff[i];
    p.ns_ws = (double *)w;
    w += ns_bytes(x.n);
    p.S = x.S ? x.S : (double *)w;
    if (!x.S) w += align256((size_t)x.n * x.n * 8);
    p.T = x.T ? x.T : (double *)w;
    if (!x.T) w += align256((size_t)x.n * x.n * 8);
    if (split) {
      p.sp = oa_split_plan_for(prec, x.n, p.K, p.tile);
      p.split_ws = w;
    } else {
      p.ovl_ws = w;
      p.ovl_bytes = oa_overlap_f64_ws_bytes(x.n, p.K);
    }
  }
  // tables
  Blob blob;
  b->seq.clear();
  std::vector<std::pair<size_t, size_t>> fix;
  std::vector<oa_split_inst_args> sa(count);
  long ncnt = 0;
  std::vector<long> cnt_off(count, 0);
  if (split) {
    for (int i = 0; i < count; i++) {
      cnt_off[i] = ncnt;
      ncnt += 2L * b->ip[i].sp.nz;
    }
  }
  const size_t off_cnt = blob.reserve((size_t)std::max(ncnt, 1L) * 8);
  const size_t off_rep = blob.reserve((size_t)count * sizeof(oa_ns_report));
  b->ncounters = ncnt;
  // counters and reports get device addresses below; argument structs need
  // them now, so size the region first: tables go after a fixed prefix whose
  // device base is the table allocation itself.
  const size_t est = 1 << 20;
  st = grow(b->tab, est);
  if (st) return st;
  char *dbase = (char *)b->tab.ptr;
  auto dcnt = [&](long k) { return (unsigned long long *)(dbase + off_cnt) + k; };
  oa_ns_report *drep = (oa_ns_report *)(dbase + off_rep);
  auto all = [](const InstPlan &) { return true; };
  if (split) {
    for (int i = 0; i < count; i++) {
      InstPlan &p = b->ip[i];
      oa_split_inst_fill(p.sp, prec, p.n, p.K, p.psi, p.ld, p.kx0_end, 2.0, p.S, p.T, 0, 0.0,
                         p.out, p.ldout, p.split_ws, dcnt(cnt_off[i]),
                         dcnt(cnt_off[i] + p.sp.nz), &sa[i]);
    }
    Launch f{};
    f.kind = LK_FILL;
    b->seq.push_back(f);
    add_group<oa_chunkmax_args>(
        blob, b->seq, LK_SPLIT, OA_SP_CHUNKMAX, 128, b->ip,
        [&](const InstPlan &, int i) { return sa[i].cm; },
        [&](const InstPlan &, int i) { return sa[i].items_cm; }, all, fix);
    if (prec & OA_PREC_ADAPTIVE)
      add_group<oa_count_args>(
          blob, b->seq, LK_SPLIT, OA_SP_COUNT, 128, b->ip,
          [&](const InstPlan &, int i) { return sa[i].cnt; },
          [&](const InstPlan &, int) { return 1; }, all, fix);
    add_group<oa_slice_args>(
        blob, b->seq, LK_SPLIT, OA_SP_SLICE, 128, b->ip,
        [&](const InstPlan &, int i) { return sa[i].sl; },
        [&](const InstPlan &, int i) { return sa[i].items_sl; }, all, fix);
    for (int tile : {128, 64}) {
      auto sel = [tile](const InstPlan &p) { return p.tile == tile; };
      add_group<oa_digit_args>(
          blob, b->seq, LK_SPLIT, OA_SP_DIGIT0, tile, b->ip,
          [&](const InstPlan &, int i) { return sa[i].d0; },
          [&](const InstPlan &, int i) { return sa[i].items_d0; }, sel, fix);
    }
    for (int tile : {128, 64}) {
      auto sel = [tile](const InstPlan &p) { return p.tile == tile; };
      add_group<oa_sreduce_args>(
          blob, b->seq, LK_SPLIT, OA_SP_REDUCE, tile, b->ip,
          [&](const InstPlan &, int i) { return sa[i].rd; },
          [&](const InstPlan &, int i) { return sa[i].items_rd; }, sel, fix);
    }
  } else {
    for (int i = 0; i < count; i++) {
      Launch L{};
      L.kind = LK_F64_OVL;
      L.inst = i;
      b->seq.push_back(L);
    }
  }
  // S->T, grouped for both precisions
  std::vector<oa_ns_args> na(count);
  for (int i = 0; i < count; i++) {
    const InstPlan &p = b->ip[i];
    const long nn = (long)p.n * p.n;
    oa_ns_args &a = na[i];
    a.S = p.S;
    a.n = p.n;
    a.nspin = p.nspin;
    a.A[0] = p.ns_ws;
    a.A[1] = p.ns_ws + nn;
    a.B[0] = p.ns_ws + 2 * nn;
    a.B[1] = p.ns_ws + 3 * nn;
    a.C = p.ns_ws + 4 * nn;
    a.partial = (double *)((char *)p.ns_ws + align256((size_t)5 * nn * 8));
    a.nwg = NS_NWG;
    a.maxit = p.maxit;
    a.tol = p.tol;
    a.T = p.T;
    a.rep = drep + i;
  }
  auto ns_group = [&](int phase) {
    add_group<oa_ns_args>(
        blob, b->seq, LK_NS, phase, 0, b->ip, [&](const InstPlan &, int i) { return na[i]; },
        [&](const InstPlan &p, int) { return oa_ns_items(phase, p.n, NS_NWG); }, all, fix);
  };
  ns_group(OA_NS_INIT);
  for (int it = 0; it < b->maxit; it++) {
    const int cur = it & 1, nxt = cur ^ 1;
    // do_iteration (ortho.C:848-870): C = 3I - A*B
    for (int wm : {4, 2}) {
      auto sel = [wm, it](const InstPlan &p) { return p.wm == wm && it < p.maxit; };
      add_group<oa_gemm_item>(
          blob, b->seq, LK_GEMM, 0, wm, b->ip,
          [&](const InstPlan &p, int i) {
            const oa_ns_args &a = na[i];
            oa_gemm_item g;
            g.p = oa_gemm_problem{p.n, p.n, p.n, a.A[cur], p.n, a.B[cur], p.n, a.C, p.n,
                                  -1.0, 0.0, 3.0};
            g.skip = &a.rep->done;
            return g;
          },
          [&](const InstPlan &p, int) { return cdivi(p.n, 32 * wm) * cdivi(p.n, 32 * wm); },
          sel, fix);
    }
    // step_2 / step_3 (ortho.C:880-929): B' = B*C/2 ; A' = C*A/2 (two
    // problems per instance in one table)
    for (int wm : {4, 2}) {
      std::vector<oa_gemm_item> g2;
      std::vector<int> prefix(1, 0);
      for (int i = 0; i < count; i++) {
        const InstPlan &p = b->ip[i];
        if (p.wm != wm || it >= p.maxit) continue;
        const oa_ns_args &a = na[i];
        const int tiles = cdivi(p.n, 32 * wm) * cdivi(p.n, 32 * wm);
        oa_gemm_item x;
        x.skip = &a.rep->done;
        x.p = oa_gemm_problem{p.n, p.n, p.n, a.B[cur], p.n, a.C, p.n, a.B[nxt], p.n, 0.5, 0.0, 0.0};
        g2.push_back(x);
        prefix.push_back(prefix.back() + tiles);
        x.p = oa_gemm_problem{p.n, p.n, p.n, a.C, p.n, a.A[cur], p.n, a.A[nxt], p.n, 0.5, 0.0, 0.0};
        g2.push_back(x);
        prefix.push_back(prefix.back() + tiles);
      }
      if (g2.empty()) continue;
      Launch L{};
      L.kind = LK_GEMM;
      L.variant = wm;
      L.g.count = (int)g2.size();
      L.g.total = prefix.back();
      L.g.args = reinterpret_cast<const void *>(blob.put(g2.data(), g2.size() * sizeof(oa_gemm_item)));
      L.g.prefix = reinterpret_cast<const int *>(blob.put(prefix.data(), prefix.size() * sizeof(int)));
      fix.push_back({b->seq.size(), 0});
      b->seq.push_back(L);
    }
    ns_group(cur == 0 ? OA_NS_RESID0 : OA_NS_RESID1);
    ns_group(OA_NS_UPDATE);
  }
  ns_group(OA_NS_FINISH);
  if (split) {
    add_group<oa_rows_args>(
        blob, b->seq, LK_SPLIT, OA_SP_ROWS, 128, b->ip,
        [&](const InstPlan &, int i) { return sa[i].rows; },
        [&](const InstPlan &, int i) { return sa[i].items_rows; }, all, fix);
    for (int tile : {128, 64}) {
      auto sel = [tile](const InstPlan &p) { return p.tile == tile; };
      add_group<oa_digit_args>(
          blob, b->seq, LK_SPLIT, OA_SP_DIGIT1, tile, b->ip,
          [&](const InstPlan &, int i) { return sa[i].d1; },
          [&](const InstPlan &, int i) { return sa[i].items_d1; }, sel, fix);
    }
  } else {
    for (int i = 0; i < count; i++) {
      Launch L{};
      L.kind = LK_F64_BACK;
      L.inst = i;
      b->seq.push_back(L);
    }
  }
  // the table region: grow (and rebuild the address-carrying structs) if the
  // estimate was short
  if (blob.bytes.size() > b->tab.bytes) {
    OA_HIP_CHECK(hipFree(b->tab.ptr));
    b->tab.ptr = nullptr;
    b->tab.bytes = 0;
    st = grow(b->tab, blob.bytes.size() * 2);
    if (st) return st;
    return plan(b, count, in, prec);
  }
  for (auto &f : fix) {
    Launch &L = b->seq[f.first];
    L.g.args = dbase + (size_t)L.g.args;
    L.g.prefix = reinterpret_cast<const int *>(dbase + (size_t)L.g.prefix);
  }
  b->counters = dcnt(0);
  b->reports = drep;
  OA_HIP_CHECK(hipMemcpy(b->tab.ptr, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice));
  return OA_SUCCESS;
}

double time_batch(oa_batch *b, int reps) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1.0;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1.0;
  }
  double ms = -1.0;
  if (run_seq(b, b->c->stream) == OA_SUCCESS && hipEventRecord(e0, b->c->stream) == hipSuccess) {
    bool ok = true;
    for (int r = 0; r < reps && ok; r++) ok = run_seq(b, b->c->stream) == OA_SUCCESS;
    float f = 0.f;
    if (ok && hipEventRecord(e1, b->c->stream) == hipSuccess &&
        hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&f, e0, e1) == hipSuccess)
      ms = f / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

}  // namespace

extern "C" {

int oa_batch_create(oa_handle_t h, oa_batch_t *out) {
  if (!h) return 1;
  if (!out) return 2;
  *out = nullptr;
  DevGuard dg(h->device);
  if (!dg.ok) return OA_ERR_NO_DEVICE;
  oa_batch *b = new oa_batch();
  b->c = h;
  if (hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    oa_batch_destroy(b);
    return OA_ERR_HIP;
  }
  *out = b;
  return OA_SUCCESS;
}

int oa_batch_destroy(oa_batch_t b) {
  if (!b) return OA_SUCCESS;
  DevGuard dg(b->c->device);
  if (b->gstream) (void)hipStreamSynchronize(b->gstream);
  (void)hipStreamSynchronize(b->c->stream);
  drop_graph(b);
  if (b->ws.ptr) (void)hipFree(b->ws.ptr);
  if (b->tab.ptr) (void)hipFree(b->tab.ptr);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->gstream) (void)hipStreamDestroy(b->gstream);
  delete b;
  return OA_SUCCESS;
}

int oa_batch_plan(oa_batch_t b, int count, const oa_instance *inst, int prec) {
  if (!b) return 1;
  if (count < 0) return 2;
  if (count > 0 && !inst) return 3;
  const int v = validate(count, inst);
  if (v) return v;
  return plan(b, count, inst, prec);
}

size_t oa_batch_footprint(oa_batch_t b) { return b ? b->ws.bytes + b->tab.bytes : 0; }

int oa_batch_reserve(oa_batch_t b, size_t workspace_bytes, size_t table_bytes) {
  if (!b) return 1;
  DevGuard dg(b->c->device);
  if (!dg.ok) return OA_ERR_NO_DEVICE;
  OA_HIP_CHECK(hipStreamSynchronize(b->gstream));
  OA_HIP_CHECK(hipStreamSynchronize(b->c->stream));
  if (workspace_bytes > b->ws.bytes || table_bytes > b->tab.bytes) {
    // the plan's addresses point into these allocations: plan again after
    drop_graph(b);
    int st = grow(b->ws, workspace_bytes);
    if (st) return st;
    st = grow(b->tab, table_bytes);
    if (st) return st;
    if (b->count > 0) return plan(b, b->count, b->inst.data(), b->prec);
  }
  return OA_SUCCESS;
}

size_t oa_instance_footprint(int n, int npw, int prec) {
  if (n < 1 || npw < 1) return 0;
  const size_t a = inst_bytes(n, npw, prec, 128, true, true);
  const size_t c = (prec & 0xff) == OA_PREC_FP64 ? a : inst_bytes(n, npw, prec, 64, true, true);
  return std::max(a, c) + 4096;
}

int oa_batch_run(oa_batch_t b) {
  if (!b) return 1;
  if (b->count == 0) return OA_SUCCESS;
  DevGuard dg(b->c->device);
  return run_seq(b, b->c->stream);
}

int oa_batch_launch(oa_batch_t b) {
  if (!b) return 1;
  if (b->count == 0) return OA_SUCCESS;
  DevGuard dg(b->c->device);
  if (!b->exec) {
    OA_HIP_CHECK(hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal));
    const int st = run_seq(b, b->gstream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->gstream, &g);
    if (st != OA_SUCCESS || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return st ? st : OA_ERR_HIP;
    }
    b->graph = g;
    OA_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  OA_HIP_CHECK(hipEventRecord(b->ev_in, b->c->stream));
  OA_HIP_CHECK(hipStreamWaitEvent(b->gstream, b->ev_in, 0));
  OA_HIP_CHECK(hipGraphLaunch(b->exec, b->gstream));
  OA_HIP_CHECK(hipEventRecord(b->ev_out, b->gstream));
  OA_HIP_CHECK(hipStreamWaitEvent(b->c->stream, b->ev_out, 0));
  return OA_SUCCESS;
}

int oa_batch_reports(oa_batch_t b, oa_ns_report *host, int count) {
  if (!b) return 1;
  if (count < 0 || count > b->count) return 3;
  if (count > 0 && !host) return 2;
  if (count == 0) return OA_SUCCESS;
  DevGuard dg(b->c->device);
  OA_HIP_CHECK(hipStreamSynchronize(b->gstream));
  OA_HIP_CHECK(hipStreamSynchronize(b->c->stream));
  OA_HIP_CHECK(hipMemcpy(host, b->reports, (size_t)count * sizeof(oa_ns_report),
                         hipMemcpyDeviceToHost));
  return OA_SUCCESS;
}

int oa_batch_config(oa_batch_t b, int i, int *tile, int *wm) {
  if (!b) return 1;
  if (i < 0 || i >= b->count) return 2;
  if (tile) *tile = b->ip[i].tile;
  if (wm) *wm = b->ip[i].wm;
  return OA_SUCCESS;
}

int oa_batch_tune(oa_batch_t b, int reps) {
  if (!b) return 1;
  if (reps < 1) return 2;
  if (b->count == 0) return OA_SUCCESS;
  DevGuard dg(b->c->device);
  const bool split = (b->prec & 0xff) != OA_PREC_FP64;
  const std::vector<oa_instance> inst = b->inst;
  const int prec = b->prec;
  // start from the current choices of every bucket in the batch
  std::map<int, std::pair<int, int>> cfg;
  for (const InstPlan &p : b->ip) cfg[p.bucket] = {p.tile, p.wm};
  auto measure = [&](double &ms) {
    b->table = cfg;
    int st = plan(b, (int)inst.size(), inst.data(), prec);
    if (st) return st;
    ms = time_batch(b, reps);
    return ms < 0 ? OA_ERR_HIP : OA_SUCCESS;
  };
  double best = 0;
  int st = measure(best);
  if (st) return st;
  for (auto &kv : cfg) {
    const int bucket = kv.first;
    std::vector<std::pair<int, int>> alts;
    const auto cur = cfg[bucket];
    if (split) alts.push_back({cur.first == 128 ? 64 : 128, cur.second});
    alts.push_back({cur.first, cur.second == 4 ? 2 : 4});
    for (const auto &alt : alts) {
      const auto keep = cfg[bucket];
      cfg[bucket] = alt;
      double ms = 0;
      st = measure(ms);
      if (st) return st;
      if (ms < best) best = ms;
      else cfg[bucket] = keep;
    }
  }
  b->table = cfg;
  return plan(b, (int)inst.size(), inst.data(), prec);
}

int oa_batch_get_table(oa_batch_t b, char *buf, int bytes) {
  if (!b) return 1;
  std::string s;
  char line[64];
  for (const auto &kv : b->table) {
    snprintf(line, sizeof line, "%d %d %d\n", kv.first, kv.second.first, kv.second.second);
    s += line;
  }
  if (!buf || bytes <= (int)s.size()) return (int)s.size() + 1;
  memcpy(buf, s.c_str(), s.size() + 1);
  return OA_SUCCESS;
}

int oa_batch_set_table(oa_batch_t b, const char *text) {
  if (!b) return 1;
  if (!text) return 2;
  std::map<int, std::pair<int, int>> t;
  const char *p = text;
  while (*p) {
    int bucket, tile, wm, used = 0;
    if (sscanf(p, " %d %d %d%n", &bucket, &tile, &wm, &used) != 3) break;
    if ((tile != 128 && tile != 64) || (wm != 2 && wm != 4) || bucket < 1) return 2;
    t[bucket] = {tile, wm};
    p += used;
  }
  b->table = t;
  if (b->count > 0) return plan(b, b->count, b->inst.data(), b->prec);
  return OA_SUCCESS;
}

}  // extern "C"
