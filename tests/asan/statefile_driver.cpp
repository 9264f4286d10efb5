// Host AddressSanitizer driver for the state-file reader and writer
// (openatom_amd/csrc/oa_statefile.hip, host code only: no GPU call is made).
//
// Built by tests/test_host_asan.py with hipcc, the sanitizer on the host side
// only, and run on the CPU.  It reads the reference's own gzip text state
// files, round-trips a gamma-point set through all four formats, and feeds
// the parsers malformed input: truncated records, a header that claims two
// billion records, a line longer than the reader's line buffer, an empty
// file, a binary file cut inside a record, capacities too small.  Every case
// must return a status, never read or write outside its buffers.
#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oa_ortho.h"

static int fails = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "statefile_driver:%d: %s\n", __LINE__, #c); \
      fails++;                                                     \
    }                                                              \
  } while (0)

static void write_text(const std::string &p, const std::string &s) {
  FILE *f = std::fopen(p.c_str(), "w");
  std::fwrite(s.data(), 1, s.size(), f);
  std::fclose(f);
}

static void write_gz(const std::string &p, const std::string &s) {
  gzFile g = gzopen((p + ".gz").c_str(), "wb");
  gzwrite(g, s.data(), (unsigned)s.size());
  gzclose(g);
}

struct Set {
  std::vector<double> coef;
  std::vector<int> kx, ky, kz;
  int npts = 0, nkx0 = 0, grid[3] = {0, 0, 0};
};

static int read_set(const std::string &p, int fmt, int dp, int cap, Set &s) {
  s.coef.assign(2 * (size_t)(cap > 0 ? cap : 1), 0.0);
  s.kx.assign(cap > 0 ? cap : 1, 0);
  s.ky.assign(cap > 0 ? cap : 1, 0);
  s.kz.assign(cap > 0 ? cap : 1, 0);
  return oa_state_read(p.c_str(), fmt, dp, cap, cap > 0 ? s.coef.data() : nullptr, s.kx.data(),
                       s.ky.data(), s.kz.data(), &s.npts, &s.nkx0, s.grid);
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: statefile_driver <golden state path (no .gz)> <tmp dir>\n");
    return 2;
  }
  const std::string golden = argv[1], tmp = std::string(argv[2]) + "/";

  // the reference's own k-point state file (gzip text), full sphere
  Set g;
  EXPECT(read_set(golden, 2, 0, 0, g) == OA_ERR_CAPACITY && g.npts > 0);
  const int np = g.npts;
  EXPECT(read_set(golden, 2, 0, np, g) == OA_SUCCESS && g.npts == np);
  EXPECT(read_set(golden, 2, 0, np - 1, g) == OA_ERR_CAPACITY);

  // a gamma-point half sphere: upper half of the kx = 0 plane, kx > 0, G = 0 last
  std::string gam = "9 4 4 4\n";
  const int pts[9][3] = {{0, 0, 1}, {0, 1, -1}, {0, 1, 0}, {0, 1, 1}, {1, -1, 0},
                         {1, 0, 0}, {1, 1, 1},  {2, 0, -1}, {0, 0, 0}};
  for (int i = 0; i < 9; i++) {
    char b[96];
    std::snprintf(b, sizeof(b), "%g %g %d %d %d \n", 0.5 + i, -0.25 * i, pts[i][0], pts[i][1],
                  pts[i][2]);
    gam += b;
  }
  write_text(tmp + "gam", gam);
  Set s;
  EXPECT(read_set(tmp + "gam", 0, 1, 64, s) == OA_SUCCESS);
  EXPECT(s.nkx0 == 2 * 4 + 1 && s.npts == 9 + 4);
  for (int fmt = 0; fmt < 4; fmt++) {
    const std::string p = tmp + "rt" + std::to_string(fmt);
    EXPECT(oa_state_write(p.c_str(), fmt, 1, s.npts, s.coef.data(), s.kx.data(), s.ky.data(),
                          s.kz.data(), s.grid) == OA_SUCCESS);
    Set r;
    EXPECT(read_set(p, fmt, 1, s.npts, r) == OA_SUCCESS && r.npts == s.npts);
    EXPECT(std::memcmp(r.coef.data(), s.coef.data(), 2 * sizeof(double) * s.npts) == 0 ||
           fmt % 2 == 0);  // text keeps %g digits only
    EXPECT(read_set(p, fmt, 1, s.npts - 1, r) == OA_ERR_CAPACITY);
  }

  // malformed input: each must give a status and touch nothing outside
  Set m;
  write_text(tmp + "trunc", "5 4 4 4\n1 2 0 0 1\n3 4 0 1");
  EXPECT(read_set(tmp + "trunc", 0, 0, 8, m) == OA_ERR_IO);
  write_text(tmp + "huge", "2000000000 4 4 4\n1 2 0 0 1\n");
  EXPECT(read_set(tmp + "huge", 0, 0, 8, m) == OA_ERR_IO);
  write_gz(tmp + "hugegz", "2000000000 4 4 4\n1 2 0 0 1\n");
  EXPECT(read_set(tmp + "hugegz", 2, 0, 8, m) == OA_ERR_IO);
  write_text(tmp + "neg", "-3 4 4 4\n");
  EXPECT(read_set(tmp + "neg", 0, 0, 8, m) == OA_ERR_IO);
  write_text(tmp + "empty", "");
  EXPECT(read_set(tmp + "empty", 0, 0, 8, m) == OA_ERR_IO);
  write_gz(tmp + "emptygz", "");
  EXPECT(read_set(tmp + "emptygz", 2, 0, 8, m) == OA_ERR_IO);
  {
    std::string longline = "2 4 4 4\n";
    longline += std::string(5000, ' ') + "1 2 0 0 1\n0.5 0.5 0 0 0\n";
    write_gz(tmp + "longgz", longline);
    (void)read_set(tmp + "longgz", 2, 1, 8, m);  // any status; no overrun
  }
  {
    // binary, cut in the middle of the second record
    const int h[4] = {3, 4, 4, 4};
    std::string b(reinterpret_cast<const char *>(h), sizeof(h));
    const double re = 1.0, im = 2.0;
    const int k3[3] = {0, 0, 1};
    for (int r = 0; r < 2; r++) {
      b.append(reinterpret_cast<const char *>(&re), 8);
      b.append(reinterpret_cast<const char *>(&im), 8);
      b.append(reinterpret_cast<const char *>(k3), r == 0 ? 12 : 5);
    }
    write_text(tmp + "bin", b);
    EXPECT(read_set(tmp + "bin", 1, 0, 8, m) == OA_ERR_IO);
    write_gz(tmp + "bin", b);
    EXPECT(read_set(tmp + "bin", 3, 0, 8, m) == OA_ERR_IO);
  }
  // a gamma-point file with no G = 0, or G = 0 not last
  write_text(tmp + "nog0", "2 4 4 4\n1 2 0 0 1\n3 4 1 0 0\n");
  EXPECT(read_set(tmp + "nog0", 0, 1, 8, m) == OA_ERR_IO);
  // argument checks
  EXPECT(oa_state_read(nullptr, 0, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr) == 1);
  EXPECT(oa_state_read("x", 7, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr) == 2);
  EXPECT(oa_state_read((tmp + "gam").c_str(), 0, 1, 4, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr) == 5);
  EXPECT(oa_state_read((tmp + "missing").c_str(), 0, 1, 0, nullptr, nullptr, nullptr, nullptr,
                       nullptr, nullptr, nullptr) == OA_ERR_IO);
  // a writer given a gamma-point set without G = 0
  {
    const double c[4] = {1, 2, 3, 4};
    const int x[2] = {0, 1}, y[2] = {0, 0}, z[2] = {1, 0}, gr[3] = {4, 4, 4};
    EXPECT(oa_state_write((tmp + "w").c_str(), 0, 1, 2, c, x, y, z, gr) == OA_ERR_IO);
    EXPECT(oa_state_write((tmp + "w").c_str(), 0, 0, 2, c, x, y, z, nullptr) == 9);
  }
  if (fails) return 1;
  std::printf("statefile_driver: ok\n");
  return 0;
}
