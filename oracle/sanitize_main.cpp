// Standalone driver of the CPU oracle for its sanitizer build (test
// infrastructure: tests/test_oracle_sanitized.py builds it with
// -fsanitize=address,undefined and compares what it computes with the
// ordinary library on the same input).
//
// Input file (little-endian):
//   orc_params, int32 precision, int32 nsteps,
//   int32 nrigid, nrigid x (int32 obj, double cm[3]),
//   int32 nset,  nset x (int32 field, int64 bytes, bytes),
//   int32 nget,  nget x (int32 field, int64 bytes)
// Output file: nget x (int32 field, int64 bytes, bytes), then double[4]
// globals.  Exit status 0 on success, 2 on a malformed input or an oracle
// error.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "oracle.h"

namespace {

template <class T>
bool rd(FILE* f, T* v, size_t n = 1) {
  return std::fread(v, sizeof(T), n, f) == n;
}

int fail(const char* what) {
  std::fprintf(stderr, "sanitize_main: %s\n", what);
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) return fail("usage: sanitize_main IN OUT");
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return fail("cannot open input");
  orc_params p;
  int32_t precision = 0, nsteps = 0, nrigid = 0, nset = 0, nget = 0;
  if (!rd(in, &p) || !rd(in, &precision) || !rd(in, &nsteps) || !rd(in, &nrigid))
    return fail("short header");
  if (nrigid < 0 || nrigid > 64) return fail("bad rigid count");
  std::vector<int32_t> robj(nrigid);
  std::vector<double> rcm(3 * (size_t)nrigid);
  for (int i = 0; i < nrigid; ++i)
    if (!rd(in, &robj[i]) || !rd(in, &rcm[3 * i], 3)) return fail("short rigid record");
  orc_set_threads(1);
  void* h = orc_create(&p, precision);
  if (!h) return fail("orc_create failed");
  int rc = 0;
  if (!rd(in, &nset) || nset < 0 || nset > 64) rc = fail("bad field count");
  for (int i = 0; i < nset && rc == 0; ++i) {
    int32_t field;
    int64_t bytes;
    if (!rd(in, &field) || !rd(in, &bytes) || bytes < 0 || bytes > (int64_t)1 << 34) {
      rc = fail("bad field record");
      break;
    }
    std::vector<unsigned char> buf((size_t)bytes);
    if (!rd(in, buf.data(), buf.size())) rc = fail("short field data");
    else if (orc_set(h, field, buf.data()) != 0) rc = fail("orc_set failed");
  }
  for (int i = 0; i < nrigid && rc == 0; ++i)
    if (orc_set_rigid_object(h, robj[i], &rcm[3 * i]) != 0) rc = fail("orc_set_rigid_object failed");
  if (rc == 0 && orc_step(h, nsteps) != 0) rc = fail("orc_step failed");
  FILE* out = rc == 0 ? std::fopen(argv[2], "wb") : nullptr;
  if (rc == 0 && !out) rc = fail("cannot open output");
  if (rc == 0 && (!rd(in, &nget) || nget < 0 || nget > 64)) rc = fail("bad output count");
  for (int i = 0; i < nget && rc == 0; ++i) {
    int32_t field;
    int64_t bytes;
    if (!rd(in, &field) || !rd(in, &bytes) || bytes < 0 || bytes > (int64_t)1 << 34) {
      rc = fail("bad output record");
      break;
    }
    std::vector<unsigned char> buf((size_t)bytes);
    if (orc_get(h, field, buf.data()) != 0) {
      rc = fail("orc_get failed");
      break;
    }
    std::fwrite(&field, 4, 1, out);
    std::fwrite(&bytes, 8, 1, out);
    std::fwrite(buf.data(), 1, buf.size(), out);
  }
  if (rc == 0) {
    double g[4];
    orc_globals(h, g);
    std::fwrite(g, 8, 4, out);
  }
  if (out) std::fclose(out);
  std::fclose(in);
  orc_destroy(h);
  return rc;
}
