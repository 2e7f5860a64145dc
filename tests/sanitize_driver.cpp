// Driver for the AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU code: the oracle
// (oracle/tomo_oracle.c) and the host build of the kernels' per-slot pairing networks and
// rasteriser coverage helpers (tests/host_slot_harness.cpp).  tests/test_sanitize.py builds all
// three with g++/gcc -fsanitize=address,undefined and runs this program on cases it writes.
//
//   sanitize_driver CASES OUT
//
// CASES holds cases back to back: int32 {nTri, nVtx, nYPR}, the orc_opts bytes, tri[3 nTri]
// (int32), vtx[3 nVtx], vnrm[3 nVtx], ypr[3 nYPR] (float32).  For each case the program runs
// orc_sweep (two threads), orc_one_rows and orc_insert_log on the first orientation, and writes
// the sweep's records to OUT.  It then pairs seeded random slots with every register network
// and with the oracle's vslotPair, and walks seeded random triangles (ordinary, sliver, tiny,
// huge) through the coverage helpers.  Exit status: 0, or 1 with a message on a mismatch; a
// sanitizer finding aborts the program.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../oracle/tomo_oracle.h"

extern "C" {
int harness_pair2(const uint32_t *, int, int, int, int, int, int, int32_t *);
int harness_pair4(const uint32_t *, int, int, int, int, int, int, int32_t *);
int harness_pair8(const uint32_t *, int, int, int, int, int, int, int32_t *);
int harness_pair16(const uint32_t *, int, int, int, int, int, int, int32_t *);
int harness_pair63(const uint32_t *, int, int, int, int, int, int, int32_t *);
uint32_t harness_make_key(int z, int nz);
int harness_accept_box(const float *tri6, int x0, int x1, int y0, int y1, float h2, int32_t *out);
int harness_bary(const float *tri6, int x0, int x1, int y0, int y1, float h2, uint32_t *out);
}

namespace {

struct Rng {  // 64-bit LCG, top bits
  uint64_t s;
  uint32_t next() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(s >> 33);
  }
  float uni() { return (float)(next() & 0xffffff) / 16777216.0f; }
};

bool read_all(FILE *f, void *dst, size_t bytes) { return fread(dst, 1, bytes, f) == bytes; }

int run_cases(const char *in_path, const char *out_path) {
  FILE *f = fopen(in_path, "rb");
  FILE *o = fopen(out_path, "wb");
  if (!f || !o) {
    fprintf(stderr, "cannot open %s or %s\n", in_path, out_path);
    return 1;
  }
  int ncase = 0;
  for (;;) {
    int32_t hdr[3];
    if (!read_all(f, hdr, sizeof hdr)) break;
    const int32_t nt = hdr[0], nv = hdr[1], ny = hdr[2];
    if (nt < 0 || nv < 0 || ny < 1) {
      fprintf(stderr, "bad case header\n");
      return 1;
    }
    orc_opts opt;
    std::vector<int32_t> tri((size_t)3 * nt);
    std::vector<float> vtx((size_t)3 * nv), vn((size_t)3 * nv), ypr((size_t)3 * ny);
    if (!read_all(f, &opt, sizeof opt) || !read_all(f, tri.data(), tri.size() * 4) ||
        !read_all(f, vtx.data(), vtx.size() * 4) || !read_all(f, vn.data(), vn.size() * 4) ||
        !read_all(f, ypr.data(), ypr.size() * 4)) {
      fprintf(stderr, "short case %d\n", ncase);
      return 1;
    }
    std::vector<orc_record> rec((size_t)ny);
    if (orc_sweep(&opt, nt, tri.data(), nv, vtx.data(), vn.data(), ny, ypr.data(), rec.data(), 2) != 0) {
      fprintf(stderr, "orc_sweep failed on case %d\n", ncase);
      return 1;
    }
    const size_t g2 = (size_t)opt.grid * (size_t)opt.grid;
    std::vector<int32_t> pre(g2), fin(g2), svo(g2), svss(g2);
    std::vector<uint64_t> rows(g2 * 64);
    orc_record one;
    if (orc_one_rows(&opt, nt, tri.data(), nv, vtx.data(), vn.data(), ypr.data(), &one, pre.data(),
                     fin.data(), svo.data(), svss.data(), rows.data()) != 0 ||
        memcmp(&one, &rec[0], sizeof one) != 0) {
      fprintf(stderr, "orc_one_rows disagrees with orc_sweep on case %d\n", ncase);
      return 1;
    }
    std::vector<int32_t> log((size_t)4 << 16);
    const int64_t nlog = orc_insert_log(&opt, nt, tri.data(), nv, vtx.data(), vn.data(), ypr.data(),
                                        (int64_t)(log.size() / 4), log.data());
    if (nlog < 0) {
      fprintf(stderr, "orc_insert_log failed on case %d\n", ncase);
      return 1;
    }
    fwrite(rec.data(), sizeof(orc_record), rec.size(), o);
    ncase++;
  }
  fclose(f);
  fclose(o);
  printf("cases %d\n", ncase);
  return 0;
}

// Every register network that can hold c keys must agree with the 63-entry one, and the Vo and
// Vss of the 63-entry one must equal the oracle's vslotPair on the same rows.
int run_pairs() {
  Rng r{12345};
  int32_t ref[14 + 63], out[14 + 63];
  int bad = 0;
  for (int it = 0; it < 20000; it++) {
    const int c = (int)(r.next() % 64u);
    int16_t z[63], nz[63];
    uint32_t keys[63];
    int n = 0;
    while (n < c) {  // distinct keys
      const int zz = (int)(r.next() % 600u) - 100, nn = (int)(r.next() % 2001u) - 1000;
      const uint32_t k = harness_make_key(zz, nn);
      bool dup = false;
      for (int i = 0; i < n; i++) dup = dup || keys[i] == k;
      if (dup) continue;
      keys[n] = k;
      z[n] = (int16_t)zz;
      nz[n] = (int16_t)nn;
      n++;
    }
    const float theta = (30.0f + (float)(r.next() % 50u)) * 3.14159265f / 180.0f;
    int32_t cs, cn;
    orc_theta_cuts(theta, &cs, &cn);
    const int ex = (int)(r.next() & 1u), bt = (int)(r.next() % 4u), br = 1 + (int)(r.next() % 3u);
    if (harness_pair63(keys, c, cs, cn, ex, bt, br, ref) != 0) return 1;
    int (*nets[4])(const uint32_t *, int, int, int, int, int, int, int32_t *) = {
        harness_pair2, harness_pair4, harness_pair8, harness_pair16};
    const int caps[4] = {2, 4, 8, 16};
    for (int k = 0; k < 4; k++) {
      if (c > caps[k]) continue;
      if (nets[k](keys, c, cs, cn, ex, bt, br, out) != 0) return 1;
      if (memcmp(out, ref, 14 * sizeof(int32_t)) != 0) bad++;
    }
    int32_t vo = 0, vss = 0;
    int16_t zo[64], nzo[64], ty[64];
    if (orc_pair_slot(c, z, nz, theta, ex, &vo, &vss, zo, nzo, ty) < 0) return 1;
    if (vo != ref[0] || vss != ref[1]) bad++;
  }
  if (bad) fprintf(stderr, "pairing mismatches: %d\n", bad);
  printf("pairs ok\n");
  return bad ? 1 : 0;
}

// Coverage helpers on random triangles: the tightened walk box must lie inside the pixel box and
// hold every pixel the exact test accepts, and no accepted pixel may fail the prefilter.
int run_coverage() {
  Rng r{777};
  std::vector<uint32_t> px((size_t)5 * 64 * 64);
  int bad = 0;
  for (int it = 0; it < 4000; it++) {
    const int kind = it % 4;
    const float scale = kind == 2 ? 0.02f : (kind == 3 ? 60.0f : 6.0f);
    const float cx = 20.0f + 40.0f * r.uni(), cy = 20.0f + 40.0f * r.uni();
    float t[6];
    for (int q = 0; q < 3; q++) {
      t[2 * q] = cx + scale * (r.uni() - 0.5f);
      t[2 * q + 1] = cy + scale * (r.uni() - 0.5f);
    }
    if (kind == 1) {  // sliver: third vertex almost on the first edge
      const float a = r.uni();
      t[4] = t[0] + a * (t[2] - t[0]) + 1e-4f * (r.uni() - 0.5f);
      t[5] = t[1] + a * (t[3] - t[1]) + 1e-4f * (r.uni() - 0.5f);
    }
    float mnx = t[0], mxx = t[0], mny = t[1], mxy = t[1];
    for (int q = 1; q < 3; q++) {
      mnx = fminf(mnx, t[2 * q]); mxx = fmaxf(mxx, t[2 * q]);
      mny = fminf(mny, t[2 * q + 1]); mxy = fmaxf(mxy, t[2 * q + 1]);
    }
    int x0 = (int)mnx, x1 = (int)mxx, y0 = (int)mny, y1 = (int)mxy;
    x1 = x1 < x0 + 63 ? x1 : x0 + 63;
    y1 = y1 < y0 + 63 ? y1 : y0 + 63;
    int32_t box[5];
    harness_accept_box(t, x0, x1, y0, y1, 0.5f, box);
    if (box[4] && (box[0] < x0 || box[1] > x1 || box[2] < y0 || box[3] > y1)) bad++;
    const int n = harness_bary(t, x0, x1, y0, y1, 0.5f, px.data());
    int i = 0;
    for (int x = x0; x <= x1; x++)
      for (int y = y0; y <= y1; y++, i++) {
        const bool cand = px[5 * i] != 0, acc = px[5 * i + 1] != 0;
        if (acc && !cand) bad++;
        if (acc && box[4] && !(x >= box[0] && x <= box[1] && y >= box[2] && y <= box[3])) bad++;
      }
    if (n != i) bad++;
  }
  if (bad) fprintf(stderr, "coverage mismatches: %d\n", bad);
  printf("coverage ok\n");
  return bad ? 1 : 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: sanitize_driver CASES OUT\n");
    return 2;
  }
  if (run_cases(argv[1], argv[2]) != 0) return 1;
  if (run_pairs() != 0) return 1;
  if (run_coverage() != 0) return 1;
  return 0;
}
