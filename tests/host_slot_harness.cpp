// Host build of the product's per-slot pairing (tomo_slot.h) for CPU unit tests.
#include <algorithm>
#include <cmath>

#include "../tomo_shell2026_amd/csrc/tomo_raster.h"
#include "../tomo_shell2026_amd/csrc/tomo_slot.h"

using namespace tomo;

template <int N>
static int run(const uint32_t *keys, int c, int cut_ssb, int cut_nvb, int explicit_ss, int bed_type,
               int bed_r, int32_t *out) {
  uint32_t k[N];
  for (int i = 0; i < N; i++) k[i] = i < c ? keys[i] : 0u;
  PairParams pp{cut_ssb, cut_nvb, explicit_ss, bed_type, bed_r};
  SlotResult r;
  pair_keys<N>(k, c, pp, r);
  out[0] = r.vo; out[1] = r.vss; out[2] = r.va; out[3] = r.vb; out[4] = r.vnv; out[5] = r.vtc;
  out[6] = r.npre; out[7] = r.n; out[8] = r.sabs; out[9] = r.last_z; out[10] = r.last_type;
  out[11] = r.has10; out[12] = r.flags; out[13] = (int32_t)bed_info(r, pp);
  for (int i = 0; i < N; i++) out[14 + i] = (int32_t)k[i];
  return 0;
}

extern "C" {
// Pair c keys with the 8-entry register network (c <= 8) or the generic network (c <= 63).
int harness_pair8(const uint32_t *keys, int c, int cs, int cn, int ex, int bt, int br, int32_t *out) {
  if (c < 0 || c > 8) return -1;
  return run<8>(keys, c, cs, cn, ex, bt, br, out);
}
int harness_pair63(const uint32_t *keys, int c, int cs, int cn, int ex, int bt, int br, int32_t *out) {
  if (c < 0 || c > 63) return -1;
  return run<63>(keys, c, cs, cn, ex, bt, br, out);
}
int harness_pair16(const uint32_t *keys, int c, int cs, int cn, int ex, int bt, int br, int32_t *out) {
  if (c < 0 || c > 16) return -1;
  return run<16>(keys, c, cs, cn, ex, bt, br, out);
}
int harness_pair4(const uint32_t *keys, int c, int cs, int cn, int ex, int bt, int br, int32_t *out) {
  if (c < 0 || c > 4) return -1;
  return run<4>(keys, c, cs, cn, ex, bt, br, out);
}
int harness_pair2(const uint32_t *keys, int c, int cs, int cn, int ex, int bt, int br, int32_t *out) {
  if (c < 0 || c > 2) return -1;
  return run<2>(keys, c, cs, cn, ex, bt, br, out);
}
uint32_t harness_make_key(int z, int nz) { return make_key(z, nz); }

// pair_two against pair_keys<2> for every ordered pair of keys from a pool whose z and nZ values
// straddle every decision the pairing makes (signs, the SS and NV cuts, the (1, 0) row, the
// int16 ends), every count 0..2, both SS modes and every bed type and radius: returns the number
// of disagreeing results (all SlotResult fields and the bed bits).
int harness_pair_two_mismatches(int cs, int cn) {
  const int zs[] = {-32768, -3, -1, 0, 1, 2, 5, 32767};
  const int ns[] = {-32768, -1000, cs - 1, cs, cs + 1, cn - 1, cn, cn + 1, -2, -1, 0, 1, 2, 999, 32767};
  uint32_t pool[sizeof zs / sizeof zs[0] * sizeof ns / sizeof ns[0]];
  int np = 0;
  for (int z : zs)
    for (int n : ns) pool[np++] = make_key(z, std::max(-32768, std::min(32767, n)));
  int bad = 0;
  for (int ex = 0; ex < 2; ex++)
    for (int bt = 0; bt < 4; bt++)
      for (int br = 0; br < 4; br += 3) {
        const PairParams pp{cs, cn, ex, bt, br};
        for (int i = 0; i < np; i++)
          for (int j = 0; j < np; j++)
            for (int c = 0; c <= 2; c++) {
              const uint32_t ka = c >= 1 ? pool[i] : 0u, kb = c >= 2 ? pool[j] : 0u;
              uint32_t k[2] = {ka, kb};
              SlotResult a, b;
              pair_keys<2>(k, c, pp, a);
              pair_two(ka, kb, c, pp, b);
              const bool same = a.vo == b.vo && a.vss == b.vss && a.va == b.va && a.vb == b.vb &&
                                a.vnv == b.vnv && a.vtc == b.vtc && a.npre == b.npre && a.n == b.n &&
                                a.sabs == b.sabs && a.last_z == b.last_z && a.last_type == b.last_type &&
                                a.has10 == b.has10 && a.flags == b.flags && bed_info(a, pp) == bed_info(b, pp);
              bad += same ? 0 : 1;
            }
      }
  return bad;
}

// Coverage test of the rasteriser over the pixel box [x0, x1] x [y0, y1] of one margined
// triangle, pixels x-major as the kernel walks them: out[i*5] = prefilter candidate,
// out[i*5+1] = exact accept, out[i*5+2..4] = u, v, w bits.
// The walk box [x0, x1] x [y0, y1] of one margined triangle after accept_range / clip_span, as
// the rasteriser tightens it: out = {x0, x1, y0, y1, ok}.
int harness_accept_box(const float *tri6, int x0, int x1, int y0, int y1, float h2, int32_t *out) {
  TriSetup s = tri_setup(tri6[0], tri6[1], tri6[2], tri6[3], tri6[4], tri6[5]);
  const float ex = std::max(std::fabs(centre_offset(x0, h2, s.v0x)), std::fabs(centre_offset(x1, h2, s.v0x)));
  const float ey = std::max(std::fabs(centre_offset(y0, h2, s.v0y)), std::fabs(centre_offset(y1, h2, s.v0y)));
  const TriScan sc = tri_scan(s, ex, ey);
  const AcceptRange ar = accept_range(s, sc);
  if (ar.ok && !tri_degenerate(s)) {
    clip_span(x0, x1, ar.lox, ar.hix, s.v0x, h2);
    clip_span(y0, y1, ar.loy, ar.hiy, s.v0y, h2);
  }
  out[0] = x0; out[1] = x1; out[2] = y0; out[3] = y1; out[4] = ar.ok ? 1 : 0;
  return 0;
}

int harness_bary(const float *tri6, int x0, int x1, int y0, int y1, float h2, uint32_t *out) {
  TriSetup s = tri_setup(tri6[0], tri6[1], tri6[2], tri6[3], tri6[4], tri6[5]);
  const float ex = std::max(std::fabs(centre_offset(x0, h2, s.v0x)), std::fabs(centre_offset(x1, h2, s.v0x)));
  const float ey = std::max(std::fabs(centre_offset(y0, h2, s.v0y)), std::fabs(centre_offset(y1, h2, s.v0y)));
  const TriScan sc = tri_scan(s, ex, ey);
  int i = 0;
  for (int x = x0; x <= x1; x++) {
    const float e2x = centre_offset(x, h2, s.v0x);
    const float rax = e2x * sc.ax, rbx = e2x * sc.bx;
    for (int y = y0; y <= y1; y++, i++) {
      float nv, nw, u = -1.f, v = -1.f, w = -1.f;
      bary_num(s, ((float)x + kEps) + h2, ((float)y + kEps) + h2, nv, nw);
      const bool cand = scan_candidate(sc, rax, rbx, centre_offset(y, h2, s.v0y));
      bool acc = !tri_degenerate(s) && bary_exact(s.den, nv, nw, u, v, w);
      if (!acc) { u = v = w = -1.f; }
      uint32_t ub, vb, wb;
      __builtin_memcpy(&ub, &u, 4); __builtin_memcpy(&vb, &v, 4); __builtin_memcpy(&wb, &w, 4);
      out[5 * i] = cand; out[5 * i + 1] = acc; out[5 * i + 2] = ub; out[5 * i + 3] = vb; out[5 * i + 4] = wb;
    }
  }
  return i;
}
}
