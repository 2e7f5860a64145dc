// tomo_raster.h -- the pixel-centre coverage test of the rasteriser, shared by the HIP kernel
// and the host unit test.
//
// The exact test is _getBaryCoord (Tomo_types.cpp:86-115) as triVoxel uses it
// (STomoSh_INT3.cpp:1163-1176): with e0 = b - a, e1 = c - a, e2 = p - a (z dropped),
//   v = (d11*d20 - d01*d21) / den,  w = (d00*d21 - d01*d20) / den,  den = d00*d11 - d01^2,
//   u = (float)((1.0 - (double)v) - (double)w)
//   accept iff |den| > eps, u >= -eps, v >= -eps, (double)v <= 1 + eps, (double)(u + v) <= 1 + eps
// with eps = 0.001f, all float operations unfused.  The numerators depend only on the pixel;
// den, d00, d01, d11 only on the triangle.
//
// The prefilter rejects pixels whose barycentric coordinates are clearly outside, without
// dividing and without the dot products of the exact test.  Over the triangle's pixel box,
// a = sign(den) * num_v and b = sign(den) * num_w are linear in the float offsets
// e2x = cx - v0x, e2y = cy - v0y of the pixel centre:
//   a = e2x * ax + e2y * ay,  ax = sign(den) (d11 e0x - d01 e1x),  ay = sign(den) (d11 e0y - d01 e1y)
//   b = e2x * bx + e2y * by,  bx = sign(den) (d00 e1x - d01 e0x),  by = sign(den) (d00 e1y - d01 e0y)
// With D = |den|, Sa = Ex (|d11 e0x| + |d01 e1x|) + Ey (|d11 e0y| + |d01 e1y|), Sb likewise for b,
// Ex, Ey the largest |e2x|, |e2y| in the box, and the slack M = 2^-12 D + 2^-16 (Sa + Sb), a
// pixel is rejected when
//   a < -eps*D - M   (v < -eps)          b < -eps*D - M   (u + v > 1 + eps)
//   a + b > (1+eps)*D + M  (u < -eps)
// The exact test's v <= 1 + eps is left to it: with b >= -eps*D - M and a + b <= (1+eps)*D + M
// it only removes pixels with v within about 2 eps above 1.  |e2x| <= Ex holds exactly
// (the float pixel-centre offset is monotone in the pixel index, so its extremes sit at the box
// ends).  The affine evaluation and the exact test's float numerators each lie within
// 4 * 2^-24 Sa (resp. Sb) of the real-number numerator of the same float inputs; the correctly
// rounded quotients and the double u add 2^-23 D + 2^-24 (Sa + Sb); the float bounds a few
// 2^-24 (D + M).  The total stays below 2^-20 (D + Sa + Sb), far inside M, so every rejected
// pixel is also rejected by the exact test (checked against it over whole pixel boxes of
// ordinary, sliver, tiny and huge triangles in the host tests).
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define TOMO_RHD __host__ __device__ __forceinline__
#else
#define TOMO_RHD inline
#endif

namespace tomo {

constexpr float kEps = 0.001f;
constexpr float kOnePlusEpsF = 1.00099992752075195f;  // largest float <= 1 + (double)0.001f

struct TriSetup {
  float v0x, v0y, e0x, e0y, e1x, e1y, d00, d01, d11, den;
};

// Per-triangle part of _getBaryCoord for margined vertices (v0, v1, v2).
TOMO_RHD TriSetup tri_setup(float v0x, float v0y, float v1x, float v1y, float v2x, float v2y) {
  TriSetup s;
  s.v0x = v0x;
  s.v0y = v0y;
  s.e0x = v1x - v0x;
  s.e0y = v1y - v0y;
  s.e1x = v2x - v0x;
  s.e1y = v2y - v0y;
  // _getBaryCoord's dot products include the z term 0 * 0 = +0.  Adding +0 changes nothing in
  // a sum of squares (never -0), so it is kept only in d01.
  s.d00 = s.e0x * s.e0x + s.e0y * s.e0y;
  s.d01 = s.e0x * s.e1x + s.e0y * s.e1y + 0.0f;
  s.d11 = s.e1x * s.e1x + s.e1y * s.e1y;
  s.den = s.d00 * s.d11 - s.d01 * s.d01;
  return s;
}

TOMO_RHD bool tri_degenerate(const TriSetup &s) {
  const float aden = (s.den >= 0.0f) ? s.den : -s.den;
  return !(aden > kEps);
}

// Numerators of v and w for the pixel centre (cx, cy).
TOMO_RHD void bary_num(const TriSetup &s, float cx, float cy, float &nv, float &nw) {
  const float e2x = cx - s.v0x, e2y = cy - s.v0y;
  const float d20 = e2x * s.e0x + e2y * s.e0y + 0.0f;
  const float d21 = e2x * s.e1x + e2y * s.e1y + 0.0f;
  nv = s.d11 * d20 - s.d01 * d21;
  nw = s.d00 * d21 - s.d01 * d20;
}

// Per-triangle constants of the prefilter for the pixel-centre offsets |e2x| <= ex, |e2y| <= ey.
struct TriScan {
  float ax, ay, bx, by;  // gradients of a and b (sign of den folded in)
  float lo, his;         // a, b >= lo;  a + b <= his
};

TOMO_RHD TriScan tri_scan(const TriSetup &s, float ex, float ey) {
  const bool pos = s.den >= 0.0f;
  const float D = pos ? s.den : -s.den;
  const float p0x = s.d11 * s.e0x, q1x = s.d01 * s.e1x, p0y = s.d11 * s.e0y, q1y = s.d01 * s.e1y;
  const float p1x = s.d00 * s.e1x, q0x = s.d01 * s.e0x, p1y = s.d00 * s.e1y, q0y = s.d01 * s.e0y;
  const float ax = p0x - q1x, ay = p0y - q1y, bx = p1x - q0x, by = p1y - q0y;
  TriScan t;
  t.ax = pos ? ax : -ax; t.ay = pos ? ay : -ay;
  t.bx = pos ? bx : -bx; t.by = pos ? by : -by;
  const float sa = ex * (__builtin_fabsf(p0x) + __builtin_fabsf(q1x)) +
                   ey * (__builtin_fabsf(p0y) + __builtin_fabsf(q1y));
  const float sb = ex * (__builtin_fabsf(p1x) + __builtin_fabsf(q0x)) +
                   ey * (__builtin_fabsf(p1y) + __builtin_fabsf(q0y));
  const float M = 0.000244140625f * D + 0.0000152587890625f * (sa + sb);
  t.lo = -kEps * D - M;
  t.his = (1.0f + kEps) * D + M;
  return t;
}

// Pixel-centre offset from v0 along one axis, as the exact test forms it.
TOMO_RHD float centre_offset(int32_t i, float h2, float v0) { return (((float)i + kEps) + h2) - v0; }

// True when the exact test may accept (never false for an accepted pixel; a NaN anywhere leaves
// the pixel to the exact test).  rax = e2x * ax and rbx = e2x * bx for the pixel's column.
TOMO_RHD bool scan_candidate(const TriScan &t, float rax, float rbx, float e2y) {
  const float a = __builtin_fmaf(e2y, t.ay, rax);
  const float b = __builtin_fmaf(e2y, t.by, rbx);
  return !(a < t.lo || b < t.lo || a + b > t.his);
}

// Upper bound on 1 / x for finite x > 0: the correctly rounded quotient (within 2^-24 of 1 / x)
// raised by 2^-10, the same operations on the host and the device, so the host tests of the
// accept box cover the kernel's arithmetic exactly.
TOMO_RHD float rcp_up(float x) { return (1.0f / x) * 1.0009765625f; }

// Offsets from v0 that an accepted pixel centre can have: a tightening of the pixel box.
// Every accepted pixel satisfies a >= lo, b >= lo and a + b <= his in real arithmetic (the
// prefilter's bounds above).  Dividing by the real determinant den_g = d00 d11 - d01^2 of the
// float Gram matrix, |den_g| >= Dl = D - Ed with Ed = 2^-19 d00 d11 + 2^-23 D (it bounds the
// rounding of den), gives v_g, w_g >= -s1 and v_g + w_g <= 1 + s1 with s1 = (eps D + M + Ed) / Dl,
// where (v_g, w_g) solves the float Gram system.  The true barycentric coordinates (v, w) of the
// centre offset (e2 = v e0 + w e1 exactly) differ from them by at most
// d = 2^-20 K (2 + 4 s1), K = (d00 + d11)^2 / Dl (the Gram matrix's rounding times its
// conditioning), so (v, w) lies in v, w >= -s, v + w <= 1 + s with s = s1 + 2 d.  That region has
// corners (-s, -s), (1 + 2s, -s), (-s, 1 + 2s), so
//   min(0, e0x, e1x) - 2 s (|e0x| + |e1x|) <= e2x <= max(0, e0x, e1x) + 2 s (|e0x| + |e1x|)
// and likewise along y.  ok is false (no tightening) for near-degenerate or non-finite input.
struct AcceptRange {
  float lox, hix, loy, hiy;
  bool ok;
};

TOMO_RHD AcceptRange accept_range(const TriSetup &s, const TriScan &t) {
  AcceptRange r;
  const float D = __builtin_fabsf(s.den);
  const float Ed = 0.0000019073486328125f * (s.d00 * s.d11) + 0.00000011920928955078125f * D;
  const float Dl = D - Ed;
  r.ok = Dl > 0.5f * D;  // false for NaN and infinities
  const float inv = rcp_up(r.ok ? Dl : 1.0f);
  const float s1 = ((-t.lo) + Ed) * inv;  // -lo = eps D + M
  const float dsum = s.d00 + s.d11;
  const float dk = 0.00000095367431640625f * (dsum * dsum * inv) * (2.0f + 4.0f * s1);
  const float sl = (s1 + 2.0f * dk) * 1.0009765625f;
  const float mx = 2.0f * sl * (__builtin_fabsf(s.e0x) + __builtin_fabsf(s.e1x)) * 1.0009765625f;
  const float my = 2.0f * sl * (__builtin_fabsf(s.e0y) + __builtin_fabsf(s.e1y)) * 1.0009765625f;
  const float nx = s.e0x < s.e1x ? s.e0x : s.e1x, xx = s.e0x > s.e1x ? s.e0x : s.e1x;
  const float ny = s.e0y < s.e1y ? s.e0y : s.e1y, xy = s.e0y > s.e1y ? s.e0y : s.e1y;
  r.lox = (nx < 0.0f ? nx : 0.0f) - mx;
  r.hix = (xx > 0.0f ? xx : 0.0f) + mx;
  r.loy = (ny < 0.0f ? ny : 0.0f) - my;
  r.hiy = (xy > 0.0f ? xy : 0.0f) + my;
  r.ok = r.ok && (r.hix - r.lox) < 1.0e6f && (r.hiy - r.loy) < 1.0e6f;
  return r;
}

// Drop from [i0, i1] the pixels whose float centre offset (centre_offset(i, h2, v0)) is surely
// below lo or above hi.  The centre offset is within 2^-22 (|i| + |v0| + h2 + 2) of
// i + eps + h2 - v0, and the bounds below are formed with errors of the same order; the margin
// covers both several times over.
TOMO_RHD void clip_span(int32_t &i0, int32_t &i1, float lo, float hi, float v0, float h2) {
  if (i1 < i0) return;
  const float fi0 = (float)i0, fi1 = (float)i1;
  const float mag = __builtin_fabsf(lo) + __builtin_fabsf(hi) + __builtin_fabsf(v0) + __builtin_fabsf(h2) +
                    __builtin_fabsf(fi0) + __builtin_fabsf(fi1) + 2.0f;
  const float margin = 0.015625f + 0.00000095367431640625f * mag;
  float ta = ((lo + v0) - (kEps + h2)) - margin;  // every i < ta is below lo
  float tb = ((hi + v0) - (kEps + h2)) + margin;  // every i > tb is above hi
  ta = ta > fi0 ? ta : fi0;                      // also clamps NaN to the box
  ta = ta < fi1 + 1.0f ? ta : fi1 + 1.0f;
  tb = tb < fi1 ? tb : fi1;
  tb = tb > fi0 - 1.0f ? tb : fi0 - 1.0f;
  i0 = (int32_t)__builtin_ceilf(ta);
  i1 = (int32_t)__builtin_floorf(tb);
}

// The exact acceptance; on success u, v, w are the barycentric weights.
TOMO_RHD bool bary_exact(float den, float nv, float nw, float &u, float &v, float &w) {
  v = nv / den;
  w = nw / den;
  u = (float)((1.0 - (double)v) - (double)w);
  return u >= -kEps && v >= -kEps && v <= kOnePlusEpsF && (u + v) <= kOnePlusEpsF;
}

}  // namespace tomo
