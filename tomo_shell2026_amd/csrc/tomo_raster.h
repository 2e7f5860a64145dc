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
//   a < -eps*D - M   (v < -eps)          a > L*D + M       (v > 1 + eps)
//   b < -eps*D - M   (u + v > 1 + eps)   a + b > (1+eps)*D + M  (u < -eps)
// where L = 1.0009999 is the largest float not above 1 + (double)eps.  |e2x| <= Ex holds exactly
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
  s.d00 = s.e0x * s.e0x + s.e0y * s.e0y + 0.0f;
  s.d01 = s.e0x * s.e1x + s.e0y * s.e1y + 0.0f;
  s.d11 = s.e1x * s.e1x + s.e1y * s.e1y + 0.0f;
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
  float lo, hia, his;    // a, b >= lo;  a <= hia;  a + b <= his
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
  t.hia = kOnePlusEpsF * D + M;
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
  return !(a < t.lo || b < t.lo || a > t.hia || a + b > t.his);
}

// The exact acceptance; on success u, v, w are the barycentric weights.
TOMO_RHD bool bary_exact(float den, float nv, float nw, float &u, float &v, float &w) {
  v = nv / den;
  w = nw / den;
  u = (float)((1.0 - (double)v) - (double)w);
  return u >= -kEps && v >= -kEps && v <= kOnePlusEpsF && (u + v) <= kOnePlusEpsF;
}

}  // namespace tomo
