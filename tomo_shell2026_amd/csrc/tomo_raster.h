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
// dividing.  With D = |den|, a = sign(den) * num_v, b = sign(den) * num_w (so v ~ a/D, w ~ b/D)
// and the slack M = 2^-12 (D + |a| + |b|), a pixel is rejected when
//   a < -eps*D - M   (v < -eps)          a > L*D + M       (v > 1 + eps)
//   a + b > (1+eps)*D + M  (u < -eps)    b < -eps*D - M    (u + v > 1 + eps, i.e. w < -(L-1))
// where L = 1.0009999 is the largest float not above 1 + (double)eps.  The correctly rounded
// quotients differ from a/D, b/D by at most 2^-24 relatively and the float evaluation of these
// bounds by a few 2^-24 (D + |a| + |b|), far inside M, so every rejected pixel is also rejected
// by the exact test (checked exhaustively against it in the host tests).
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

// True when the exact test may accept (never false for an accepted pixel).
TOMO_RHD bool bary_candidate(float den, float nv, float nw) {
  const float D = den >= 0.0f ? den : -den;
  const float a = den >= 0.0f ? nv : -nv;
  const float b = den >= 0.0f ? nw : -nw;
  const float M = 0.000244140625f * (D + (a >= 0.0f ? a : -a) + (b >= 0.0f ? b : -b));
  const float lo = -kEps * D - M;
  if (a < lo || b < lo) return false;
  if (a > kOnePlusEpsF * D + M) return false;
  if (a + b > (1.0f + kEps) * D + M) return false;
  return true;
}

// The exact acceptance; on success u, v, w are the barycentric weights.
TOMO_RHD bool bary_exact(float den, float nv, float nw, float &u, float &v, float &w) {
  v = nv / den;
  w = nw / den;
  u = (float)((1.0 - (double)v) - (double)w);
  return u >= -kEps && v >= -kEps && v <= kOnePlusEpsF && (u + v) <= kOnePlusEpsF;
}

}  // namespace tomo
