// tomo_kernels.hip -- the per-orientation support-tomography sweep for gfx950.
//
// One persistent work-group (4 waves of 64) owns one orientation at a time and walks the
// orientation list with a grid stride.  Per orientation it
//   (1) rotates every vertex and min-reduces the rotated AABB          (STomoSh_Base.cpp:45-109)
//   (2) rasterises every triangle against pixel centres and appends     (STomoSh_INT3.cpp:1138-1209,
//       (z, nZ) keys to its private slot table in HBM                    STomoVoxelSpaceInfo.cpp:175-230)
//   (3) pairs each touched slot: <= 8 entries in registers, larger slots (STomoSh_INT3.cpp:1041-1062)
//       one wave per slot with an LDS bitonic sort and ballot scans
//   (4) lays the raft/brim/skirt bed with the reference's raster-order   (STomoSh_INT3.cpp:1331-1443)
//       semantics and
//   (5) reduces the totals with wave shuffles into one record            (STomoSh_INT3.cpp:1211-1255)
// There is no MFMA: this is rasterisation, sorting and integer reduction.  All float arithmetic
// follows the reference's operation order without contraction (built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include "tomo_device.h"

namespace tomo {

namespace {

constexpr float kMargin = 0.001f;           // g_fMARGIN (Tomo_types.h:40)
constexpr float kNormFactor = 1000.0f;      // g_fNORMALFACTOR (Tomo_types.h:39)

struct alignas(16) Shared {
  uint64_t sortbuf[kWaves][kBigCap];
  int64_t acc[kWaves][20];
  float red[kWaves][3];
  int32_t bbox[4];
  uint32_t big_top, pool_top;
  int32_t flags;
  uint32_t scan[kWaves];
};

// Release own global writes, meet, then drop possibly stale L1 lines before reading what other
// waves of the work-group wrote (slot tables live in global memory).
__device__ __forceinline__ void wg_sync_global() {
  __threadfence();
  __syncthreads();
  __threadfence();
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// SMatrix4f::Dot row (SMatrix4f.cpp:242-262): r = D0*a0; r += D1*a1; r += D2*a2; r += 0*1.
__device__ __forceinline__ float dot_row(float r0, float r1, float r2, float a0, float a1, float a2) {
  float s = r0 * a0;
  s = s + r1 * a1;
  s = s + r2 * a2;
  s = s + 0.0f;
  return s;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ int msb64(uint64_t m) { return 63 - __clzll((long long)m); }

struct Tables {
  uint32_t *count, *keys, *ranks, *bedinfo, *off, *cursor, *pool_slot, *pool_key, *pool_rank,
      *cmp_key, *cmp_rank;
  int32_t *big;
  int32_t pool;
};

struct Acc {
  int64_t Va, Vb, Vo, Vss, Vnv, Vtc, nbed, nent, nslot, nfinal;
  uint64_t chkpre, chkfinal;
  int64_t aVo, aVss, aVa, aVb, aVnv, aVtc;
  int32_t flags;
};

__device__ __forceinline__ int64_t iabs64(int32_t v) { return v < 0 ? -(int64_t)v : (int64_t)v; }

__device__ __forceinline__ void account(Acc &a, const SlotResult &r, int32_t s, const SweepParams &p) {
  a.Va += r.va; a.Vb += r.vb; a.Vo += r.vo; a.Vss += r.vss; a.Vnv += r.vnv; a.Vtc += r.vtc;
  a.aVo += iabs64(r.vo); a.aVss += iabs64(r.vss); a.aVa += iabs64(r.va); a.aVb += iabs64(r.vb);
  a.aVnv += iabs64(r.vnv); a.aVtc += iabs64(r.vtc);
  a.nent += r.npre;
  a.chkpre += (uint64_t)(s + 1) * (uint64_t)r.npre;
  a.flags |= r.flags;
  if (r.sabs > 32767) a.flags |= F_INT16_WRAP;
  if (p.dbg_pre) {
    p.dbg_pre[s] = r.npre;
    p.dbg_vo[s] = r.vo;
    p.dbg_vss[s] = r.vss;
  }
}

// float -> int16 as SLOT_BUFFER_TYPE(f) stores it; flags values outside int16.
__device__ __forceinline__ int32_t to_i16(float f, int32_t &flags) {
  int32_t i = (int32_t)f;
  if (!(f > -2147483648.0f && f < 2147483648.0f) || i != (int32_t)(int16_t)i) flags |= F_INT16_WRAP;
  return (int32_t)(int16_t)i;
}

struct Emit {
  int32_t G, G2;
  Tables t;
  Shared *sh;
  int32_t bx0, bx1, by0, by1;
  int32_t flags;

  // SetBit's slot addressing (coord2vID, STomoVoxelSpaceInfo.cpp:97-105) + append to the slot.
  __device__ __forceinline__ int32_t put(float px, float py, float pz, float nz, uint32_t rank) {
    int32_t sx = (int32_t)px, sy = (int32_t)py;
    if (!(px > -2147483648.0f && px < 2147483648.0f && py > -2147483648.0f && py < 2147483648.0f) ||
        sx < 0 || sy < 0 || sx >= G || sy >= G) {
      flags |= F_FOOTPRINT_OOB;
      return -1;
    }
    int32_t z16 = to_i16(pz, flags);
    int32_t nz16 = to_i16(nz * kNormFactor, flags);
    uint32_t key = make_key(z16, nz16);
    int32_t s = sx * G + sy;
    bx0 = min(bx0, sx); bx1 = max(bx1, sx);
    by0 = min(by0, sy); by1 = max(by1, sy);
    uint32_t k = atomicAdd(&t.count[s], 1u);
    if (k < (uint32_t)kInline) {
      t.keys[(int64_t)k * G2 + s] = key;
      t.ranks[(int64_t)k * G2 + s] = rank;
    } else {
      uint32_t j = atomicAdd(&sh->pool_top, 1u);
      if (j < (uint32_t)t.pool) {
        t.pool_slot[j] = (uint32_t)s;
        t.pool_key[j] = key;
        t.pool_rank[j] = rank;
      }
    }
    if (k == (uint32_t)kSmall) {
      uint32_t b = atomicAdd(&sh->big_top, 1u);
      t.big[b] = s;
    }
    return s;
  }
};

// _getBaryCoord (Tomo_types.cpp:86-115) split into a per-triangle and a per-pixel part, and
// triVoxel (STomoSh_INT3.cpp:1138-1209).  a*: placed vertices; n*z: rotated normal z.
__device__ __forceinline__ void raster_tri(Emit &em, const SweepParams &p, int32_t t, float a0x,
                                           float a0y, float a0z, float a1x, float a1y, float a1z,
                                           float a2x, float a2y, float a2z, float n0z, float n1z,
                                           float n2z, float vlim) {
  const float v0x = a0x + kMargin, v0y = a0y + kMargin, v0z = a0z + kMargin;
  const float v1x = a1x + kMargin, v1y = a1y + kMargin, v1z = a1z + kMargin;
  const float v2x = a2x + kMargin, v2y = a2y + kMargin, v2z = a2z + kMargin;
  float mnx = (v0x <= v1x) ? v0x : v1x; mnx = (mnx <= v2x) ? mnx : v2x;
  float mny = (v0y <= v1y) ? v0y : v1y; mny = (mny <= v2y) ? mny : v2y;
  float mxx = (v0x >= v1x) ? v0x : v1x; mxx = (mxx >= v2x) ? mxx : v2x;
  float mxy = (v0y >= v1y) ? v0y : v1y; mxy = (mxy >= v2y) ? mxy : v2y;
  int32_t x0 = (int32_t)mnx, y0 = (int32_t)mny, x1 = (int32_t)mxx, y1 = (int32_t)mxy;
  // Pixels far outside the grid can only yield dropped (out-of-grid) entries: bound the loop so
  // a wild vertex cannot stall the work-group, and report the footprint overflow.
  const int32_t lo = -64, hi = p.grid + 64;
  if (x0 < lo || y0 < lo || x1 > hi || y1 > hi) {
    em.flags |= F_FOOTPRINT_OOB;
    x0 = max(x0, lo); y0 = max(y0, lo); x1 = min(x1, hi); y1 = min(y1, hi);
  }
  const float e0x = v1x - v0x, e0y = v1y - v0y;
  const float e1x = v2x - v0x, e1y = v2y - v0y;
  const float d00 = e0x * e0x + e0y * e0y + 0.0f;
  const float d01 = e0x * e1x + e0y * e1y + 0.0f;
  const float d11 = e1x * e1x + e1y * e1y + 0.0f;
  const float den = d00 * d11 - d01 * d01;
  const float aden = (den >= 0.0f) ? den : -den;
  int32_t hits = 0;
  const uint32_t rank = 2u * (uint32_t)t;
  if (aden > kMargin) {
    for (int32_t x = x0; x <= x1; x++) {
      const float cx = (float)x + kMargin + p.h2;
      const float e2x = cx - v0x;
      for (int32_t y = y0; y <= y1; y++) {
        const float cy = (float)y + kMargin + p.h2;
        const float e2y = cy - v0y;
        const float d20 = e2x * e0x + e2y * e0y + 0.0f;
        const float d21 = e2x * e1x + e2y * e1y + 0.0f;
        const float v = (d11 * d20 - d01 * d21) / den;
        const float w = (d00 * d21 - d01 * d20) / den;
        const float u = (float)((1.0 - (double)v) - (double)w);
        // (double)v <= 1. + g_fMARGIN  <=>  v <= vlim (largest float not above 1 + (double)0.001f)
        if (!(u >= -kMargin && v >= -kMargin && v <= vlim && (u + v) <= vlim)) continue;
        const float px = v0x * u + v1x * v + v2x * w;
        const float py = v0y * u + v1y * v + v2y * w;
        float pz = v0z * u + v1z * v + v2z * w;
        float nz = n0z * u + n1z * v + n2z * w;
        int32_t s = em.put(px, py, pz, nz, rank);
        if (s >= 0 && s != x * p.grid + y) em.flags |= F_RANK_TIE;
        hits++;
        if (p.do_shell) {
          pz = pz + ((nz >= 0.0f) ? -p.shell_off : p.shell_off);
          nz = -nz;
          em.put(px, py, pz, nz, rank + 1u);
          hits++;
        }
      }
    }
  }
  if (hits == 0) {  // very small triangle: the un-margined centroid (:1190-1206)
    const float tx = (a0x + a1x + a2x) * 0.333333f + kMargin;
    const float ty = (a0y + a1y + a2y) * 0.333333f + kMargin;
    const float tz = (a0z + a1z + a2z) * 0.333333f + kMargin;
    em.put(tx, ty, tz, n0z, rank);
    if (p.do_shell) {
      const float sz = tz + ((n0z >= 0.0f) ? -p.shell_off : p.shell_off);
      em.put(tx, ty, sz, -n0z, rank + 1u);
    }
  }
}

// Bitonic sort of P (power of two) 64-bit values in this wave's LDS buffer, descending.
__device__ void wave_sort_desc(uint64_t *buf, int P, int lane) {
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = lane; i < (P >> 1); i += 64) {
        int lo = 2 * i - (i & (stride - 1));
        int hi = lo + stride;
        uint64_t a = buf[lo], b = buf[hi];
        bool sw = ((lo & size) == 0) ? (a < b) : (a > b);
        if (sw) {
          buf[lo] = b;
          buf[hi] = a;
        }
      }
      wave_sync();
    }
  }
}

__device__ __forceinline__ int pow2_at_least(int n) {
  int P = 64;
  while (P < n) P <<= 1;
  return P;
}

// Pair D (<= 63) distinct keys held one per lane, sorted descending (lanes >= D invalid).
// Same semantics as pair_keys in tomo_slot.h, with the sequential state machines turned into
// "last event before this lane" ballot scans.  The result is valid in every lane.
__device__ void wave_pair(uint32_t key, int D, int lane, const PairParams &pp, SlotResult &r) {
  const bool valid = lane < D;
  const int32_t z = key_z(key), nz = key_nz(key);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const uint32_t k10 = make_key(1, 0);
  r.npre = D;
  r.sabs = wave_sum(valid ? (z < 0 ? -z : z) : 0);
  r.has10 = __ballot(valid && key == k10) != 0ull;
  r.flags = 0;
  r.vo = r.vss = r.va = r.vb = r.vnv = r.vtc = 0;
  if (D <= 1) {
    r.n = D;
    r.last_z = D ? __shfl(z, 0) : 0;
    r.last_type = 0;
    return;
  }
  const bool isA = valid && nz >= 1, isB = valid && nz <= -1;
  const uint64_t mNZ = __ballot(isA || isB), mA = __ballot(isA);
  const uint64_t pv = mNZ & lt;
  const bool prevA = pv ? (((mA >> msb64(pv)) & 1ull) != 0) : false;
  const bool kA = isA && !prevA, kB = isB && prevA;
  const bool bs = kB && nz <= pp.cut_ssb, bn = kB && nz >= pp.cut_nvb;
  const uint64_t mKA = __ballot(kA), mBS = __ballot(bs), mBN = __ballot(bn);
  const uint64_t evS = (mBS | mKA), evN = (mBN | mKA);
  const bool prevS = (evS & lt) ? (((mBS >> msb64(evS & lt)) & 1ull) != 0) : false;
  const bool prevN = (evN & lt) ? (((mBN >> msb64(evN & lt)) & 1ull) != 0) : false;
  const bool ssb = bs && !prevS, ssa = kA && prevS;
  const bool nvb = bn && !prevN, nva = kA && prevN;
  const bool exOpen = evS && (((mBS >> msb64(evS)) & 1ull) != 0);
  const bool imOpen = evN && (((mBN >> msb64(evN)) & 1ull) != 0);
  const bool tc = !pp.explicit_ss && kA && mKA && lane == (int)__ffsll((long long)mKA) - 1;
  int32_t ty = (kA ? T_AL : 0) | (kB ? T_BE : 0) | (ssb ? T_SSB : 0) | (ssa ? T_SSA : 0) |
               (nvb ? T_NVB : 0) | (nva ? T_NVA : 0) | (tc ? T_TC : 0);
  const int32_t al = wave_sum(kA ? z : 0), be = wave_sum(kB ? z : 0);
  const int32_t sSSB = wave_sum(ssb ? z : 0), sSSA = wave_sum(ssa ? z : 0);
  const int32_t sNVB = wave_sum(nvb ? z : 0), sNVA = wave_sum(nva ? z : 0);
  const int32_t sTC = wave_sum(tc ? z : 0);
  int32_t last_z = __shfl(z, D - 1), last_type = __shfl(ty, D - 1);
  int n = D;
  if (exOpen) {
    if (n < kMaxPix) { n++; last_z = 0; last_type = T_SSA; } else r.flags |= F_SENTINEL_DROP;
  }
  if (imOpen) {
    if (n < kMaxPix) { n++; last_z = 0; last_type = T_NVA; } else r.flags |= F_SENTINEL_DROP;
  }
  r.n = n;
  r.last_z = last_z;
  r.last_type = last_type;
  r.vo = al > be ? al - be : 0;
  r.va = al;
  r.vb = be;
  r.vnv = sNVB - sNVA;
  if (pp.explicit_ss) {
    r.vss = sSSB - sSSA;
  } else {
    r.vss = -al + be + sTC - sNVB + sNVA;
    r.vtc = sTC;
  }
}

// One slot with more than kSmall entries, handled by a whole wave.  Restores SetBit's exact
// kept set: the first 63 distinct keys in insertion order, insertion order being (triangle,
// main-before-sub) = ascending rank (STomoVoxelSpaceInfo.cpp:200-220).
__device__ void big_slot(int32_t s, uint32_t c, const Tables &t, int32_t G2, Shared &sh, int wave,
                         int lane, const PairParams &pp, SlotResult &r, int32_t &flags) {
  uint64_t *buf = sh.sortbuf[wave];
  int32_t fl = 0;
  if (c > (uint32_t)kBigCap) {
    fl |= F_UNRESOLVED;
    c = kBigCap;
  }
  const uint32_t off = (c > (uint32_t)kInline) ? t.off[s] : 0u;
  const int P = pow2_at_least((int)c);
  for (int i = lane; i < P; i += 64) {
    uint64_t v = 0;
    if (i < (int)c) {
      uint32_t key, rank;
      if (i < kInline) {
        key = t.keys[(int64_t)i * G2 + s];
        rank = t.ranks[(int64_t)i * G2 + s];
      } else {
        uint32_t j = off + (uint32_t)(i - kInline);
        if (j < (uint32_t)t.pool) {
          key = t.cmp_key[j];
          rank = t.cmp_rank[j];
        } else {
          key = 0u;
          rank = 0xffffffffu;
          fl |= F_UNRESOLVED;
        }
      }
      // descending (key, then ascending rank)
      v = ((uint64_t)key << 32) | (uint64_t)(~rank);
    }
    buf[i] = v;
  }
  wave_sync();
  wave_sort_desc(buf, P, lane);
  // Compact the first entry of every key group (it carries the smallest rank) in place.
  int D = 0;
  uint32_t carry = 0u;  // key of the element just before this chunk
  for (int base = 0; base < (int)c; base += 64) {
    const int i = base + lane;
    const uint64_t v = (i < (int)c) ? buf[i] : 0ull;
    const uint32_t vkey = (uint32_t)(v >> 32);
    uint32_t pk = __shfl_up(vkey, 1);
    if (lane == 0) pk = carry;
    carry = __shfl(vkey, 63);
    const bool first = (i < (int)c) && (i == 0 || vkey != pk);
    const uint64_t m = __ballot(first);
    const int pos = D + __popcll(m & ((lane == 0) ? 0ull : ((~0ull) >> (64 - lane))));
    wave_sync();
    if (first) {
      const uint32_t key = (uint32_t)(v >> 32), rank = ~(uint32_t)v;
      buf[pos] = ((uint64_t)rank << 32) | key;  // (rank, key) for the cap selection below
    }
    wave_sync();
    D += __popcll(m);
  }
  uint32_t mykey = 0;
  if (D > kMaxPix) {
    // More than 63 distinct keys: SetBit keeps the 63 that arrived first.
    fl |= F_CAP_DROP;
    const int P2 = pow2_at_least(D);
    for (int i = D + lane; i < P2; i += 64) buf[i] = ~0ull;
    wave_sync();
    // ascending by rank == descending by complement
    for (int i = lane; i < P2; i += 64) buf[i] = ~buf[i];
    wave_sync();
    wave_sort_desc(buf, P2, lane);
    uint64_t v = (lane < kMaxPix) ? ~buf[lane] : 0ull;
    wave_sync();
    buf[lane] = (lane < kMaxPix) ? (uint64_t)(uint32_t)v : 0ull;  // keys only, pad 0
    wave_sync();
    wave_sort_desc(buf, 64, lane);
    D = kMaxPix;
    mykey = (uint32_t)buf[lane];
  } else {
    mykey = (lane < D) ? (uint32_t)buf[lane] : 0u;
  }
  wave_sync();
  wave_pair(mykey, D, lane, pp, r);
  r.flags |= fl;
  flags |= fl;
}

}  // namespace

__global__ __launch_bounds__(kThreads) void sweep_kernel(SweepParams p) {
  __shared__ Shared sh;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t G = p.grid, G2 = G * G;
  uint8_t *base = p.scratch + (int64_t)blockIdx.x * p.lay.bytes;
  Tables t;
  t.count = (uint32_t *)(base + p.lay.count);
  t.keys = (uint32_t *)(base + p.lay.keys);
  t.ranks = (uint32_t *)(base + p.lay.ranks);
  t.bedinfo = (uint32_t *)(base + p.lay.bedinfo);
  t.big = (int32_t *)(base + p.lay.big);
  t.off = (uint32_t *)(base + p.lay.off);
  t.cursor = (uint32_t *)(base + p.lay.cursor);
  t.pool_slot = (uint32_t *)(base + p.lay.pool_slot);
  t.pool_key = (uint32_t *)(base + p.lay.pool_key);
  t.pool_rank = (uint32_t *)(base + p.lay.pool_rank);
  t.cmp_key = (uint32_t *)(base + p.lay.cmp_key);
  t.cmp_rank = (uint32_t *)(base + p.lay.cmp_rank);
  t.pool = p.lay.pool;
  // (double)v <= 1. + g_fMARGIN, as a float bound
  const float vlim = __uint_as_float(0x3f8020c4u);

  for (int32_t o = blockIdx.x; o < p.n_orient; o += gridDim.x) {
    const float *Rg = p.rot + (int64_t)o * 9;
    const float r0 = Rg[0], r1 = Rg[1], r2 = Rg[2], r3 = Rg[3], r4 = Rg[4], r5 = Rg[5],
                r6 = Rg[6], r7 = Rg[7], r8 = Rg[8];
    // ---- (1) rotated AABB minimum (STomoSh_Base.cpp:77-109) --------------------------------
    float mx = 1e5f, my = 1e5f, mz = 1e5f;
    for (int32_t v = tid; v < p.n_vtx; v += kThreads) {
      const float t0 = p.vx[v] - p.cx, t1 = p.vy[v] - p.cy, t2 = p.vz[v] - p.cz;
      mx = fminf(mx, dot_row(r0, r1, r2, t0, t1, t2));
      my = fminf(my, dot_row(r3, r4, r5, t0, t1, t2));
      mz = fminf(mz, dot_row(r6, r7, r8, t0, t1, t2));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mx = fminf(mx, __shfl_xor(mx, off));
      my = fminf(my, __shfl_xor(my, off));
      mz = fminf(mz, __shfl_xor(mz, off));
    }
    if (lane == 0) {
      sh.red[wave][0] = mx;
      sh.red[wave][1] = my;
      sh.red[wave][2] = mz;
    }
    if (tid == 0) {
      sh.bbox[0] = G; sh.bbox[1] = -1; sh.bbox[2] = G; sh.bbox[3] = -1;
      sh.big_top = 0; sh.pool_top = 0; sh.flags = 0;
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      mx = fminf(mx, sh.red[w][0]);
      my = fminf(my, sh.red[w][1]);
      mz = fminf(mz, sh.red[w][2]);
    }
    const float kx = mx - p.bed_out, ky = my - p.bed_out, kz = mz;

    // ---- (2) rasterise (Pixelize, STomoSh_INT3.cpp:285-364) --------------------------------
    Emit em;
    em.G = G; em.G2 = G2; em.t = t; em.sh = &sh;
    em.bx0 = G; em.bx1 = -1; em.by0 = G; em.by1 = -1; em.flags = 0;
    for (int32_t tr = tid; tr < p.n_tri; tr += kThreads) {
      const int32_t i0 = p.tri[3 * tr + 0], i1 = p.tri[3 * tr + 1], i2 = p.tri[3 * tr + 2];
      float a[3][3], nzv[3];
      const int32_t ids[3] = {i0, i1, i2};
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int32_t i = ids[q];
        const float t0 = p.vx[i] - p.cx, t1 = p.vy[i] - p.cy, t2 = p.vz[i] - p.cz;
        a[q][0] = dot_row(r0, r1, r2, t0, t1, t2) - kx;
        a[q][1] = dot_row(r3, r4, r5, t0, t1, t2) - ky;
        a[q][2] = dot_row(r6, r7, r8, t0, t1, t2) - kz;
        nzv[q] = dot_row(r6, r7, r8, p.nx[i], p.ny[i], p.nz[i]);
      }
      raster_tri(em, p, tr, a[0][0], a[0][1], a[0][2], a[1][0], a[1][1], a[1][2], a[2][0],
                 a[2][1], a[2][2], nzv[0], nzv[1], nzv[2], vlim);
    }
    // bounding box of touched slots and flags
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      em.bx0 = min(em.bx0, __shfl_xor(em.bx0, off));
      em.bx1 = max(em.bx1, __shfl_xor(em.bx1, off));
      em.by0 = min(em.by0, __shfl_xor(em.by0, off));
      em.by1 = max(em.by1, __shfl_xor(em.by1, off));
      em.flags |= __shfl_xor(em.flags, off);
    }
    if (lane == 0) {
      atomicMin(&sh.bbox[0], em.bx0);
      atomicMax(&sh.bbox[1], em.bx1);
      atomicMin(&sh.bbox[2], em.by0);
      atomicMax(&sh.bbox[3], em.by1);
      atomicOr(&sh.flags, em.flags);
    }
    wg_sync_global();
    const int32_t bx0 = sh.bbox[0], bx1 = sh.bbox[1], by0 = sh.bbox[2], by1 = sh.bbox[3];
    const uint32_t nbig = sh.big_top;
    const uint32_t npool = sh.pool_top;
    Acc acc = {};
    if (tid == 0 && npool > (uint32_t)t.pool) acc.flags |= F_UNRESOLVED;

    // ---- overflow compaction: entries beyond kInline, grouped per slot ---------------------
    if (npool > 0) {
      uint32_t running = 0;
      for (uint32_t b0 = 0; b0 < nbig; b0 += kThreads) {
        const uint32_t b = b0 + tid;
        int32_t s = -1;
        uint32_t ex = 0;
        if (b < nbig) {
          s = t.big[b];
          const uint32_t c = t.count[s];
          ex = c > (uint32_t)kInline ? c - (uint32_t)kInline : 0u;
        }
        // work-group exclusive scan of ex
        uint32_t incl = ex;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          uint32_t u = __shfl_up(incl, off);
          if (lane >= off) incl += u;
        }
        if (lane == 63) sh.scan[wave] = incl;
        __syncthreads();
        uint32_t wbase = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
          if (w < wave) wbase += sh.scan[w];
          total += sh.scan[w];
        }
        if (s >= 0 && ex > 0) {
          t.off[s] = running + wbase + incl - ex;
          t.cursor[s] = 0u;
        }
        running += total;
        __syncthreads();
      }
      wg_sync_global();
      const uint32_t np = npool < (uint32_t)t.pool ? npool : (uint32_t)t.pool;
      for (uint32_t j = tid; j < np; j += kThreads) {
        const uint32_t s = t.pool_slot[j];
        const uint32_t q = atomicAdd(&t.cursor[s], 1u);
        const uint32_t d = t.off[s] + q;
        if (d < (uint32_t)t.pool) {
          t.cmp_key[d] = t.pool_key[j];
          t.cmp_rank[d] = t.pool_rank[j];
        }
      }
      wg_sync_global();
    }

    // ---- (3a) pair small slots in registers ------------------------------------------------
    const bool any = bx1 >= bx0;
    const int32_t wy = by1 - by0 + 1;
    const int32_t nbb = any ? (bx1 - bx0 + 1) * wy : 0;
    for (int32_t i = tid; i < nbb; i += kThreads) {
      const int32_t sx = bx0 + i / wy, sy = by0 + i % wy;
      const int32_t s = sx * G + sy;
      const uint32_t c = t.count[s];
      if (c == 0u) {
        t.bedinfo[s] = BI_POSS;
        continue;
      }
      if (c > (uint32_t)kSmall) continue;  // handled by a wave below
      t.count[s] = 0u;
      acc.nslot++;
      uint32_t k[kSmall];
#pragma unroll
      for (int q = 0; q < kSmall; q++) k[q] = (q < (int)c) ? t.keys[(int64_t)q * G2 + s] : 0u;
      SlotResult r;
      pair_keys<kSmall>(k, (int)c, p.pp, r);
      account(acc, r, s, p);
      t.bedinfo[s] = bed_info(r, p.pp);
    }
    // Large slots reset their counters below; every wave must be done reading counters above.
    __syncthreads();
    // ---- (3b) pair large slots, one wave each ----------------------------------------------
    for (uint32_t b = wave; b < nbig; b += kWaves) {
      const int32_t s = t.big[b];
      const uint32_t c = t.count[s];
      SlotResult r;
      int32_t fl = 0;
      big_slot(s, c, t, G2, sh, wave, lane, p.pp, r, fl);
      if (lane == 0) {
        t.count[s] = 0u;
        acc.nslot++;
        account(acc, r, s, p);
        t.bedinfo[s] = bed_info(r, p.pp);
      }
    }
    wg_sync_global();

    // ---- (4) bed (GenerateBed, STomoSh_INT3.cpp:1375-1443) and final per-slot counts -------
    const int32_t br = p.pp.bed_r;
    const bool bed_on = any && p.pp.bed_type != 0 && br >= 1;
    const int32_t m = bed_on ? br + 1 : 0;
    const int32_t rx0 = any ? max(bx0 - m, 0) : 0, rx1 = any ? min(bx1 + m + 1, G) : 0;
    const int32_t ry0 = any ? max(by0 - m, 0) : 0, ry1 = any ? min(by1 + m + 1, G) : 0;
    const int32_t rwy = ry1 - ry0;
    const int32_t nreg = (rx1 - rx0) * (rwy > 0 ? rwy : 0);
    for (int32_t i = tid; i < nreg; i += kThreads) {
      const int32_t X = rx0 + i / rwy, Y = ry0 + i % rwy;
      const int32_t P = X * G + Y;
      const bool inb = X >= bx0 && X <= bx1 && Y >= by0 && Y <= by1;
      const uint32_t bi = inb ? t.bedinfo[P] : BI_POSS;
      const int32_t n = (int32_t)(bi & 127u);
      int32_t add = 0;
      if (bed_on && (bi & BI_POSS)) {
        float md = 1e5f;
        const int32_t xa = max(X - br, 0), xb = min(X + br, G);
        const int32_t ya = max(Y - br, 0), yb = min(Y + br, G);
        for (int32_t sx = xa; sx < xb; sx++) {
          if (sx < bx0 || sx > bx1) continue;
          for (int32_t sy = ya; sy < yb; sy++) {
            if (sy < by0 || sy > by1) continue;
            const uint32_t bs = t.bedinfo[sx * G + sy];
            if (!(bs & BI_SRC)) continue;
            const bool before = (sx < X) || (sx == X && sy < Y);
            if ((bs & BI_CONV) && before) continue;  // already covered by its own bed row
            md = fminf(md, p.bed_dist[(sx - X + br) * (2 * br) + (sy - Y + br)]);
          }
        }
        const bool bed = !(md > p.bed_out) &&
                         (p.pp.bed_type == 1 || p.pp.bed_type == 2 || (p.pp.bed_type == 3 && md > p.bed_in));
        if (bed) {
          if (bi & BI_HAS10) {
            acc.nbed++;
          } else if (n < kMaxPix) {
            acc.nbed++;
            add = 1;
          } else {
            acc.flags |= F_CAP_DROP;
          }
        }
      }
      const int32_t nf = n + add;
      acc.nfinal += nf;
      acc.chkfinal += (uint64_t)(P + 1) * (uint64_t)nf;
      if (p.dbg_final) p.dbg_final[P] = nf;
    }

    // ---- (5) reduce and write the record ---------------------------------------------------
    int64_t v[19] = {acc.Va, acc.Vb, acc.Vo, acc.Vss, acc.Vnv, acc.Vtc, acc.nbed, acc.nent,
                     acc.nslot, acc.nfinal, (int64_t)acc.chkpre, (int64_t)acc.chkfinal, acc.aVo,
                     acc.aVss, acc.aVa, acc.aVb, acc.aVnv, acc.aVtc, (int64_t)acc.flags};
#pragma unroll
    for (int q = 0; q < 18; q++) v[q] = wave_sum(v[q]);
    int32_t f = (int32_t)v[18];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < 18; q++) sh.acc[wave][q] = v[q];
      sh.acc[wave][18] = f;
    }
    __syncthreads();
    if (tid == 0) {
      int64_t s[19] = {};
      for (int w = 0; w < kWaves; w++) {
        for (int q = 0; q < 18; q++) s[q] += sh.acc[w][q];
        s[18] |= sh.acc[w][18];
      }
      DevRecord &R = p.rec[o];
      R.Va = s[0]; R.Vb = s[1]; R.Vo = s[2]; R.Vss = s[3]; R.Vnv = s[4]; R.Vtc = s[5];
      R.Vbed = s[6]; R.n_entries = s[7]; R.n_slots = s[8]; R.n_final = s[9];
      R.chk_pre = (uint64_t)s[10]; R.chk_final = (uint64_t)s[11];
      R.aVo = s[12]; R.aVss = s[13]; R.aVa = s[14]; R.aVb = s[15]; R.aVnv = s[16]; R.aVtc = s[17];
      R.flags = (int32_t)s[18] | sh.flags;
      R.n_bed = (int32_t)s[6];
      R.kx = kx; R.ky = ky; R.kz = kz;
      R.pad = 0;
    }
    wg_sync_global();
  }
}

hipError_t launch_sweep(const SweepParams &p, int n_groups, hipStream_t stream) {
  hipLaunchKernelGGL(sweep_kernel, dim3(n_groups), dim3(kThreads), 0, stream, p);
  return hipGetLastError();
}

}  // namespace tomo
