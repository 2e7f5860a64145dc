// tomo_kernels.hip -- the per-orientation support-tomography sweep for gfx950.
//
// Orientations are processed in batches; each orientation of a batch owns a slot table in HBM.
// Five kernels per batch, each with its own register budget and many short work-groups:
//   aabb_kernel     rotate every vertex, min-reduce the rotated AABB     (STomoSh_Base.cpp:45-109)
//   raster_kernel   rasterise triangles against pixel centres and append  (STomoSh_INT3.cpp:1138-1209,
//                   (z, nZ) keys to the orientation's slot table            STomoVoxelSpaceInfo.cpp:175-230)
//   compact_kernel  group entries beyond the inline layers per slot
//   pair_kernel     pair each touched slot: <= 8 entries in registers,    (STomoSh_INT3.cpp:1041-1062,
//                   9..63 with a wave-wide shuffle bitonic sort, more with  1211-1255)
//                   an LDS sort and rank selection; totals by atomics
//   bed_kernel      raft/brim/skirt bed with the reference's raster-order (STomoSh_INT3.cpp:1331-1443)
//                   semantics, from LDS-staged neighbour flags
// There is no MFMA: this is rasterisation, sorting and integer reduction.  All float arithmetic
// follows the reference's operation order without contraction (built with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "tomo_device.h"
#include "tomo_raster.h"

// Slot tables and orientation headers of tomo_run_static, the entry point that allocates
// nothing: kStaticScratchBytes on every device the library runs on, held until the process
// ends.  Host code never passes this name to a kernel (on the host it is a stand-in, not a
// device address); it uses the address static_scratch_ptr looks up per device.
static __device__ __align__(256) uint8_t scratch[tomo::kStaticScratchBytes];

namespace tomo {

namespace {

constexpr float kNormFactor = 1000.0f;      // g_fNORMALFACTOR (Tomo_types.h:39)

constexpr int kBedLdsSlots = 13 * 1024;  // bed band + halo slot info ((rows + 2r) x (width + 2r))
constexpr int kBedWords = 1024;          // halo source bit words ((rows + 2r) x ceil((width + 2r) / 32))
constexpr int kBedTabR = 3;              // largest bed radius r served by the bit-window search

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

// dot_row without its trailing "+ 0": the two differ only in the sign of a zero result.  The
// rasteriser offsets every rotated coordinate (by the placement, then by the margin) before any
// other use, and reads the rotated normal only through the sign test n.z >= 0 and the 16-bit
// conversion, both of which treat -0 as +0; its outputs are the same either way.
__device__ __forceinline__ float dot_row_nz(float r0, float r1, float r2, float a0, float a1, float a2) {
  float s = r0 * a0;
  s = s + r1 * a1;
  s = s + r2 * a2;
  return s;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ int msb64(uint64_t m) { return 63 - __clzll((long long)m); }

// Mask of the active lanes for which c holds, taken straight from the compare that forms c
// (HIP's __ballot takes an int, which costs a select and a second compare per call).
__device__ __forceinline__ uint64_t ballot(bool c) { return __builtin_amdgcn_ballot_w64(c); }

// Compare predicates of __builtin_amdgcn_fcmpf / __builtin_amdgcn_uicmp (LLVM's CmpInst codes):
// each call is one vector compare whose result is the wave's lane mask.
constexpr int kFcmpUge = 11, kFcmpUle = 13, kIcmpSgt = 38;

// scan_candidate (tomo_raster.h) as a lane mask: a >= lo, b >= lo and a + b <= his, each also
// true when unordered (a NaN leaves the pixel to the exact test), with a and b formed as there.
__device__ __forceinline__ uint64_t scan_mask(const TriScan &t, float rax, float rbx, float e2y) {
  const float a = __builtin_fmaf(e2y, t.ay, rax);
  const float b = __builtin_fmaf(e2y, t.by, rbx);
  return __builtin_amdgcn_fcmpf(a, t.lo, kFcmpUge) & __builtin_amdgcn_fcmpf(b, t.lo, kFcmpUge) &
         __builtin_amdgcn_fcmpf(a + b, t.his, kFcmpUle);
}

// Index of the calling wave in its work-group, as a wave-uniform (scalar) value: state derived
// from it (turn counters, LDS row bases) then stays in scalar registers.
__device__ __forceinline__ int wave_index() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Number of lanes below the caller's whose bit is set in m.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

struct Tables {
  uint2 *ent;  // [inl][G2] {key, rank}, layer-major
  uint16_t *bedinfo;  // [G2] bed bookkeeping (bed_info bits, 11 used)
  uint32_t *count, *off, *cursor, *pool_slot, *pool_key, *pool_rank,
      *cmp_key, *cmp_rank;
  int32_t *big;
  int32_t *sv;  // [6][G2] per-slot Vo, Vss, Va, Vb, Vnv, Vtc (kept only when totals may round)
  uint32_t *spill_slot, *spill_key, *spill_rank;
  int32_t *huge;  // slots with more than kBigCap entries
  int32_t pool;
  int32_t inl;    // inline entries per slot (>= kSmall)
  int32_t spill;  // spill list capacity
  int32_t nhuge;  // huge list capacity
};

// Per-lane totals.  T = int32_t where a lane accounts for a bounded number of slots: a slot's
// values stay below 2^22 in magnitude (at most 64 rows of |z| <= 2^15), so up to 2^9 slots per
// lane fit; the work-group reduction widens them to 64 bits.
template <typename T>
struct AccT {
  T Va, Vb, Vo, Vss, Vnv, Vtc, nbed, nent, nslot, nfinal;
  uint64_t chkpre, chkfinal;
  T aVo, aVss, aVa, aVb, aVnv, aVtc;
  int32_t flags;
};
using Acc = AccT<int64_t>;

__device__ __forceinline__ int64_t iabs64(int32_t v) { return v < 0 ? -(int64_t)v : (int64_t)v; }

// The six per-slot totals of one slot, kept for the ordered float re-summation.
__device__ __forceinline__ void put_sv(int32_t *sv, int32_t G2, int32_t s, const SlotResult &r) {
  sv[s] = r.vo;
  sv[G2 + s] = r.vss;
  sv[2 * G2 + s] = r.va;
  sv[3 * G2 + s] = r.vb;
  sv[4 * G2 + s] = r.vnv;
  sv[5 * G2 + s] = r.vtc;
}

// Every float total of an orientation is exact when the sum of |per-slot value| stays within
// 2^24.  zabs (sum of |z| over all emitted entries, a superset of the distinct rows) bounds
// those sums; the implicit Vss (-alpha + beta + TC - NVB + NVA) can count a row twice.
__device__ __forceinline__ bool totals_may_round(const OrientHdr &h, const PairParams &pp) {
  return (pp.explicit_ss ? h.zabs : 2ull * h.zabs) > (1ull << 24);
}

// The sums of |per-slot value| only matter when they can pass 2^24 (the float totals may round,
// and the record says so); otherwise they are left at zero (track false).
template <bool DBG = true, typename T>
__device__ __forceinline__ void account(AccT<T> &a, const SlotResult &r, int32_t s, const SweepParams &p, bool track) {
  a.Va += r.va; a.Vb += r.vb; a.Vo += r.vo; a.Vss += r.vss; a.Vnv += r.vnv; a.Vtc += r.vtc;
  if (track) {
    a.aVo += (T)iabs64(r.vo); a.aVss += (T)iabs64(r.vss); a.aVa += (T)iabs64(r.va); a.aVb += (T)iabs64(r.vb);
    a.aVnv += (T)iabs64(r.vnv); a.aVtc += (T)iabs64(r.vtc);
  }
  a.nent += r.npre;
  a.chkpre += (uint64_t)(s + 1) * (uint64_t)r.npre;
  a.flags |= r.flags;
  if (r.sabs > 32767) a.flags |= F_INT16_WRAP;
  if (DBG && p.dbg_pre) {
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

// Entry of slot s at position k: inline layer or overflow pool.
__device__ __forceinline__ void place_entry(const Tables &t, int32_t G2, uint32_t *pool_top, int32_t s,
                                            uint32_t k, uint32_t key, uint32_t rank) {
  if (k < (uint32_t)t.inl) {
    t.ent[k * (uint32_t)G2 + (uint32_t)s] = make_uint2(key, rank);  // < 32 * 4096^2 + 4096^2
  } else {
    const uint32_t j = atomicAdd(pool_top, 1u);
    if (j < (uint32_t)t.pool) {
      t.pool_slot[j] = (uint32_t)s;
      t.pool_key[j] = key;
      t.pool_rank[j] = rank;
    }
  }
}

// SetBit's slot append (STomoVoxelSpaceInfo.cpp:175-230) for one rasteriser tile.  The tile's
// slot counters live in LDS, 16 bits each, two to a word; an entry whose slot lies outside the
// tile (the interpolated point can leave the pixel it was tested at, and a centroid pixel can
// leave the owning tile) goes to the orientation's spill list, merged after the rasteriser.
struct Emit {
  int32_t G, G2;
  Tables t;
  uint32_t *pool_top, *spill_top;  // this orientation's counters (global)
  uint32_t *lcnt;                  // the tile's LDS counters
  int32_t xa, xb, ya, yb, yshift;  // the tile's slots: [xa, xb) x [ya, yb)
  int32_t bx0, bx1, by0, by1;
  int32_t flags;
  uint64_t zabs;  // sum of |z| of this lane's entries

  // coord2vID (STomoVoxelSpaceInfo.cpp:97-105) with the footprint check: true when (px, py) lies
  // on the grid.  The touched box grows only for such points (the selects keep it otherwise).
  __device__ __forceinline__ bool slot_xy(float px, float py, int32_t &sx, int32_t &sy) {
    sx = (int32_t)px;
    sy = (int32_t)py;
    const bool ok = __builtin_fabsf(px) < 2147483648.0f && __builtin_fabsf(py) < 2147483648.0f &&
                    (uint32_t)sx < (uint32_t)G && (uint32_t)sy < (uint32_t)G;
    bx0 = min(bx0, ok ? sx : G); bx1 = max(bx1, ok ? sx : -1);
    by0 = min(by0, ok ? sy : G); by1 = max(by1, ok ? sy : -1);
    return ok;
  }

  __device__ __forceinline__ void spill(int32_t s, uint32_t key, uint32_t rank) {
    const uint32_t j = atomicAdd(spill_top, 1u);
    if (j < (uint32_t)t.spill) {
      t.spill_slot[j] = (uint32_t)s;
      t.spill_key[j] = key;
      t.spill_rank[j] = rank;
    }
  }

  // Reserve n positions in an in-tile slot's LDS counter; returns the first.
  __device__ __forceinline__ uint32_t reserve(int32_t sx, int32_t sy, uint32_t n) {
    const uint32_t li = ((uint32_t)(sx - xa) << yshift) + (uint32_t)(sy - ya);
    const uint32_t sh = (li & 1u) << 4;
    const uint32_t old = atomicAdd(&lcnt[li >> 1], n << sh);
    const uint32_t k = (old >> sh) & 0xffffu;
    if (k > 0xffffu - n) flags |= F_UNRESOLVED;  // the 16-bit counter would carry
    return k;
  }

  __device__ __forceinline__ bool in_tile(int32_t sx, int32_t sy) const {
    return sx >= xa && sx < xb && sy >= ya && sy < yb;
  }

  // An entry for the point (px, py, pz) with normal z nz.  Points off the grid only raise the
  // footprint flag; the arithmetic runs for them too (its flags and sums are dropped), so that
  // no lane takes a path of its own before the append.
  __device__ __forceinline__ void put(float px, float py, float pz, float nz, uint32_t rank) {
    int32_t sx, sy, f = 0;
    const bool ok = slot_xy(px, py, sx, sy);
    const int32_t z16 = to_i16(pz, f);
    const int32_t nz16 = to_i16(nz * kNormFactor, f);
    flags |= ok ? f : F_FOOTPRINT_OOB;
    zabs += ok ? (uint32_t)(z16 < 0 ? -z16 : z16) : 0u;
    const uint32_t key = make_key(z16, nz16);
    const int32_t s = sx * G + sy;
    if (ok) {
      if (in_tile(sx, sy))
        place_entry(t, G2, pool_top, s, reserve(sx, sy, 1u), key, rank);
      else
        spill(s, key, rank);
    }
  }

  // A shell pixel and its sub-pixel (same slot, ranks r and r + 1) with one counter update.
  __device__ __forceinline__ void put2(float px, float py, float pz, float nz, float pz2, float nz2,
                                       uint32_t rank) {
    int32_t sx, sy, f = 0;
    const bool ok = slot_xy(px, py, sx, sy);
    const int32_t z1 = to_i16(pz, f), n1 = to_i16(nz * kNormFactor, f);
    const int32_t z2 = to_i16(pz2, f), n2 = to_i16(nz2 * kNormFactor, f);
    flags |= ok ? f : F_FOOTPRINT_OOB;
    zabs += ok ? (uint32_t)(z1 < 0 ? -z1 : z1) + (uint32_t)(z2 < 0 ? -z2 : z2) : 0u;
    const int32_t s = sx * G + sy;
    const uint32_t k1 = make_key(z1, n1), k2 = make_key(z2, n2);
    if (ok) {
      if (in_tile(sx, sy)) {
        const uint32_t k = reserve(sx, sy, 2u);
        place_entry(t, G2, pool_top, s, k, k1, rank);
        place_entry(t, G2, pool_top, s, k + 1u, k2, rank + 1u);
      } else {
        spill(s, k1, rank);
        spill(s, k2, rank + 1u);
      }
    }
  }
};


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

// Sum over the W-lane segment holding this lane (W = 32 or 64).
template <int W>
__device__ __forceinline__ int32_t seg_sum(int32_t v) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Pair the distinct keys of the lanes set in V (at most 63), in descending key order from low
// to high lane.  Same semantics as pair_keys in tomo_slot.h, with the sequential state machines
// turned into "last event before this lane" ballot scans.  With W = 32 the wave pairs two
// slots, one per 32-lane half; the result is valid in every lane of its segment.
template <int W = 64>
__device__ void wave_pair(uint32_t key, uint64_t V, int lane, const PairParams &pp, SlotResult &r,
                          uint64_t *rows) {
  const int base = lane & ~(W - 1);
  const uint64_t seg = (W == 64) ? ~0ull : (0xffffffffull << base);
  V &= seg;
  const bool valid = ((V >> lane) & 1ull) != 0;
  const int D = __popcll(V);
  const int32_t z = key_z(key), nz = key_nz(key);
  const uint64_t lt = ((lane == 0) ? 0ull : ((~0ull) >> (64 - lane))) & seg;
  const uint32_t k10 = make_key(1, 0);
  r.npre = D;
  r.sabs = seg_sum<W>(valid ? (z < 0 ? -z : z) : 0);
  r.has10 = (ballot(valid && key == k10) & seg) != 0ull;
  r.flags = 0;
  r.vo = r.vss = r.va = r.vb = r.vnv = r.vtc = 0;
  if (ballot(D > 1) == 0ull) {  // every segment has at most one row
    const int32_t zl = __shfl(z, V ? msb64(V) : base);
    r.n = D;
    r.last_z = D ? zl : 0;
    r.last_type = 0;
    if (rows && valid) rows[0] = make_row(0, key);
    return;
  }
  const bool isA = valid && nz >= 1, isB = valid && nz <= -1;
  const uint64_t mNZ = ballot(isA || isB), mA = ballot(isA);
  const uint64_t pv = mNZ & lt;
  const bool prevA = pv ? (((mA >> msb64(pv)) & 1ull) != 0) : false;
  const bool kA = isA && !prevA, kB = isB && prevA;
  const bool bs = kB && nz <= pp.cut_ssb, bn = kB && nz >= pp.cut_nvb;
  const uint64_t mKA = ballot(kA) & seg, mBS = ballot(bs) & seg, mBN = ballot(bn) & seg;
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
  const int32_t al = seg_sum<W>(kA ? z : 0), be = seg_sum<W>(kB ? z : 0);
  const int32_t sSSB = seg_sum<W>(ssb ? z : 0), sSSA = seg_sum<W>(ssa ? z : 0);
  const int32_t sNVB = seg_sum<W>(nvb ? z : 0), sNVA = seg_sum<W>(nva ? z : 0);
  const int32_t sTC = seg_sum<W>(tc ? z : 0);
  const int lastl = V ? msb64(V) : base;
  int32_t last_z = __shfl(z, lastl), last_type = __shfl(ty, lastl);
  if (D <= 1) {  // this segment's slot has at most one row (Pairing skips it)
    if (rows && valid) rows[0] = make_row(0, key);
    r.n = D;
    r.last_z = D ? last_z : 0;
    r.last_type = 0;
    return;
  }
  if (rows && valid) rows[__popcll(V & lt)] = make_row(ty, key);
  int n = D;
  if (exOpen) {
    if (n < kMaxPix) {
      if (rows && lane == base) rows[n] = make_row(T_SSA, make_key(0, 0));
      n++; last_z = 0; last_type = T_SSA;
    } else r.flags |= F_SENTINEL_DROP;
  }
  if (imOpen) {
    if (n < kMaxPix) {
      if (rows && lane == base) rows[n] = make_row(T_NVA, make_key(0, 0));
      n++; last_z = 0; last_type = T_NVA;
    } else r.flags |= F_SENTINEL_DROP;
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

constexpr uint32_t kNoSeq = 0xffffffffu;

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)v, o), hi = __shfl_xor((uint32_t)(v >> 32), o);
    const uint64_t w = ((uint64_t)hi << 32) | lo;
    v = w < v ? w : v;
  }
  return v;
}

// Entry i of slot s (inline layer or its compacted pool run starting at off).
__device__ __forceinline__ void slot_entry(const Tables &t, int32_t G2, int32_t s, uint32_t off, uint32_t i,
                                           uint32_t &key, uint32_t &rank, int32_t &fl) {
  if (i < (uint32_t)t.inl) {
    const uint2 e = t.ent[(int64_t)i * G2 + s];
    key = e.x;
    rank = e.y;
  } else {
    const uint32_t j = off + (i - (uint32_t)t.inl);
    if (j < (uint32_t)t.pool) {
      key = t.cmp_key[j];
      rank = t.cmp_rank[j];
    } else {  // the pool overflowed: the host sweeps this orientation again with a larger one
      key = 0u;
      rank = 0xffffffffu;
      fl |= F_RETRY;
    }
  }
}

// Insertion order of one triangle's rows in one slot.  triVoxel walks a triangle's pixel box x
// outer, y inner and appends each accepted pixel's row and then its shell sub-pixel row
// (STomoSh_INT3.cpp:1163-1186); SetBit keeps a slot's first 63 distinct keys in that order
// (STomoVoxelSpaceInfo.cpp:200-220).  Entries carry the rank 2 * triangle + sub, which orders
// the rows of different triangles.  Two pixels of one triangle can land in one slot (the
// interpolated point of a sliver can leave the pixel it was tested at); their order is found
// here by walking the triangle's box again with the rasteriser's arithmetic, the wave testing
// 64 consecutive pixels per step.  For this lane's key gk (if gv) returns the first sequence
// number 2 * (pixel index in the box, x-major) + sub at which the caller's triangle t appends
// gk to slot s of orientation b, or kNoSeq if it never does (its rows are then its zero-hit
// centroid rows, which the rank already orders main before sub).
__device__ uint32_t tri_key_order(const SweepParams &p, int32_t b, uint32_t t, int32_t s, uint32_t gk, bool gv,
                                  int lane) {
  const OrientHdr &h = p.hdr[b];
  const float kx = h.kx, ky = h.ky, kz = h.kz;
  const float *Rg = p.rot + (int64_t)b * 9;
  const float R0 = Rg[0], R1 = Rg[1], R2 = Rg[2], R3 = Rg[3], R4 = Rg[4], R5 = Rg[5], R6 = Rg[6],
              R7 = Rg[7], R8 = Rg[8];
  const int4 ix = p.tidx[p.tpos[t]];
  const int32_t ids[3] = {ix.x, ix.y, ix.z};
  float m[3][3], nzv[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {  // as raster_kernel: rotate, place, then the margin
    const float4 va = p.vaos[2 * ids[q]], vb = p.vaos[2 * ids[q] + 1];
    m[q][0] = (dot_row_nz(R0, R1, R2, va.x, va.y, va.z) - kx) + kEps;
    m[q][1] = (dot_row_nz(R3, R4, R5, va.x, va.y, va.z) - ky) + kEps;
    m[q][2] = (dot_row_nz(R6, R7, R8, va.x, va.y, va.z) - kz) + kEps;
    nzv[q] = dot_row_nz(R6, R7, R8, va.w, vb.x, vb.y);
  }
  const float v0x = m[0][0], v0y = m[0][1], v1x = m[1][0], v1y = m[1][1], v2x = m[2][0], v2y = m[2][1];
  float mnx = (v0x <= v1x) ? v0x : v1x; mnx = (mnx <= v2x) ? mnx : v2x;
  float mny = (v0y <= v1y) ? v0y : v1y; mny = (mny <= v2y) ? mny : v2y;
  float mxx = (v0x >= v1x) ? v0x : v1x; mxx = (mxx >= v2x) ? mxx : v2x;
  float mxy = (v0y >= v1y) ? v0y : v1y; mxy = (mxy >= v2y) ? mxy : v2y;
  const int32_t G = p.grid, lo = -64, hi = G + 64;
  const int32_t x0 = max((int32_t)mnx, lo), y0 = max((int32_t)mny, lo);
  const int32_t x1 = min((int32_t)mxx, hi), y1 = min((int32_t)mxy, hi);
  const TriSetup ts = tri_setup(v0x, v0y, v1x, v1y, v2x, v2y);
  uint32_t best = kNoSeq;
  if (tri_degenerate(ts) || x1 < x0 || y1 < y0) return best;
  const int32_t hgt = y1 - y0 + 1;
  const int64_t npx = (int64_t)(x1 - x0 + 1) * hgt;
  for (int64_t i0 = 0; i0 < npx; i0 += 64) {
    const int64_t i = i0 + lane;
    bool hit = false;
    uint32_t k1 = 0u, k2 = 0u;
    if (i < npx) {
      const int32_t x = x0 + (int32_t)(i / hgt), y = y0 + (int32_t)(i % hgt);
      float nv, nw, u, v, w;
      bary_num(ts, ((float)x + kEps) + p.h2, ((float)y + kEps) + p.h2, nv, nw);
      if (bary_exact(ts.den, nv, nw, u, v, w)) {
        const float px = m[0][0] * u + m[1][0] * v + m[2][0] * w;
        const float py = m[0][1] * u + m[1][1] * v + m[2][1] * w;
        const float pz = m[0][2] * u + m[1][2] * v + m[2][2] * w;
        const float nz = nzv[0] * u + nzv[1] * v + nzv[2] * w;
        const int32_t sx = (int32_t)px, sy = (int32_t)py;
        const bool in = px > -2147483648.0f && px < 2147483648.0f && py > -2147483648.0f &&
                        py < 2147483648.0f && sx >= 0 && sy >= 0 && sx < G && sy < G;
        if (in && sx * G + sy == s) {
          int32_t fl = 0;
          hit = true;
          k1 = make_key(to_i16(pz, fl), to_i16(nz * kNormFactor, fl));
          if (p.do_shell) {
            const float pz2 = pz + ((nz >= 0.0f) ? -p.shell_off : p.shell_off);
            const float nz2 = -nz;
            k2 = make_key(to_i16(pz2, fl), to_i16(nz2 * kNormFactor, fl));
          }
        }
      }
    }
    uint64_t mh = ballot(hit);
    while (mh) {
      const int j = __ffsll((unsigned long long)mh) - 1;
      mh &= mh - 1ull;
      const uint32_t ka = __shfl(k1, j), kb = __shfl(k2, j);
      const uint32_t seq = 2u * (uint32_t)(i0 + j);
      if (gv && gk == ka) best = min(best, seq);
      if (p.do_shell && gv && gk == kb) best = min(best, seq + 1u);
    }
  }
  return best;
}

// Of the distinct keys buf[a..e) (low words), all first appended to slot s by the caller's
// triangle T and listed in ascending rank order, the `need` that T appends first: lane
// lane0 + k receives the k-th of them in mykey.  Orders by (tri_key_order, list position), so
// keys T appends only as centroid rows keep their rank order.  Overwrites the high words.
__device__ void pick_first_of_triangle(const SweepParams &p, int32_t b, int32_t s, uint32_t T, uint64_t *buf,
                                       int a, int e, int need, int lane0, int lane, uint32_t &mykey) {
  for (int g0 = a; g0 < e; g0 += 64) {
    const int gi = g0 + lane;
    const bool gv = gi < e;
    const uint32_t gk = gv ? (uint32_t)buf[gi] : 0u;
    const uint32_t seq = tri_key_order(p, b, T, s, gk, gv, lane);
    wave_sync();
    if (gv) buf[gi] = ((uint64_t)seq << 32) | gk;
  }
  wave_sync();
  uint64_t last = 0ull;
  for (int it = 0; it < need; it++) {
    uint64_t best = ~0ull;
    for (int g0 = a; g0 < e; g0 += 64) {
      const int gi = g0 + lane;
      if (gi < e) {
        const uint64_t c = (buf[gi] & 0xffffffff00000000ull) | (uint32_t)gi;
        if ((it == 0 || c > last) && c < best) best = c;
      }
    }
    best = wave_min_u64(best);
    if (best == ~0ull) break;  // callers ask for fewer keys than the list holds
    last = best;
    const uint32_t wk = (uint32_t)buf[(uint32_t)best];
    if (lane == lane0 + it) mykey = wk;
  }
  wave_sync();
}

// The 63 keys SetBit keeps of a slot with D > 63 distinct keys, from buf[0..D) holding each
// key's smallest rank as (rank << 32) | key in ascending order: the first 63, except that when
// the 63rd and 64th belong to one triangle its keys are ordered by pick_first_of_triangle.
// Returns the kept key of lanes 0..62 (0 in lane 63); may overwrite buf.
__device__ uint32_t cap_select(const SweepParams &p, int32_t b, int32_t s, uint64_t *buf, int D, int lane) {
  uint32_t mykey = lane < kMaxPix ? (uint32_t)buf[lane] : 0u;
  const uint32_t t62 = (uint32_t)(buf[kMaxPix - 1] >> 33), t63 = (uint32_t)(buf[kMaxPix] >> 33);
  if (t62 != t63) return mykey;
  // the triangle's keys occupy positions [a, e), a <= 62 < 63 < e
  const int a = __ffsll((unsigned long long)ballot((uint32_t)(buf[lane] >> 33) == t62)) - 1;
  int e = 64;
  for (;;) {
    const int i = e + lane;
    const uint64_t outside = ballot(!(i < D && (uint32_t)(buf[i < D ? i : 0] >> 33) == t62));
    if (outside) {
      e += __ffsll((unsigned long long)outside) - 1;
      break;
    }
    e += 64;
  }
  wave_sync();
  pick_first_of_triangle(p, b, s, t62, buf, a, e, kMaxPix - a, a, lane, mykey);
  return mykey;
}

// One slot with kSmall < c <= kBigCap entries, handled by a whole wave.  Restores SetBit's kept
// set: the first 63 distinct keys in insertion order (triangle, then pixel, then main before
// sub; STomoVoxelSpaceInfo.cpp:200-220).  Returns false, leaving r unset, when the 63rd and
// 64th keys come from one triangle: their order needs cap_select, and the slot goes to
// pair_huge_kernel (this kernel's register budget stays that of its common path).
__device__ bool big_slot(const SweepParams &p, int32_t b, int32_t s, uint32_t c, const Tables &t, int32_t G2,
                         uint64_t *buf, int lane, SlotResult &r, uint64_t *rows) {
  int32_t fl = 0;
  const uint32_t off = (c > (uint32_t)t.inl) ? t.off[s] : 0u;
  const int P = pow2_at_least((int)c);
  for (int i = lane; i < P; i += 64) {
    uint64_t v = 0;
    if (i < (int)c) {
      uint32_t key, rank;
      slot_entry(t, G2, s, off, (uint32_t)i, key, rank, fl);
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
    const uint64_t m = ballot(first);
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
  bool tie = false;
  if (D > kMaxPix) {
    // More than 63 distinct keys: SetBit keeps the 63 that arrived first.
    fl |= F_CAP_DROP;
    const int P2 = pow2_at_least(D);
    for (int i = D + lane; i < P2; i += 64) buf[i] = ~0ull;
    wave_sync();
    // ascending by (rank, key) == descending by complement
    for (int i = lane; i < P2; i += 64) buf[i] = ~buf[i];
    wave_sync();
    wave_sort_desc(buf, P2, lane);
    // the 63rd and 64th by rank are ~buf[62] and ~buf[63]
    tie = (uint32_t)(~buf[kMaxPix - 1] >> 33) == (uint32_t)(~buf[kMaxPix] >> 33) && !(fl & F_RETRY);
    const uint64_t v = (lane < kMaxPix) ? ~buf[lane] : 0ull;
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
  wave_pair(mykey, (D >= 64) ? ~0ull : ((1ull << D) - 1ull), lane, p.pp, r, rows);
  r.flags |= fl;
  return !tie;
}

// A slot with kSmall < c < W entries in a W-lane segment (W = 64: the whole wave; W = 32: two
// slots per wave): one key per lane, register bitonic sort over the segment with lane
// shuffles, duplicates masked out.  No ranks are needed: at most 63 distinct keys.
// This lane's key of such a slot (0 past its count).
template <int W = 64>
__device__ __forceinline__ uint32_t mid_key(int32_t s, uint32_t c, const Tables &t, int32_t G2, int lane) {
  const int i = lane & (W - 1);
  uint32_t key = 0u;
  if (i < (int)c) {
    const uint32_t off = (c > (uint32_t)t.inl) ? t.off[s] : 0u;
    key = (i < t.inl) ? t.ent[(int64_t)i * G2 + s].x : t.cmp_key[off + (uint32_t)(i - t.inl)];
  }
  return key;
}

template <int W = 64>
__device__ void mid_sort(uint32_t key, uint32_t c, int lane, const PairParams &pp, SlotResult &r, uint64_t *rows) {
  const int i = lane & (W - 1);
#pragma unroll
  for (int size = 2; size <= W; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      const uint32_t other = __shfl_xor(key, stride);
      // runs of length `size` alternate direction; each segment's final merge is descending
      const bool desc = size == W || (i & size) == 0;
      const bool keep_max = desc == ((i & stride) == 0);
      key = keep_max ? max(key, other) : min(key, other);
    }
  }
  const uint32_t prev = __shfl_up(key, 1);
  const bool first = i < (int)c && (i == 0 || key != prev);
  wave_pair<W>(key, ballot(first), lane, pp, r, rows);
}

template <int W = 64>
__device__ __forceinline__ void mid_slot(int32_t s, uint32_t c, const Tables &t, int32_t G2, int lane,
                                         const PairParams &pp, SlotResult &r, uint64_t *rows) {
  mid_sort<W>(mid_key<W>(s, c, t, G2, lane), c, lane, pp, r, rows);
}

// Work-groups are dispatched to the 8 XCDs round-robin by block index.  Map block i to
// orientation b = (i mod 8) + 8 * ((i / 8) / per) and part (i / 8) mod per, so every part of an
// orientation runs on the same XCD (its slot table stays in that XCD's L2) and each XCD walks
// its orientations one after another.  The grid is padded to a multiple of 8 orientations.
__device__ __forceinline__ bool xcd_map(int32_t blk, int32_t per, int32_t nb, int32_t &b, int32_t &part) {
  const int32_t x = blk & 7, j = blk >> 3;
  b = x + 8 * (j / per);
  part = j % per;
  return b < nb;
}

__device__ __forceinline__ Tables tables_for(const SweepParams &p, int32_t b) {
  uint8_t *base = p.scratch + (int64_t)b * p.lay.bytes;
  Tables t;
  t.count = (uint32_t *)(base + p.lay.count);
  t.ent = (uint2 *)(base + p.lay.ent);
  t.bedinfo = (uint16_t *)(base + p.lay.bedinfo);
  t.big = (int32_t *)(base + p.lay.big);
  t.off = (uint32_t *)(base + p.lay.off);
  t.cursor = (uint32_t *)(base + p.lay.cursor);
  t.pool_slot = (uint32_t *)(base + p.lay.pool_slot);
  t.pool_key = (uint32_t *)(base + p.lay.pool_key);
  t.pool_rank = (uint32_t *)(base + p.lay.pool_rank);
  t.cmp_key = (uint32_t *)(base + p.lay.cmp_key);
  t.cmp_rank = (uint32_t *)(base + p.lay.cmp_rank);
  t.sv = (int32_t *)(base + p.lay.sv);
  t.spill_slot = (uint32_t *)(base + p.lay.spill_slot);
  t.spill_key = (uint32_t *)(base + p.lay.spill_key);
  t.spill_rank = (uint32_t *)(base + p.lay.spill_rank);
  t.huge = (int32_t *)(base + p.lay.huge);
  t.pool = p.lay.pool;
  t.nhuge = p.lay.nhuge;
  t.inl = p.lay.inl;
  t.spill = p.lay.spill;
  return t;
}

// Sum N per-thread values over the work-group and add them to the N consecutive 64-bit
// counters at dst (zero sums are skipped).
template <int N, int W = kWaves>
__device__ __forceinline__ void group_add(int64_t (&v)[N], int64_t (*red)[N], int lane, int wave,
                                          int tid, unsigned long long *dst) {
  // one total at a time, each stored as soon as it is reduced: few registers stay live
#pragma unroll
  for (int q = 0; q < N; q++) {
    const int64_t w = wave_sum(v[q]);
    if (lane == 0) red[wave][q] = w;
  }
  __syncthreads();
  if (tid < N) {
    int64_t s = 0;
#pragma unroll
    for (int w = 0; w < W; w++) s += red[w][tid];
    if (s != 0) atomicAdd(dst + tid, (unsigned long long)s);
  }
}

// DevRecord's 64-bit totals, in field order from Va.
static_assert(offsetof(DevRecord, Vbed) == 6 * 8 && offsetof(DevRecord, n_final) == 9 * 8 &&
                  offsetof(DevRecord, chk_final) == 11 * 8 && offsetof(DevRecord, aVtc) == 17 * 8,
              "DevRecord totals must be consecutive 64-bit fields");

template <int W = kWaves>
__device__ __forceinline__ int32_t group_or(int32_t f, int32_t *red, int lane, int wave) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off);
  if (lane == 0) red[wave] = f;
  __syncthreads();
  int32_t r = 0;
#pragma unroll
  for (int w = 0; w < W; w++) r |= red[w];
  return r;
}

}  // namespace

// ---- aabb_kernel: one work-group per orientation ------------------------------------------
// Rotates every vertex and reduces the rotated AABB minimum (STomoSh_Base.cpp:77-109) and
// resets the orientation's header and record.  The slot counters need no reset: the rasteriser
// writes every counter of its tile, zeros included.
__device__ __forceinline__ void aabb_body(const SweepParams &p, int32_t blk) {
  __shared__ float red[kWaves][3];
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_index();
  const int32_t b = blk;
  if (b >= p.n_batch) return;
  const float *Rg = p.rot + (int64_t)b * 9;
  const float r0 = Rg[0], r1 = Rg[1], r2 = Rg[2], r3 = Rg[3], r4 = Rg[4], r5 = Rg[5], r6 = Rg[6],
              r7 = Rg[7], r8 = Rg[8];
  OrientHdr &h = p.hdr[b];
  float mx = 1e5f, my = 1e5f, mz = 1e5f;
  for (int32_t v = tid; v < p.n_vtx; v += kThreads) {
    const float t0 = p.vx[v] - p.cx, t1 = p.vy[v] - p.cy, t2 = p.vz[v] - p.cz;
    const float x = dot_row(r0, r1, r2, t0, t1, t2);
    const float y = dot_row(r3, r4, r5, t0, t1, t2);
    const float z = dot_row(r6, r7, r8, t0, t1, t2);
    mx = fminf(mx, x);
    my = fminf(my, y);
    mz = fminf(mz, z);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mx = fminf(mx, __shfl_xor(mx, off));
    my = fminf(my, __shfl_xor(my, off));
    mz = fminf(mz, __shfl_xor(mz, off));
  }
  if (lane == 0) {
    red[wave][0] = mx;
    red[wave][1] = my;
    red[wave][2] = mz;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 0; w < kWaves; w++) {
      mx = fminf(mx, red[w][0]);
      my = fminf(my, red[w][1]);
      mz = fminf(mz, red[w][2]);
    }
    const float kx = mx - p.bed_out, ky = my - p.bed_out, kz = mz;
    h.bbox[0] = p.grid; h.bbox[1] = -1; h.bbox[2] = p.grid; h.bbox[3] = -1;
    h.big_top = 0u; h.pool_top = 0u; h.flags = 0; h.spill_top = 0u; h.huge_top = 0u;
    h.kx = kx; h.ky = ky; h.kz = kz;
    h.zabs = 0ull;
    DevRecord &R = p.rec[b];
    R = DevRecord{};
    R.kx = kx; R.ky = ky; R.kz = kz;
  }
}

// ---- raster_kernel: one tile of slots of one orientation per work-group -------------------
// Pixelize + triVoxel (STomoSh_INT3.cpp:285-364, 1138-1209).  The work-group owns the slot
// counters of a tile (in LDS) and rasterises every triangle whose pixel box meets the tile:
// the work-group culls the chunks (one per thread), dropping those whose projected box misses
// the tile, into an LDS list, and each wave takes listed chunks from a shared counter, one
// triangle per lane.  A lane walks its pixel box (clipped to the tile and to the columns and rows
// that can hold an accepted pixel) with the affine prefilter (tomo_raster.h); surviving pixels
// are queued in LDS and the exact coverage test, interpolation and slot append run on compacted
// lanes.  The kernel is bound by vector instruction issue; it is held to 80 VGPRs so that two
// 80 KB work-groups (6 waves per SIMD) stay resident per CU.  Every pixel is
// tested by exactly one tile; the edge tiles also take the pixels beyond the grid on their side.
// A triangle's zero-hit fallback (its centroid pixel) is decided by its owner tile, the one
// holding its clamped box corner (x0, y0), which walks the whole box to count hits.
// Work-groups follow xcd_map so an orientation's tiles share one XCD's L2.
enum : int { TR_A0 = 0, TR_A1 = 3, TR_A2 = 6, TR_NZ = 9, TR_STRIDE = 13 };
constexpr int kLaneBox = 48;  // pixel boxes walked lane by lane; larger ones by the whole wave
constexpr int kCullCap = 896;  // chunk list entries per window (LDS)

struct RasterLds {
  // placed (un-margined) vertices a0, a1, a2 (x, y, z) and n0z, n1z, n2z; everything else the
  // exact test needs is recomputed from them bit for bit.  Odd stride: lanes reading different
  // owners hit different banks.  First in the block, so a record's words are within reach of
  // the immediate offsets of LDS loads and stores.
  float tri[kTileWaves][64][TR_STRIDE];
  uint32_t cnt[kTileSlots / 2];  // the tile's slot counters, 16 bits each
  uint32_t qmeta[kTileWaves][128];  // owner lane | (x + 256) << 6 | (y + 256) << 19
  uint8_t hits[kTileWaves][64];
  int32_t red[kTileWaves][5];
  uint64_t zred[kTileWaves];
  uint16_t clist[kCullCap];  // kept chunks of the current window (offsets from its first chunk)
  uint32_t wcnt[kTileWaves];
  uint32_t take;             // next list position to hand out
  int32_t cut;
};
static_assert(sizeof(RasterLds) <= 80 * 1024, "two rasteriser work-groups must fit one CU's LDS");

// Conservative pixel range along one axis of a chunk's vertices after rotation (row r of the
// orientation matrix about the centre c) and placement (minus k): the float rotation and
// placement of each vertex stay within 1e-6 (|r| . |t| + |k|) of the real-number values, and
// the rasteriser's pixel box of a triangle lies within [floor(min), floor(max) + 1].
__device__ __forceinline__ void chunk_span(const float *bmin, const float *bmax, const float *c,
                                           float r0, float r1, float r2, float k, int32_t &lo,
                                           int32_t &hi) {
  const float r[3] = {r0, r1, r2};
  float mn = 0.0f, mx = 0.0f, mag = fabsf(k);
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const float a = r[q] * (bmin[q] - c[q]), b = r[q] * (bmax[q] - c[q]);
    mn += fminf(a, b);
    mx += fmaxf(a, b);
    mag += fmaxf(fabsf(a), fabsf(b));
  }
  const float pad = 2.0f + 1e-6f * mag;
  const float flo = fminf(fmaxf(mn - k - pad, -1e6f), 1e6f);
  const float fhi = fminf(fmaxf(mx - k + pad, -1e6f), 1e6f);
  lo = (int32_t)floorf(flo);
  hi = (int32_t)floorf(fhi) + 1;
  if (!(mn == mn) || !(mx == mx) || !(mag <= 3.0e38f)) {  // NaN or overflow: keep the chunk
    lo = -(1 << 30);
    hi = 1 << 30;
  }
}

__device__ __forceinline__ void raster_body(const SweepParams &p, int32_t blk) {
  __shared__ RasterLds L;
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_index();
  int32_t b, tile;
  if (!xcd_map(blk, p.tiles, p.n_batch, b, tile)) return;
  const int32_t G = p.grid;
  const int32_t ti = tile / p.tiles_y, tj = tile - ti * p.tiles_y;
  const int32_t xa = ti * p.tile_x, xb = min(xa + p.tile_x, G);
  const int32_t ya = tj << p.tile_y_shift, yb = min(ya + (1 << p.tile_y_shift), G);
  // pixels this tile tests: its slots, and beyond the grid on the sides where it is the edge
  const int32_t exlo = ti == 0 ? INT32_MIN : xa, exhi = ti == p.tiles_x - 1 ? INT32_MAX : xb - 1;
  const int32_t eylo = tj == 0 ? INT32_MIN : ya, eyhi = tj == p.tiles_y - 1 ? INT32_MAX : yb - 1;
  for (int i = tid; i < kTileSlots / 2; i += kTileThreads) L.cnt[i] = 0u;
  __syncthreads();
  OrientHdr &h = p.hdr[b];
  const float kx = h.kx, ky = h.ky, kz = h.kz;
  const float *Rg = p.rot + (int64_t)b * 9;
  const float R0 = Rg[0], R1 = Rg[1], R2 = Rg[2], R3 = Rg[3], R4 = Rg[4], R5 = Rg[5], R6 = Rg[6],
              R7 = Rg[7], R8 = Rg[8];
  const float ctr[3] = {p.cx, p.cy, p.cz};
  Emit em;
  em.G = G; em.G2 = G * G; em.t = tables_for(p, b);
  em.pool_top = &h.pool_top; em.spill_top = &h.spill_top;
  em.lcnt = L.cnt;
  em.xa = xa; em.xb = xb; em.ya = ya; em.yb = yb; em.yshift = p.tile_y_shift;
  em.bx0 = G; em.bx1 = -1; em.by0 = G; em.by1 = -1; em.flags = 0; em.zabs = 0ull;
  float (*tri)[TR_STRIDE] = L.tri[wave];
  // Chunks in windows: the work-group culls chunks together (one per thread) and lists the kept
  // ones in LDS, in chunk order, until the list is full; then each wave takes the listed chunks
  // one at a time from a shared LDS counter (the waves stay balanced whatever each chunk costs).
  int32_t next = 0;  // first chunk not yet culled (the same in every thread)
  while (next < p.n_chunks) {
    const int32_t wbase = next;
    uint32_t n = 0;  // listed chunks of this window
    if (tid == 0) L.take = 0u;
    while (next < p.n_chunks && n < (uint32_t)kCullCap && next - wbase <= 65535 - kTileThreads) {
      const int32_t ch = next + tid;
      bool keep = false;
      if (ch < p.n_chunks) {
        const float *bx = p.chunk_box + (int64_t)ch * 6;
        const float bmin[3] = {bx[0], bx[1], bx[2]}, bmax[3] = {bx[3], bx[4], bx[5]};
        int32_t lx, hx, ly, hy;
        chunk_span(bmin, bmax, ctr, R0, R1, R2, kx, lx, hx);
        chunk_span(bmin, bmax, ctr, R3, R4, R5, ky, ly, hy);
        keep = hx >= exlo && lx <= exhi && hy >= eylo && ly <= eyhi;
      }
      const uint64_t km = ballot(keep);
      if (lane == 0) L.wcnt[wave] = (uint32_t)__popcll(km);
      __syncthreads();
      uint32_t before = 0u, total = 0u;
#pragma unroll
      for (int w2 = 0; w2 < kTileWaves; w2++) {
        const uint32_t c = L.wcnt[w2];
        before += w2 < wave ? c : 0u;
        total += c;
      }
      const uint32_t pos = n + before + lanes_below(km);
      if (keep && pos < (uint32_t)kCullCap) L.clist[pos] = (uint16_t)(ch - wbase);
      if (keep && pos == (uint32_t)kCullCap) L.cut = ch;  // the first kept chunk that did not fit
      __syncthreads();
      if (n + total <= (uint32_t)kCullCap) {
        next = min(next + kTileThreads, p.n_chunks);
        n += total;
      } else {
        next = L.cut;
        n = (uint32_t)kCullCap;
      }
    }
    for (;;) {
      uint32_t idx = 0u;
      if (lane == 0) idx = atomicAdd(&L.take, 1u);
      idx = (uint32_t)__builtin_amdgcn_readlane((int)idx, 0);
      if (idx >= n) break;
      const int32_t ch = wbase + __builtin_amdgcn_readfirstlane((int)L.clist[idx]);
      const int32_t tr = ch * kChunk + lane;
      const bool active = tr < p.n_tri;
      // Lanes past the last triangle (only in the last chunk) set up that chunk's last triangle
      // again and walk nothing: no branch around the setup.
      int32_t x0, x1, y0, y1;
      bool owner;
      TriScan sc;
      float v0x, v0y;
      uint32_t my_rank;
      {
        float a[3][3], nzv[3];
        const int4 ix = p.tidx[active ? tr : p.n_tri - 1];
        my_rank = 2u * (uint32_t)ix.w;
        const int32_t ids[3] = {ix.x, ix.y, ix.z};
        // Rotate the three vertices (stored relative to the rotation centre, so shared by every
        // orientation and resident in L2) as aabb_kernel rotates them (STomoSh_Base.cpp:85-101),
        // then place them.
        float4 va[3], vb[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
          va[q] = p.vaos[2 * ids[q]];
          vb[q] = p.vaos[2 * ids[q] + 1];
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
          a[q][0] = dot_row_nz(R0, R1, R2, va[q].x, va[q].y, va[q].z) - kx;
          a[q][1] = dot_row_nz(R3, R4, R5, va[q].x, va[q].y, va[q].z) - ky;
          a[q][2] = dot_row_nz(R6, R7, R8, va[q].x, va[q].y, va[q].z) - kz;
          nzv[q] = dot_row_nz(R6, R7, R8, va[q].w, vb[q].x, vb[q].y);
        }
        v0x = a[0][0] + kEps; v0y = a[0][1] + kEps;
        const float v1x = a[1][0] + kEps, v1y = a[1][1] + kEps;
        const float v2x = a[2][0] + kEps, v2y = a[2][1] + kEps;
        float mnx = (v0x <= v1x) ? v0x : v1x; mnx = (mnx <= v2x) ? mnx : v2x;
        float mny = (v0y <= v1y) ? v0y : v1y; mny = (mny <= v2y) ? mny : v2y;
        float mxx = (v0x >= v1x) ? v0x : v1x; mxx = (mxx >= v2x) ? mxx : v2x;
        float mxy = (v0y >= v1y) ? v0y : v1y; mxy = (mxy >= v2y) ? mxy : v2y;
        x0 = (int32_t)mnx; y0 = (int32_t)mny; x1 = (int32_t)mxx; y1 = (int32_t)mxy;
        // Pixels far outside the grid only yield dropped entries: bound the loop so a wild
        // vertex cannot stall the group, and report the footprint overflow.
        const int32_t lo = -64, hi = G + 64;
        if (x0 < lo || y0 < lo || x1 > hi || y1 > hi) {
          em.flags |= F_FOOTPRINT_OOB;
          x0 = max(x0, lo); y0 = max(y0, lo); x1 = min(x1, hi); y1 = min(y1, hi);
        }
        const int32_t ox = min(max(x0, 0), G - 1), oy = min(max(y0, 0), G - 1);
        owner = active && ox >= xa && ox < xb && oy >= ya && oy < yb;
        // The owner walks the whole box when it leaves the tile (hits decide the fallback);
        // every other walk is clipped to the pixels this tile tests.
        if (!(owner && (x0 < exlo || x1 > exhi || y0 < eylo || y1 > eyhi))) {
          x0 = max(x0, exlo); x1 = min(x1, exhi); y0 = max(y0, eylo); y1 = min(y1, eyhi);
        }
        // The chunk's culling box is padded: when no triangle of it reaches the tile (no box
        // left and no owner), the rest of the set-up and the walk have nothing to do.
        if (!ballot(active && (owner || (x1 >= x0 && y1 >= y0)))) continue;
        const TriSetup ts = tri_setup(v0x, v0y, v1x, v1y, v2x, v2y);
        if (tri_degenerate(ts)) { x1 = x0 - 1; }  // no pixel can pass
        const float ex = fmaxf(fabsf(centre_offset(x0, p.h2, v0x)), fabsf(centre_offset(x1, p.h2, v0x)));
        const float ey = fmaxf(fabsf(centre_offset(y0, p.h2, v0y)), fabsf(centre_offset(y1, p.h2, v0y)));
        sc = tri_scan(ts, ex, ey);
        // Walk only the columns and rows that can hold an accepted pixel.
        const AcceptRange ar = accept_range(ts, sc);
        if (ar.ok) {
          clip_span(x0, x1, ar.lox, ar.hix, v0x, p.h2);
          clip_span(y0, y1, ar.loy, ar.hiy, v0y, p.h2);
        }
        float *rec = tri[lane];
#pragma unroll
        for (int q = 0; q < 3; q++) {
          rec[TR_A0 + q] = a[0][q];
          rec[TR_A1 + q] = a[1][q];
          rec[TR_A2 + q] = a[2][q];
          rec[TR_NZ + q] = nzv[q];
        }
      }
      L.hits[wave][lane] = 0;
      int32_t x = x0;
      bool done = !active || x1 < x0 || y1 < y0;
      // Boxes of more than kLaneBox pixels are left out of the lane-by-lane walk and walked
      // afterwards by the whole wave, 64 pixels per step, so that one large box does not hold
      // the wave for its whole length.
      const bool big = !done && (int64_t)(x1 - x0 + 1) * (int64_t)(y1 - y0 + 1) > kLaneBox;
      done = done || big;
      uint64_t bigm = ballot(big);
      uint32_t qn = 0;
      // Exact test, interpolation and slot append of the first min(qn, 64) queued candidates.
      auto flush = [&]() {
        wave_sync();
        const uint32_t np = qn < 64u ? qn : 64u;
        const uint32_t meta = (uint32_t)lane < np ? L.qmeta[wave][lane] : 0u;
        const int32_t owner_lane = (int32_t)(meta & 63u);
        const uint32_t rank = __shfl(my_rank, owner_lane);
        if ((uint32_t)lane < np) {
          const int32_t px_i = (int32_t)((meta >> 6) & 8191u) - 256;
          const int32_t py_i = (int32_t)(meta >> 19) - 256;
          const float *rec = tri[owner_lane];
          // margined vertices and the exact test's numerators, formed as triVoxel and
          // _getBaryCoord form them (the same operations as the setup above)
          float m[3][3];
#pragma unroll
          for (int q = 0; q < 3; q++) {
            m[0][q] = rec[TR_A0 + q] + kEps;
            m[1][q] = rec[TR_A1 + q] + kEps;
            m[2][q] = rec[TR_A2 + q] + kEps;
          }
          // tri_setup and bary_num without their "+ 0" z terms: those only turn a -0 dot product
          // into +0, and the sign of a zero there reaches no decision and no stored value (the
          // quotients' zeros enter sums, comparisons and truncations to integers only).
          const float e0x = m[1][0] - m[0][0], e0y = m[1][1] - m[0][1];
          const float e1x = m[2][0] - m[0][0], e1y = m[2][1] - m[0][1];
          const float d00 = e0x * e0x + e0y * e0y, d01 = e0x * e1x + e0y * e1y, d11 = e1x * e1x + e1y * e1y;
          const float den = d00 * d11 - d01 * d01;
          const float e2x = (((float)px_i + kEps) + p.h2) - m[0][0], e2y = (((float)py_i + kEps) + p.h2) - m[0][1];
          const float d20 = e2x * e0x + e2y * e0y, d21 = e2x * e1x + e2y * e1y;
          const float nv = d11 * d20 - d01 * d21, nw = d00 * d21 - d01 * d20;
          float u, v, w;
          if (bary_exact(den, nv, nw, u, v, w)) {
            L.hits[wave][owner_lane] = 1;
            if (px_i >= exlo && px_i <= exhi && py_i >= eylo && py_i <= eyhi) {
              // _bary_product (Tomo_types.cpp:75-84) of the margined vertices and the normals
              const float px = m[0][0] * u + m[1][0] * v + m[2][0] * w;
              const float py = m[0][1] * u + m[1][1] * v + m[2][1] * w;
              const float pz = m[0][2] * u + m[1][2] * v + m[2][2] * w;
              const float nz = rec[TR_NZ + 0] * u + rec[TR_NZ + 1] * v + rec[TR_NZ + 2] * w;
              if (p.do_shell) {  // with its shell sub-pixel (STomoSh_INT3.cpp:1179-1185)
                const float pz2 = pz + ((nz >= 0.0f) ? -p.shell_off : p.shell_off);
                em.put2(px, py, pz, nz, pz2, -nz, rank);
              } else {
                em.put(px, py, pz, nz, rank);
              }
            }
          }
        }
        wave_sync();
        if (qn > 64u) {
          const uint32_t rest = qn - 64u;
          uint32_t mm = 0u;
          if ((uint32_t)lane < rest) mm = L.qmeta[wave][64 + lane];
          wave_sync();
          if ((uint32_t)lane < rest) L.qmeta[wave][lane] = mm;
          wave_sync();
        }
        qn -= np;
      };
      wave_sync();
      // Column-major walk of the box, written without divergent branches: every lane advances
      // every step (a finished lane's position runs on past x1 and is never tested).  The row
      // is carried as yf == (float)y (exact: |y| < 2^24), so its centre offset is formed with
      // the same float operations as centre_offset(y, ...); meta is the queue word of (x, y),
      // lane | (x + 256) << 6 | (y + 256) << 19, advanced in place.
      {
        const float y0f = (float)y0, y1f = (float)y1;
        const uint32_t mcol = (uint32_t)lane + ((uint32_t)(y0 + 256) << 19) + (256u << 6);
        const uint64_t lanes = __builtin_amdgcn_read_exec();
        uint64_t dm = ballot(done);  // lanes whose walk has ended
        float yf = y0f;
        uint32_t meta = mcol + ((uint32_t)x << 6);
        float e2x = centre_offset(x, p.h2, v0x);
        float rax = e2x * sc.ax, rbx = e2x * sc.bx;
        for (;;) {
          const float e2y = ((yf + kEps) + p.h2) - v0y;
          const uint64_t m = scan_mask(sc, rax, rbx, e2y) & ~dm;
          const bool wrap = yf >= y1f;
          x += wrap ? 1 : 0;
          yf = wrap ? y0f : yf + 1.0f;
          const uint32_t mnext = wrap ? mcol + ((uint32_t)x << 6) : meta + (1u << 19);
          dm |= __builtin_amdgcn_uicmp((uint32_t)x, (uint32_t)x1, kIcmpSgt);
          e2x = centre_offset(x, p.h2, v0x);
          rax = e2x * sc.ax;
          rbx = e2x * sc.bx;
          if (__builtin_amdgcn_inverse_ballot_w64(m)) L.qmeta[wave][qn + lanes_below(m)] = meta;
          meta = mnext;
          qn += (uint32_t)__popcll(m);
          if (qn >= 64u) flush();
          if ((lanes & ~dm) == 0ull) break;
        }
      }
      // The large boxes, one at a time: their triangle's walk state is read from its lane into
      // wave-uniform values and the wave takes 64 consecutive pixels (column-major) per step,
      // with the same prefilter arithmetic as the lane-by-lane walk.
      while (bigm) {
        const int j = __ffsll((unsigned long long)bigm) - 1;
        bigm &= bigm - 1ull;
        const int32_t jx0 = __builtin_amdgcn_readlane(x0, j), jx1 = __builtin_amdgcn_readlane(x1, j);
        const int32_t jy0 = __builtin_amdgcn_readlane(y0, j), jy1 = __builtin_amdgcn_readlane(y1, j);
        const float jv0x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v0x), j));
        const float jv0y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v0y), j));
        TriScan js;
        js.ax = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc.ax), j));
        js.ay = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc.ay), j));
        js.bx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc.bx), j));
        js.by = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc.by), j));
        js.lo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc.lo), j));
        js.his = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sc.his), j));
        const int32_t hgt = jy1 - jy0 + 1, npx = (jx1 - jx0 + 1) * hgt;
        const int32_t dq = 64 / hgt, dr = 64 - dq * hgt;
        const int32_t col = lane / hgt;
        int32_t wx = jx0 + col, wy = jy0 + (lane - col * hgt);
        for (int32_t k0 = 0; k0 < npx; k0 += 64) {
          const float e2 = centre_offset(wx, p.h2, jv0x);
          const uint64_t m = scan_mask(js, e2 * js.ax, e2 * js.bx, centre_offset(wy, p.h2, jv0y)) &
                             ballot(k0 + lane < npx);
          if (__builtin_amdgcn_inverse_ballot_w64(m))
            L.qmeta[wave][qn + lanes_below(m)] = (uint32_t)j | ((uint32_t)(wx + 256) << 6) | ((uint32_t)(wy + 256) << 19);
          qn += (uint32_t)__popcll(m);
          if (qn >= 64u) flush();
          wy += dr;
          wx += dq;
          if (wy > jy1) {
            wy -= hgt;
            wx += 1;
          }
        }
      }
      while (qn > 0u) flush();
      wave_sync();
      if (active && owner && L.hits[wave][lane] == 0) {
        // very small triangle: one pixel at the un-margined centroid (STomoSh_INT3.cpp:1190-1206)
        const float *rec = tri[lane];
        float c3[3];
#pragma unroll
        for (int q = 0; q < 3; q++) c3[q] = (rec[TR_A0 + q] + rec[TR_A1 + q] + rec[TR_A2 + q]) * 0.333333f + kEps;
        const float tx = c3[0], ty = c3[1], tz = c3[2], n0 = rec[TR_NZ + 0];
        if (p.do_shell) {
          const float sz = tz + ((n0 >= 0.0f) ? -p.shell_off : p.shell_off);
          em.put2(tx, ty, tz, n0, sz, -n0, my_rank);
        } else {
          em.put(tx, ty, tz, n0, my_rank);
        }
      }
      wave_sync();
    }
    __syncthreads();  // every listed chunk is done before the list is refilled
  }
  // Write every counter of the tile, zeros included (so no later kernel has to clear the
  // global table), and list the slots with more than kSmall entries.
  {
    const Tables &t = em.t;
    const int32_t yshift = p.tile_y_shift, ymask = (1 << yshift) - 1;
    const int32_t nwords = ((xb - xa) << yshift) >> 1;
    for (int32_t w0 = 0; w0 < nwords; w0 += kTileThreads) {
      const int32_t wi = w0 + tid;
      const uint32_t word = wi < nwords ? L.cnt[wi] : 0u;
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const uint32_t c = (word >> (16 * half)) & 0xffffu;
        const int32_t li = 2 * wi + half;
        const int32_t sx = xa + (li >> yshift), sy = ya + (li & ymask);
        const int32_t s = sx * G + sy;
        // only this tile's slots: threads past its last word, and the part of the last tile
        // row beyond the grid, write nothing
        if (wi < nwords && sy < yb) t.count[s] = c;
        const bool big = c > (uint32_t)kSmall;
        const uint64_t bm = ballot(big);
        if (bm) {
          uint32_t base = 0u;
          if (lane == 0) base = atomicAdd(&h.big_top, (uint32_t)__popcll(bm));
          base = __shfl(base, 0);
          if (big) t.big[base + lanes_below(bm)] = s;
        }
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    em.bx0 = min(em.bx0, __shfl_xor(em.bx0, off));
    em.bx1 = max(em.bx1, __shfl_xor(em.bx1, off));
    em.by0 = min(em.by0, __shfl_xor(em.by0, off));
    em.by1 = max(em.by1, __shfl_xor(em.by1, off));
    em.flags |= __shfl_xor(em.flags, off);
  }
  const uint64_t zw = wave_sum(em.zabs);
  if (lane == 0) {
    L.red[wave][0] = em.bx0; L.red[wave][1] = em.bx1; L.red[wave][2] = em.by0; L.red[wave][3] = em.by1;
    L.red[wave][4] = em.flags;
    L.zred[wave] = zw;
  }
  __syncthreads();
  if (tid == 0) {
    int32_t v0 = G, v1 = -1, v2 = G, v3 = -1, f = 0;
    for (int w = 0; w < kTileWaves; w++) {
      v0 = min(v0, L.red[w][0]); v1 = max(v1, L.red[w][1]);
      v2 = min(v2, L.red[w][2]); v3 = max(v3, L.red[w][3]);
      f |= L.red[w][4];
    }
    if (v1 >= v0) {
      atomicMin(&h.bbox[0], v0); atomicMax(&h.bbox[1], v1);
      atomicMin(&h.bbox[2], v2); atomicMax(&h.bbox[3], v3);
    }
    if (f) atomicOr(&h.flags, f);
    uint64_t z = 0;
    for (int w = 0; w < kTileWaves; w++) z += L.zred[w];
    if (z) atomicAdd(&h.zabs, (unsigned long long)z);
  }
}

// ---- compact_kernel: one work-group per orientation ----------------------------------------
// First appends the rasteriser's spilled entries to their slots (global counters: every tile
// has written its counters back).  Entries beyond the inline layers were appended to a pool in
// arrival order; group them per slot (exclusive scan over the large-slot list, then scatter).
constexpr int kCompactCap = 4096;  // large slots whose overflow cursors fit in LDS

__device__ __forceinline__ void compact_body(const SweepParams &p, int32_t blk) {
  __shared__ uint32_t scan[kWaves];
  __shared__ uint32_t lcur[kCompactCap];  // overflow cursor of each large-slot list entry
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_index();
  const int32_t b = blk;
  if (b >= p.n_batch) return;
  OrientHdr &h = p.hdr[b];
  const Tables t = tables_for(p, b);
  const uint32_t nspill = h.spill_top;
  if (nspill > 0u) {
    const int32_t G2 = p.grid * p.grid;
    const uint32_t ns = nspill < (uint32_t)t.spill ? nspill : (uint32_t)t.spill;
    for (uint32_t j = tid; j < ns; j += kThreads) {
      const int32_t s = (int32_t)t.spill_slot[j];
      const uint32_t k = atomicAdd(&t.count[s], 1u);
      place_entry(t, G2, &h.pool_top, s, k, t.spill_key[j], t.spill_rank[j]);
      if (k == (uint32_t)kSmall) t.big[atomicAdd(&h.big_top, 1u)] = s;
    }
    __syncthreads();
  }
  const uint32_t npool = __hip_atomic_load(&h.pool_top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  const uint32_t nbig = __hip_atomic_load(&h.big_top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (npool == 0u) return;
  // With at most kCompactCap large slots their cursors live in LDS (many entries of one slot
  // would otherwise serialise on one memory-side atomic) and cursor[s] holds s's list index.
  const bool lcnt = nbig <= (uint32_t)kCompactCap;
  if (lcnt)
    for (uint32_t i = tid; i < nbig; i += kThreads) lcur[i] = 0u;
  uint32_t running = 0;
  for (uint32_t b0 = 0; b0 < nbig; b0 += kThreads) {
    const uint32_t j = b0 + tid;
    int32_t s = -1;
    uint32_t ex = 0;
    if (j < nbig) {
      s = t.big[j];
      const uint32_t c = t.count[s];
      ex = c > (uint32_t)t.inl ? c - (uint32_t)t.inl : 0u;
    }
    uint32_t incl = ex;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t u = __shfl_up(incl, off);
      if (lane >= off) incl += u;
    }
    if (lane == 63) scan[wave] = incl;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      if (w < wave) wbase += scan[w];
      total += scan[w];
    }
    if (s >= 0 && ex > 0) {
      t.off[s] = running + wbase + incl - ex;
      t.cursor[s] = lcnt ? j : 0u;
    }
    running += total;
    __syncthreads();
  }
  const uint32_t np = npool < (uint32_t)t.pool ? npool : (uint32_t)t.pool;
  for (uint32_t j = tid; j < np; j += kThreads) {
    const uint32_t s = t.pool_slot[j];
    const uint32_t q = lcnt ? atomicAdd(&lcur[t.cursor[s]], 1u) : atomicAdd(&t.cursor[s], 1u);
    const uint32_t d = t.off[s] + q;
    if (d < (uint32_t)t.pool) {
      t.cmp_key[d] = t.pool_key[j];
      t.cmp_rank[d] = t.pool_rank[j];
    }
  }
}

// ---- pair_kernel: pair_span slots of the touched box per work-group -----------------------
// Pairing (STomoSh_INT3.cpp:366-395) and the per-slot sums of Calculate (:1211-1255).
// pair_kernel (PART 0) pairs the slots of at most kSmall entries found in the work-group's part
// of the touched box; pair_big_kernel (PART 1, a second launch with its own register budget)
// pairs the large-slot list.
template <int PART, bool DBG>
__device__ __forceinline__ void pair_body(const SweepParams &p, int32_t blk) {
  __shared__ uint64_t sortbuf[kWaves][kBigCap];
  __shared__ uint32_t pairq[2][kWaves][128];  // queued slots of 3 .. kSmall entries: i - i0 | entries << 24
  // pair_kernel's bed bookkeeping for its part of the box, two 16-bit values a word (0xffff: a
  // slot of pair_big_kernel's), written out in slot order at the end
  __shared__ uint32_t bwin2[PART == 0 ? kPairSpanMax / 2 : 1];
  __shared__ int64_t red[kWaves][18];
  __shared__ int32_t fred[kWaves];
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_index();
  int32_t b, tile;
  const int32_t groups = PART == 0 ? p.pair_tiles : p.big_groups;  // work-groups per orientation
  if (!xcd_map(blk, groups, p.n_batch, b, tile)) return;
  const int32_t G = p.grid, G2 = G * G;
  const OrientHdr &h = p.hdr[b];
  const int32_t bx0 = h.bbox[0], bx1 = h.bbox[1], by0 = h.bbox[2], by1 = h.bbox[3];
  const uint32_t nbig = h.big_top, npool = h.pool_top;
  const bool any = bx1 >= bx0;
  const int32_t wy = by1 - by0 + 1;
  const int32_t nbb = any ? (bx1 - bx0 + 1) * wy : 0;
  const int32_t i0 = tile * p.pair_span;
  const int32_t nwaves = groups * kWaves;
  if (PART == 0 && tile == 0 && tid == 0) {
    // rasteriser flags, and pool or spill list overflow
    const int32_t f0 = h.flags | (npool > (uint32_t)p.lay.pool || h.spill_top > (uint32_t)p.lay.spill ? F_RETRY : 0);
    if (f0) atomicOr(&p.rec[b].flags, f0);
  }
  if (PART == 0 ? i0 >= nbb : (uint32_t)(tile * kWaves) >= nbig) return;  // nothing for this group
  const Tables t = tables_for(p, b);
  const bool final_here = !p.bed_on;  // without a bed the final row count is known here
  const bool track = totals_may_round(h, p.pp);
  int32_t *const sv = (p.exact_fsum && track) ? t.sv : nullptr;
  // pair_kernel: a wave's span holds at most 2,048 slots (32 steps); a lane pairs at most one
  // slot of at most 2 entries per step and one slot per batch of the two queued classes, so at
  // most 2^7 slots of |value| < 2^22: 32-bit totals do not overflow.  pair_big_kernel's lanes
  // can account for a whole wave's large slots and keep 64-bit totals.
  AccT<typename std::conditional<PART == 0, int32_t, int64_t>::type> acc = {};
  // The debug dumps (rows, final row counts, per-slot sums) exist only in the DBG instances,
  // which the host launches when a dump was asked for: the others keep no registers for them.
  uint64_t *const dbg_rows = DBG ? p.dbg_rows : nullptr;
  int32_t *const dbg_final = DBG ? p.dbg_final : nullptr;
  // Pair one slot of at most N entries in registers (one lane).  Slots of at most 2 entries
  // come with their keys (ka, kb: 0 past the count); larger ones read theirs here.
  uint16_t *const bwin = (uint16_t *)bwin2;
  auto pair_small = [&](auto nt, int32_t s, int32_t wi, uint32_t c, uint32_t ka, uint32_t kb) {
    constexpr int N = decltype(nt)::value;
    acc.nslot++;
    uint32_t k[N];
    if constexpr (N == 2) {
      k[0] = ka;
      k[1] = kb;
    } else {
#pragma unroll
      for (int q = 0; q < N; q++) k[q] = (q < (int)c) ? t.ent[(int64_t)q * G2 + s].x : 0u;
    }
    SlotResult r;
    if (N == 2 && !dbg_rows) pair_two(k[0], k[1], (int)c, p.pp, r);  // straight-line form
    else pair_keys<N>(k, (int)c, p.pp, r, dbg_rows ? dbg_rows + (int64_t)s * 64 : nullptr);
    account<DBG>(acc, r, s, p, track);
    bwin[wi] = (uint16_t)bed_info(r, p.pp);
    if (sv) put_sv(sv, G2, s, r);
    if (final_here) {
      acc.nfinal += r.n;
      acc.chkfinal += (uint64_t)(s + 1) * (uint64_t)r.n;
      if (dbg_final) dbg_final[s] = r.n;
    }
  };
  // The touched box is walked 64 slots per wave step.  A step issues the loads of the next
  // step's counters and of its own slots' keys when they have at most 2 entries (most slots),
  // then pairs the previous step's slots of at most 2 entries (their keys have arrived).  Slots
  // of 3 .. kSmall entries are queued by size class (<= 4, <= 8) and paired 64 at a time with
  // that class's register network, so no wave runs a larger network than its slots need.  One
  // extra step with no slots pairs what remains.  The bed bookkeeping goes to LDS and leaves in
  // slot order once the work-group is done: stored as each slot is paired, the queued slots'
  // values reached a line long after its neighbours' and L2 wrote it out in pieces, three times
  // the bytes.
  if constexpr (PART == 0) {
  uint32_t n4 = 0u, n8 = 0u;  // queue fills (wave-uniform)
  uint32_t *const q4 = pairq[0][wave], *const q8 = pairq[1][wave];
  const int32_t i1 = min(i0 + p.pair_span, nbb);
  __syncthreads();  // a work-queue launch reuses the window for its next unit
  for (int32_t j = tid; j < (i1 - i0 + 1) / 2; j += kThreads) bwin2[j] = ~0u;
  __syncthreads();
  if (i0 + wave * 64 < i1) {
    // slot i of the box is (bx0 + i / wy, by0 + i % wy); step kThreads slots at a time
    const int32_t first = i0 + wave * 64 + lane;
    int32_t sx = bx0 + first / wy, sy = by0 + first % wy;
    const int32_t dq = kThreads / wy, dr = kThreads % wy;
    uint32_t c_next = first < i1 ? t.count[sx * G + sy] : 0u;
    // the previous step's slot of this lane when it has at most 2 entries (pc = 3: none)
    int32_t ps = 0;
    uint32_t pc = 3u, pka = 0u, pkb = 0u;
    for (int32_t base = i0 + wave * 64;; base += kThreads) {
      const bool tail = base >= i1;  // the extra step
      const int32_t i = base + lane;
      const int32_t s = sx * G + sy;
      const uint32_t c = i < i1 ? c_next : 0u;
      sx += dq;
      sy += dr;
      if (sy > by1) {
        sy -= wy;
        sx += 1;
      }
      if (i + kThreads < i1) c_next = t.count[sx * G + sy];
      uint32_t ka = 0u, kb = 0u;
      if (c - 1u < 2u) {
        ka = t.ent[s].x;
        if (c == 2u) kb = t.ent[(int64_t)G2 + s].x;
      }
      const int32_t pwi = base - kThreads + lane - i0;  // the previous step's slot in the window
      if (pc == 0u) {
        bwin[pwi] = (uint16_t)BI_POSS;
        if (sv) {
          const SlotResult z = {};
          put_sv(sv, G2, ps, z);
        }
      } else if (pc <= 2u) {
        pair_small(std::integral_constant<int, 2>{}, ps, pwi, pc, pka, pkb);
      }
      // Pair full batches (in the extra step, everything queued), one class at a time: a single
      // call site per network keeps the code (and the register demand) small.
      for (;;) {
        const uint32_t lim = tail ? 1u : 64u;
        const int fc = n4 >= lim ? 0 : (n8 >= lim ? 1 : -1);
        if (fc < 0) break;
        uint32_t *const q = fc == 0 ? q4 : q8;
        const uint32_t n = fc == 0 ? n4 : n8;
        const uint32_t np = n < 64u ? n : 64u;
        wave_sync();
        if ((uint32_t)lane < np) {
          const uint32_t qe = q[lane];
          const int32_t qwi = (int32_t)(qe & 0xffffffu), qi = i0 + qwi;
          const int32_t qs = (bx0 + qi / wy) * G + by0 + qi % wy;
          const uint32_t qc = qe >> 24;
          if (fc == 0) pair_small(std::integral_constant<int, 4>{}, qs, qwi, qc, 0u, 0u);
          else pair_small(std::integral_constant<int, kSmall>{}, qs, qwi, qc, 0u, 0u);
        }
        wave_sync();
        if (n > 64u) {
          const uint32_t rest = n - 64u;
          uint32_t mm = 0u;
          if ((uint32_t)lane < rest) mm = q[64 + lane];
          wave_sync();
          if ((uint32_t)lane < rest) q[lane] = mm;
          wave_sync();
        }
        if (fc == 0) n4 = n - np;
        else n8 = n - np;
      }
      if (tail) break;
      // this step's slots: at most 2 entries wait for the next step, larger ones are queued
      // (more than kSmall: paired from the large-slot list)
      ps = s;
      pc = i < i1 && c <= 2u ? c : 3u;
      pka = ka;
      pkb = kb;
      int cls = -1;
      uint32_t e = 0u;
      if (i < i1 && c > 2u && c <= (uint32_t)kSmall) {
        cls = c <= 4u ? 0 : 1;
        e = (uint32_t)(i - i0) | (c << 24);
      }
      {
        const uint64_t m = ballot(cls == 0);
        if (cls == 0) q4[n4 + lanes_below(m)] = e;
        n4 += (uint32_t)__popcll(m);
      }
      {
        const uint64_t m = ballot(cls == 1);
        if (cls == 1) q8[n8 + lanes_below(m)] = e;
        n8 += (uint32_t)__popcll(m);
      }
    }
  }
  __syncthreads();
  {
    int32_t wsx = bx0 + (i0 + tid) / wy, wsy = by0 + (i0 + tid) % wy;
    const int32_t dq = kThreads / wy, dr = kThreads % wy;
    for (int32_t wi = tid; wi < i1 - i0; wi += kThreads) {
      const uint32_t v = bwin[wi];
      if (v != 0xffffu) t.bedinfo[wsx * G + wsy] = (uint16_t)v;
      wsx += dq;
      wsy += dr;
      if (wsy > by1) {
        wsy -= wy;
        wsx += 1;
      }
    }
  }
  } else {
  // Those with at most 16 entries, all inline: one lane each, 16-key register network.
  const uint32_t lane_cap = t.inl < 16 ? (uint32_t)t.inl : 16u;
  for (uint32_t j = (uint32_t)(tile * kThreads + tid); j < nbig; j += (uint32_t)(groups * kThreads)) {
    const int32_t s = t.big[j];
    const uint32_t c = t.count[s];
    if (c > lane_cap) continue;
    uint32_t k[16];
#pragma unroll
    for (int q = 0; q < 16; q++) k[q] = (q < (int)c) ? t.ent[(int64_t)q * G2 + s].x : 0u;
    SlotResult r;
    pair_keys<16>(k, (int)c, p.pp, r, dbg_rows ? dbg_rows + (int64_t)s * 64 : nullptr);
    acc.nslot++;
    account<DBG>(acc, r, s, p, track);
    t.bedinfo[s] = (uint16_t)bed_info(r, p.pp);
    if (sv) put_sv(sv, G2, s, r);
    if (final_here) {
      acc.nfinal += r.n;
      acc.chkfinal += (uint64_t)(s + 1) * (uint64_t)r.n;
      if (dbg_final) dbg_final[s] = r.n;
    }
  }
  // Larger ones.  Each wave reads its share of the big list 64 entries at a time; slots of at
  // most 32 entries are paired two per wave (one per 32-lane half), the rest one per wave.
  auto finish_slot = [&](int32_t s, const SlotResult &r) {
    acc.nslot++;
    account<DBG>(acc, r, s, p, track);
    t.bedinfo[s] = (uint16_t)bed_info(r, p.pp);
    if (sv) put_sv(sv, G2, s, r);
    if (final_here) {
      acc.nfinal += r.n;
      acc.chkfinal += (uint64_t)(s + 1) * (uint64_t)r.n;
      if (dbg_final) dbg_final[s] = r.n;
    }
  };
  // A slot left to pair_huge_kernel (see big_slot).
  auto defer = [&](int32_t s) {
    const uint32_t q = atomicAdd(&p.hdr[b].huge_top, 1u);
    if (q < (uint32_t)t.nhuge) t.huge[q] = s;
    else acc.flags |= F_RETRY;  // only when the pool overflowed as well
  };
  const bool pool_ok = npool <= (uint32_t)t.pool;
  for (uint32_t j0 = (uint32_t)(tile * kWaves + wave); j0 < nbig; j0 += (uint32_t)nwaves * 64u) {
    const uint32_t j = j0 + (uint32_t)nwaves * (uint32_t)lane;
    int32_t ls = 0;
    uint32_t lc = 0u;
    if (j < nbig) {
      ls = t.big[j];
      lc = t.count[ls];
    }
    const bool fits = lc <= (uint32_t)t.inl || pool_ok;
    const bool half = lc > lane_cap && lc <= 32u && fits;
    uint64_t m32 = ballot(half);
    uint64_t m64 = ballot(lc > lane_cap && !half);
    // Two slots per step, one per 32-lane half; the next step's keys are loaded before this
    // step sorts, so their latency overlaps the shuffles.
    auto take2 = [&](int32_t &s, uint32_t &c) {
      const int la = __ffsll((unsigned long long)m32) - 1;
      m32 &= m32 - 1ull;
      int lb = -1;
      if (m32) {
        lb = __ffsll((unsigned long long)m32) - 1;
        m32 &= m32 - 1ull;
      }
      const bool upper = lane >= 32;
      const int src = (upper && lb >= 0) ? lb : la;
      s = __shfl(ls, src);
      const uint32_t cs = __shfl(lc, src);
      c = (upper && lb < 0) ? 0u : cs;  // an idle upper half pairs nothing
    };
    int32_t s = 0;
    uint32_t c = 0u, key = 0u;
    bool have = m32 != 0ull;
    if (have) {
      take2(s, c);
      key = mid_key<32>(s, c, t, G2, lane);
    }
    while (have) {
      int32_t s2 = 0;
      uint32_t c2 = 0u, key2 = 0u;
      const bool have2 = m32 != 0ull;
      if (have2) {
        take2(s2, c2);
        key2 = mid_key<32>(s2, c2, t, G2, lane);
      }
      SlotResult r;
      mid_sort<32>(key, c, lane, p.pp, r, (dbg_rows && c) ? dbg_rows + (int64_t)s * 64 : nullptr);
      if ((lane & 31) == 0 && c) finish_slot(s, r);
      s = s2;
      c = c2;
      key = key2;
      have = have2;
    }
    while (m64) {
      const int la = __ffsll((unsigned long long)m64) - 1;
      m64 &= m64 - 1ull;
      const int32_t s = __shfl(ls, la);
      const uint32_t c = __shfl(lc, la);
      if (c > (uint32_t)kBigCap) {
        if (lane == 0) defer(s);
        continue;
      }
      uint64_t *const rows = dbg_rows ? dbg_rows + (int64_t)s * 64 : nullptr;
      SlotResult r;
      if (c < 64u && (c <= (uint32_t)t.inl || pool_ok)) {
        mid_slot<64>(s, c, t, G2, lane, p.pp, r, rows);
      } else if (!big_slot(p, b, s, c, t, G2, sortbuf[wave], lane, r, rows)) {
        if (lane == 0) defer(s);
        continue;
      }
      if (lane == 0) finish_slot(s, r);
    }
  }
  }
  DevRecord &R = p.rec[b];
  // Va, Vb, Vo, Vss, Vnv, Vtc, Vbed, n_entries, n_slots, n_final, chk_pre, chk_final, aVo .. aVtc
  int64_t v[18] = {acc.Va, acc.Vb, acc.Vo, acc.Vss, acc.Vnv, acc.Vtc, 0, acc.nent, acc.nslot,
                   acc.nfinal, (int64_t)acc.chkpre, (int64_t)acc.chkfinal, acc.aVo, acc.aVss,
                   acc.aVa, acc.aVb, acc.aVnv, acc.aVtc};
  group_add<18>(v, red, lane, wave, tid, (unsigned long long *)&R.Va);
  const int32_t f = group_or(acc.flags, fred, lane, wave);
  if (tid == 0 && f) atomicOr(&R.flags, f);
}


// ---- pair_huge_kernel: one work-group per orientation -------------------------------------
// Pairs the slots pair_big_kernel listed: those with more than kBigCap entries (rows from many
// triangles over one pixel: vertical walls of finely meshed parts put thousands there) and those
// whose 63rd and 64th distinct keys come from one triangle (cap_select orders them).  Up to
// kHugeLds entries are sorted in LDS by the whole work-group and the first entry of each key
// keeps the key's smallest rank, as in big_slot; beyond that the kept keys are chosen in passes
// over the entries in global memory, one triangle (the one with the smallest rank not yet kept)
// per pass.  Either way the 63-row cap keeps SetBit's first 63 distinct keys
// (STomoVoxelSpaceInfo.cpp:200-220), with one triangle's keys ordered by pick_first_of_triangle.
// Most orientations have no such slot and return at once.
__device__ void group_sort_desc(uint64_t *buf, int P, int tid) {
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < (P >> 1); i += kThreads) {
        const int lo = 2 * i - (i & (stride - 1));
        const int hi = lo + stride;
        const uint64_t a = buf[lo], c = buf[hi];
        const bool sw = ((lo & size) == 0) ? (a < c) : (a > c);
        if (sw) {
          buf[lo] = c;
          buf[hi] = a;
        }
      }
      __syncthreads();
    }
  }
}

// Exclusive count of `flag` over the work-group's threads before this one; *total gets the sum.
__device__ __forceinline__ uint32_t group_scan(bool flag, uint32_t *wcnt, int lane, int wave, uint32_t &total) {
  const uint64_t m = ballot(flag);
  if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t before = 0u;
  total = 0u;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    const uint32_t c = wcnt[w];
    before += w < wave ? c : 0u;
    total += c;
  }
  __syncthreads();
  return before + lanes_below(m);
}

__device__ __forceinline__ uint32_t group_min(uint32_t v, uint32_t *wv, int lane, int wave) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o));
  if (lane == 0) wv[wave] = v;
  __syncthreads();
  uint32_t r = wv[0];
#pragma unroll
  for (int w = 1; w < kWaves; w++) r = min(r, wv[w]);
  __syncthreads();
  return r;
}

// key in the ascending list sel[0..n)
__device__ __forceinline__ bool in_sorted(const uint32_t *sel, uint32_t n, uint32_t key) {
  uint32_t lo = 0u, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sel[mid] < key) lo = mid + 1u;
    else hi = mid;
  }
  return lo < n && sel[lo] == key;
}

// Wave: buf[0..n) = (key << 32 | ~rank) entries (n <= P, P a power of two >= 64): distinct keys
// with their smallest rank, as (rank << 32) | key in ascending (rank, key) order, in buf[0..D).
__device__ int wave_distinct_by_rank(uint64_t *buf, int n, int lane) {
  const int P = pow2_at_least(n);
  for (int i = n + lane; i < P; i += 64) buf[i] = 0ull;
  wave_sync();
  wave_sort_desc(buf, P, lane);
  int D = 0;
  uint32_t carry = 0u;
  for (int base = 0; base < n; base += 64) {
    const int i = base + lane;
    const uint64_t v = (i < n) ? buf[i] : 0ull;
    const uint32_t vkey = (uint32_t)(v >> 32);
    uint32_t pk = __shfl_up(vkey, 1);
    if (lane == 0) pk = carry;
    carry = __shfl(vkey, 63);
    const bool first = (i < n) && (i == 0 || vkey != pk);
    const uint64_t m = ballot(first);
    const int pos = D + (int)lanes_below(m);
    wave_sync();
    if (first) buf[pos] = ((uint64_t)(~(uint32_t)v) << 32) | vkey;
    wave_sync();
    D += __popcll(m);
  }
  const int P2 = pow2_at_least(D);
  for (int i = D + lane; i < P2; i += 64) buf[i] = ~0ull;
  wave_sync();
  for (int i = lane; i < P2; i += 64) buf[i] = ~buf[i];
  wave_sync();
  wave_sort_desc(buf, P2, lane);
  for (int i = lane; i < P2; i += 64) buf[i] = ~buf[i];
  wave_sync();
  return D;
}

__device__ __forceinline__ void pair_huge_body(const SweepParams &p, int32_t blk) {
  __shared__ uint64_t hbuf[kHugeLds];
  __shared__ uint32_t wcnt[kWaves];
  __shared__ uint32_t sel[64];  // kept keys, ascending (selection in passes)
  __shared__ uint32_t lcount, nsel;
  __shared__ int32_t capdrop;
  __shared__ int64_t red[kWaves][18];
  __shared__ int32_t fred[kWaves];
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_index();
  const int32_t b = blk;
  if (b >= p.n_batch) return;
  const OrientHdr &h = p.hdr[b];
  const uint32_t nh = min(h.huge_top, (uint32_t)p.lay.nhuge);
  if (nh == 0u) return;
  const Tables t = tables_for(p, b);
  const int32_t G2 = p.grid * p.grid;
  const bool final_here = !p.bed_on;
  const bool track = totals_may_round(h, p.pp);
  int32_t *const sv = (p.exact_fsum && track) ? t.sv : nullptr;
  Acc acc = {};
  for (uint32_t k = 0; k < nh; k++) {
    const int32_t s = t.huge[k];
    const uint32_t c = t.count[s];
    const uint32_t off = t.off[s];  // c > 63 > inline capacity: the slot has a pool run
    int32_t fl = 0;
    uint32_t mykey = 0u;  // wave 0: the kept keys, descending from lane 0
    int D = 0;
    if (c <= (uint32_t)kHugeLds) {
      const int P = pow2_at_least((int)c);
      for (int i = tid; i < P; i += kThreads) {
        uint64_t v = 0ull;
        if (i < (int)c) {
          uint32_t key, rank;
          slot_entry(t, G2, s, off, (uint32_t)i, key, rank, fl);
          v = ((uint64_t)key << 32) | (uint64_t)(~rank);
        }
        hbuf[i] = v;
      }
      __syncthreads();
      group_sort_desc(hbuf, P, tid);
      // first entry of each key (its smallest rank), compacted in place as (rank << 32) | key
      uint32_t nd = 0u, carry = 0u;
      for (int base = 0; base < (int)c; base += kThreads) {
        const int i = base + tid;
        const uint64_t v = i < (int)c ? hbuf[i] : 0ull;
        const uint32_t vkey = (uint32_t)(v >> 32);
        const uint32_t pk = tid == 0 ? carry : (uint32_t)(hbuf[i - 1] >> 32);
        const uint32_t next_carry = (uint32_t)(hbuf[min(base + kThreads, (int)c) - 1] >> 32);
        const bool first = i < (int)c && (i == 0 || vkey != pk);
        uint32_t total;
        const uint32_t pos = nd + group_scan(first, wcnt, lane, wave, total);
        if (first) hbuf[pos] = ((uint64_t)(~(uint32_t)v) << 32) | vkey;
        __syncthreads();
        carry = next_carry;
        nd += total;
      }
      D = (int)nd;
      if (D > kMaxPix) {
        fl |= F_CAP_DROP;
        const int P2 = pow2_at_least(D);
        for (int i = D + tid; i < P2; i += kThreads) hbuf[i] = ~0ull;
        __syncthreads();
        for (int i = tid; i < P2; i += kThreads) hbuf[i] = ~hbuf[i];
        __syncthreads();
        group_sort_desc(hbuf, P2, tid);
        for (int i = tid; i < P2; i += kThreads) hbuf[i] = ~hbuf[i];
        __syncthreads();
        if (wave == 0) {
          const uint32_t kept = cap_select(p, b, s, hbuf, D, lane);
          wave_sync();
          hbuf[lane] = kept;
          wave_sync();
          wave_sort_desc(hbuf, 64, lane);
          mykey = (uint32_t)hbuf[lane];
        }
        D = kMaxPix;
      } else if (wave == 0) {
        mykey = lane < D ? (uint32_t)hbuf[lane] : 0u;
      }
    } else {
      // Passes over global memory: each adds the not yet kept keys of the triangle holding the
      // smallest rank among entries whose key is not kept.
      if (tid == 0) {
        nsel = 0u;
        capdrop = 0;
      }
      __syncthreads();
      for (;;) {
        const uint32_t ns = nsel;
        uint32_t rmin = 0xffffffffu;
        for (uint32_t i = tid; i < c; i += kThreads) {
          uint32_t key, rank;
          slot_entry(t, G2, s, off, i, key, rank, fl);
          if (!in_sorted(sel, ns, key)) rmin = min(rmin, rank);
        }
        rmin = group_min(rmin, wcnt, lane, wave);
        if (rmin == 0xffffffffu) break;
        if (ns == (uint32_t)kMaxPix) {
          if (tid == 0) capdrop = 1;
          break;
        }
        const uint32_t T = rmin >> 1;
        if (tid == 0) lcount = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < c; i += kThreads) {
          uint32_t key, rank;
          slot_entry(t, G2, s, off, i, key, rank, fl);
          if ((rank >> 1) == T && !in_sorted(sel, ns, key)) {
            const uint32_t q = atomicAdd(&lcount, 1u);
            if (q < (uint32_t)kHugeList) hbuf[q] = ((uint64_t)key << 32) | (uint64_t)(~rank);
          }
        }
        __syncthreads();
        if (wave == 0) {
          uint32_t nl = lcount;
          if (nl > (uint32_t)kHugeList) {
            fl |= F_UNRESOLVED;  // one triangle with more than kHugeList rows in one slot
            nl = kHugeList;
          }
          const int u = wave_distinct_by_rank(hbuf, (int)nl, lane);
          const int room = kMaxPix - (int)ns;
          uint32_t add = 0u;
          if (u <= room) {
            add = lane < u ? (uint32_t)hbuf[lane] : 0u;
          } else {
            pick_first_of_triangle(p, b, s, T, hbuf, 0, u, room, 0, lane, add);
            if (lane == 0) capdrop = 1;
          }
          const int na = u < room ? u : room;
          if (lane < na) sel[ns + lane] = add;
          wave_sync();
          // keep sel ascending: one lane inserts the new keys (at most 63 in all)
          if (lane == 0) {
            for (int q = (int)ns; q < (int)ns + na; q++) {
              const uint32_t v = sel[q];
              int j = q - 1;
              while (j >= 0 && sel[j] > v) {
                sel[j + 1] = sel[j];
                j--;
              }
              sel[j + 1] = v;
            }
            nsel = ns + (uint32_t)na;
          }
          wave_sync();
        }
        __syncthreads();
        if (capdrop) break;
      }
      __syncthreads();
      const uint32_t ns = nsel;
      D = (int)ns;
      if (capdrop) fl |= F_CAP_DROP;
      if (wave == 0) mykey = lane < D ? sel[D - 1 - lane] : 0u;
    }
    fl = group_or(fl, fred, lane, wave);
    if (wave == 0) {
      uint64_t *const rows = p.dbg_rows ? p.dbg_rows + (int64_t)s * 64 : nullptr;
      SlotResult r;
      wave_pair(mykey, (D >= 64) ? ~0ull : ((1ull << D) - 1ull), lane, p.pp, r, rows);
      r.flags |= fl;
      if (lane == 0) {
        acc.nslot++;
        account(acc, r, s, p, track);
        t.bedinfo[s] = (uint16_t)bed_info(r, p.pp);
        if (sv) put_sv(sv, G2, s, r);
        if (final_here) {
          acc.nfinal += r.n;
          acc.chkfinal += (uint64_t)(s + 1) * (uint64_t)r.n;
          if (p.dbg_final) p.dbg_final[s] = r.n;
        }
      }
    }
    __syncthreads();
  }
  DevRecord &R = p.rec[b];
  int64_t v[18] = {acc.Va, acc.Vb, acc.Vo, acc.Vss, acc.Vnv, acc.Vtc, 0, acc.nent, acc.nslot,
                   acc.nfinal, (int64_t)acc.chkpre, (int64_t)acc.chkfinal, acc.aVo, acc.aVss,
                   acc.aVa, acc.aVb, acc.aVnv, acc.aVtc};
  group_add<18>(v, red, lane, wave, tid, (unsigned long long *)&R.Va);
  const int32_t f = group_or(acc.flags, fred, lane, wave);
  if (tid == 0 && f) atomicOr(&R.flags, f);
}

// ---- bed_kernel: band_rows grid rows of the bed scan region per work-group ----------------
// GenerateBed / IsBedCandidate (STomoSh_INT3.cpp:1331-1443) with the raster-order effect: a raft
// source slot that takes its own bed row stops being a source for slots scanned after it.
__device__ __forceinline__ void bed_body(const SweepParams &p, int32_t blk) {
  __shared__ uint16_t nb[kBedLdsSlots];  // bed bookkeeping (bed_info bits) of the band and halo
  // Source bits of the halo rows (one bit per column, filled by ballots), then their OR over
  // each column's half-open window [y - r, y + r): an exact "no source can be near" test.
  __shared__ uint32_t srcw[kBedWords], rowany[kBedWords], convw[kBedWords];
  // dtab[dx][m]: the least bed_dist[dx][j] over the window columns j set in m (1e5 if none)
  __shared__ float dtab[2 * kBedTabR][1 << (2 * kBedTabR)];
  __shared__ int64_t red[kBedWaves][6];
  __shared__ int32_t fred[kBedWaves];
  const int tid = threadIdx.x, lane = tid & 63, wave = wave_index();
  int32_t b, band;
  if (!xcd_map(blk, p.bed_bands, p.n_batch, b, band)) return;
  const int32_t G = p.grid;
  const OrientHdr &h = p.hdr[b];
  const int32_t bx0 = h.bbox[0], bx1 = h.bbox[1], by0 = h.bbox[2], by1 = h.bbox[3];
  if (bx1 < bx0) return;
  const int32_t br = p.pp.bed_r;
  const int32_t m = br + 1;
  const int32_t rx0 = max(bx0 - m, 0), rx1 = min(bx1 + m + 1, G);
  const int32_t ry0 = max(by0 - m, 0), ry1 = min(by1 + m + 1, G);
  const int32_t X0 = rx0 + band * p.band_rows;
  if (X0 >= rx1) return;
  const int32_t X1 = min(X0 + p.band_rows, rx1);
  const Tables t = tables_for(p, b);
  // neighbour flags of rows [X0 - r, X1 + r) x columns [ry0 - r, ry1 + r), clipped to the grid
  const int32_t hx0 = max(X0 - br, 0), hx1 = min(X1 + br, G);
  const int32_t hy0 = max(ry0 - br, 0), hy1 = min(ry1 + br, G);
  const int32_t hw = hy1 - hy0;
  const int32_t nr = hx1 - hx0;
  const int32_t nw = (hw + 31) >> 5;
  const bool lds = (int64_t)nr * hw <= kBedLdsSlots;
  const bool bits = lds && nr * nw <= kBedWords;
  const bool wtab = bits && br >= 1 && br <= kBedTabR;  // bit-window neighbour search
  if (wtab) {
    const int32_t wd = 2 * br, nm = 1 << wd;
    for (int32_t i = tid; i < wd * nm; i += kBedThreads) {
      const int32_t dx = i / nm, m = i - dx * nm;
      float md = 1e5f;
      for (int32_t j = 0; j < wd; j++)
        if ((m >> j) & 1) md = fminf(md, p.bed_dist[dx * wd + j]);
      dtab[dx][m] = md;
    }
  }
  if (lds) {
    // One halo row per wave step, its loads (up to 8 column chunks of 64) issued together.
    constexpr int kChunks = 8;
    for (int32_t r0 = wave; r0 < nr; r0 += kBedWaves) {
      const int32_t x = hx0 + r0;
      const bool xin = x >= bx0 && x <= bx1;
      for (int32_t c00 = 0; c00 < hw; c00 += 64 * kChunks) {
        uint32_t v[kChunks];
#pragma unroll
        for (int q = 0; q < kChunks; q++) {
          const int32_t yc = c00 + 64 * q + lane, y = hy0 + yc;
          v[q] = BI_POSS;  // outside the touched box: an empty slot
          if (yc < hw && xin && y >= by0 && y <= by1) v[q] = t.bedinfo[x * G + y];
        }
#pragma unroll
        for (int q = 0; q < kChunks; q++) {
          const int32_t c0 = c00 + 64 * q;
          if (c0 >= hw) break;
          const int32_t yc = c0 + lane;
          if (yc < hw) nb[r0 * hw + yc] = (uint16_t)v[q];
          const uint64_t sm = ballot(yc < hw && (v[q] & BI_SRC) != 0u);
          const uint64_t cm = ballot(yc < hw && (v[q] & BI_CONV) != 0u);
          if (bits && lane == 0) {
            const int32_t w0 = c0 >> 5;
            srcw[r0 * nw + w0] = (uint32_t)sm;
            convw[r0 * nw + w0] = (uint32_t)cm;
            if (w0 + 1 < nw) {
              srcw[r0 * nw + w0 + 1] = (uint32_t)(sm >> 32);
              convw[r0 * nw + w0 + 1] = (uint32_t)(cm >> 32);
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if (bits) {
    for (int32_t i = tid; i < nr * nw; i += kBedThreads) {
      const int32_t w0 = i % nw;
      const uint32_t cur = srcw[i];
      const uint32_t prev = w0 > 0 ? srcw[i - 1] : 0u;
      const uint32_t next = w0 + 1 < nw ? srcw[i + 1] : 0u;
      uint32_t o = cur;  // d = 0
      for (int32_t d = 1; d < br; d++) o |= (cur >> d) | (next << (32 - d));   // columns y + d
      for (int32_t e = 1; e <= br; e++) o |= (cur << e) | (prev >> (32 - e));  // columns y - e
      rowany[i] = o;
    }
    __syncthreads();
  }
  const int32_t rwy = ry1 - ry0;
  int64_t nbed = 0, nfinal = 0;
  uint64_t chk = 0;
  int32_t flags = 0;
  for (int32_t xr = wave; xr < X1 - X0; xr += kBedWaves)
  for (int32_t c0 = 0; c0 < rwy; c0 += 64) {
    if (c0 + lane >= rwy) continue;
    const int32_t X = X0 + xr, Y = ry0 + c0 + lane;
    const int32_t P = X * G + Y;
    const bool inb = X >= bx0 && X <= bx1 && Y >= by0 && Y <= by1;
    const uint32_t bi = lds ? (uint32_t)nb[(X - hx0) * hw + (Y - hy0)] : (inb ? t.bedinfo[P] : BI_POSS);
    const int32_t nrow = (int32_t)(bi & 127u);
    int32_t add = 0;
    bool near = true;
    if ((bi & BI_POSS) && bits) {
      const int32_t xa = max(X - br, 0), xb = min(X + br, G);
      const int32_t yy = Y - hy0;
      uint32_t any = 0u;
      for (int32_t sx = xa; sx < xb; sx++) any |= rowany[(sx - hx0) * nw + (yy >> 5)];
      near = ((any >> (yy & 31)) & 1u) != 0u;
    }
    if ((bi & BI_POSS) && near) {
      float md = 1e5f;
      const int32_t xa = max(X - br, 0), xb = min(X + br, G);
      const int32_t ya = max(Y - br, 0), yb = min(Y + br, G);
      if (wtab) {
        // The window's source bits per row (columns Y - r .. Y + r - 1, bit j <-> column
        // Y - r + j), minus the converted sources scanned before (Y, X), looked up in dtab.
        const int32_t wd = 2 * br;
        const uint32_t wmask = (1u << wd) - 1u, lowm = (1u << br) - 1u;
        const int32_t idx = Y - br - hy0;  // bit position of the window's first column
        const int32_t w0 = idx >> 5, sh = idx & 31;
        for (int32_t sx = xa; sx < xb; sx++) {
          const int32_t row = (sx - hx0) * nw;
          const uint32_t s0 = (w0 >= 0 && w0 < nw) ? srcw[row + w0] : 0u;
          const uint32_t s1 = (w0 + 1 >= 0 && w0 + 1 < nw) ? srcw[row + w0 + 1] : 0u;
          const uint32_t c0w = (w0 >= 0 && w0 < nw) ? convw[row + w0] : 0u;
          const uint32_t c1w = (w0 + 1 >= 0 && w0 + 1 < nw) ? convw[row + w0 + 1] : 0u;
          uint32_t sb = (uint32_t)((((uint64_t)s1 << 32) | s0) >> sh) & wmask;
          const uint32_t cb = (uint32_t)((((uint64_t)c1w << 32) | c0w) >> sh) & wmask;
          if (sx < X) sb &= ~cb;
          else if (sx == X) sb &= ~(cb & lowm);
          md = fminf(md, dtab[sx - X + br][sb]);
        }
      } else
      for (int32_t sx = xa; sx < xb; sx++) {
        for (int32_t sy = ya; sy < yb; sy++) {
          uint32_t bs = 0u;
          if (lds)
            bs = nb[(sx - hx0) * hw + (sy - hy0)];
          else if (sx >= bx0 && sx <= bx1 && sy >= by0 && sy <= by1)
            bs = t.bedinfo[sx * G + sy];
          const uint32_t f = ((bs & BI_SRC) ? 1u : 0u) | ((bs & BI_CONV) ? 2u : 0u);
          if (!(f & 1u)) continue;
          const bool before = (sx < X) || (sx == X && sy < Y);
          if ((f & 2u) && before) continue;  // already covered by its own bed row
          md = fminf(md, p.bed_dist[(sx - X + br) * (2 * br) + (sy - Y + br)]);
        }
      }
      const bool bed = !(md > p.bed_out) &&
                       (p.pp.bed_type == 1 || p.pp.bed_type == 2 || (p.pp.bed_type == 3 && md > p.bed_in));
      if (bed) {
        const uint32_t k10 = make_key(1, 0);
        if (bi & BI_HAS10) {
          nbed++;
          if (p.dbg_rows) {  // the bed bit merges into the existing (1, 0) row
            uint64_t *rw = p.dbg_rows + (int64_t)P * 64;
            for (int q = 0; q < nrow; q++)
              if ((uint32_t)rw[q] == k10) rw[q] |= (uint64_t)T_BED << 32;
          }
        } else if (nrow < kMaxPix) {
          nbed++;
          add = 1;
          if (p.dbg_rows) p.dbg_rows[(int64_t)P * 64 + nrow] = make_row(T_BED, k10);
        } else {
          flags |= F_CAP_DROP;
        }
      }
    }
    const int32_t nf = nrow + add;
    nfinal += nf;
    chk += (uint64_t)(P + 1) * (uint64_t)nf;
    if (p.dbg_final) p.dbg_final[P] = nf;
  }
  DevRecord &R = p.rec[b];
  // Vbed, n_entries, n_slots, n_final, chk_pre, chk_final
  int64_t v[6] = {nbed, 0, 0, nfinal, 0, (int64_t)chk};
  group_add<6, kBedWaves>(v, red, lane, wave, tid, (unsigned long long *)&R.Vbed);
  const int32_t f = group_or<kBedWaves>(flags, fred, lane, wave);
  if (tid == 0 && f) atomicOr(&R.flags, f);
}

// ---- fsum_kernel: one 64-lane work-group per orientation ----------------------------------
// The reference accumulates per-slot integer totals into float running sums in ascending slot
// order: Vo and Vss in Pairing (STomoSh_INT3.cpp:386-392), Va, Vb, Vnv and Vtc in Calculate
// (:1233-1240).  A running sum is exact, and so equal to the integer prefix sum, as long as the
// sum of |slot value| so far stays within 2^24.  A total whose sum of |slot value| stays within
// 2^24 is therefore float(total).  For the others, a first pass sums whole blocks of 512 slots
// exactly in integers, lane-parallel, while the bound holds for all of them, and the serial
// replay (lane q owns total q) starts at the first block where it fails for one, from the exact
// sums before it.
__device__ __forceinline__ void fsum_body(const SweepParams &p, int32_t blk) {
  constexpr int kAhead = 8;                 // chunks of 64 slots in flight per lane
  constexpr int kBlock = 8;                 // chunks per block of the exact first pass
  __shared__ int32_t vals[6][65];           // padded: lanes 0..5 read one row each
  const int lane = threadIdx.x;
  const int32_t b = blk;
  if (b >= p.n_batch) return;
  const OrientHdr &h = p.hdr[b];
  DevRecord &R = p.rec[b];
  const int64_t lim = (int64_t)1 << 24;
  if (!totals_may_round(h, p.pp)) return;
  const int64_t tot[6] = {R.Vo, R.Vss, R.Va, R.Vb, R.Vnv, R.Vtc};
  const int64_t absv[6] = {R.aVo, R.aVss, R.aVa, R.aVb, R.aVnv, R.aVtc};
  if (!(absv[0] > lim || absv[1] > lim || absv[2] > lim || absv[3] > lim || absv[4] > lim || absv[5] > lim)) return;
  const Tables t = tables_for(p, b);
  const int32_t G = p.grid, G2 = G * G;
  const int32_t bx0 = h.bbox[0], bx1 = h.bbox[1], by0 = h.bbox[2], by1 = h.bbox[3];
  const int32_t wy = by1 - by0 + 1;
  const int32_t nbb = bx1 >= bx0 ? (bx1 - bx0 + 1) * wy : 0;
  const int32_t nchunk = (nbb + 63) >> 6;
  // Totals whose sum of |value| passes 2^24 need the replay; only their per-slot values are read.
  bool need[6];
#pragma unroll
  for (int k = 0; k < 6; k++) need[k] = absv[k] > lim;
  // Slots in ascending id order: row-major over the bounding box (id = x * G + y).
  auto load = [&](int32_t c, int32_t (&v)[6]) {
    const int32_t i = (c << 6) + lane;
    const bool in = c < nchunk && i < nbb;
    const int32_t s = in ? (bx0 + i / wy) * G + by0 + i % wy : 0;
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = (in && need[k]) ? t.sv[(int64_t)k * G2 + s] : 0;
  };
  // First pass (wave-uniform): exact integer sums S[k] of whole blocks while, for every total
  // that needs the replay, the sum of |value| so far stays within 2^24 (then every partial sum
  // before the block is exact); cmin is the first block where that fails for some total.
  int64_t S[6], A[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    S[k] = 0;
    A[k] = 0;
  }
  int32_t cmin = nchunk;
  for (int32_t c0 = 0; c0 < nchunk; c0 += kBlock) {
    int64_t ba[6] = {0, 0, 0, 0, 0, 0}, bs[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < kBlock; j++) {
      int32_t v[6];
      load(c0 + j, v);
#pragma unroll
      for (int k = 0; k < 6; k++) {
        ba[k] += v[k] < 0 ? -(int64_t)v[k] : (int64_t)v[k];
        bs[k] += v[k];
      }
    }
    bool fits = true;
    int64_t BS[6], BA[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
      BS[k] = BA[k] = 0;
      if (!need[k]) continue;
      BS[k] = wave_sum(bs[k]);
      BA[k] = wave_sum(ba[k]);
      fits = fits && A[k] + BA[k] <= lim;
    }
    if (!fits) {
      cmin = c0;
      break;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
      A[k] += BA[k];
      S[k] += BS[k];
    }
  }
  int32_t mystart = nchunk;
  int64_t myS = 0, mytot = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (lane == k) {
      mystart = need[k] ? cmin : nchunk;
      myS = S[k];
      mytot = tot[k];
    }
  }
  // |myS| <= A <= 2^24: the float running sum before chunk mystart is exactly myS
  float acc = mystart < nchunk ? (float)myS : (float)mytot;
  // Second pass: the serial replay from chunk cmin, each lane adding from its own start.
  const int q = lane < 6 ? lane : 0;
  int32_t buf[kAhead][6];
#pragma unroll
  for (int j = 0; j < kAhead; j++) load(cmin + j, buf[j]);
  for (int32_t c0 = cmin; c0 < nchunk; c0 += kAhead) {
#pragma unroll
    for (int j = 0; j < kAhead; j++) {
      const int32_t c = c0 + j;
      if (c < nchunk) {
#pragma unroll
        for (int k = 0; k < 6; k++) vals[k][lane] = buf[j][k];
        __syncthreads();
        load(c + kAhead, buf[j]);
        if (lane < 6 && c >= mystart) {
          const int32_t n = min(64, nbb - (c << 6));
          if (n == 64) {
            // a whole chunk: the 64 values are read and converted first, so the LDS reads are
            // in flight together and only the additions form the serial chain
            float f[64];
#pragma unroll
            for (int i = 0; i < 64; i++) f[i] = (float)vals[q][i];
#pragma unroll
            for (int i = 0; i < 64; i++) acc = acc + f[i];
          } else {
            for (int32_t i = 0; i < n; i++) acc = acc + (float)vals[q][i];
          }
        }
        __syncthreads();
      }
    }
  }
  if (lane < 6) R.fsum[lane] = acc;
  if (lane == 0) R.fsum_done = 1;
}

// ---- kernel entry points ------------------------------------------------------------------
// One sweep slice per launch: work-group i serves block i.
__global__ __launch_bounds__(kThreads) void aabb_kernel(SweepParams p) { aabb_body(p, (int32_t)blockIdx.x); }
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(6))) void raster_kernel(SweepParams p) {
  raster_body(p, (int32_t)blockIdx.x);
}
__global__ __launch_bounds__(kThreads) void compact_kernel(SweepParams p) { compact_body(p, (int32_t)blockIdx.x); }
// Six waves per SIMD (80 VGPRs): the few registers spilled are the accumulators' epilogue, after
// the step loop; measured 4.5% faster than five waves without spills.
__global__ __launch_bounds__(kThreads, 6) void pair_kernel(SweepParams p) { pair_body<0, false>(p, (int32_t)blockIdx.x); }
// Four waves per SIMD (128 VGPRs, a few spilled) beat three without spills: the kernel waits on
// memory far more than it computes.
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void pair_big_kernel(SweepParams p) {
  pair_body<1, false>(p, (int32_t)blockIdx.x);
}
// The same two with the debug dumps, launched only when one was asked for.
__global__ __launch_bounds__(kThreads, 6) void pair_kernel_dbg(SweepParams p) { pair_body<0, true>(p, (int32_t)blockIdx.x); }
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void pair_big_kernel_dbg(SweepParams p) {
  pair_body<1, true>(p, (int32_t)blockIdx.x);
}
__global__ __launch_bounds__(kThreads) void pair_huge_kernel(SweepParams p) { pair_huge_body(p, (int32_t)blockIdx.x); }
__global__ __launch_bounds__(kBedThreads) void bed_kernel(SweepParams p) { bed_body(p, (int32_t)blockIdx.x); }
__global__ __launch_bounds__(64) void fsum_kernel(SweepParams p) { fsum_body(p, (int32_t)blockIdx.x); }

// Persistent form: each work-group takes items from its XCD's queue (thread 0 counts them off
// with one atomic, the unit is found by binary search over first[]) until the queue is empty.
// Every work-group leaves the loop when its queue runs out; no item waits on another.  The unit
// table comes in as a read-only, unaliased kernel argument, so its fields are read with scalar
// loads wherever the body needs them, as from a kernel's own parameter block.
#define TOMO_QUEUE_ARGS                                                                              \
  const SweepParams *__restrict__ units, const int32_t *__restrict__ first, int32_t n_units, \
      uint32_t *__restrict__ heads
template <typename Body>
__device__ __forceinline__ void drain(const SweepParams *__restrict__ units, const int32_t *__restrict__ first,
                                      int32_t n_units, uint32_t *__restrict__ heads, Body body) {
  __shared__ int32_t s_item, s_unit;
  const int32_t x = (int32_t)blockIdx.x & 7;
  const int32_t items = first[n_units];
  for (;;) {
    __syncthreads();  // the previous item is done in every thread (its LDS and s_item are free)
    if (threadIdx.x == 0) {
      const int64_t g = (int64_t)x + 8 * (int64_t)atomicAdd(&heads[x], 1u);
      int32_t u = -1;
      if (g < items) {  // the last unit whose first item is at most g (empty units are skipped)
        int32_t lo = 0, hi = n_units - 1;
        while (lo < hi) {
          const int32_t mid = (lo + hi + 1) >> 1;
          if (first[mid] <= g) lo = mid;
          else hi = mid - 1;
        }
        u = lo;
      }
      s_item = u >= 0 ? (int32_t)g : -1;
      s_unit = u;
    }
    __syncthreads();
    const int32_t g = __builtin_amdgcn_readfirstlane(s_item), u = __builtin_amdgcn_readfirstlane(s_unit);
    if (g < 0) break;
    // The table stays unchanged while the kernel runs: read it through the constant address
    // space, as the parameter block is, so its fields load into scalar registers after barriers.
    const __attribute__((address_space(4))) SweepParams *cu =
        (const __attribute__((address_space(4))) SweepParams *)units;
    body(*(const SweepParams *)&cu[u], g - first[u]);
  }
}

__global__ __launch_bounds__(kThreads) void aabb_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { aabb_body(p, blk); });
}
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(6))) void raster_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { raster_body(p, blk); });
}
__global__ __launch_bounds__(kThreads) void compact_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { compact_body(p, blk); });
}
__global__ __launch_bounds__(kThreads, 5) void pair_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { pair_body<0, false>(p, blk); });
}
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void pair_big_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { pair_body<1, false>(p, blk); });
}
__global__ __launch_bounds__(kThreads) void pair_huge_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { pair_huge_body(p, blk); });
}
__global__ __launch_bounds__(kBedThreads) void bed_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { bed_body(p, blk); });
}
__global__ __launch_bounds__(64) void fsum_queue(TOMO_QUEUE_ARGS) {
  drain(units, first, n_units, heads, [](const SweepParams &p, int32_t blk) { fsum_body(p, blk); });
}

// Zeroes n queue counters (a kernel, not a memset node, so a captured graph replays it exactly).
__global__ void clear_heads_kernel(uint32_t *heads, int32_t n) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) heads[i] = 0u;
}

// Device address of `scratch` on `device`, which must be the current device the first time it
// is asked for: looked up once per device with hipGetSymbolAddress and kept (every GPU holds its
// own copy of the array at its own address).
hipError_t static_scratch_ptr(int device, uint8_t **out) {
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *out = (uint8_t *)addr[device];
  return hipSuccess;
}

static const void *queue_kernel(int k) {
  switch (k) {
    case QK_AABB: return (const void *)aabb_queue;
    case QK_RASTER: return (const void *)raster_queue;
    case QK_COMPACT: return (const void *)compact_queue;
    case QK_PAIR: return (const void *)pair_queue;
    case QK_BIG: return (const void *)pair_big_queue;
    case QK_HUGE: return (const void *)pair_huge_queue;
    case QK_BED: return (const void *)bed_queue;
    case QK_FSUM: return (const void *)fsum_queue;
    default: return nullptr;
  }
}

int queue_threads(int k) {
  return k == QK_RASTER ? kTileThreads : (k == QK_FSUM ? 64 : (k == QK_BED ? kBedThreads : kThreads));
}

// Items of unit p for queue kernel k (xcd_map pads the orientation count to whole XCD rounds).
int32_t queue_items(int k, const SweepParams &p) {
  const int32_t nb8 = (p.n_batch + 7) / 8 * 8;
  switch (k) {
    case QK_AABB: case QK_COMPACT: case QK_HUGE: return nb8;
    case QK_RASTER: return nb8 * p.tiles;
    case QK_PAIR: return nb8 * p.pair_tiles;
    case QK_BIG: return nb8 * p.big_groups;
    case QK_BED: return p.bed_on ? nb8 * p.bed_bands : 0;
    case QK_FSUM: return p.exact_fsum ? nb8 : 0;
    default: return 0;
  }
}

// Work-groups of queue kernel k resident on one CU at once.
hipError_t queue_occupancy(int k, int *per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, queue_kernel(k), queue_threads(k), 0);
}

hipError_t launch_queue(int k, const UnitQueue &q, int32_t groups, hipStream_t stream) {
  const void *f = queue_kernel(k);
  if (!f || groups < 8 || groups % 8 != 0) return hipErrorInvalidValue;
  const SweepParams *units = q.units;
  const int32_t *first = q.first;
  int32_t n_units = q.n_units;
  uint32_t *heads = q.heads;
  void *args[] = {(void *)&units, (void *)&first, (void *)&n_units, (void *)&heads};
  return hipLaunchKernel(f, dim3((unsigned)groups), dim3((unsigned)queue_threads(k)), args, 0, stream);
}

hipError_t launch_clear_heads(uint32_t *heads, int32_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(clear_heads_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, heads, n);
  return hipGetLastError();
}

// Stage k of one sweep slice: 0 aabb, 1 raster, 2 compact, 3 pair (small, large and huge
// slots), 4 bed (skipped without a bed), 5 ordered float totals (skipped unless kept).
hipError_t launch_stage(int stage, const SweepParams &p, hipStream_t stream, hipStream_t side, hipEvent_t fork,
                        hipEvent_t join) {
  const int32_t nb = p.n_batch;
  const int32_t nb8 = (nb + 7) / 8 * 8;  // xcd_map pads the orientation count to whole XCD rounds
  switch (stage) {
    case 0: hipLaunchKernelGGL(aabb_kernel, dim3(nb), dim3(kThreads), 0, stream, p); break;
    case 1: hipLaunchKernelGGL(raster_kernel, dim3(nb8 * p.tiles), dim3(kTileThreads), 0, stream, p); break;
    case 2: hipLaunchKernelGGL(compact_kernel, dim3(nb), dim3(kThreads), 0, stream, p); break;
    case 3: {
      const bool dbg = p.dbg_pre || p.dbg_final || p.dbg_rows;
      // The small and the large slots are disjoint sets and both kernels add their totals
      // atomically: with a side stream the small-slot pairing runs beside the large-slot one,
      // joined before pair_huge_kernel.  The large-slot kernel goes on the main stream, so its
      // work-groups (few, long and latency-bound) are dispatched first and do not trail behind
      // the small-slot kernel's: that order took 0.5 ms off a 3.7 ms pairing stage.
      hipStream_t ss = stream;
      if (side) {
        hipError_t e = hipEventRecord(fork, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
        if (e != hipSuccess) return e;
        ss = side;
      }
      hipLaunchKernelGGL(dbg ? pair_big_kernel_dbg : pair_big_kernel, dim3(nb8 * p.big_groups), dim3(kThreads), 0,
                         stream, p);
      hipLaunchKernelGGL(dbg ? pair_kernel_dbg : pair_kernel, dim3(nb8 * p.pair_tiles), dim3(kThreads), 0, ss, p);
      if (side) {
        hipError_t e = hipEventRecord(join, side);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, join, 0);
        if (e != hipSuccess) return e;
      }
      hipLaunchKernelGGL(pair_huge_kernel, dim3(nb), dim3(kThreads), 0, stream, p);
      break;
    }
    case 4:
      if (p.bed_on) hipLaunchKernelGGL(bed_kernel, dim3(nb8 * p.bed_bands), dim3(kBedThreads), 0, stream, p);
      break;
    case 5:
      if (p.exact_fsum) hipLaunchKernelGGL(fsum_kernel, dim3(nb), dim3(64), 0, stream, p);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tomo
