// tomo_slot.h -- per-slot sort / pair / shadow logic shared by the HIP kernels and the host
// unit test (compiled for both: TOMO_HD expands to __host__ __device__ under hipcc).
//
// A slot entry is a packed 32-bit key: ((z + 32768) << 16) | (nZ + 32768) with z and nZ the
// int16 values SetBit stores (STomoVoxelSpaceInfo.cpp:179-180).  Descending key order is the
// order sortSlotByZ produces: z descending, ties by nZ descending (STomoSh_INT3.cpp:398-446).
//
// Semantics restated from the brief vslotPair (STomoSh_INT3.cpp:1041-1062):
//   splitAlBe (:493-513)            alpha iff nZ >= 1, beta iff nZ <= -1
//   matchAlBePairBriefly (:583-611) an alpha is kept iff the previous alpha/beta row is a beta
//                                   (or none); a beta is kept iff the previous one is an alpha
//   createVoPixels (:934-960)       Vo = max(sum z(alpha) - sum z(beta), 0)
//   createShadowBriefly (:749-812)  SSB/NVB open on kept betas below/above the theta cut,
//                                   SSA/NVA close on the next kept alpha; open shadows get a
//                                   z = 0 floor row while the slot has fewer than 63 rows
//   createVss_Explicit (:962-987)   Vss = sum z(SSB) - sum z(SSA)
//   createTCPixels + createVss_Implicit (:734-747, :989-1025)
//   Calculate / sumType (:1211-1255, :679-694) per-slot Va, Vb, Vnv, Vtc.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define TOMO_HD __host__ __device__ __forceinline__
#else
#define TOMO_HD inline
#endif

namespace tomo {

enum : int32_t {
  T_AL = 1, T_BE = 2, T_SSB = 4, T_SSA = 8, T_SS = 16, T_BED = 32, T_VO = 64, T_VSS = 128,
  T_TC = 256, T_NVB = 512, T_NVA = 1024
};

// Per-orientation flags (the public TOMO_F_* values).  F_RETRY is internal: the overflow pool
// or the spill list of the orientation's slot table filled up, and the host sweeps the
// orientation again with larger ones before it reports the record.
enum : int32_t {
  F_FP32_INEXACT = 1, F_INT16_WRAP = 2, F_FOOTPRINT_OOB = 4, F_CAP_DROP = 8,
  F_SENTINEL_DROP = 16, F_UNRESOLVED = 64, F_RETRY = 128
};

constexpr int kMaxPix = 63;  // pixel rows per slot (nSlotCapacityHeight - 1)

TOMO_HD uint32_t make_key(int32_t z16, int32_t nz16) {
  return ((uint32_t)(z16 + 32768) << 16) | (uint32_t)(nz16 + 32768);
}
TOMO_HD int32_t key_z(uint32_t k) { return (int32_t)(k >> 16) - 32768; }
TOMO_HD int32_t key_nz(uint32_t k) { return (int32_t)(k & 0xffffu) - 32768; }

struct PairParams {
  int32_t cut_ssb;   // beta rows with nZ <= cut_ssb are support-structure casters
  int32_t cut_nvb;   // beta rows with nZ >= cut_nvb are non-vertical casters
  int32_t explicit_ss;
  int32_t bed_type;  // 0 none, 1 brim, 2 raft, 3 skirt
  int32_t bed_r;     // int(BedOuterBound)
};

// Per-slot result.
struct SlotResult {
  int32_t vo, vss, va, vb, vnv, vtc;  // pairing and Calculate sums for this slot
  int32_t npre;      // distinct rows after Pixelize
  int32_t n;         // rows after Pairing (distinct + floor rows)
  int32_t sabs;      // sum |z| over those rows
  int32_t last_z, last_type;
  int32_t has10;     // a (z = 1, nZ = 0) row exists (a bed row would merge into it)
  int32_t flags;
};

// Bed bookkeeping packed per slot: n (7 bits) | has10 | src | possible | conv.
enum : uint32_t { BI_HAS10 = 1u << 7, BI_SRC = 1u << 8, BI_POSS = 1u << 9, BI_CONV = 1u << 10 };

// GenerateBed / IsBedCandidate slot tests (STomoSh_INT3.cpp:1331-1373, 1408-1428):
//  src      : the bottom row is at z = 0 and carries an alpha/beta/SSA/SS bit;
//  possible : empty slot, or bottom row at z = 0 that is exactly NVA/Vo/Vss, or (raft) a
//             beta/SSA bottom;
//  conv     : a raft source that will take a bed row on its own turn and so stops being a
//             source for every slot scanned after it.
// Written with bitwise operators and no branches (slot pairing calls it once per slot, lane by
// lane, and a chain of ifs there compiles to a branch tree); the truth table is the one above.
TOMO_HD uint32_t bed_info(const SlotResult &r, const PairParams &pp) {
  const int32_t lt = r.last_type;
  const int32_t solid = T_AL | T_BE | T_SSA | T_SS;
  const bool n0 = r.n == 0, lz = r.last_z == 0;
  const bool lsolid = (lt & solid) != 0, lbs = (lt & (T_BE | T_SSA)) != 0, bt2 = pp.bed_type == 2;
  const bool open = (lt == T_NVA) | (lt == T_VO) | (lt == T_VSS);
  const bool src = !n0 & lz & lsolid;
  // possible: an empty slot; a bottom row at z 0 that is open (NVA/Vo/Vss, no solid bit) or, on
  // a raft, a beta/SSA row; nothing else
  const bool poss = n0 | (lz & ((open & !lsolid) | (bt2 & lbs)));
  const bool conv = src & bt2 & lbs & (pp.bed_r >= 1) & (r.n < kMaxPix) & !r.has10;
  return (uint32_t)r.n | (r.has10 ? BI_HAS10 : 0u) | (src ? BI_SRC : 0u) | (poss ? BI_POSS : 0u) |
         (conv ? BI_CONV : 0u);
}

// A final slot row for the render export: (type << 32) | key.
TOMO_HD uint64_t make_row(int32_t type, uint32_t key) { return ((uint64_t)(uint32_t)type << 32) | key; }

// Pair N register-resident keys, the first c valid (c <= N <= 63), any order, duplicates
// allowed.  Fully unrolled so the keys stay in registers.  rows (optional, 64 entries) receives
// the slot's final rows in slot order, floor rows included.
template <int N>
TOMO_HD void pair_keys(uint32_t (&k)[N], int c, const PairParams &pp, SlotResult &r,
                       uint64_t *rows = nullptr) {
  // Sort descending: optimal networks for N = 2 (1 comparator), 4 (5) and 8 (19), a bitonic
  // network for N = 16 (80); other N an odd-even transposition network (the host unit test's
  // N = 63).
  if constexpr (N == 2) {
    const uint32_t hi = k[0] > k[1] ? k[0] : k[1];
    const uint32_t lo = k[0] > k[1] ? k[1] : k[0];
    k[0] = hi;
    k[1] = lo;
  } else if constexpr (N == 4) {
#define TOMO_CX4(a, b)                                  \
  {                                                     \
    uint32_t hi = k[a] > k[b] ? k[a] : k[b];            \
    uint32_t lo = k[a] > k[b] ? k[b] : k[a];            \
    k[a] = hi;                                          \
    k[b] = lo;                                          \
  }
    TOMO_CX4(0, 1) TOMO_CX4(2, 3) TOMO_CX4(0, 2) TOMO_CX4(1, 3) TOMO_CX4(1, 2)
#undef TOMO_CX4
  } else if constexpr (N == 16) {
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
        for (int i = 0; i < N; i++) {
          const int j = i ^ stride;
          if (j > i) {
            // runs of length `size` alternate direction; the final merge is descending
            const bool desc = (i & size) == 0;
            const uint32_t a = k[i], b = k[j];
            const bool sw = desc ? (a < b) : (a > b);
            k[i] = sw ? b : a;
            k[j] = sw ? a : b;
          }
        }
      }
    }
  } else if constexpr (N == 8) {
#define TOMO_CX(a, b)                                   \
  {                                                     \
    uint32_t hi = k[a] > k[b] ? k[a] : k[b];            \
    uint32_t lo = k[a] > k[b] ? k[b] : k[a];            \
    k[a] = hi;                                          \
    k[b] = lo;                                          \
  }
    TOMO_CX(0, 2) TOMO_CX(1, 3) TOMO_CX(4, 6) TOMO_CX(5, 7)
    TOMO_CX(0, 4) TOMO_CX(1, 5) TOMO_CX(2, 6) TOMO_CX(3, 7)
    TOMO_CX(0, 1) TOMO_CX(2, 3) TOMO_CX(4, 5) TOMO_CX(6, 7)
    TOMO_CX(2, 4) TOMO_CX(3, 5)
    TOMO_CX(1, 4) TOMO_CX(3, 6)
    TOMO_CX(1, 2) TOMO_CX(3, 4) TOMO_CX(5, 6)
#undef TOMO_CX
  } else {
#pragma unroll
    for (int pass = 0; pass < N; pass++) {
#pragma unroll
      for (int i = pass & 1; i + 1 < N; i += 2) {
        uint32_t hi = k[i] > k[i + 1] ? k[i] : k[i + 1];
        uint32_t lo = k[i] > k[i + 1] ? k[i + 1] : k[i];
        k[i] = hi;
        k[i + 1] = lo;
      }
    }
  }
  if constexpr (N <= 4) {
    // The rows in descending order, written with selects only (no divergent branches): a row
    // is distinct (d) when it is valid and differs from the one before.
    const uint32_t k10 = make_key(1, 0);
    int npre = 0;
    int32_t sabs = 0;
    bool has10 = false;
    bool started = false, ex = false, im = false, tc_done = false;
    int32_t al = 0, be = 0, ssb = 0, ssa = 0, nvb = 0, nva = 0, tc = 0;
    int32_t last_z = 0, last_type = 0;
    const bool imp = !pp.explicit_ss;
  #pragma unroll
    for (int i = 0; i < N; i++) {
      const bool d = (i < c) && (i == 0 || k[i] != k[i - 1]);
      const int32_t z = key_z(k[i]), nz = key_nz(k[i]);
      // an alpha is kept when no kept alpha is open, a beta when one is
      const bool kA = d && !started && nz >= 1;
      const bool kB = d && started && nz <= -1;
      started = (started || kA) && !kB;
      // a kept beta opens SSB / NVB below / above the cuts; the next kept alpha closes them
      const bool sb = kB && !ex && nz <= pp.cut_ssb, nb = kB && !im && nz >= pp.cut_nvb;
      const bool sa = kA && ex, na = kA && im;
      const bool tcb = kA && imp && !tc_done;  // the first kept alpha is the TC row
      ex = (ex || sb) && !sa;
      im = (im || nb) && !na;
      tc_done = tc_done || tcb;
      al += kA ? z : 0;
      be += kB ? z : 0;
      ssb += sb ? z : 0;
      nvb += nb ? z : 0;
      ssa += sa ? z : 0;
      nva += na ? z : 0;
      tc += tcb ? z : 0;
      const int32_t ty = (kA ? T_AL : 0) | (kB ? T_BE : 0) | (sb ? T_SSB : 0) | (nb ? T_NVB : 0) |
                         (sa ? T_SSA : 0) | (na ? T_NVA : 0) | (tcb ? T_TC : 0);
      if (rows) {
        if (d) rows[npre] = make_row(ty, k[i]);
      }
      npre += d ? 1 : 0;
      sabs += d ? (z < 0 ? -z : z) : 0;
      has10 = has10 || (d && k[i] == k10);
      last_z = d ? z : last_z;
      last_type = d ? ty : last_type;
    }
    r.npre = npre;
    r.sabs = sabs;
    r.has10 = has10 ? 1 : 0;
    // Pairing skips slots with a single pixel (STomoSh_INT3.cpp:386): no types, no sums.
    const bool paired = npre >= 2;
    ex = ex && paired;
    im = im && paired;
    // open shadows get a z = 0 floor row (SSA, then NVA) while the slot has fewer than 63 rows
    const bool fa = ex && npre < kMaxPix, fb = im && npre + (fa ? 1 : 0) < kMaxPix;
    if (rows) {
      if (!paired && npre) rows[0] = make_row(0, k[0]);
      if (fa) rows[npre] = make_row(T_SSA, make_key(0, 0));
      if (fb) rows[npre + (fa ? 1 : 0)] = make_row(T_NVA, make_key(0, 0));
    }
    r.flags = ((ex && !fa) || (im && !fb)) ? F_SENTINEL_DROP : 0;
    r.n = npre + (fa ? 1 : 0) + (fb ? 1 : 0);
    r.last_z = (fa || fb) ? 0 : (npre ? last_z : 0);
    r.last_type = fb ? T_NVA : (fa ? T_SSA : (paired ? last_type : 0));
    al = paired ? al : 0;
    be = paired ? be : 0;
    r.vo = al > be ? al - be : 0;
    r.va = al;
    r.vb = be;
    r.vnv = paired ? nvb - nva : 0;
    r.vss = !paired ? 0 : (pp.explicit_ss ? ssb - ssa : -al + be + tc - nvb + nva);
    r.vtc = (paired && imp) ? tc : 0;

  } else {
  // Larger networks keep the branches: the selects of the form above hold every row's
  // conditions as lane masks at once, and at 8 rows and more they spill.
    const uint32_t k10 = make_key(1, 0);
    int npre = 0;
    int32_t sabs = 0, has10 = 0;
  #pragma unroll
    for (int i = 0; i < N; i++) {
      bool d = (i < c) && (i == 0 || k[i] != k[i - 1]);
      if (d) {
        npre++;
        int32_t z = key_z(k[i]);
        sabs += z < 0 ? -z : z;
        if (k[i] == k10) has10 = 1;
      }
    }
    r.npre = npre;
    r.sabs = sabs;
    r.has10 = has10;
    r.flags = 0;
    r.vo = r.vss = r.va = r.vb = r.vnv = r.vtc = 0;
    if (npre <= 1) {  // Pairing skips slots with a single pixel (STomoSh_INT3.cpp:386)
      r.n = npre;
      r.last_z = npre ? key_z(k[0]) : 0;
      r.last_type = 0;
      if (rows && npre) rows[0] = make_row(0, k[0]);
      return;
    }
    bool started = false, ex = false, im = false, tc_done = false;
    int32_t al = 0, be = 0, ssb = 0, ssa = 0, nvb = 0, nva = 0, tc = 0;
    int32_t last_z = 0, last_type = 0;
    int row = 0;
  #pragma unroll
    for (int i = 0; i < N; i++) {
      bool d = (i < c) && (i == 0 || k[i] != k[i - 1]);
      if (d) {
        int32_t z = key_z(k[i]), nz = key_nz(k[i]);
        bool isA = nz >= 1, isB = nz <= -1;
        bool kA = false, kB = false;
        if (!started && isA) { kA = true; started = true; }
        else if (started && isB) { kB = true; started = false; }
        int32_t ty = kA ? T_AL : (kB ? T_BE : 0);
        if (kA) al += z;
        if (kB) be += z;
        if (kB) {
          if (!ex && nz <= pp.cut_ssb) { ty |= T_SSB; ex = true; ssb += z; }
          if (!im && nz >= pp.cut_nvb) { ty |= T_NVB; im = true; nvb += z; }
        } else if (kA) {
          if (ex) { ty |= T_SSA; ex = false; ssa += z; }
          if (im) { ty |= T_NVA; im = false; nva += z; }
          if (!pp.explicit_ss && !tc_done) { ty |= T_TC; tc_done = true; tc += z; }
        }
        last_z = z;
        last_type = ty;
        if (rows) rows[row] = make_row(ty, k[i]);
        row++;
      }
    }
    int n = npre;
    if (ex) {
      if (n < kMaxPix) {
        if (rows) rows[n] = make_row(T_SSA, make_key(0, 0));
        n++; last_z = 0; last_type = T_SSA;
      } else r.flags |= F_SENTINEL_DROP;
    }
    if (im) {
      if (n < kMaxPix) {
        if (rows) rows[n] = make_row(T_NVA, make_key(0, 0));
        n++; last_z = 0; last_type = T_NVA;
      } else r.flags |= F_SENTINEL_DROP;
    }
    r.n = n;
    r.last_z = last_z;
    r.last_type = last_type;
    r.vo = al > be ? al - be : 0;
    r.va = al;
    r.vb = be;
    r.vnv = nvb - nva;
    if (pp.explicit_ss) {
      r.vss = ssb - ssa;
      r.vtc = 0;
    } else {
      r.vss = -al + be + tc - nvb + nva;
      r.vtc = tc;
    }

  }
}

// pair_keys<2> without branches, for the slots of at most two entries (most slots of a closed
// surface): the keys ka, kb (0 past the count c <= 2), the same SlotResult.  With two distinct
// rows, the top one (larger key) is a kept alpha when nZ >= 1; the bottom one is then a kept
// beta when nZ <= -1 (opening SSB / NVB at the cuts, each closed by a z = 0 floor row, SSA
// before NVA), or else a kept alpha when the top one was not.  The first kept alpha is the TC
// row of implicit SS.
TOMO_HD void pair_two(uint32_t ka, uint32_t kb, int c, const PairParams &pp, SlotResult &r) {
  const uint32_t k0 = ka > kb ? ka : kb, k1 = ka > kb ? kb : ka;
  const uint32_t k10 = make_key(1, 0);
  const bool v0 = c >= 1, two = c >= 2 && k1 != k0;  // row 0 exists; row 1 is distinct
  const int32_t z0 = key_z(k0), n0 = key_nz(k0), z1 = key_z(k1), n1 = key_nz(k1);
  r.npre = (v0 ? 1 : 0) + (two ? 1 : 0);
  r.sabs = (v0 ? (z0 < 0 ? -z0 : z0) : 0) + (two ? (z1 < 0 ? -z1 : z1) : 0);
  r.has10 = ((v0 && k0 == k10) || (two && k1 == k10)) ? 1 : 0;
  r.flags = 0;
  const bool a0 = two && n0 >= 1;          // row 0 a kept alpha
  const bool a1 = two && !a0 && n1 >= 1;   // row 1 a kept alpha
  const bool b1 = a0 && n1 <= -1;          // row 1 a kept beta
  const bool ssb = b1 && n1 <= pp.cut_ssb, nvb = b1 && n1 >= pp.cut_nvb;
  const int32_t al = a0 ? z0 : (a1 ? z1 : 0);
  const int32_t be = b1 ? z1 : 0;
  const int32_t tc = pp.explicit_ss ? 0 : al;
  const int32_t tc_bit = pp.explicit_ss ? 0 : T_TC;
  const int32_t ty1 = a1 ? (T_AL | tc_bit) : (b1 ? (T_BE | (ssb ? T_SSB : 0) | (nvb ? T_NVB : 0)) : 0);
  r.n = two ? 2 + (ssb ? 1 : 0) + (nvb ? 1 : 0) : r.npre;
  r.last_z = two ? ((ssb || nvb) ? 0 : z1) : (v0 ? z0 : 0);
  r.last_type = two ? (nvb ? T_NVA : (ssb ? T_SSA : ty1)) : 0;
  const int32_t zssb = ssb ? z1 : 0, znvb = nvb ? z1 : 0;
  r.vo = al > be ? al - be : 0;
  r.va = al;
  r.vb = be;
  r.vnv = znvb;
  r.vss = pp.explicit_ss ? zssb : -al + be + tc - znvb;
  r.vtc = tc;
}

}  // namespace tomo
