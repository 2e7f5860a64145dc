// tomo_device.h -- launch parameters and device records shared by the HIP kernels and the host
// side of libtomo_hip.so.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "tomo_slot.h"

namespace tomo {

constexpr int kThreads = 256;          // work-group size: 4 waves of 64
constexpr int kWaves = kThreads / 64;
constexpr int kSmall = 8;              // slots with <= 8 entries are paired in registers
constexpr int kInlineDefault = 16;     // entries stored inline per slot (layer-major); tunable
constexpr int kInlineMax = 32;         // in [kSmall, kInlineMax]
constexpr int kBigCap = 512;           // entries of a slot sorted by one wave in pair_big_kernel
constexpr int kPairSpanMax = 8192;     // touched-box slots per pair_kernel work-group (LDS window)
constexpr int kHugeLds = 8192;         // entries of a slot sorted in LDS by pair_huge_kernel; more
                                       // are selected in passes over global memory
constexpr int kHugeList = 1024;        // one triangle's entries of one slot (pair_huge_kernel)
constexpr int kMaxBedR = 16;           // int(BedOuterBound) supported by the bed kernel
constexpr int kBedThreads = 512;       // bed work-group: 8 waves; three 40 KB groups per CU give 6 waves
                                       // per SIMD (measured faster than 4 or 16 waves per group)
constexpr int kBedWaves = kBedThreads / 64;
constexpr int kTileThreads = 768;      // rasteriser work-group: 12 waves
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kTileSlots = 16384;      // rasteriser tile: slots with a 16-bit LDS counter each
constexpr int kTileYDefault = 128;     // tile extent along y (power of two); x extent = kTileSlots / y
constexpr int kChunk = 64;             // triangles per culling chunk (one per lane)
constexpr size_t kStaticScratchBytes = (size_t)256 << 20;  // tomo_run_static's array, per device
constexpr int kMaxDevices = 64;

// Device record, one per orientation.  Totals are exact 64-bit integers accumulated with
// atomics by the pairing and bed kernels; the host turns them into the reference's float
// totals and masses.
struct DevRecord {
  int64_t Va, Vb, Vo, Vss, Vnv, Vtc, Vbed;
  int64_t n_entries, n_slots, n_final;
  uint64_t chk_pre, chk_final;
  int64_t aVo, aVss, aVa, aVb, aVnv, aVtc;  // sums of |per-slot value| (float exactness test)
  int32_t flags, n_bed;
  float kx, ky, kz;                          // placement offsets (rotated AABB min - Bo, - Bo, 0)
  int32_t fsum_done;                         // fsum[] holds the reference's float running totals
  float fsum[6];                             // Vo, Vss, Va, Vb, Vnv, Vtc summed in slot order
  int32_t pad[2];
};

// Per in-flight orientation bookkeeping (reset by the AABB kernel).
struct OrientHdr {
  int32_t bbox[4];       // touched slots: x0, x1, y0, y1
  uint32_t big_top;      // slots with more than kSmall entries
  uint32_t pool_top;     // entries beyond the inline layers
  int32_t flags;
  float kx, ky, kz;
  unsigned long long zabs;  // sum of |z| over emitted entries: bounds every sum of |slot value|
  uint32_t spill_top;       // entries whose slot lies outside the tile that produced them
  uint32_t huge_top;        // slots with more than kBigCap entries (pair_huge_kernel's list)
  int32_t pad[2];
};

// Per in-flight orientation scratch, carved from one device allocation.  G2 = grid * grid.
struct ScratchLayout {
  int64_t count, ent, bedinfo, big, off, cursor, pool_slot, pool_key, pool_rank, cmp_key,
      cmp_rank, sv, spill_slot, spill_key, spill_rank, huge;  // byte offsets inside one orientation's region
  int64_t bytes; // bytes per region
  int32_t pool;  // overflow pool capacity (entries)
  int32_t inl;   // inline entries per slot
  int32_t spill; // spill list capacity (entries)
  int32_t nhuge; // huge-slot list capacity (slots)
};

struct SweepParams {
  // mesh (raw coordinates in voxel units); triangles in culling order (Morton order of centroids)
  const float *chunk_box;   // [n_chunks][6] raw-coordinate AABB (min xyz, max xyz) of kChunk triangles
  const float *vx, *vy, *vz;  // [n_vtx] raw vertex positions (structure of arrays)
  const int4 *tidx;         // [n_tri] vertex indices i0, i1, i2 and the caller's triangle index
  const int32_t *tpos;      // [n_tri] position in tidx of the caller's triangle t
  const float4 *vaos;       // [n_vtx][2] (x - cx, y - cy, z - cz, nx), (ny, nz, 0, 0)
  int32_t n_tri, n_vtx, n_chunks;
  float cx, cy, cz;        // raw AABB centre (STomoSh_Base.cpp:75-82)
  // orientations of this batch: rot/rec already offset to the batch's first orientation
  const float *rot;        // [n_batch][9] row-major rotation matrices
  int32_t n_batch;
  int32_t tiles;           // rasteriser work-groups (slot tiles) per orientation
  int32_t tiles_x, tiles_y;
  int32_t tile_x;          // tile extent along x (slots)
  int32_t tile_y_shift;    // log2 of the tile extent along y
  int32_t pair_tiles;      // pairing work-groups per orientation (touched box)
  int32_t pair_span;       // touched-box slots per pairing work-group
  int32_t big_groups;      // pairing work-groups per orientation (large-slot list)
  int32_t bed_bands;       // bed work-groups per orientation
  int32_t band_rows;       // bed: grid rows per work-group
  // options
  int32_t grid;
  float bed_out, bed_in, h2, shell_off;
  int32_t do_shell;
  int32_t bed_on;          // bed type != none and int(BedOuterBound) >= 1
  PairParams pp;
  const float *bed_dist;   // [2r][2r] float distances sqrtf(dx^2 + dy^2), dx,dy in [-r, r)
  // scratch and outputs
  uint8_t *scratch;
  OrientHdr *hdr;          // [n_batch]
  ScratchLayout lay;
  DevRecord *rec;          // [n_batch]
  // optional per-slot dumps for a one-orientation batch (grid*grid each, zero-filled by host)
  int32_t *dbg_pre, *dbg_final, *dbg_vo, *dbg_vss;
  uint64_t *dbg_rows;      // [grid*grid][64] final rows (type << 32 | key) for the render export
  int32_t exact_fsum;      // keep per-slot values and rebuild float totals when they round
};

// The stage kernels of the batched launch as persistent work-queue kernels: one launch per stage
// serves every unit (a slice of one mesh's orientations) of a round.  An item is one work-group's
// share of one unit (an orientation-tile for the rasteriser, an orientation for aabb), numbered
// unit after unit from first[u]; every first[u] is a multiple of 8 and item g belongs to the
// queue of XCD g & 7 (the unit's xcd_map), taken by the work-groups with blockIdx & 7 == g & 7.
enum QueueKernel : int {
  QK_AABB = 0, QK_RASTER, QK_COMPACT, QK_PAIR, QK_BIG, QK_HUGE, QK_BED, QK_FSUM, QK_COUNT
};

struct UnitQueue {
  const SweepParams *units;  // [n_units] each unit's parameters (device memory)
  const int32_t *first;      // [n_units + 1] first item of each unit; first[n_units] = items
  int32_t n_units;
  uint32_t *heads;           // [8] items taken per XCD queue; zero when the launch starts
};

}  // namespace tomo
