/*
 * tomo_hip.h -- C ABI of libtomo_hip.so, the MI355X (gfx950) support-tomography sweep.
 *
 * Two entry-point families:
 *
 *  1. The drop-in legacy shim.  Same ten-pointer signature, header-array layout, return value
 *     and getter set as the reference DLL that cfms_tomo binds with ctypes
 *     (cfms_tomo/Tomo_Shell_vc_src/Tomo_Shell2026.h:19-35; bound at
 *     cfms_tomo/tomo_dll_interface.py:24-35, called at :85-97).  Outputs are owned by the library and stay valid
 *     until the next TomoSh_HIP call or OnDestroy, as in the reference (Tomo_Shell2026.cpp:19-35,
 *     73-87).  Not re-entrant (global state), like the reference.
 *
 *  2. A handle API (re-entrant per handle, one HIP stream per handle, caller-owned outputs,
 *     status codes instead of exit()).  It replaces the same reference entry point
 *     (_TomoSh_Function_Call, Tomo_Shell2026.cpp:119-249) split into set-up, sweep and readback.
 *
 * Units and options follow the reference: Float32Info[12] and Int32Info[11] in the order packed
 * by tomo_dll_interface.py:73-82 and decoded by S3DPrinterInfo::Set (S3DPrinterInfo.cpp:124-145);
 * vertex coordinates are voxel indices (STomoVoxelSpaceInfo.cpp:97-105).
 */
#ifndef TOMO_HIP_H
#define TOMO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------- */
#define TOMO_OK 0
#define TOMO_E_INVALID_ARG (-1)
#define TOMO_E_HIP (-2)
#define TOMO_E_OOM (-3)
#define TOMO_E_NO_DEVICE (-4)
#define TOMO_E_STATE (-5)

/* ---- per-orientation flags (tomo_record.flags) ------------------------------------------- */
#define TOMO_F_FP32_INEXACT 1   /* a float32 running total may round: |terms| sum > 2^24      */
#define TOMO_F_INT16_WRAP 2     /* per-slot |z| sum > 32767 or a z/nZ outside int16 (the      */
                                /* pre-fix reference wrapped there)                           */
#define TOMO_F_FOOTPRINT_OOB 4  /* an entry fell outside the grid and was dropped (the        */
                                /* reference writes out of bounds: undefined behaviour)       */
#define TOMO_F_CAP_DROP 8       /* a slot held 63 distinct rows and dropped one               */
#define TOMO_F_SENTINEL_DROP 16 /* a floor row was skipped because its slot was full          */
                                /* (bit 32 is unused)                                         */
#define TOMO_F_UNRESOLVED 64    /* more than 65,535 rows fell on one slot (a slot counter is  */
                                /* 16 bits) or one triangle put more than 1,024 rows into a   */
                                /* slot of over 8,192 rows: the record is not exact           */

/* Bed types (Tomo_types.h:64-69). */
#define TOMO_BED_NONE 0
#define TOMO_BED_BRIM 1
#define TOMO_BED_RAFT 2
#define TOMO_BED_SKIRT 3

/* Per-orientation result.  Integer totals are exact; the f* fields are the reference's float32
 * running totals (STomoVolMassInfo.h:24-30), identical to them unless TOMO_F_FP32_INEXACT. */
typedef struct {
  int64_t Va, Vb, Vo, Vss, Vnv, Vtc, Vbed; /* Vbed = number of bed rows (each at z = 1)        */
  int64_t n_entries;  /* pixel rows after rasterisation, summed over slots                    */
  int64_t n_slots;    /* slots touched by rasterisation                                       */
  int64_t n_final;    /* rows after pairing floor rows and the bed                            */
  uint64_t chk_pre;   /* sum over slots of (slot id + 1) * rows after rasterisation           */
  uint64_t chk_final; /* same after pairing and bed                                          */
  int32_t flags;
  int32_t n_bed;
  float fVa, fVb, fVtc, fVnv, fVss, fVo, fVbed;
  float Mo, Mss, Mtotal; /* VolToMass (STomoVolMassInfo.cpp:42-81)                           */
  float mat4x4[16];      /* placement transform, row-major (STomoSh_Base.cpp:111-123)        */
} tomo_record;

typedef struct tomo_handle tomo_handle;

/* ---- handle API ---------------------------------------------------------------------------- */

/* Create a sweep context on the current HIP device.  grid: slot grid edge (the reference CPU
 * engine is fixed at 256, STomoVoxelSpaceInfo.h:10; 0 means Int32Info[3] = nVoxel).  Replaces
 * the option unpacking of S3DPrinterInfo::Set (S3DPrinterInfo.cpp:124-145). */
int tomo_create(const float *float32_info_x12, const int32_t *int32_info_x11, int32_t grid,
                tomo_handle **out);

/* Upload a mesh (host pointers, copied): tri[nTri*3] int32, vtx[nVtx*3], vtx_nrm[nVtx*3] float32.
 * Replaces the borrowed rpTri0/rpVtx0/rpVtxNrm0 pointers of S3DPrinterInfo::Set
 * (S3DPrinterInfo.cpp:179-182). */
int tomo_set_mesh(tomo_handle *h, int32_t n_tri, const int32_t *tri, int32_t n_vtx,
                  const float *vtx, const float *vtx_nrm);

/* Upload orientations: ypr[n*3] radians (yaw, pitch, roll) as TomoSh_INT3 takes them
 * (Tomo_Shell2026.cpp:104-106); rotation matrices are built on the host with cosf/sinf
 * (SMatrix4f.cpp:213-240). */
int tomo_set_orientations(tomo_handle *h, int32_t n, const float *ypr);

/* Enqueue the sweep of orientations [begin, end) on the handle's stream (asynchronous). */
int tomo_run(tomo_handle *h, int32_t begin, int32_t end);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): tomo_run for
 * orientations [begin, end), with the slot tables and orientation headers in a static device
 * array of the library, TOMO_STATIC_SCRATCH_BYTES on every device it runs on, held until the
 * process ends, instead of allocated scratch.  It calls no allocation routine.  The mesh and
 * orientations come from tomo_set_mesh / tomo_set_orientations (set-up calls, which allocate)
 * and the records are read with tomo_get_records as after tomo_run.  Orientations run in batches
 * of as many as the array holds at the handle's grid and inline capacity (about 20 at a 256
 * grid with 16 inline entries, 5 at 512; tomo_get_info reports it).  Fails with
 * TOMO_E_INVALID_ARG when one orientation's slot table exceeds the array (grids above 1,024 at
 * 16 inline entries).  All calls on one device share the array, which keeps the last call's
 * contents (nothing in it needs clearing: a sweep writes every header, counter and table before
 * it reads them); calls that overlap on one device race.  An orientation whose overflow pool
 * fills up is reported with TOMO_F_UNRESOLVED instead of being swept again. */
#define TOMO_STATIC_SCRATCH_BYTES (256u << 20)
int tomo_run_static(tomo_handle *h, int32_t begin, int32_t end);

/* Wait for the handle's stream. */
int tomo_sync(tomo_handle *h);

/* Copy records [begin, end) of the last run to out[end-begin] and finish them on the host
 * (float totals, masses, transform).  Synchronises. */
int tomo_get_records(tomo_handle *h, int32_t begin, int32_t end, tomo_record *out);

/* set_orientations + run + get_records for all n orientations. */
int tomo_sweep(tomo_handle *h, int32_t n, const float *ypr, tomo_record *out);

/* One orientation with per-slot dumps (host arrays of grid*grid int32, slot id x*grid + y; any
 * may be NULL): rows after rasterisation, rows at the end, slot Vo and Vss.  For parity tests
 * against the reference's per-slot state (STomoSh_INT3.cpp:366-395). */
int tomo_debug_slots(tomo_handle *h, const float ypr[3], tomo_record *rec, int32_t *cnt_pre,
                     int32_t *cnt_final, int32_t *slot_vo, int32_t *slot_vss);

/* Final rows of every slot of one orientation (the render export's input): rows[grid*grid*64]
 * as (type << 32 | key) with key = ((z + 32768) << 16) | (nZ + 32768) and type the pixel-type
 * bits (Tomo_types.h:72-82); nrows[grid*grid] rows per slot.  Replaces reading the slot buffer
 * of the optimum (Tomo_Shell2026.cpp:219-236). */
int tomo_debug_rows(tomo_handle *h, const float ypr[3], tomo_record *rec, uint64_t *rows, int32_t *nrows);

/* Orientations in flight per batch for tomo_run (0 = automatic; each needs one slot-table
 * region of scratch memory).  Launch-shape knob for tuning. */
int tomo_set_launch(tomo_handle *h, int32_t batch);

/* Launch-shape knobs for tuning (values take effect at the next tomo_run):
 * TOMO_TUNE_BATCH           orientations in flight (0 = automatic), same as tomo_set_launch;
 * TOMO_TUNE_TILE_Y          rasteriser tile extent along y in slots: a power of two 32..512,
 *                           or 0 (default: 128 up to a 256 grid, 64 above); a tile holds
 *                           16384 slots, so its x extent is 16384 / value;
 * TOMO_TUNE_INLINE          slot entries stored inline per slot (8..32, default 16); the rest go
 *                           to an overflow pool (more inline = fewer pool passes, more memory);
 * TOMO_TUNE_EXACT_FSUM      1 (default): where a float total can round (TOMO_F_FP32_INEXACT),
 *                           replay the reference's ascending-slot float accumulation so fVo,
 *                           fVss, fVa, fVb, fVnv, fVtc and the masses equal the reference's;
 *                           0: report float(integer total) instead (cheaper, may differ there). */
#define TOMO_TUNE_BATCH 0
#define TOMO_TUNE_TILE_Y 1
#define TOMO_TUNE_EXACT_FSUM 2
#define TOMO_TUNE_INLINE 3
int tomo_set_tuning(tomo_handle *h, int32_t knob, int32_t value);

/* Duration in milliseconds of the last tomo_run's kernel, from HIP events on the handle's
 * stream (negative if unavailable). */
float tomo_last_kernel_ms(tomo_handle *h);

/* Enable (1) or disable (0) per-stage timing (HIP events between kernels) in later tomo_run
 * calls, and read the last run's milliseconds summed over batches for: [0] rotated-AABB
 * reduction, [1] rasterisation, [2] overflow compaction, [3] pairing, [4] bed, [5] ordered float
 * totals; [6..7] = 0. */
int tomo_set_phase_profile(tomo_handle *h, int32_t enable);
int tomo_phase_profile(tomo_handle *h, double *ms8);

/* Grid edge, batch size of the last run, and scratch bytes held. */
int tomo_get_info(tomo_handle *h, int32_t *grid, int32_t *batch, int64_t *scratch_bytes);

int tomo_destroy(tomo_handle *h);

/* Message for the last error on this thread (never NULL). */
const char *tomo_last_error(void);

/* First strict minimum, initial value 1e5; 0 for n <= 1 (_find1stOptimal,
 * Tomo_Shell2026.cpp:89-97). */
int32_t tomo_find_first_optimal(int32_t n, const float *mss);

/* ---- batched sweeps: many meshes in one graph launch --------------------------------------
 * Replaces calling the reference entry point once per cut part (TomoManager.get_best_orientation,
 * cfms_tomo/__init__.py:26-30, one TomoSh_INT3 per part): every member mesh keeps its own options,
 * orientations and records; all members' sweeps are captured into one hipGraph on one stream
 * and run by one launch.  Members share the slot-table scratch, sized for the largest. */
typedef struct tomo_batch tomo_batch;

int tomo_batch_create(tomo_batch **out);

/* Add a member (host arrays, copied): options and grid as tomo_create, mesh as tomo_set_mesh,
 * orientations as tomo_set_orientations.  *member receives its index.  Invalidates a capture. */
int tomo_batch_add(tomo_batch *b, const float *float32_info_x12, const int32_t *int32_info_x11,
                   int32_t grid, int32_t n_tri, const int32_t *tri, int32_t n_vtx, const float *vtx,
                   const float *vtx_nrm, int32_t n_ypr, const float *ypr, int32_t *member);

/* Autotuner: per group of members with the same grid and ceil(log2(nTri)), time `sample`
 * orientations of the group's largest member over the rasteriser tile shape, inline slot
 * capacity and orientations in flight, and keep the fastest shape that resolves every slot. */
int tomo_batch_tune(tomo_batch *b, int32_t sample);

/* The launch shape chosen for a member (defaults before tomo_batch_tune) and its measured
 * microseconds per orientation (negative if not tuned). */
int tomo_batch_get_tuning(tomo_batch *b, int32_t member, int32_t *batch, int32_t *tile_y,
                          int32_t *inline_capacity, float *us_per_orientation);

/* Set a member's launch shape by hand, as tomo_batch_tune would: orientations in flight
 * (0 = the default, min(n_ypr, 2048)), rasteriser tile extent along y (0 or a power of two
 * 32..512) and inline slot capacity (8..32).  Lets a caller reuse shapes tuned in an earlier
 * process instead of tuning again.  Invalidates a capture.  No reference counterpart (the
 * reference sweeps parts one call at a time). */
int tomo_batch_set_tuning(tomo_batch *b, int32_t member, int32_t batch, int32_t tile_y,
                          int32_t inline_capacity);

/* How tomo_batch_capture builds the graph.  TOMO_BATCH_QUEUE (the default): the members'
 * orientations are cut into units (a member's orientations in flight), consecutive units are
 * packed into rounds whose slot tables fit the shared block side by side, and each round runs as
 * one persistent work-queue launch per stage kernel over all its units' (mesh, orientation-tile)
 * items.  TOMO_BATCH_CHAIN: every member's own chain of stage launches, member after member.
 * Results are the same.  Invalidates a capture.  No reference counterpart (the reference sweeps
 * parts one call at a time). */
#define TOMO_BATCH_QUEUE 0
#define TOMO_BATCH_CHAIN 1
int tomo_batch_set_mode(tomo_batch *b, int32_t mode);

/* Rounds and units of the last persistent capture (0 and 0 for the chained form). */
int tomo_batch_get_plan(tomo_batch *b, int32_t *rounds, int32_t *units);

/* Allocate the shared scratch and capture every member's sweep into one graph (no launch). */
int tomo_batch_capture(tomo_batch *b);

/* Launch the graph (capturing first if needed); asynchronous on the batch's stream. */
int tomo_batch_launch(tomo_batch *b);
int tomo_batch_sync(tomo_batch *b);

/* Milliseconds of the last graph launch (HIP events on the batch's stream; negative if none). */
float tomo_batch_last_ms(tomo_batch *b);

int32_t tomo_batch_size(tomo_batch *b);

/* Records of one member's orientations (out[n_ypr of that member]); synchronises. */
int tomo_batch_records(tomo_batch *b, int32_t member, tomo_record *out);

int tomo_batch_destroy(tomo_batch *b);

/* ---- continuous support volume (differentiable form) -------------------------------------
 * Vss_c(yaw, pitch) = sum over triangles of sigma(-r.n - sin(theta_c)) A (-r.n) (r.c - zlow),
 * r the build direction (third row of the reference's Rz(roll) Ry(pitch) Rx(yaw),
 * SMatrix4f.cpp:213-240), n / A / c the unit face normal, area and centroid, sigma a C1
 * smoothstep of width tau around the critical angle (tau = 0: the hard step) and zlow the soft
 * minimum -temp log sum_v exp(-r.v / temp) of the vertex heights (temp = 0: the lowest vertex).
 * The per-triangle precedent is getCHVss (STomoNV_CvxH.cpp:97-174); there is no reference entry
 * point for it.  Roll does not enter (d/d roll = 0).  Device pointers and the stream (a
 * hipStream_t, NULL for the null stream) are the caller's; calls are asynchronous on it. */
typedef struct tomo_cvss tomo_cvss;

/* Mesh as for tomo_set_mesh (host arrays; vertex normals are not used), theta_c from
 * f32x12[1]; tau in [0, 2] (units of r.n), temp >= 0 (mesh units). */
int tomo_cvss_create(const float *f32x12, int32_t n_tri, const int32_t *tri, int32_t n_vtx,
                     const float *vtx, float tau, float temp, tomo_cvss **out);

/* n orientations ypr_dev[n][3] (radians) -> vss_dev[n].  aux_dev: double[n][12] workspace that
 * keeps each orientation's lowest point and sums for the backward pass. */
int tomo_cvss_forward(tomo_cvss *c, int32_t n, const float *ypr_dev, float *vss_dev, double *aux_dev,
                      void *stream);

/* grad_ypr_dev[n][3] = grad_vss_dev[n] * d Vss_c / d (yaw, pitch, roll), using the aux_dev the
 * forward call filled for the same ypr_dev (its slots 6..8 are overwritten). */
int tomo_cvss_backward(tomo_cvss *c, int32_t n, const float *ypr_dev, double *aux_dev,
                       const float *grad_vss_dev, float *grad_ypr_dev, void *stream);

int tomo_cvss_destroy(tomo_cvss *c);

/* ---- legacy drop-in shim (Tomo_Shell2026.h:19-35) ----------------------------------------- */

/* Same arguments and return (optimal orientation index) as TomoSh_INT3 / TomoSh_CUDA
 * (Tomo_Shell2026.h:33-35).  Convex-hull and triangle-normal arrays are accepted and unused, as
 * in the INT3 engine.  Returns a negative status code on error instead of exiting. */
int32_t TomoSh_HIP(float *float32_info_x12, int32_t *int32_info_x11, float *YPR, int32_t *tri,
                   float *vtx, float *vtx_nrm, float *tri_nrm, int32_t *chull_tri,
                   float *chull_vtx, float *chull_trinrm);

float *getMss(void);         /* float32[nYPR] (Tomo_Shell2026.cpp:67)        */
float *getMo(void);          /* float32[nYPR] (Tomo_Shell2026.cpp:68)        */
float *getVtc(void);         /* float32[nYPR] (Tomo_Shell2026.cpp:69)        */
float *getVolMassInfo(void); /* float32[22] of the optimum (:70, :244)      */
float *getMat4x4(void);      /* float32[16] of the optimum (:71, :239-242)  */
/* Render export of the optimum's pixel groups (Tomo_Shell2026.cpp:65-66, 221-236):
 * rows {x, y, z, nx, ny, nz} int16 per pixel. */
int32_t getnData2i(int16_t pixel_type);
int16_t *getpData2i(int16_t pixel_type);
void OnDestroy(void); /* frees the shim's outputs (Tomo_Shell2026.cpp:73-87) */

#ifdef __cplusplus
}
#endif
#endif
