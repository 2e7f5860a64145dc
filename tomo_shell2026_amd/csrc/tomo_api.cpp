// tomo_api.cpp -- host side of libtomo_hip.so: option decoding, device buffers, launches,
// record finishing (float totals, VolToMass, transform) and the legacy drop-in shim.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/tomo_hip.h"
#include "tomo_device.h"

namespace tomo {
hipError_t launch_stage(int stage, const SweepParams &p, hipStream_t stream, hipStream_t side = nullptr,
                        hipEvent_t fork = nullptr, hipEvent_t join = nullptr);
hipError_t static_scratch_ptr(int device, uint8_t **out);
int32_t queue_items(int k, const SweepParams &p);
hipError_t queue_occupancy(int k, int *per_cu);
hipError_t launch_queue(int k, const UnitQueue &q, int32_t groups, hipStream_t stream);
hipError_t launch_clear_heads(uint32_t *heads, int32_t n, hipStream_t stream);
struct CvssMeshDev {
  const float *nx, *ny, *nz, *area, *cx, *cy, *cz, *vx, *vy, *vz;
  int32_t n_tri, n_vtx;
  float s, tau, temp;
};
hipError_t cvss_forward(const CvssMeshDev &m, int32_t n, const float *ypr, float *vss, double *aux,
                        hipStream_t st);
hipError_t cvss_backward(const CvssMeshDev &m, int32_t n, const float *ypr, double *aux,
                         const float *grad_vss, float *grad_ypr, hipStream_t st);
}

using namespace tomo;

namespace {

thread_local std::string g_err = "";

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(TOMO_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr float kMargin = 0.001f;

struct Options {
  float f32[12];
  int32_t i32[11];
  int32_t grid;
  float dvoxel, theta_c, area, wall, fcore, fclad, fss, css, rho, bed_in, bed_out, bed_thick;
  int explicit_ss, shell_mesh, bed_type, bed_r, do_shell;
  float shell_t, shell_off, h2;
  int32_t cut_ssb, cut_nvb;
};

// S3DPrinterInfo::Set (S3DPrinterInfo.cpp:124-145) and the constants triVoxel derives
// (STomoSh_INT3.cpp:1160-1163).
int decode_options(const float *f32, const int32_t *i32, int32_t grid, Options &o) {
  if (!f32 || !i32) return fail(TOMO_E_INVALID_ARG, "null option array");
  std::memcpy(o.f32, f32, sizeof(o.f32));
  std::memcpy(o.i32, i32, sizeof(o.i32));
  o.dvoxel = f32[0]; o.theta_c = f32[1]; o.area = f32[2]; o.wall = f32[3];
  o.fcore = f32[4]; o.fclad = f32[5]; o.fss = f32[6]; o.css = f32[7]; o.rho = f32[8];
  o.bed_in = f32[9]; o.bed_out = f32[10]; o.bed_thick = f32[11];
  o.explicit_ss = (float)i32[1] > kMargin;
  o.shell_mesh = (float)i32[2] > kMargin;
  o.shell_t = (float)i32[10] * 0.001f;
  o.bed_type = i32[9];
  o.grid = grid > 0 ? grid : (i32[3] > 0 ? i32[3] : 256);
  if (o.grid < 8 || o.grid > 4096) return fail(TOMO_E_INVALID_ARG, "grid edge out of range [8, 4096]");
  if (o.bed_type < 0 || o.bed_type > 3) return fail(TOMO_E_INVALID_ARG, "BedType must be 0..3");
  o.bed_r = (int)o.bed_out;
  if (o.bed_type != 0 && o.bed_r > kMaxBedR)
    return fail(TOMO_E_INVALID_ARG, "BedOuterBound above the supported maximum (16)");
  o.h2 = o.dvoxel * 0.5f;
  o.shell_off = (float)(int)(o.shell_t + 0.999f);
  o.do_shell = o.shell_mesh && o.shell_t > kMargin;
  // createShadowBriefly's float tests (STomoSh_INT3.cpp:755-776) as integer cut points.
  const float thr = (float)(-1.0 * std::sin((double)o.theta_c));
  int32_t cs = -32769, cn = 32768;
  for (int32_t n = -32768; n <= 32767; n++) {
    const float f = (float)n / 1000.0f;
    if (f < thr) cs = n;
    if (f > thr && n < cn) cn = n;
  }
  o.cut_ssb = cs;
  o.cut_nvb = cn;
  return TOMO_OK;
}

// SMatrix4f::YPR (SMatrix4f.cpp:213-240), host libm.
void ypr_matrix(float yaw, float pitch, float roll, float *R) {
  const float ca = cosf(roll), cb = cosf(pitch), cg = cosf(yaw);
  const float sa = sinf(roll), sb = sinf(pitch), sg = sinf(yaw);
  R[0] = ca * cb;
  R[1] = ca * sb * sg - sa * cg;
  R[2] = ca * sb * cg + sa * sg;
  R[3] = sa * cb;
  R[4] = sa * sb * sg + ca * cg;
  R[5] = sa * sb * cg - ca * sg;
  R[6] = -sb;
  R[7] = cb * sg;
  R[8] = cb * cg;
}

// STomoVolMassInfo::VolToMass (STomoVolMassInfo.cpp:42-81); out[22] in the struct's order
// (STomoVolMassInfo.h:24-30).
void vol_to_mass(const Options &o, float *vm) {
  float &Vss = vm[4], &Vss_clad = vm[5], &Vss_core = vm[6], &Ass_clad = vm[7];
  float &Vo = vm[8], &Vo_clad = vm[9], &Vo_core = vm[10];
  float &Vbed = vm[11], &Mbed = vm[12], &Mss = vm[13], &Mss_clad = vm[14], &Mss_core = vm[15];
  float &Mo = vm[16], &Mo_clad = vm[17], &Mo_core = vm[18], &Mtotal = vm[19];
  Vo_clad = Vo_core = Mo_clad = Mo_core = Mo = 0.f;
  Vss_clad = Vss_core = Ass_clad = Mss_clad = Mss_core = Mss = 0.f;
  Mbed = Mtotal = 0.f;
  Vo_clad = o.area * o.wall;
  Vo_core = Vo - Vo_clad;
  Mo_clad = Vo_clad * o.fclad * o.rho;
  Mo_core = Vo_core * o.fcore * o.rho;
  Mo = Mo_core + Mo_clad;
  if (o.explicit_ss) {
    if (Vss > kMargin) {
      const float radius = powf(Vss / (4.f / 3.f * 3.141592f), 1.f / 3.f);
      Ass_clad = (float)(4. * 3.141592 * (double)radius * (double)radius);
      Vss_clad = Ass_clad * o.wall;
      if (Vss_clad > Vss) Vss_clad = Vss;
      Vss_core = Vss - Vss_clad;
      Mss_clad = Vss_clad * o.fclad * o.rho;
      Mss_core = Vss_core * o.fcore * o.rho;
      Mss = Mss_core + Mss_clad;
    }
  } else {
    Mss = Vss * o.fss * o.css * o.rho;
  }
  Mbed = Vbed * o.wall * o.rho;
  Mss += Mbed;
  Mtotal = Mo + Mss;
}

// One orientation's slot-table region.  inl entries per slot are stored inline; the overflow
// pool grows as inl shrinks (2 * max(G^2, 65536) entries at inl = 8).  scale multiplies the
// pool and the spill list (orientations that overflow them are swept again with scale 4, 16, 64).
ScratchLayout make_layout(int32_t G, int32_t inl, int32_t scale = 1) {
  ScratchLayout L{};
  const int64_t G2 = (int64_t)G * G;
  const int64_t base = G2 > 65536 ? G2 : 65536;
  const int64_t pool64 = base * kInlineDefault / (inl < kInlineDefault ? inl : kInlineDefault) * scale;
  const int32_t pool = (int32_t)std::min<int64_t>(pool64, (int64_t)1 << 30);
  int64_t at = 0;
  auto take = [&](int64_t bytes) {
    const int64_t here = at;
    at += (bytes + 255) & ~int64_t(255);
    return here;
  };
  L.count = take(4 * G2);
  L.ent = take(8 * G2 * inl);
  L.bedinfo = take(2 * G2);
  L.big = take(4 * G2);
  L.off = take(4 * G2);
  L.cursor = take(4 * G2);
  L.pool_slot = take(4 * (int64_t)pool);
  L.pool_key = take(4 * (int64_t)pool);
  L.pool_rank = take(4 * (int64_t)pool);
  L.cmp_key = take(4 * (int64_t)pool);
  L.cmp_rank = take(4 * (int64_t)pool);
  L.sv = take(4 * 6 * G2);
  const int32_t spill = (16 * G + 4096) * scale;
  L.spill_slot = take(4 * (int64_t)spill);
  L.spill_key = take(4 * (int64_t)spill);
  L.spill_rank = take(4 * (int64_t)spill);
  // slots of more than kBigCap entries each hold over kBigCap - inl pool entries
  const int32_t nhuge = pool / 256 + 64;
  L.huge = take(4 * (int64_t)nhuge);
  L.bytes = at;
  L.pool = pool;
  L.inl = inl;
  L.spill = spill;
  L.nhuge = nhuge;
  return L;
}

}  // namespace

// Legacy shim outputs (Tomo_Shell2026.cpp:19-35).
static std::vector<float> g_mss, g_mo, g_vtc;
static float g_vm[22];
static float g_mat[16];
static bool g_have = false;
static std::vector<std::vector<int16_t>> g_pix(11);

struct tomo_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipStream_t side = nullptr;  // second stream: large-slot pairing beside the small-slot pairing
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_pair = nullptr, ev_bed = nullptr;  // a batch's pairing done; its bed and sums done
  hipEvent_t ev_aabb = nullptr;                     // the next batch's placement done
  bool timed = false;
  Options opt{};
  // mesh
  int32_t n_tri = 0, n_vtx = 0;
  float center[3] = {0, 0, 0};
  float *d_chunk = nullptr;     // [n_chunks][6] raw AABB of each kChunk triangles
  int32_t n_chunks = 0;
  float *d_soa = nullptr;  // 3 * n_vtx: x, y, z
  int4 *d_tidx = nullptr;    // n_tri: vertex indices + caller's index (culling order)
  int32_t *d_tpos = nullptr; // n_tri: position in d_tidx of each caller's triangle
  float4 *d_vaos = nullptr;  // 2 * n_vtx: (x, y, z, nx), (ny, nz, 0, 0)
  float *d_bed_dist = nullptr;
  // orientations
  int32_t n_orient = 0;
  std::vector<float> rot;
  float *d_rot = nullptr;
  DevRecord *d_rec = nullptr;
  int32_t rec_cap = 0;
  // scratch: one slot-table region per in-flight orientation
  ScratchLayout lay{};
  uint8_t *d_scratch = nullptr;
  size_t scratch_cap = 0;      // bytes of the block behind d_scratch
  OrientHdr *d_hdr = nullptr;
  int32_t scratch_batch = 0;
  int64_t scratch_region = 0;  // bytes per orientation region of the current allocation
  // Shared mode (a tomo_batch member): stream, scratch and headers belong to the batch;
  // own_stream is the handle's original stream, kept for destruction.
  bool shared = false;
  hipStream_t own_stream = nullptr;
  int32_t batch_req = 0;
  int32_t tile_y = 0;  // 0: kTileYDefault up to a 256 grid, half that above
  int32_t exact_fsum = 1;
  int32_t last_batch = 0;
  int32_t last_begin = 0, last_end = 0;
  // tomo_run_static: this device's static slot-table array (nullptr if the lookup failed) and
  // whether the last run used it (its records are not swept again on overflow)
  uint8_t *d_static = nullptr;
  hipError_t static_err = hipSuccess;
  bool last_static = false;
  std::vector<float> vbed_seq;
  // per-stage timing (events around every stage of every batch)
  bool phases = false;
  std::vector<hipEvent_t> stage_ev;
  std::vector<int> stage_of;
};

namespace {

void free_stage_events(tomo_handle *h) {
  for (hipEvent_t e : h->stage_ev) (void)hipEventDestroy(e);
  h->stage_ev.clear();
  h->stage_of.clear();
}

// Slot-table scratch is large (tens of GB at the default batch) and hipMalloc / hipFree of such
// blocks take about a second each, so the most recently released block is kept for the next
// handle that needs one (any size up to 4x larger than asked for) instead of being freed.
std::mutex g_scratch_mu;
struct {
  void *ptr = nullptr;
  size_t bytes = 0;
  int device = -1;
} g_scratch_cache;

hipError_t scratch_alloc(size_t bytes, int device, uint8_t **out, size_t *cap) {
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_cache.ptr && g_scratch_cache.device == device && g_scratch_cache.bytes >= bytes &&
        g_scratch_cache.bytes / 4 <= bytes) {
      *out = (uint8_t *)g_scratch_cache.ptr;
      *cap = g_scratch_cache.bytes;
      g_scratch_cache.ptr = nullptr;
      return hipSuccess;
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {  // the cached block may be what stands in the way
    void *old = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_scratch_mu);
      old = g_scratch_cache.ptr;
      g_scratch_cache.ptr = nullptr;
    }
    if (!old) return e;
    (void)hipGetLastError();
    (void)hipFree(old);
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e;
  }
  *out = (uint8_t *)p;
  *cap = bytes;
  return hipSuccess;
}

size_t scratch_cached(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (g_scratch_cache.ptr && g_scratch_cache.device == device) ? g_scratch_cache.bytes : 0;
}

// Give a block back (its stream must be idle): it replaces the cached one, which is freed.
void scratch_release(uint8_t *p, size_t cap, int device) {
  if (!p) return;
  void *old = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    old = g_scratch_cache.ptr;
    g_scratch_cache.ptr = p;
    g_scratch_cache.bytes = cap;
    g_scratch_cache.device = device;
  }
  if (old) (void)hipFree(old);
}

int ensure_scratch(tomo_handle *h, int32_t batch) {
  if (batch <= h->scratch_batch && h->lay.bytes == h->scratch_region) return TOMO_OK;
  if (h->shared) return fail(TOMO_E_STATE, "shared scratch does not fit this launch shape");
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (h->d_scratch) { scratch_release(h->d_scratch, h->scratch_cap, h->device); h->d_scratch = nullptr; }
  if (h->d_hdr) { HIP_TRY(hipFree(h->d_hdr)); h->d_hdr = nullptr; }
  h->scratch_batch = 0;
  const size_t bytes = (size_t)h->lay.bytes * (size_t)batch;
  if (scratch_alloc(bytes, h->device, &h->d_scratch, &h->scratch_cap) != hipSuccess) {
    h->d_scratch = nullptr;
    return fail(TOMO_E_OOM, "scratch allocation failed");
  }
  // two header blocks: consecutive batches alternate between them (enqueue)
  if (hipMalloc(&h->d_hdr, sizeof(OrientHdr) * 2 * (size_t)batch) != hipSuccess) {
    h->d_hdr = nullptr;
    return fail(TOMO_E_OOM, "header allocation failed");
  }
  // Defined contents from the start (each sweep writes every counter and header it reads).
  HIP_TRY(hipMemsetAsync(h->d_scratch, 0, bytes, h->stream));
  HIP_TRY(hipMemsetAsync(h->d_hdr, 0, sizeof(OrientHdr) * 2 * (size_t)batch, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->scratch_batch = batch;
  h->scratch_region = h->lay.bytes;
  return TOMO_OK;
}

// Orientations in flight per batch: as requested; else up to 8192 within a quarter of free
// memory (counting the handle's current scratch and the cached block as free), split evenly over
// the launches so the last one is not a short tail.
int32_t auto_batch(tomo_handle *h, int32_t n) {
  const bool requested = h->batch_req > 0;
  int32_t g = requested ? h->batch_req : 8192;
  size_t freeb = 0, total = 0;
  if (g > h->scratch_batch && hipMemGetInfo(&freeb, &total) == hipSuccess && h->lay.bytes > 0) {
    const size_t mine = (h->shared ? 0 : h->scratch_cap) + scratch_cached(h->device);
    const size_t cap = ((freeb + mine) / 4) / (size_t)h->lay.bytes;
    if (cap >= 1 && (size_t)g > cap) g = (int32_t)cap;
  }
  if (g > n) g = n;
  if (g < 1) g = 1;
  if (!requested) {
    const int32_t launches = (n + g - 1) / g;
    g = (n + launches - 1) / launches;
  }
  return g;
}

void finish_record(const tomo_handle *h, const DevRecord &d, const float *R, tomo_record &r) {
  std::memset(&r, 0, sizeof(r));
  r.Va = d.Va; r.Vb = d.Vb; r.Vo = d.Vo; r.Vss = d.Vss; r.Vnv = d.Vnv; r.Vtc = d.Vtc;
  r.Vbed = d.Vbed; r.n_entries = d.n_entries; r.n_slots = d.n_slots; r.n_final = d.n_final;
  r.chk_pre = d.chk_pre; r.chk_final = d.chk_final; r.n_bed = (int32_t)d.Vbed;
  int32_t flags = d.flags;
  if (flags & F_RETRY) flags = (flags & ~F_RETRY) | TOMO_F_UNRESOLVED;  // still overflowing
  const int64_t lim = (int64_t)1 << 24;
  if (d.aVo > lim || d.aVss > lim || d.aVa > lim || d.aVb > lim || d.aVnv > lim || d.aVtc > lim)
    flags |= TOMO_F_FP32_INEXACT;
  r.flags = flags;
  // Integer partial sums below 2^24 are exact in float32, so the reference's running totals
  // equal the integers (Pairing :386-392 and Calculate :1233-1240 accumulate ints into floats).
  r.fVa = (float)d.Va; r.fVb = (float)d.Vb; r.fVtc = (float)d.Vtc; r.fVnv = (float)d.Vnv;
  r.fVss = (float)d.Vss; r.fVo = (float)d.Vo;
  if (d.fsum_done) {  // replayed on the device in the reference's accumulation order
    r.fVo = d.fsum[0]; r.fVss = d.fsum[1]; r.fVa = d.fsum[2]; r.fVb = d.fsum[3];
    r.fVnv = d.fsum[4]; r.fVtc = d.fsum[5];
  }
  const size_t nb = (size_t)(d.Vbed < 0 ? 0 : d.Vbed);
  r.fVbed = nb < h->vbed_seq.size() ? h->vbed_seq[nb] : (float)nb * h->opt.bed_thick;
  float vm[22] = {};
  vm[0] = r.fVa; vm[1] = r.fVb; vm[2] = r.fVtc; vm[3] = r.fVnv; vm[4] = r.fVss; vm[8] = r.fVo;
  vm[11] = r.fVbed;
  vol_to_mass(h->opt, vm);
  r.Mo = vm[16]; r.Mss = vm[13]; r.Mtotal = vm[19];
  // mat4x4 (STomoSh_Base.cpp:111-123)
  const float m = -h->center[0], n = -h->center[1], o = -h->center[2];
  const float tr[3] = {-d.kx, -d.ky, -d.kz};
  for (int row = 0; row < 3; row++) {
    r.mat4x4[row * 4 + 0] = R[row * 3 + 0];
    r.mat4x4[row * 4 + 1] = R[row * 3 + 1];
    r.mat4x4[row * 4 + 2] = R[row * 3 + 2];
    r.mat4x4[row * 4 + 3] = tr[row] + R[row * 3 + 0] * m + R[row * 3 + 1] * n + R[row * 3 + 2] * o;
  }
  r.mat4x4[15] = 1.0f;
}

void vm_from_record(const tomo_handle *h, const tomo_record &r, float *vm) {
  std::memset(vm, 0, sizeof(float) * 22);
  vm[0] = r.fVa; vm[1] = r.fVb; vm[2] = r.fVtc; vm[3] = r.fVnv; vm[4] = r.fVss; vm[8] = r.fVo;
  vm[11] = r.fVbed;
  vol_to_mass(h->opt, vm);
}

SweepParams base_params(tomo_handle *h) {
  SweepParams p{};
  p.chunk_box = h->d_chunk;
  p.n_chunks = h->n_chunks;
  p.tidx = h->d_tidx;
  p.tpos = h->d_tpos;
  p.vaos = h->d_vaos;
  p.vx = h->d_soa;
  p.vy = h->d_soa + h->n_vtx;
  p.vz = h->d_soa + 2 * (size_t)h->n_vtx;
  p.n_tri = h->n_tri;
  p.n_vtx = h->n_vtx;
  p.cx = h->center[0]; p.cy = h->center[1]; p.cz = h->center[2];
  p.grid = h->opt.grid;
  p.bed_out = h->opt.bed_out;
  p.bed_in = h->opt.bed_in;
  p.h2 = h->opt.h2;
  p.shell_off = h->opt.shell_off;
  p.do_shell = h->opt.do_shell;
  p.pp.cut_ssb = h->opt.cut_ssb;
  p.pp.cut_nvb = h->opt.cut_nvb;
  p.pp.explicit_ss = h->opt.explicit_ss;
  p.pp.bed_type = h->opt.bed_type;
  p.pp.bed_r = h->opt.bed_r;
  p.bed_dist = h->d_bed_dist;
  p.bed_on = (h->opt.bed_type != 0 && h->opt.bed_r >= 1) ? 1 : 0;
  p.lay = h->lay;
  p.hdr = h->d_hdr;
  p.scratch = h->d_scratch;
  const int32_t g2 = h->opt.grid * h->opt.grid;
  const int32_t tile_y = h->tile_y > 0 ? h->tile_y : (h->opt.grid > 256 ? kTileYDefault / 2 : kTileYDefault);
  int32_t ys = 0;
  while ((1 << (ys + 1)) <= tile_y) ys++;
  p.tile_y_shift = ys;
  p.tile_x = kTileSlots >> ys;
  p.tiles_x = (h->opt.grid + p.tile_x - 1) / p.tile_x;
  p.tiles_y = (h->opt.grid + (1 << ys) - 1) >> ys;
  p.tiles = p.tiles_x * p.tiles_y;
  // Pairing work-groups: 8,192 touched-box slots each up to a 256 grid, 1,024 above (measured on
  // the 69k-triangle sphere at 256 against 4,096 and 16,384: 3% faster than either; on the
  // thin-shell manikin at 512, 3% faster than 2,048).  The large-slot list gets a quarter as
  // many (two per orientation at 256: 10% faster than one, four and eight were slower; on the
  // manikin 1% faster than an eighth, a sixteenth was 3% slower).
  p.pair_span = h->opt.grid <= 256 ? kPairSpanMax : 1024;
  p.pair_tiles = (g2 + p.pair_span - 1) / p.pair_span;
  p.big_groups = p.pair_tiles >= 4 ? p.pair_tiles / 4 : 1;
  // Bed bands of 40 rows up to a 256 grid, where a band and its halo fit the bed kernel's LDS
  // staging for bed radii up to 6 (on the 69k-triangle sphere 15% faster than 16 rows; 32 and
  // 48 were slower), 16 rows above.
  p.band_rows = h->opt.grid <= 256 ? 40 : 16;
  p.bed_bands = (h->opt.grid + p.band_rows - 1) / p.band_rows;
  p.exact_fsum = h->exact_fsum;
  return p;
}

// Enqueue the five stages for orientations [begin, end) in batches of `batch`.
// scr / hdr, when given, replace the handle's slot-table scratch and orientation headers.
int enqueue(tomo_handle *h, int32_t begin, int32_t end, int32_t batch, int32_t *dbg4, size_t g2,
            uint64_t *dbg_rows = nullptr, uint8_t *scr = nullptr, OrientHdr *hdr = nullptr) {
  // The side stream is left out while the stream is being captured into a graph.
  hipStream_t side = h->shared ? nullptr : h->side;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (side && (hipStreamIsCapturing(h->stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone))
    side = nullptr;
  // With the side stream, a batch's bed and ordered sums run on it after the batch's pairing,
  // beside the next batch's placement, rasteriser and compaction (which touch neither the bed
  // bookkeeping nor the per-slot sums, and whose headers are the other of two blocks); the next
  // batch's pairing waits for them.  Not for debug dumps, stage timing or the static array.
  const bool overlap = side && !scr && !dbg4 && !dbg_rows && !h->phases;
  int32_t k = 0;
  for (int32_t b0 = begin; b0 < end; b0 += batch, k++) {
    SweepParams p = base_params(h);
    if (scr) {
      p.scratch = scr;
      p.hdr = hdr;
    } else if (overlap) {
      p.hdr = h->d_hdr + (size_t)(k & 1) * (size_t)batch;
    }
    p.n_batch = (end - b0) < batch ? (end - b0) : batch;
    p.rot = h->d_rot + (size_t)b0 * 9;
    p.rec = h->d_rec + b0;
    if (dbg4) {
      p.dbg_pre = dbg4;
      p.dbg_final = dbg4 + g2;
      p.dbg_vo = dbg4 + 2 * g2;
      p.dbg_vss = dbg4 + 3 * g2;
    }
    p.dbg_rows = dbg_rows;
    if (overlap) {
      // placement: the first batch's here; each later one's ran on the side stream during the
      // previous batch's pairing
      if (k == 0) HIP_TRY(launch_stage(0, p, h->stream));
      else HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_aabb, 0));
      for (int stage = 1; stage < 3; stage++) HIP_TRY(launch_stage(stage, p, h->stream));
      if (k > 0) HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_bed, 0));
      if (b0 + batch < end) {
        // the next batch's placement beside this batch's pairing, before the small-slot pairing
        // on the side stream (the large-slot kernel on the main stream is the longer of the two);
        // its headers are the block this batch's predecessor used, whose bed is done
        SweepParams pn = base_params(h);
        pn.hdr = h->d_hdr + (size_t)((k + 1) & 1) * (size_t)batch;
        pn.n_batch = (end - b0 - batch) < batch ? (end - b0 - batch) : batch;
        pn.rot = h->d_rot + (size_t)(b0 + batch) * 9;
        pn.rec = h->d_rec + b0 + batch;
        HIP_TRY(hipEventRecord(h->ev_pair, h->stream));
        HIP_TRY(hipStreamWaitEvent(side, h->ev_pair, 0));
        HIP_TRY(launch_stage(0, pn, side));
        HIP_TRY(hipEventRecord(h->ev_aabb, side));
      }
      HIP_TRY(launch_stage(3, p, h->stream, side, h->ev_fork, h->ev_join));
      HIP_TRY(hipEventRecord(h->ev_pair, h->stream));
      HIP_TRY(hipStreamWaitEvent(side, h->ev_pair, 0));
      HIP_TRY(launch_stage(4, p, side));
      HIP_TRY(launch_stage(5, p, side));
      HIP_TRY(hipEventRecord(h->ev_bed, side));
      continue;
    }
    for (int stage = 0; stage < 6; stage++) {
      if (h->phases) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventRecord(e, h->stream));
        h->stage_ev.push_back(e);
        h->stage_of.push_back(stage);
      }
      HIP_TRY(launch_stage(stage, p, h->stream, side, h->ev_fork, h->ev_join));
    }
  }
  if (overlap && k > 0) HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_bed, 0));
  if (h->phases) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(e, h->stream));
    h->stage_ev.push_back(e);
    h->stage_of.push_back(-1);
  }
  return TOMO_OK;
}

// Device copy of a host array (hipErrorInvalidValue-free for empty arrays: they hold one
// element).  The type of *dst may differ from the element type (same bytes).
template <typename D, typename T>
hipError_t upload(D **dst, const std::vector<T> &v) {
  *dst = nullptr;
  const size_t bytes = v.size() * sizeof(T);
  hipError_t e = hipMalloc((void **)dst, bytes);
  if (e != hipSuccess) {
    *dst = nullptr;
    return e;
  }
  return hipMemcpy(*dst, v.data(), bytes, hipMemcpyHostToDevice);
}

void drop_mesh(tomo_handle *h) {
  if (h->d_tidx) (void)hipFree(h->d_tidx);
  if (h->d_tpos) (void)hipFree(h->d_tpos);
  if (h->d_vaos) (void)hipFree(h->d_vaos);
  if (h->d_chunk) (void)hipFree(h->d_chunk);
  if (h->d_soa) (void)hipFree(h->d_soa);
  h->d_tidx = nullptr; h->d_tpos = nullptr; h->d_vaos = nullptr; h->d_chunk = nullptr; h->d_soa = nullptr;
  h->n_tri = h->n_vtx = h->n_chunks = 0;
}

// Orientations whose overflow pool or spill list filled up (F_RETRY in d, the records of
// orientations [begin, begin + d.size())) are swept again one at a time in a private slot table
// whose pool and spill list are 4x larger, then 16x and 64x while some still overflow.  The
// handle's own scratch (possibly shared with a batch) is not touched.
int retry_overflowed(tomo_handle *h, int32_t begin, std::vector<DevRecord> &d) {
  std::vector<int32_t> idx;
  for (size_t i = 0; i < d.size(); i++)
    if (d[i].flags & F_RETRY) idx.push_back((int32_t)i);
  for (int32_t scale = 4; scale <= 64 && !idx.empty(); scale *= 4) {
    const ScratchLayout lay = make_layout(h->opt.grid, h->lay.inl, scale);
    uint8_t *scr = nullptr;
    OrientHdr *hdr = nullptr;
    if (hipMalloc(&scr, (size_t)lay.bytes) != hipSuccess || hipMalloc(&hdr, sizeof(OrientHdr)) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(scr);
      return fail(TOMO_E_OOM, "slot table for an overflowing orientation");
    }
    hipError_t e = hipSuccess;
    for (int32_t i : idx) {
      SweepParams p = base_params(h);
      p.lay = lay;
      p.scratch = scr;
      p.hdr = hdr;
      p.n_batch = 1;
      p.rot = h->d_rot + (size_t)(begin + i) * 9;
      p.rec = h->d_rec + begin + i;
      for (int stage = 0; stage < 6 && e == hipSuccess; stage++) e = launch_stage(stage, p, h->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(&d[(size_t)i], p.rec, sizeof(DevRecord), hipMemcpyDeviceToHost, h->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(scr);
    (void)hipFree(hdr);
    if (e != hipSuccess) return fail(TOMO_E_HIP, std::string("overflow re-sweep: ") + hipGetErrorString(e));
    std::vector<int32_t> still;
    for (int32_t i : idx)
      if (d[(size_t)i].flags & F_RETRY) still.push_back(i);
    idx.swap(still);
  }
  return TOMO_OK;
}

void drop_orientations(tomo_handle *h) {
  if (h->d_rec) (void)hipFree(h->d_rec);
  if (h->d_rot) (void)hipFree(h->d_rot);
  h->d_rec = nullptr;
  h->d_rot = nullptr;
  h->rec_cap = 0;
  h->n_orient = 0;
  h->rot.clear();
}

}  // namespace

extern "C" {

const char *tomo_last_error(void) { return g_err.c_str(); }

int32_t tomo_find_first_optimal(int32_t n, const float *mss) {
  if (n <= 1 || !mss) return 0;
  int32_t best = 0;
  float mv = 1e5f;
  for (int32_t i = 0; i < n; i++)
    if (mss[i] < mv) {
      mv = mss[i];
      best = i;
    }
  return best;
}

int tomo_create(const float *f32, const int32_t *i32, int32_t grid, tomo_handle **out) {
  if (!out) return fail(TOMO_E_INVALID_ARG, "null output handle");
  *out = nullptr;
  Options o;
  int rc = decode_options(f32, i32, grid, o);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TOMO_E_NO_DEVICE, "no HIP device");
  tomo_handle *h = new (std::nothrow) tomo_handle();
  if (!h) return fail(TOMO_E_OOM, "handle allocation failed");
  h->opt = o;
  if (hipGetDevice(&h->device) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_pair, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_bed, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->ev_aabb, hipEventDisableTiming) != hipSuccess) {
    tomo_destroy(h);
    return fail(TOMO_E_HIP, "stream/event creation failed");
  }
  h->lay = make_layout(o.grid, kInlineDefault);
  // tomo_run_static's array on this (current) device; a failed lookup only disables that entry.
  h->static_err = static_scratch_ptr(h->device, &h->d_static);
  if (h->static_err != hipSuccess) {
    (void)hipGetLastError();
    h->d_static = nullptr;
  }
  // IsBedCandidate distances dist2D (STomoSh_INT3.cpp:1326-1329) for dx, dy in [-r, r).
  const int r = o.bed_r > 0 ? o.bed_r : 1;
  std::vector<float> dist((size_t)(2 * r) * (2 * r));
  for (int dx = -r; dx < r; dx++)
    for (int dy = -r; dy < r; dy++) {
      const float fx = (float)dx, fy = (float)dy;
      dist[(size_t)(dx + r) * (2 * r) + (dy + r)] = sqrtf(fx * fx + fy * fy);
    }
  if (hipMalloc(&h->d_bed_dist, dist.size() * sizeof(float)) != hipSuccess ||
      hipMemcpy(h->d_bed_dist, dist.data(), dist.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    tomo_destroy(h);
    return fail(TOMO_E_HIP, "bed table upload failed");
  }
  // Calculate adds sumType(Bed) * BedThickness slot by slot (STomoSh_INT3.cpp:1242): the float
  // total after k bed rows is a k-fold running sum.
  const size_t g2 = (size_t)o.grid * o.grid;
  h->vbed_seq.resize(g2 + 1);
  float acc = 0.0f;
  h->vbed_seq[0] = 0.0f;
  for (size_t k = 1; k <= g2; k++) {
    acc += (float)1 * o.bed_thick;
    h->vbed_seq[k] = acc;
  }
  *out = h;
  return TOMO_OK;
}

int tomo_set_mesh(tomo_handle *h, int32_t n_tri, const int32_t *tri, int32_t n_vtx,
                  const float *vtx, const float *vtx_nrm) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  if (n_tri < 0 || n_vtx < 0 || (n_tri > 0 && (!tri || !vtx || !vtx_nrm)) || (n_vtx > 0 && (!vtx || !vtx_nrm)))
    return fail(TOMO_E_INVALID_ARG, "bad mesh arguments");
  for (int64_t i = 0; i < (int64_t)n_tri * 3; i++)
    if (tri[i] < 0 || tri[i] >= n_vtx) return fail(TOMO_E_INVALID_ARG, "triangle index out of range");
  if (n_tri >= (1 << 30)) return fail(TOMO_E_INVALID_ARG, "too many triangles");
  HIP_TRY(hipStreamSynchronize(h->stream));
  drop_mesh(h);  // a failure below leaves the handle without a mesh (tomo_run: TOMO_E_STATE)
  // raw AABB centre (STomoAABB3Df.cpp:42-72): float add, then *0.5 in double
  float mn[3] = {1e5f, 1e5f, 1e5f}, mx[3] = {-1e5f, -1e5f, -1e5f};
  for (int32_t v = 0; v < n_vtx; v++)
    for (int k = 0; k < 3; k++) {
      const float c = vtx[(size_t)v * 3 + k];
      mn[k] = (mn[k] <= c) ? mn[k] : c;
      mx[k] = (mx[k] >= c) ? mx[k] : c;
    }
  for (int k = 0; k < 3; k++) h->center[k] = (float)((double)(mn[k] + mx[k]) * 0.5);
  std::vector<float> soa((size_t)3 * (n_vtx > 0 ? n_vtx : 1));
  for (int32_t v = 0; v < n_vtx; v++)
    for (int k = 0; k < 3; k++) soa[(size_t)k * n_vtx + v] = vtx[(size_t)v * 3 + k];
  // Culling order: triangles sorted by the Morton code of their centroid, so each run of kChunk
  // triangles is compact in space and its AABB bounds where the rasteriser can touch it.
  std::vector<uint64_t> code((size_t)n_tri);
  {
    float span = 0.0f;
    for (int k = 0; k < 3; k++) span = std::max(span, mx[k] - mn[k]);
    const double scale = span > 0.0f ? 1023.0 / (double)span : 0.0;
    auto spread = [](uint64_t v) {  // 10 bits -> every third bit
      v &= 1023u;
      v = (v | (v << 16)) & 0x030000FFull;
      v = (v | (v << 8)) & 0x0300F00Full;
      v = (v | (v << 4)) & 0x030C30C3ull;
      v = (v | (v << 2)) & 0x09249249ull;
      return v;
    };
    for (int32_t t = 0; t < n_tri; t++) {
      uint64_t q[3];
      for (int k = 0; k < 3; k++) {
        double c = 0.0;
        for (int j = 0; j < 3; j++) c += vtx[(size_t)tri[(size_t)t * 3 + j] * 3 + k];
        double u = (c / 3.0 - (double)mn[k]) * scale;
        u = u < 0.0 ? 0.0 : (u > 1023.0 ? 1023.0 : u);
        q[k] = (uint64_t)u;
      }
      code[(size_t)t] = ((spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2)) << 32) | (uint64_t)t;
    }
    std::sort(code.begin(), code.end());
  }
  std::vector<int32_t> ord_tri((size_t)3 * (n_tri > 0 ? n_tri : 1)), ord_id((size_t)(n_tri > 0 ? n_tri : 1));
  for (int32_t i = 0; i < n_tri; i++) {
    const int32_t t = (int32_t)(code[(size_t)i] & 0xffffffffu);
    ord_id[(size_t)i] = t;
    for (int j = 0; j < 3; j++) ord_tri[(size_t)i * 3 + j] = tri[(size_t)t * 3 + j];
  }
  const int32_t n_chunks = (n_tri + kChunk - 1) / kChunk;
  std::vector<float> box((size_t)6 * (n_chunks > 0 ? n_chunks : 1));
  for (int32_t c = 0; c < n_chunks; c++) {
    float b[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
    for (int32_t i = c * kChunk; i < std::min(n_tri, (c + 1) * kChunk); i++)
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) {
          const float v = vtx[(size_t)ord_tri[(size_t)i * 3 + j] * 3 + k];
          b[k] = std::min(b[k], v);
          b[3 + k] = std::max(b[3 + k], v);
        }
    for (int k = 0; k < 6; k++) box[(size_t)c * 6 + k] = b[k];
  }
  // One 16-byte index record per triangle in culling order (its three vertex indices and the
  // caller's index) and two float4 per vertex (position relative to the rotation centre, formed
  // by the float subtraction the rotation starts with, and normal): 32 bytes per vertex plus 16
  // per triangle stay resident in an XCD's L2 while every orientation reads them.
  std::vector<int32_t> tix((size_t)4 * (n_tri > 0 ? n_tri : 1), 0);
  std::vector<int32_t> tpos((size_t)(n_tri > 0 ? n_tri : 1), 0);
  for (int32_t i = 0; i < n_tri; i++) {
    for (int j = 0; j < 3; j++) tix[(size_t)i * 4 + j] = ord_tri[(size_t)i * 3 + j];
    tix[(size_t)i * 4 + 3] = ord_id[(size_t)i];
    tpos[(size_t)ord_id[(size_t)i]] = i;
  }
  std::vector<float> aos((size_t)8 * (n_vtx > 0 ? n_vtx : 1), 0.0f);
  for (int32_t v = 0; v < n_vtx; v++) {
    for (int k = 0; k < 3; k++) aos[(size_t)v * 8 + k] = vtx[(size_t)v * 3 + k] - h->center[k];
    for (int k = 0; k < 3; k++) aos[(size_t)v * 8 + 3 + k] = vtx_nrm[(size_t)v * 3 + k];
  }
  float *d_soa = nullptr, *d_chunk = nullptr;
  int4 *d_tidx = nullptr;
  int32_t *d_tpos = nullptr;
  float4 *d_vaos = nullptr;
  hipError_t e = upload(&d_soa, soa);
  if (e == hipSuccess) e = upload(&d_tidx, tix);
  if (e == hipSuccess) e = upload(&d_tpos, tpos);
  if (e == hipSuccess) e = upload(&d_vaos, aos);
  if (e == hipSuccess) e = upload(&d_chunk, box);
  if (e != hipSuccess) {
    (void)hipFree(d_soa); (void)hipFree(d_tidx); (void)hipFree(d_tpos);
    (void)hipFree(d_vaos); (void)hipFree(d_chunk);
    return fail(e == hipErrorOutOfMemory ? TOMO_E_OOM : TOMO_E_HIP,
                std::string("mesh upload: ") + hipGetErrorString(e));
  }
  h->d_soa = d_soa; h->d_tidx = d_tidx; h->d_tpos = d_tpos; h->d_vaos = d_vaos; h->d_chunk = d_chunk;
  h->n_tri = n_tri;
  h->n_vtx = n_vtx;
  h->n_chunks = n_chunks;
  return TOMO_OK;
}

int tomo_set_orientations(tomo_handle *h, int32_t n, const float *ypr) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  if (n < 0 || (n > 0 && !ypr)) return fail(TOMO_E_INVALID_ARG, "bad orientation arguments");
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->n_orient = 0;  // a failure below leaves no orientations (tomo_run: range error)
  std::vector<float> rot((size_t)9 * (n > 0 ? n : 1), 0.0f);
  for (int32_t i = 0; i < n; i++)
    ypr_matrix(ypr[(size_t)i * 3 + 0], ypr[(size_t)i * 3 + 1], ypr[(size_t)i * 3 + 2], &rot[(size_t)i * 9]);
  const int32_t need = n > 0 ? n : 1;
  if (need > h->rec_cap || !h->d_rec || !h->d_rot) {
    drop_orientations(h);
    if (hipMalloc(&h->d_rec, sizeof(DevRecord) * (size_t)need) != hipSuccess ||
        hipMalloc(&h->d_rot, sizeof(float) * 9 * (size_t)need) != hipSuccess) {
      (void)hipGetLastError();
      drop_orientations(h);
      return fail(TOMO_E_OOM, "orientation buffers");
    }
    h->rec_cap = need;
  }
  if (n > 0 && hipMemcpy(h->d_rot, rot.data(), sizeof(float) * 9 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
    drop_orientations(h);
    return fail(TOMO_E_HIP, "rotation upload failed");
  }
  h->rot.swap(rot);
  h->n_orient = n;
  return TOMO_OK;
}

int tomo_set_launch(tomo_handle *h, int32_t batch) {
  if (!h || batch < 0) return fail(TOMO_E_INVALID_ARG, "bad launch arguments");
  h->batch_req = batch;
  return TOMO_OK;
}

int tomo_set_tuning(tomo_handle *h, int32_t knob, int32_t value) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  switch (knob) {
    case TOMO_TUNE_BATCH:
      if (value < 0) return fail(TOMO_E_INVALID_ARG, "batch must be >= 0");
      h->batch_req = value;
      return TOMO_OK;
    case TOMO_TUNE_TILE_Y:
      if (value != 0 && (value < 32 || value > 512 || (value & (value - 1))))
        return fail(TOMO_E_INVALID_ARG, "tile extent along y must be 0 or a power of two in 32..512");
      h->tile_y = value;
      return TOMO_OK;
    case TOMO_TUNE_INLINE:
      if (value < kSmall || value > kInlineMax) return fail(TOMO_E_INVALID_ARG, "inline slot capacity must be 8..32");
      if (value != h->lay.inl) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->lay = make_layout(h->opt.grid, value);
      }
      return TOMO_OK;
    case TOMO_TUNE_EXACT_FSUM:
      if (value != 0 && value != 1) return fail(TOMO_E_INVALID_ARG, "exact float totals must be 0 or 1");
      h->exact_fsum = value;
      return TOMO_OK;
    default:
      return fail(TOMO_E_INVALID_ARG, "unknown tuning knob");
  }
}

int tomo_run(tomo_handle *h, int32_t begin, int32_t end) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  if (begin < 0 || end > h->n_orient || begin > end) return fail(TOMO_E_INVALID_ARG, "bad orientation range");
  if (!h->d_soa || !h->d_tidx || !h->d_tpos || !h->d_vaos || !h->d_chunk) return fail(TOMO_E_STATE, "no mesh");
  if (!h->d_rot || !h->d_rec) return fail(TOMO_E_STATE, "no orientations");
  h->last_begin = begin;
  h->last_end = end;
  h->timed = false;
  h->last_static = false;
  const int32_t n = end - begin;
  if (n == 0) return TOMO_OK;
  const int32_t batch = auto_batch(h, n);
  int rc = ensure_scratch(h, batch);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h->stream));
  free_stage_events(h);
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  rc = enqueue(h, begin, end, batch, nullptr, 0);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  h->timed = true;
  h->last_batch = batch;
  return TOMO_OK;
}

int tomo_run_static(tomo_handle *h, int32_t begin, int32_t end) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  if (begin < 0 || end > h->n_orient || begin > end) return fail(TOMO_E_INVALID_ARG, "bad orientation range");
  if (!h->d_soa || !h->d_tidx || !h->d_tpos || !h->d_vaos || !h->d_chunk) return fail(TOMO_E_STATE, "no mesh");
  if (!h->d_rot || !h->d_rec) return fail(TOMO_E_STATE, "no orientations");
  if (h->shared) return fail(TOMO_E_STATE, "a batch member runs through tomo_batch_launch");
  if (!h->d_static)
    return fail(TOMO_E_STATE, std::string("static slot-table array unavailable: ") + hipGetErrorString(h->static_err));
  // Headers first (256-byte aligned), then one slot-table region per orientation in flight.
  const size_t region = (size_t)h->lay.bytes, cap = kStaticScratchBytes;
  auto fits = [&](size_t b) { return ((b * sizeof(OrientHdr) + 255) & ~(size_t)255) + b * region <= cap; };
  size_t batch = cap / (region + sizeof(OrientHdr));
  while (batch > 0 && !fits(batch)) batch--;
  if (batch == 0)
    return fail(TOMO_E_INVALID_ARG, "one orientation's slot table (" + std::to_string(region) + " bytes at grid " +
                                        std::to_string(h->opt.grid) + ", " + std::to_string(h->lay.inl) +
                                        " inline entries) exceeds the static array (" + std::to_string(cap) +
                                        " bytes)");
  h->last_begin = begin;
  h->last_end = end;
  h->timed = false;
  h->last_static = true;
  const int32_t n = end - begin;
  if (n == 0) return TOMO_OK;
  const int32_t b = (int32_t)std::min<size_t>(batch, (size_t)n);
  HIP_TRY(hipStreamSynchronize(h->stream));
  const bool was = h->phases;  // per-stage timing creates events: off here
  h->phases = false;
  OrientHdr *hdr = (OrientHdr *)h->d_static;
  uint8_t *scr = h->d_static + ((b * sizeof(OrientHdr) + 255) & ~(size_t)255);
  hipError_t e = hipEventRecord(h->ev0, h->stream);
  int rc = e == hipSuccess ? enqueue(h, begin, end, b, nullptr, 0, nullptr, scr, hdr) : TOMO_OK;
  if (e == hipSuccess && rc == TOMO_OK) e = hipEventRecord(h->ev1, h->stream);
  h->phases = was;
  if (rc) return rc;
  if (e != hipSuccess) return fail(TOMO_E_HIP, std::string("tomo_run_static: ") + hipGetErrorString(e));
  h->timed = true;
  h->last_batch = b;
  return TOMO_OK;
}

int tomo_sync(tomo_handle *h) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return TOMO_OK;
}

int tomo_get_records(tomo_handle *h, int32_t begin, int32_t end, tomo_record *out) {
  if (!h || (!out && end > begin)) return fail(TOMO_E_INVALID_ARG, "null argument");
  if (begin < 0 || end > h->n_orient || begin > end) return fail(TOMO_E_INVALID_ARG, "bad orientation range");
  const int32_t n = end - begin;
  if (n == 0) return TOMO_OK;
  std::vector<DevRecord> d((size_t)n);
  HIP_TRY(hipMemcpyAsync(d.data(), h->d_rec + begin, sizeof(DevRecord) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (!h->last_static) {  // tomo_run_static allocates nothing, so its overflows stay unresolved
    const int rc = retry_overflowed(h, begin, d);
    if (rc) return rc;
  }
  for (int32_t i = 0; i < n; i++) finish_record(h, d[(size_t)i], &h->rot[(size_t)(begin + i) * 9], out[i]);
  return TOMO_OK;
}

int tomo_sweep(tomo_handle *h, int32_t n, const float *ypr, tomo_record *out) {
  int rc = tomo_set_orientations(h, n, ypr);
  if (rc) return rc;
  rc = tomo_run(h, 0, n);
  if (rc) return rc;
  return tomo_get_records(h, 0, n, out);
}

float tomo_last_kernel_ms(tomo_handle *h) {
  if (!h || !h->timed) return -1.0f;
  float ms = -1.0f;
  if (hipEventSynchronize(h->ev1) != hipSuccess) return -1.0f;
  if (hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) return -1.0f;
  return ms;
}

int tomo_get_info(tomo_handle *h, int32_t *grid, int32_t *batch, int64_t *scratch_bytes) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  if (grid) *grid = h->opt.grid;
  if (batch) *batch = h->last_batch;
  if (scratch_bytes) *scratch_bytes = (int64_t)h->lay.bytes * h->scratch_batch;
  return TOMO_OK;
}

// One orientation with per-slot dumps and optionally the final rows of every slot.
static int debug_run(tomo_handle *h, const float ypr[3], tomo_record *rec, int32_t *cnt_pre,
                     int32_t *cnt_final, int32_t *slot_vo, int32_t *slot_vss, uint64_t *rows) {
  if (!h || !ypr || !rec) return fail(TOMO_E_INVALID_ARG, "null argument");
  if (!h->d_soa || !h->d_tidx) return fail(TOMO_E_STATE, "no mesh");
  int rc = tomo_set_orientations(h, 1, ypr);
  if (rc) return rc;
  const size_t g2 = (size_t)h->opt.grid * h->opt.grid;
  int32_t *d = nullptr;
  uint64_t *drows = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(int32_t) * 4 * g2));
  hipError_t e = rows ? hipMalloc(&drows, sizeof(uint64_t) * 64 * g2) : hipSuccess;
  const bool was = h->phases;
  h->phases = false;
  // The dumps must come from the pass whose record is reported: an orientation that overflows
  // the pool or the spill list is dumped again with larger ones.
  const ScratchLayout lay0 = h->lay;
  for (int32_t scale = 1; scale <= 64 && e == hipSuccess && rc == TOMO_OK; scale *= 4) {
    if (scale > 1) h->lay = make_layout(h->opt.grid, lay0.inl, scale);
    rc = ensure_scratch(h, 1);
    if (rc) break;
    e = hipMemsetAsync(d, 0, sizeof(int32_t) * 4 * g2, h->stream);
    if (e == hipSuccess && drows) e = hipMemsetAsync(drows, 0, sizeof(uint64_t) * 64 * g2, h->stream);
    if (e != hipSuccess) break;
    rc = enqueue(h, 0, 1, 1, d, g2, drows);
    if (rc) break;
    DevRecord dr;
    e = hipMemcpyAsync(&dr, h->d_rec, sizeof(DevRecord), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess || !(dr.flags & F_RETRY)) break;
  }
  h->lay = lay0;
  h->phases = was;
  int32_t *outs[4] = {cnt_pre, cnt_final, slot_vo, slot_vss};
  for (int k = 0; k < 4 && e == hipSuccess && rc == TOMO_OK; k++)
    if (outs[k]) e = hipMemcpy(outs[k], d + k * g2, sizeof(int32_t) * g2, hipMemcpyDeviceToHost);
  if (e == hipSuccess && rc == TOMO_OK && rows)
    e = hipMemcpy(rows, drows, sizeof(uint64_t) * 64 * g2, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (drows) (void)hipFree(drows);
  if (rc) return rc;
  if (e != hipSuccess) return fail(TOMO_E_HIP, std::string("debug sweep: ") + hipGetErrorString(e));
  h->last_begin = 0;
  h->last_end = 1;
  return tomo_get_records(h, 0, 1, rec);
}

int tomo_debug_slots(tomo_handle *h, const float ypr[3], tomo_record *rec, int32_t *cnt_pre,
                     int32_t *cnt_final, int32_t *slot_vo, int32_t *slot_vss) {
  return debug_run(h, ypr, rec, cnt_pre, cnt_final, slot_vo, slot_vss, nullptr);
}

int tomo_debug_rows(tomo_handle *h, const float ypr[3], tomo_record *rec, uint64_t *rows, int32_t *nrows) {
  if (!rows || !nrows) return fail(TOMO_E_INVALID_ARG, "null argument");
  return debug_run(h, ypr, rec, nullptr, nrows, nullptr, nullptr, rows);
}

int tomo_set_phase_profile(tomo_handle *h, int32_t enable) {
  if (!h) return fail(TOMO_E_INVALID_ARG, "null handle");
  h->phases = enable != 0;
  return TOMO_OK;
}

int tomo_phase_profile(tomo_handle *h, double *ms8) {
  if (!h || !ms8) return fail(TOMO_E_INVALID_ARG, "null argument");
  if (h->stage_ev.size() < 2) return fail(TOMO_E_STATE, "no stage timing recorded");
  HIP_TRY(hipStreamSynchronize(h->stream));
  for (int k = 0; k < 8; k++) ms8[k] = 0.0;
  for (size_t i = 0; i + 1 < h->stage_ev.size(); i++) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, h->stage_ev[i], h->stage_ev[i + 1]));
    const int st = h->stage_of[i];
    if (st >= 0 && st < 8) ms8[st] += ms;
  }
  return TOMO_OK;
}

int tomo_destroy(tomo_handle *h) {
  if (!h) return TOMO_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_stage_events(h);
  if (h->shared) {  // buffers and stream belong to the batch
    h->d_hdr = nullptr;
    h->d_scratch = nullptr;
    h->stream = h->own_stream;
  }
  if (h->d_hdr) (void)hipFree(h->d_hdr);
  if (h->d_scratch) scratch_release(h->d_scratch, h->scratch_cap, h->device);
  if (h->d_rec) (void)hipFree(h->d_rec);
  if (h->d_rot) (void)hipFree(h->d_rot);
  drop_mesh(h);
  if (h->d_bed_dist) (void)hipFree(h->d_bed_dist);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->ev_pair) (void)hipEventDestroy(h->ev_pair);
  if (h->ev_bed) (void)hipEventDestroy(h->ev_bed);
  if (h->ev_aabb) (void)hipEventDestroy(h->ev_aabb);
  if (h->side) {
    (void)hipStreamSynchronize(h->side);
    (void)hipStreamDestroy(h->side);
  }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return TOMO_OK;
}


// ---------------------------------------------------------------------------------------------
// Batched sweeps: many meshes, each with its own orientations and options, in one captured
// hipGraph launch on one stream.  Members keep their own mesh, rotations and records; they share
// the batch's stream, slot-table scratch and headers, sized for the largest member.  A member
// lays the regions out for its own grid and inline capacity over the previous member's data:
// every table a sweep reads is written by that sweep first (the rasteriser writes all counters).

struct tomo_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<tomo_handle *> members;
  struct Tune {
    int32_t batch = 0, tile_y = 0, inl = kInlineDefault;
    float us_per_orientation = -1.0f;
    int32_t tuned = 0;
  };
  std::vector<Tune> tune;
  uint8_t *d_scratch = nullptr;
  size_t scratch_cap = 0;      // bytes of the block behind d_scratch
  OrientHdr *d_hdr = nullptr;
  size_t scratch_bytes = 0, hdr_count = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool timed = false;
  int32_t mode = TOMO_BATCH_QUEUE;
  // persistent launch: every unit's parameters, the per-round per-kernel item tables and the
  // queue counters (device), and the number of rounds captured
  SweepParams *d_units = nullptr;
  int32_t *d_first = nullptr;
  uint32_t *d_heads = nullptr;
  int32_t rounds = 0, units = 0;
};

static void batch_drop_graph(tomo_batch *b) {
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}

static void release_private_scratch(tomo_handle *h) {
  (void)hipStreamSynchronize(h->stream);
  if (h->d_scratch) { scratch_release(h->d_scratch, h->scratch_cap, h->device); h->d_scratch = nullptr; }
  if (h->d_hdr) { (void)hipFree(h->d_hdr); h->d_hdr = nullptr; }
  h->scratch_batch = 0;
  h->scratch_region = 0;
}

// Detach members from the shared buffers (before re-tuning or destruction).
static void batch_unshare(tomo_batch *b) {
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (tomo_handle *h : b->members)
    if (h->shared) {
      h->shared = false;
      h->stream = h->own_stream;
      h->d_scratch = nullptr;
      h->d_hdr = nullptr;
      h->scratch_batch = 0;
      h->scratch_region = 0;
    }
  if (b->d_scratch) { scratch_release(b->d_scratch, b->scratch_cap, b->device); b->d_scratch = nullptr; }
  if (b->d_hdr) { (void)hipFree(b->d_hdr); b->d_hdr = nullptr; }
  if (b->d_units) { (void)hipFree(b->d_units); b->d_units = nullptr; }
  if (b->d_first) { (void)hipFree(b->d_first); b->d_first = nullptr; }
  if (b->d_heads) { (void)hipFree(b->d_heads); b->d_heads = nullptr; }
  b->scratch_bytes = b->hdr_count = 0;
  b->rounds = b->units = 0;
}

int tomo_batch_create(tomo_batch **out) {
  if (!out) return fail(TOMO_E_INVALID_ARG, "null output handle");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TOMO_E_NO_DEVICE, "no HIP device");
  tomo_batch *b = new (std::nothrow) tomo_batch();
  if (!b) return fail(TOMO_E_OOM, "batch allocation failed");
  if (hipGetDevice(&b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    tomo_batch_destroy(b);
    return fail(TOMO_E_HIP, "stream/event creation failed");
  }
  *out = b;
  return TOMO_OK;
}

int tomo_batch_add(tomo_batch *b, const float *f32, const int32_t *i32, int32_t grid, int32_t n_tri,
                   const int32_t *tri, int32_t n_vtx, const float *vtx, const float *vtx_nrm, int32_t n_ypr,
                   const float *ypr, int32_t *member) {
  if (!b) return fail(TOMO_E_INVALID_ARG, "null batch");
  if (n_ypr < 1) return fail(TOMO_E_INVALID_ARG, "a member needs at least one orientation");
  batch_unshare(b);
  tomo_handle *h = nullptr;
  int rc = tomo_create(f32, i32, grid, &h);
  if (rc) return rc;
  rc = tomo_set_mesh(h, n_tri, tri, n_vtx, vtx, vtx_nrm);
  if (!rc) rc = tomo_set_orientations(h, n_ypr, ypr);
  if (rc) {
    tomo_destroy(h);
    return rc;
  }
  h->own_stream = h->stream;
  b->members.push_back(h);
  tomo_batch::Tune t;
  t.batch = n_ypr < 2048 ? n_ypr : 2048;
  b->tune.push_back(t);
  if (member) *member = (int32_t)b->members.size() - 1;
  return TOMO_OK;
}

static void apply_tune(tomo_handle *h, const tomo_batch::Tune &t) {
  h->tile_y = t.tile_y;
  h->batch_req = t.batch;
  if (h->lay.inl != t.inl) h->lay = make_layout(h->opt.grid, t.inl);
}

// Time one launch shape on member h over orientations [b0, b0 + n): microseconds per
// orientation, or a negative value if the shape cannot run or leaves a slot unresolved.
static float time_shape(tomo_handle *h, const tomo_batch::Tune &t, int32_t b0, int32_t n) {
  apply_tune(h, t);
  if (tomo_run(h, b0, b0 + n) != TOMO_OK || tomo_sync(h) != TOMO_OK) return -1.0f;  // warm-up
  if (tomo_run(h, b0, b0 + n) != TOMO_OK || tomo_sync(h) != TOMO_OK) return -1.0f;
  const float ms = tomo_last_kernel_ms(h);
  std::vector<DevRecord> d((size_t)n);
  if (hipMemcpy(d.data(), h->d_rec + b0, sizeof(DevRecord) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
    return -1.0f;
  for (const DevRecord &r : d)
    if (r.flags & (TOMO_F_UNRESOLVED | F_RETRY)) return -1.0f;
  return ms < 0.0f ? -1.0f : ms * 1000.0f / (float)n;
}

// Autotuner.  Members are grouped by (grid, ceil(log2(nTri))); each group's largest member is
// timed on `sample` orientations from the middle of its list with a coordinate descent over
// rasteriser tile extent along y {32, 64, 128, 256}, inline slot capacity {8, 16, 24} and
// orientations in flight {1024, 2048, 4096, 8192} (within a quarter of free memory); shapes that leave
// a slot unresolved are rejected.  The group's winner applies to all its members.
int tomo_batch_tune(tomo_batch *b, int32_t sample) {
  if (!b) return fail(TOMO_E_INVALID_ARG, "null batch");
  if (sample < 256) sample = 256;
  batch_unshare(b);
  const size_t M = b->members.size();
  std::vector<int64_t> key(M);
  for (size_t m = 0; m < M; m++) {
    const tomo_handle *h = b->members[m];
    int lg = 0;
    while ((1 << lg) < (h->n_tri > 1 ? h->n_tri : 1)) lg++;
    key[m] = (int64_t)h->opt.grid * 64 + lg;
  }
  std::vector<bool> done(M, false);
  for (size_t m0 = 0; m0 < M; m0++) {
    if (done[m0]) continue;
    size_t rep = m0;
    for (size_t m = m0; m < M; m++)
      if (key[m] == key[m0] && b->members[m]->n_tri > b->members[rep]->n_tri) rep = m;
    tomo_handle *h = b->members[rep];
    const int32_t n = h->n_orient;
    const int32_t ns = n < sample ? n : sample;
    const int32_t b0 = (n - ns) / 2;
    tomo_batch::Tune best;
    best.batch = ns < 2048 ? ns : 2048;
    float best_us = time_shape(h, best, b0, ns);
    auto consider = [&](tomo_batch::Tune t) {
      const float us = time_shape(h, t, b0, ns);
      if (us > 0.0f && (best_us < 0.0f || us < best_us)) {
        best_us = us;
        best = t;
      }
    };
    const int32_t ty0 = h->opt.grid > 256 ? kTileYDefault / 2 : kTileYDefault;  // the default
    for (int32_t ty : {32, 64, 128, 256}) {
      if (ty == ty0) continue;
      tomo_batch::Tune t = best;
      t.tile_y = ty;
      consider(t);
    }
    for (int32_t inl : {8, 24}) {
      tomo_batch::Tune t = best;
      t.inl = inl;
      consider(t);
    }
    size_t freeb = 0, total = 0;
    (void)hipMemGetInfo(&freeb, &total);
    for (int32_t bs : {1024, 4096, 8192}) {
      tomo_batch::Tune t = best;
      t.batch = bs < ns ? bs : ns;
      const size_t need = (size_t)make_layout(h->opt.grid, t.inl).bytes * (size_t)t.batch;
      if (t.batch != best.batch && need < freeb / 4) consider(t);
    }
    release_private_scratch(h);
    if (best_us < 0.0f) return fail(TOMO_E_STATE, "no launch shape ran for a batch member");
    best.us_per_orientation = best_us;
    best.tuned = 1;
    for (size_t m = m0; m < M; m++)
      if (key[m] == key[m0]) {
        tomo_batch::Tune t = best;
        const int32_t nm = b->members[m]->n_orient;
        if (t.batch > nm) t.batch = nm;
        b->tune[m] = t;
        done[m] = true;
      }
  }
  return TOMO_OK;
}

int tomo_batch_get_tuning(tomo_batch *b, int32_t member, int32_t *batch, int32_t *tile_y, int32_t *inl,
                          float *us_per_orientation) {
  if (!b || member < 0 || member >= (int32_t)b->members.size()) return fail(TOMO_E_INVALID_ARG, "bad member");
  const tomo_batch::Tune &t = b->tune[(size_t)member];
  if (batch) *batch = t.batch;
  if (tile_y) *tile_y = t.tile_y;
  if (inl) *inl = t.inl;
  if (us_per_orientation) *us_per_orientation = t.us_per_orientation;
  return TOMO_OK;
}

int tomo_batch_set_tuning(tomo_batch *b, int32_t member, int32_t batch, int32_t tile_y, int32_t inl) {
  if (!b || member < 0 || member >= (int32_t)b->members.size()) return fail(TOMO_E_INVALID_ARG, "bad member");
  if (batch < 0) return fail(TOMO_E_INVALID_ARG, "batch must be >= 0");
  if (tile_y != 0 && (tile_y < 32 || tile_y > 512 || (tile_y & (tile_y - 1))))
    return fail(TOMO_E_INVALID_ARG, "tile extent along y must be 0 or a power of two in 32..512");
  if (inl < kSmall || inl > kInlineMax) return fail(TOMO_E_INVALID_ARG, "inline slot capacity must be 8..32");
  batch_unshare(b);
  const int32_t n = b->members[(size_t)member]->n_orient;
  tomo_batch::Tune t;
  t.batch = batch > 0 ? std::min(batch, n) : std::min(n, 2048);
  t.tile_y = tile_y;
  t.inl = inl;
  b->tune[(size_t)member] = t;
  return TOMO_OK;
}

// Attach the members to the batch's stream and shared buffers.
static void batch_share(tomo_batch *b) {
  for (size_t m = 0; m < b->members.size(); m++) {
    tomo_handle *h = b->members[m];
    h->shared = true;
    h->stream = b->stream;
    h->d_scratch = b->d_scratch;
    h->d_hdr = b->d_hdr;
    h->scratch_batch = b->tune[m].batch;
    h->scratch_region = h->lay.bytes;
    h->phases = false;
    free_stage_events(h);
  }
}

// Per-member capture: each member's sweep as its own chain of stage launches, member after
// member, every member's tables laid over the shared block from its start.
static int batch_capture_chain(tomo_batch *b) {
  size_t sbytes = 0, hcount = 0;
  for (size_t m = 0; m < b->members.size(); m++) {
    tomo_handle *h = b->members[m];
    const size_t bs = (size_t)b->tune[m].batch;
    sbytes = std::max(sbytes, (size_t)h->lay.bytes * bs);
    hcount = std::max(hcount, bs);
  }
  if (scratch_alloc(sbytes, b->device, &b->d_scratch, &b->scratch_cap) != hipSuccess) {
    b->d_scratch = nullptr;
    return fail(TOMO_E_OOM, "shared scratch allocation failed");
  }
  if (hipMalloc(&b->d_hdr, sizeof(OrientHdr) * hcount) != hipSuccess) {
    batch_unshare(b);
    return fail(TOMO_E_OOM, "shared buffer allocation failed");
  }
  b->scratch_bytes = sbytes;
  b->hdr_count = hcount;
  HIP_TRY(hipMemsetAsync(b->d_scratch, 0, sbytes, b->stream));
  HIP_TRY(hipMemsetAsync(b->d_hdr, 0, sizeof(OrientHdr) * hcount, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  batch_share(b);
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = TOMO_OK;
  for (size_t m = 0; m < b->members.size() && rc == TOMO_OK; m++) {
    tomo_handle *h = b->members[m];
    const int32_t bs = b->tune[m].batch;
    rc = enqueue(h, 0, h->n_orient, bs, nullptr, 0, nullptr);
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(TOMO_E_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
  b->graph = g;
  return TOMO_OK;
}

// Persistent capture.  Every member's orientations are cut into units of its launch shape's
// orientations in flight; consecutive units are packed into rounds whose slot tables lie side by
// side in the shared block (which is as large as the largest single unit needs, so no round asks
// for more memory than the per-member form).  Each round is swept by one persistent work-queue
// launch per stage kernel over all its units' items (mesh, orientation-tile), and the counters
// of every queue are cleared by one kernel at the start of the graph.
static int batch_capture_queue(tomo_batch *b) {
  struct Unit {
    size_t m;
    int32_t b0, n;
    size_t bytes;
  };
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  std::vector<Unit> units;
  size_t budget = 0;
  for (size_t m = 0; m < b->members.size(); m++) {
    tomo_handle *h = b->members[m];
    const int32_t bs = std::max<int32_t>(1, b->tune[m].batch);
    budget = std::max(budget, align((size_t)h->lay.bytes * (size_t)bs));
    for (int32_t b0 = 0; b0 < h->n_orient; b0 += bs) {
      const int32_t n = std::min(bs, h->n_orient - b0);
      units.push_back({m, b0, n, align((size_t)h->lay.bytes * (size_t)n)});
    }
  }
  std::vector<std::pair<size_t, size_t>> rounds;  // [first unit, end)
  size_t hmax = 0;
  for (size_t i = 0; i < units.size();) {
    size_t j = i, bytes = 0, hc = 0;
    while (j < units.size() && (j == i || bytes + units[j].bytes <= budget)) {
      bytes += units[j].bytes;
      hc += (size_t)units[j].n;
      j++;
    }
    rounds.push_back({i, j});
    hmax = std::max(hmax, hc);
    i = j;
  }
  // work-groups of each persistent kernel resident on the whole device
  int ncu = 0;
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
  int32_t resident[QK_COUNT];
  for (int k = 0; k < QK_COUNT; k++) {
    int per = 0;
    HIP_TRY(queue_occupancy(k, &per));
    resident[k] = std::max(8, (std::max(per, 1) * ncu) / 8 * 8);
  }
  if (scratch_alloc(budget, b->device, &b->d_scratch, &b->scratch_cap) != hipSuccess) {
    b->d_scratch = nullptr;
    return fail(TOMO_E_OOM, "shared scratch allocation failed");
  }
  const size_t nfirst = (units.size() + rounds.size()) * QK_COUNT;
  const size_t nheads = rounds.size() * QK_COUNT * 8;
  if (hipMalloc(&b->d_hdr, sizeof(OrientHdr) * hmax) != hipSuccess ||
      hipMalloc(&b->d_units, sizeof(SweepParams) * units.size()) != hipSuccess ||
      hipMalloc(&b->d_first, sizeof(int32_t) * nfirst) != hipSuccess ||
      hipMalloc(&b->d_heads, sizeof(uint32_t) * nheads) != hipSuccess) {
    (void)hipGetLastError();
    batch_unshare(b);
    return fail(TOMO_E_OOM, "shared buffer allocation failed");
  }
  b->scratch_bytes = budget;
  b->hdr_count = hmax;
  // unit parameters and, per round and kernel, the first item of every unit
  std::vector<SweepParams> P(units.size());
  std::vector<int32_t> F;
  F.reserve(nfirst);
  std::vector<size_t> foff(rounds.size() * QK_COUNT);
  std::vector<int32_t> items(rounds.size() * QK_COUNT);
  for (size_t r = 0; r < rounds.size(); r++) {
    size_t off = 0, hoff = 0;
    for (size_t i = rounds[r].first; i < rounds[r].second; i++) {
      const Unit &u = units[i];
      tomo_handle *h = b->members[u.m];
      SweepParams p = base_params(h);
      p.scratch = b->d_scratch + off;
      p.hdr = b->d_hdr + hoff;
      p.n_batch = u.n;
      p.rot = h->d_rot + (size_t)u.b0 * 9;
      p.rec = h->d_rec + u.b0;
      P[i] = p;
      off += u.bytes;
      hoff += (size_t)u.n;
    }
    for (int k = 0; k < QK_COUNT; k++) {
      foff[r * QK_COUNT + k] = F.size();
      int64_t acc = 0;
      F.push_back(0);
      for (size_t i = rounds[r].first; i < rounds[r].second; i++) {
        acc += queue_items(k, P[i]);
        if (acc > INT32_MAX) {
          batch_unshare(b);
          return fail(TOMO_E_INVALID_ARG, "too many work items in one round");
        }
        F.push_back((int32_t)acc);
      }
      items[r * QK_COUNT + k] = (int32_t)acc;
    }
  }
  HIP_TRY(hipMemcpy(b->d_units, P.data(), sizeof(SweepParams) * P.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->d_first, F.data(), sizeof(int32_t) * F.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemsetAsync(b->d_scratch, 0, budget, b->stream));
  HIP_TRY(hipMemsetAsync(b->d_hdr, 0, sizeof(OrientHdr) * hmax, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  batch_share(b);
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_clear_heads(b->d_heads, (int32_t)nheads, b->stream);
  for (size_t r = 0; r < rounds.size() && e == hipSuccess; r++) {
    const int32_t nu = (int32_t)(rounds[r].second - rounds[r].first);
    for (int k = 0; k < QK_COUNT && e == hipSuccess; k++) {
      const int32_t it = items[r * QK_COUNT + k];
      if (it == 0) continue;
      UnitQueue q;
      q.units = b->d_units + rounds[r].first;
      q.first = b->d_first + foff[r * QK_COUNT + k];
      q.n_units = nu;
      q.heads = b->d_heads + (r * QK_COUNT + (size_t)k) * 8;
      const int32_t groups = std::min(resident[k], (it + 7) / 8 * 8);
      e = launch_queue(k, q, groups, b->stream);
    }
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(TOMO_E_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
  }
  b->graph = g;
  b->rounds = (int32_t)rounds.size();
  b->units = (int32_t)units.size();
  return TOMO_OK;
}

int tomo_batch_capture(tomo_batch *b) {
  if (!b) return fail(TOMO_E_INVALID_ARG, "null batch");
  if (b->members.empty()) return fail(TOMO_E_STATE, "empty batch");
  batch_unshare(b);
  for (size_t m = 0; m < b->members.size(); m++) {
    tomo_handle *h = b->members[m];
    release_private_scratch(h);
    apply_tune(h, b->tune[m]);
  }
  const int rc = b->mode == TOMO_BATCH_QUEUE ? batch_capture_queue(b) : batch_capture_chain(b);
  if (rc) return rc;
  HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  for (tomo_handle *h : b->members) {
    h->last_begin = 0;
    h->last_end = h->n_orient;
    h->timed = false;
  }
  return TOMO_OK;
}

int tomo_batch_set_mode(tomo_batch *b, int32_t mode) {
  if (!b) return fail(TOMO_E_INVALID_ARG, "null batch");
  if (mode != TOMO_BATCH_QUEUE && mode != TOMO_BATCH_CHAIN) return fail(TOMO_E_INVALID_ARG, "bad batch mode");
  if (mode != b->mode) {
    batch_unshare(b);
    b->mode = mode;
  }
  return TOMO_OK;
}

int tomo_batch_get_plan(tomo_batch *b, int32_t *rounds, int32_t *units) {
  if (!b) return fail(TOMO_E_INVALID_ARG, "null batch");
  if (rounds) *rounds = b->rounds;
  if (units) *units = b->units;
  return TOMO_OK;
}

int tomo_batch_launch(tomo_batch *b) {
  if (!b) return fail(TOMO_E_INVALID_ARG, "null batch");
  if (!b->exec) {
    int rc = tomo_batch_capture(b);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(b->ev0, b->stream));
  HIP_TRY(hipGraphLaunch(b->exec, b->stream));
  HIP_TRY(hipEventRecord(b->ev1, b->stream));
  b->timed = true;
  return TOMO_OK;
}

int tomo_batch_sync(tomo_batch *b) {
  if (!b) return fail(TOMO_E_INVALID_ARG, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return TOMO_OK;
}

float tomo_batch_last_ms(tomo_batch *b) {
  if (!b || !b->timed) return -1.0f;
  float ms = -1.0f;
  if (hipEventSynchronize(b->ev1) != hipSuccess) return -1.0f;
  if (hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess) return -1.0f;
  return ms;
}

int32_t tomo_batch_size(tomo_batch *b) { return b ? (int32_t)b->members.size() : 0; }

int tomo_batch_records(tomo_batch *b, int32_t member, tomo_record *out) {
  if (!b || member < 0 || member >= (int32_t)b->members.size()) return fail(TOMO_E_INVALID_ARG, "bad member");
  HIP_TRY(hipStreamSynchronize(b->stream));
  tomo_handle *h = b->members[(size_t)member];
  return tomo_get_records(h, 0, h->n_orient, out);
}

int tomo_batch_destroy(tomo_batch *b) {
  if (!b) return TOMO_OK;
  batch_unshare(b);
  for (tomo_handle *h : b->members) tomo_destroy(h);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return TOMO_OK;
}

// ---------------------------------------------------------------------------------------------
// Legacy shim: global outputs, freed on the next call or OnDestroy (Tomo_Shell2026.cpp:19-87).

// Pixel groups of one orientation for rendering: slotsToPxls for every pixel type, then the
// support columns of GetSSPixels in place of the SS group (STomoSh_INT3.cpp:1103-1135,
// 1257-1318; Tomo_Shell2026.cpp:221-236).  Rows {x, y, z, nx, ny, nz} as int16.
static int render_export(tomo_handle *h, const float *ypr) {
  const int32_t G = h->opt.grid;
  const size_t g2 = (size_t)G * G;
  std::vector<uint64_t> rows(g2 * 64);
  std::vector<int32_t> nrows(g2);
  tomo_record r;
  int rc = tomo_debug_rows(h, ypr, &r, rows.data(), nrows.data());
  if (rc) return rc;
  for (auto &v : g_pix) v.clear();
  auto push = [](std::vector<int16_t> &dst, int x, int y, int z, int nz) {
    dst.push_back((int16_t)x); dst.push_back((int16_t)y); dst.push_back((int16_t)z);
    dst.push_back(0); dst.push_back(0); dst.push_back((int16_t)nz);
  };
  for (int x = 0; x < G; x++)
    for (int y = 0; y < G; y++) {
      const size_t s = (size_t)x * G + y;
      const uint64_t *rw = &rows[s * 64];
      for (int q = 0; q < nrows[s] && q < 64; q++) {
        const uint32_t type = (uint32_t)(rw[q] >> 32), key = (uint32_t)rw[q];
        for (int t = 0; t < 11; t++)
          if (type & (1u << t)) push(g_pix[(size_t)t], x, y, key_z(key), key_nz(key));
      }
    }
  std::vector<int16_t> &ss = g_pix[4];
  ss.clear();
  for (int x = 0; x < G; x++)
    for (int y = 0; y < G; y++) {
      const size_t s = (size_t)x * G + y;
      const uint64_t *rw = &rows[s * 64];
      int start = -1, end = -1;
      for (int q = 0; q < nrows[s] && q < 64; q++) {
        const uint32_t type = (uint32_t)(rw[q] >> 32);
        const int z = key_z((uint32_t)rw[q]);
        if (type & T_SSB) start = z;
        if (type & T_SSA) end = z;
        if (start > -1 && end > -1) {
          for (int zz = start; zz >= end; zz--) push(ss, x, y, zz, 0);
          start = end = -1;
        }
      }
    }
  return TOMO_OK;
}

int32_t TomoSh_HIP(float *f32, int32_t *i32, float *YPR, int32_t *tri, float *vtx, float *vtx_nrm,
                   float *tri_nrm, int32_t *chull_tri, float *chull_vtx, float *chull_trinrm) {
  (void)tri_nrm; (void)chull_tri; (void)chull_vtx; (void)chull_trinrm;
  OnDestroy();
  if (!f32 || !i32) return fail(TOMO_E_INVALID_ARG, "null header array");
  const int32_t n_tri = i32[4], n_vtx = i32[5], n_ypr = i32[6];
  if (n_ypr < 0 || (n_ypr > 0 && !YPR)) return fail(TOMO_E_INVALID_ARG, "bad YPR");
  tomo_handle *h = nullptr;
  int rc = tomo_create(f32, i32, 0, &h);
  if (rc) return rc;
  rc = tomo_set_mesh(h, n_tri, tri, n_vtx, vtx, vtx_nrm);
  std::vector<tomo_record> rec((size_t)(n_ypr > 0 ? n_ypr : 1));
  if (!rc) rc = tomo_sweep(h, n_ypr, YPR, rec.data());
  if (rc) {
    tomo_destroy(h);
    return rc;
  }
  g_mss.assign((size_t)n_ypr + 2, 0.0f);
  g_mo.assign((size_t)n_ypr + 2, 0.0f);
  g_vtc.assign((size_t)n_ypr + 2, 0.0f);
  for (int32_t i = 0; i < n_ypr; i++) {
    g_mss[i] = rec[i].Mss;
    g_mo[i] = rec[i].Mo;
    g_vtc[i] = rec[i].fVtc;
  }
  const int32_t opt = tomo_find_first_optimal(n_ypr, g_mss.data());
  if (n_ypr > 0) {
    vm_from_record(h, rec[opt], g_vm);
    std::memcpy(g_mat, rec[opt].mat4x4, sizeof(g_mat));
    rc = render_export(h, &YPR[(size_t)opt * 3]);
    if (rc) {
      tomo_destroy(h);
      return rc;
    }
  }
  g_have = true;
  tomo_destroy(h);
  return opt;
}

float *getMss(void) { return g_have ? g_mss.data() : nullptr; }
float *getMo(void) { return g_have ? g_mo.data() : nullptr; }
float *getVtc(void) { return g_have ? g_vtc.data() : nullptr; }
float *getVolMassInfo(void) { return g_vm; }
float *getMat4x4(void) { return g_have ? g_mat : nullptr; }

int32_t getnData2i(int16_t t) {
  if (t < 0 || t >= 11) return 0;
  return (int32_t)(g_pix[(size_t)t].size() / 6);
}

int16_t *getpData2i(int16_t t) {
  if (t < 0 || t >= 11 || g_pix[(size_t)t].empty()) return nullptr;
  return g_pix[(size_t)t].data();
}

// ---- continuous support volume ----------------------------------------------------------

struct tomo_cvss {
  int device = 0;
  float *d = nullptr;  // 7 * n_tri + 3 * n_vtx floats
  CvssMeshDev m{};
};

int tomo_cvss_create(const float *f32, int32_t n_tri, const int32_t *tri, int32_t n_vtx, const float *vtx,
                     float tau, float temp, tomo_cvss **out) {
  if (!out) return fail(TOMO_E_INVALID_ARG, "null output handle");
  *out = nullptr;
  if (!f32 || n_tri < 0 || n_vtx < 0 || (n_tri > 0 && (!tri || !vtx)) || (n_vtx > 0 && !vtx))
    return fail(TOMO_E_INVALID_ARG, "bad mesh arguments");
  if (n_tri > 0 && n_vtx == 0) return fail(TOMO_E_INVALID_ARG, "triangles without vertices");
  for (int64_t i = 0; i < (int64_t)n_tri * 3; i++)
    if (tri[i] < 0 || tri[i] >= n_vtx) return fail(TOMO_E_INVALID_ARG, "triangle index out of range");
  const float theta_c = f32[1];
  if (!(theta_c >= 0.0f && theta_c <= 3.14159265f)) return fail(TOMO_E_INVALID_ARG, "theta_c out of [0, pi]");
  if (!(tau >= 0.0f && tau <= 2.0f) || !(temp >= 0.0f && temp < 1e6f))
    return fail(TOMO_E_INVALID_ARG, "smoothing width / temperature out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TOMO_E_NO_DEVICE, "no HIP device");
  // Geometry in double on the host, relative to the AABB centre, stored as float.
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int32_t v = 0; v < n_vtx; v++)
    for (int k = 0; k < 3; k++) {
      const double x = vtx[(size_t)v * 3 + k];
      if (v == 0 || x < lo[k]) lo[k] = x;
      if (v == 0 || x > hi[k]) hi[k] = x;
    }
  const double c0[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
  const size_t T = (size_t)n_tri, V = (size_t)n_vtx;
  std::vector<float> h(7 * T + 3 * V + 1, 0.0f);
  for (size_t t = 0; t < T; t++) {
    double p[3][3];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) p[j][k] = (double)vtx[(size_t)tri[t * 3 + j] * 3 + k] - c0[k];
    const double e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
    const double e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; k++) h[k * T + t] = len > 0.0 ? (float)(n[k] / len) : 0.0f;
    h[3 * T + t] = (float)(0.5 * len);
    for (int k = 0; k < 3; k++) h[(4 + k) * T + t] = (float)((p[0][k] + p[1][k] + p[2][k]) / 3.0);
  }
  for (size_t v = 0; v < V; v++)
    for (int k = 0; k < 3; k++) h[7 * T + k * V + v] = (float)((double)vtx[v * 3 + k] - c0[k]);
  tomo_cvss *c = new (std::nothrow) tomo_cvss();
  if (!c) return fail(TOMO_E_OOM, "handle allocation failed");
  if (hipGetDevice(&c->device) != hipSuccess || hipMalloc(&c->d, sizeof(float) * h.size()) != hipSuccess) {
    delete c;
    return fail(TOMO_E_OOM, "device mesh allocation failed");
  }
  if (hipMemcpy(c->d, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(c->d);
    delete c;
    return fail(TOMO_E_HIP, "device mesh upload failed");
  }
  float *d = c->d;
  c->m = CvssMeshDev{d, d + T, d + 2 * T, d + 3 * T, d + 4 * T, d + 5 * T, d + 6 * T,
                     d + 7 * T, d + 7 * T + V, d + 7 * T + 2 * V, n_tri, n_vtx, sinf(theta_c), tau, temp};
  *out = c;
  return TOMO_OK;
}

static int cvss_check(const tomo_cvss *c, int32_t n, const void *a, const void *b) {
  if (!c || n < 0 || (n > 0 && (!a || !b))) return fail(TOMO_E_INVALID_ARG, "bad arguments");
  const int64_t chunk = 256 * 8;
  const int64_t blocks = (int64_t)n * ((std::max<int64_t>(c->m.n_tri, c->m.n_vtx) + chunk - 1) / chunk);
  if (blocks >= ((int64_t)1 << 31)) return fail(TOMO_E_INVALID_ARG, "batch too large for one launch");
  return TOMO_OK;
}

int tomo_cvss_forward(tomo_cvss *c, int32_t n, const float *ypr, float *vss, double *aux, void *stream) {
  int rc = cvss_check(c, n, ypr, vss);
  if (rc) return rc;
  if (n > 0 && !aux) return fail(TOMO_E_INVALID_ARG, "null workspace");
  HIP_TRY(cvss_forward(c->m, n, ypr, vss, aux, (hipStream_t)stream));
  return TOMO_OK;
}

int tomo_cvss_backward(tomo_cvss *c, int32_t n, const float *ypr, double *aux, const float *grad_vss,
                       float *grad_ypr, void *stream) {
  int rc = cvss_check(c, n, ypr, grad_ypr);
  if (rc) return rc;
  if (n > 0 && (!aux || !grad_vss)) return fail(TOMO_E_INVALID_ARG, "null workspace or gradient");
  HIP_TRY(cvss_backward(c->m, n, ypr, aux, grad_vss, grad_ypr, (hipStream_t)stream));
  return TOMO_OK;
}

int tomo_cvss_destroy(tomo_cvss *c) {
  if (!c) return TOMO_OK;
  if (c->d) (void)hipFree(c->d);
  delete c;
  return TOMO_OK;
}

void OnDestroy(void) {
  g_mss.clear();
  g_mo.clear();
  g_vtc.clear();
  std::memset(g_vm, 0, sizeof(g_vm));
  std::memset(g_mat, 0, sizeof(g_mat));
  for (auto &v : g_pix) v.clear();
  g_have = false;
}

}  // extern "C"
