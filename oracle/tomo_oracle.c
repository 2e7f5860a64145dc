/*
 * tomo_oracle.c -- CPU restatement of the reference's INT3 support-tomography path.
 *
 * TEST INFRASTRUCTURE ONLY (see tomo_oracle.h).  Written from the reference's behaviour;
 * every function cites the reference code it restates (paths relative to
 * cfms_tomo/Tomo_Shell_vc_src/).  Build with -ffp-contract=off: the reference's float
 * arithmetic is sequential IEEE single precision without fused multiply-add, with one step in
 * double (the barycentric u).
 *
 * Deliberate, documented deviations from the reference (which has undefined behaviour there):
 *  - entries whose slot lies outside the grid are dropped and flagged (SetBit has no bounds
 *    check, STomoVoxelSpaceInfo.cpp:175-230);
 *  - the grid edge is a parameter (the reference hard-wires 256, STomoVoxelSpaceInfo.h:10).
 */
#include "tomo_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ORC_ROWS 64            /* nSlotCapacityHeight (STomoVoxelSpaceInfo.cpp:48-50)   */
#define ORC_MAXPIX (ORC_ROWS - 1) /* pixel rows per slot; row 0 is the count             */
#define ORC_MARGIN 0.001f      /* g_fMARGIN (Tomo_types.h:40)                             */
#define ORC_NF 1000.0f         /* g_fNORMALFACTOR (Tomo_types.h:39)                       */

typedef struct {
  /* options decoded the way S3DPrinterInfo::Set does (S3DPrinterInfo.cpp:124-145) */
  int grid;
  float dvoxel, theta_c, area, wall, fcore, fclad, fss, css, rho, bed_in, bed_out, bed_thick;
  int explicit_ss, shell_mesh, bed_type;
  float shell_t;
  float center[3];
} orc_ctx;

typedef struct {
  int grid;
  int16_t *buf;        /* [grid*grid][ORC_ROWS][3] = {type, z, nZ}; row 0 col 0 = count */
  int32_t *dirty;      /* slots that went empty -> non-empty (STomoVoxelSpaceInfo.h:43-50) */
  int64_t ndirty, capdirty;
  float *vx, *vy, *vz, *nz; /* placed vertices and rotated normal z */
  int32_t flags;
  /* optional log of Pixelize's SetBit calls (test diagnostics): {triangle, seq, slot, key} */
  int32_t *log;
  int64_t nlog, caplog;
  int32_t cur_tri, cur_seq;
} orc_ws;

static inline float orc_min(float a, float b) { return (a <= b) ? a : b; } /* Tomo_types.cpp:29-32 */
static inline float orc_max(float a, float b) { return (a >= b) ? a : b; } /* Tomo_types.cpp:39-42 */

/* SMatrix4f::YPR (SMatrix4f.cpp:213-240): ga=yaw, be=pitch, al=roll, R = Rz(al) Ry(be) Rx(ga). */
void orc_ypr_matrix(float yaw, float pitch, float roll, float R[9]) {
  float ca = cosf(roll), cb = cosf(pitch), cg = cosf(yaw);
  float sa = sinf(roll), sb = sinf(pitch), sg = sinf(yaw);
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

/* SMatrix4f::Dot (SMatrix4f.cpp:242-262): r = D0*a0; r += D1*a1; r += D2*a2; r += D3*1 (D3 = 0). */
static inline float orc_dot_row(const float *row, float a0, float a1, float a2) {
  float r = row[0] * a0;
  r += row[1] * a1;
  r += row[2] * a2;
  r += 0.0f * 1.0f;
  return r;
}

void orc_theta_cuts(float theta_c, int32_t *cut_ssb, int32_t *cut_nvb) {
  /* threshold = FLOAT32(-1. * sin(theta_c)); beta rows: SSB iff nZ/1000.f < thr,
   * NVB iff nZ/1000.f > thr (STomoSh_INT3.cpp:755-776). */
  float thr = (float)(-1.0 * sin((double)theta_c));
  int32_t cs = -32769, cn = 32768;
  for (int32_t n = -32768; n <= 32767; n++) {
    float f = (float)n / ORC_NF;
    if (f < thr) cs = n;
    if (f > thr && n < cn) cn = n;
  }
  *cut_ssb = cs;
  *cut_nvb = cn;
}

static void orc_ctx_init(orc_ctx *c, const orc_opts *o, int32_t nVtx, const float *vtx) {
  c->grid = o->grid > 0 ? o->grid : 256;
  c->dvoxel = o->f32[0]; c->theta_c = o->f32[1]; c->area = o->f32[2]; c->wall = o->f32[3];
  c->fcore = o->f32[4]; c->fclad = o->f32[5]; c->fss = o->f32[6]; c->css = o->f32[7];
  c->rho = o->f32[8]; c->bed_in = o->f32[9]; c->bed_out = o->f32[10]; c->bed_thick = o->f32[11];
  /* (_int32_x11[k] > g_fMARGIN) and FLOAT32(shell_um) * FLOAT32(0.001): S3DPrinterInfo.cpp:127-131 */
  c->explicit_ss = ((float)o->i32[1] > ORC_MARGIN);
  c->shell_mesh = ((float)o->i32[2] > ORC_MARGIN);
  c->shell_t = (float)o->i32[10] * 0.001f;
  c->bed_type = o->i32[9];
  /* STomoAABB3Df::Set + GetCenter over the raw vertices (STomoAABB3Df.cpp:42-72). */
  float mn[3] = {1e5f, 1e5f, 1e5f}, mx[3] = {-1e5f, -1e5f, -1e5f};
  for (int32_t v = 0; v < nVtx; v++)
    for (int k = 0; k < 3; k++) {
      mn[k] = orc_min(mn[k], vtx[v * 3 + k]);
      mx[k] = orc_max(mx[k], vtx[v * 3 + k]);
    }
  for (int k = 0; k < 3; k++) c->center[k] = (float)((double)(mn[k] + mx[k]) * 0.5);
}

static int orc_ws_init(orc_ws *w, int grid, int32_t nVtx) {
  memset(w, 0, sizeof(*w));
  w->grid = grid;
  w->buf = (int16_t *)calloc((size_t)grid * grid * ORC_ROWS * 3, sizeof(int16_t));
  w->capdirty = 1024;
  w->dirty = (int32_t *)malloc(sizeof(int32_t) * w->capdirty);
  size_t nv = (size_t)(nVtx > 0 ? nVtx : 1);
  w->vx = (float *)malloc(sizeof(float) * nv);
  w->vy = (float *)malloc(sizeof(float) * nv);
  w->vz = (float *)malloc(sizeof(float) * nv);
  w->nz = (float *)malloc(sizeof(float) * nv);
  return (w->buf && w->dirty && w->vx && w->vy && w->vz && w->nz) ? 0 : -1;
}

static void orc_ws_free(orc_ws *w) {
  free(w->buf); free(w->dirty); free(w->vx); free(w->vy); free(w->vz); free(w->nz);
}

static inline int16_t *orc_slot(orc_ws *w, int64_t s) { return w->buf + s * (ORC_ROWS * 3); }

/* ClearDirtySlots (STomoVoxelSpaceInfo.cpp:62-80): zero every slot touched last time. */
static void orc_clear(orc_ws *w) {
  for (int64_t i = 0; i < w->ndirty; i++)
    memset(orc_slot(w, w->dirty[i]), 0, sizeof(int16_t) * ORC_ROWS * 3);
  w->ndirty = 0;
}

static void orc_push_dirty(orc_ws *w, int32_t s) {
  if (w->ndirty == w->capdirty) {
    w->capdirty *= 2;
    w->dirty = (int32_t *)realloc(w->dirty, sizeof(int32_t) * w->capdirty);
  }
  w->dirty[w->ndirty++] = s;
}

/* float -> int16 the way SLOT_BUFFER_TYPE(f) behaves on x86 (truncate to int32, keep 16 bits). */
static inline int16_t orc_to_i16(orc_ws *w, float f) {
  if (!(f > -2147483648.0f && f < 2147483648.0f)) { w->flags |= ORC_F_INT16_WRAP; return 0; }
  int32_t i = (int32_t)f;
  if (i != (int16_t)i) w->flags |= ORC_F_INT16_WRAP;
  return (int16_t)i;
}

/* STomoVoxelSpaceInfo::SetBit (STomoVoxelSpaceInfo.cpp:175-230). */
static void orc_setbit(orc_ws *w, float px, float py, float pz, float nzf, int type) {
  int x = (int)px, y = (int)py;                       /* coord2vID: int(coord), :97-105 */
  if (px <= -2147483648.0f || px >= 2147483648.0f || py <= -2147483648.0f || py >= 2147483648.0f ||
      x < 0 || y < 0 || x >= w->grid || y >= w->grid) {
    w->flags |= ORC_F_FOOTPRINT_OOB;
    return;
  }
  int16_t z = orc_to_i16(w, pz);
  int16_t nz = orc_to_i16(w, nzf * ORC_NF);
  int32_t s = x * w->grid + y;
  if (w->log && type == 0) {
    if (w->nlog < w->caplog) {
      int32_t *e = w->log + 4 * w->nlog;
      e[0] = w->cur_tri;
      e[1] = w->cur_seq;
      e[2] = s;
      e[3] = (int32_t)(((uint32_t)(z + 32768) << 16) | (uint32_t)(nz + 32768));
    }
    w->nlog++;
  }
  int16_t *slot = orc_slot(w, s);
  int n = slot[0];
  int id = -1;
  for (int p = 0; p < n; p++)
    if (slot[(p + 1) * 3 + 1] == z && slot[(p + 1) * 3 + 2] == nz) { id = p; break; }
  if (id < 0) {
    if (n >= ORC_MAXPIX) { w->flags |= ORC_F_CAP_DROP; return; }
    if (n == 0) orc_push_dirty(w, s);
    id = n;
    slot[0] = (int16_t)(n + 1);
  }
  slot[(id + 1) * 3 + 0] |= (int16_t)type;
  slot[(id + 1) * 3 + 1] = z;
  slot[(id + 1) * 3 + 2] = nz;
}

/* _getBaryCoord (Tomo_types.cpp:86-115) with _dot (Tomo_types.cpp:14-17) on z=0 vectors.
 * part[] (may be NULL) receives den, v, w, u (v, w, u only when |den| > eps, else 0). */
static inline int orc_bary_full(float px, float py, const float *a, const float *b, const float *c,
                                float *uo, float *vo, float *wo, float *part) {
  const float zz = 0.0f * 0.0f;
  float e0x = b[0] - a[0], e0y = b[1] - a[1];
  float e1x = c[0] - a[0], e1y = c[1] - a[1];
  float e2x = px - a[0], e2y = py - a[1];
  float d00 = e0x * e0x + e0y * e0y + zz;
  float d01 = e0x * e1x + e0y * e1y + zz;
  float d11 = e1x * e1x + e1y * e1y + zz;
  float d20 = e2x * e0x + e2y * e0y + zz;
  float d21 = e2x * e1x + e2y * e1y + zz;
  float den = d00 * d11 - d01 * d01;
  float aden = (den >= 0.0f) ? den : -den;
  if (part) { part[0] = den; part[1] = part[2] = part[3] = 0.0f; }
  if (aden > ORC_MARGIN) {
    float v = (d11 * d20 - d01 * d21) / den;
    float w = (d00 * d21 - d01 * d20) / den;
    float u = (float)((1.0 - (double)v) - (double)w);
    if (part) { part[1] = v; part[2] = w; part[3] = u; }
    if (u >= -ORC_MARGIN && v >= -ORC_MARGIN && (double)v <= 1.0 + (double)ORC_MARGIN &&
        (double)(u + v) <= 1.0 + (double)ORC_MARGIN) {
      *uo = u; *vo = v; *wo = w;
      return 1;
    }
  }
  return 0;
}

static inline int orc_bary(float px, float py, const float *a, const float *b, const float *c,
                           float *uo, float *vo, float *wo) {
  return orc_bary_full(px, py, a, b, c, uo, vo, wo, NULL);
}

/* STomoSh_INT3::triVoxel (STomoSh_INT3.cpp:1138-1209).  a*: placed vertices, n*z: rotated normal z. */
static void orc_trivoxel(orc_ws *w, const orc_ctx *c, const float *a0, const float *a1,
                         const float *a2, float n0z, float n1z, float n2z) {
  float v0[3], v1[3], v2[3];
  for (int i = 0; i < 3; i++) {
    v0[i] = a0[i] + ORC_MARGIN;
    v1[i] = a1[i] + ORC_MARGIN;
    v2[i] = a2[i] + ORC_MARGIN;
  }
  int x0 = (int)orc_min(orc_min(v0[0], v1[0]), v2[0]);
  int y0 = (int)orc_min(orc_min(v0[1], v1[1]), v2[1]);
  int x1 = (int)orc_max(orc_max(v0[0], v1[0]), v2[0]);
  int y1 = (int)orc_max(orc_max(v0[1], v1[1]), v2[1]);
  float h2 = c->dvoxel * 0.5f;
  float off = (float)(int)(c->shell_t + 0.999f);
  int do_shell = c->shell_mesh && c->shell_t > ORC_MARGIN;
  int hits = 0;
  for (int x = x0; x <= x1; x++) {
    float cx = (float)x + ORC_MARGIN + h2;
    for (int y = y0; y <= y1; y++) {
      float cy = (float)y + ORC_MARGIN + h2;
      float u, v, wv;
      if (!orc_bary(cx, cy, v0, v1, v2, &u, &v, &wv)) continue;
      w->cur_seq = 2 * ((x - x0) * (y1 - y0 + 1) + (y - y0)); /* insertion order within the triangle */
      /* _bary_product (Tomo_types.cpp:75-84): p = p0*u + p1*v + p2*w */
      float px = v0[0] * u + v1[0] * v + v2[0] * wv;
      float py = v0[1] * u + v1[1] * v + v2[1] * wv;
      float pz = v0[2] * u + v1[2] * v + v2[2] * wv;
      float nz = n0z * u + n1z * v + n2z * wv;
      orc_setbit(w, px, py, pz, nz, 0);
      hits++;
      if (do_shell) { /* shell sub-pixel: z moved against the normal, normal flipped (:1179-1185) */
        pz += (nz >= 0.0f) ? -off : off;
        nz = -nz;
        w->cur_seq++;
        orc_setbit(w, px, py, pz, nz, 0);
        hits++;
      }
    }
  }
  if (hits == 0) { /* tiny triangle: one pixel at the un-margined centroid (:1190-1206) */
    float t[3];
    for (int i = 0; i < 3; i++) t[i] = (a0[i] + a1[i] + a2[i]) * 0.333333f + ORC_MARGIN;
    w->cur_seq = 0;
    orc_setbit(w, t[0], t[1], t[2], n0z, 0);
    if (do_shell) {
      float sz = t[2] + ((n0z >= 0.0f) ? -off : off);
      w->cur_seq = 1;
      orc_setbit(w, t[0], t[1], sz, -n0z, 0);
    }
  }
}

/* STomoSh_Base::Rotate (STomoSh_Base.cpp:45-129): centre by raw AABB, rotate, then put the
 * rotated AABB corner at (Bo, Bo, 0).  Also the drop-in transform mat4x4 (:111-123). */
static void orc_rotate(orc_ws *w, const orc_ctx *c, const float R[9], int32_t nVtx,
                       const float *vtx, const float *vnrm, float mat[16]) {
  float mn[3] = {1e5f, 1e5f, 1e5f};
  for (int32_t v = 0; v < nVtx; v++) {
    float t0 = vtx[v * 3 + 0] - c->center[0];
    float t1 = vtx[v * 3 + 1] - c->center[1];
    float t2 = vtx[v * 3 + 2] - c->center[2];
    float x = orc_dot_row(R + 0, t0, t1, t2);
    float y = orc_dot_row(R + 3, t0, t1, t2);
    float z = orc_dot_row(R + 6, t0, t1, t2);
    w->vx[v] = x; w->vy[v] = y; w->vz[v] = z;
    w->nz[v] = orc_dot_row(R + 6, vnrm[v * 3 + 0], vnrm[v * 3 + 1], vnrm[v * 3 + 2]);
    mn[0] = orc_min(mn[0], x);
    mn[1] = orc_min(mn[1], y);
    mn[2] = orc_min(mn[2], z);
  }
  float kx = mn[0] - c->bed_out, ky = mn[1] - c->bed_out, kz = mn[2];
  for (int32_t v = 0; v < nVtx; v++) {
    w->vx[v] -= kx;
    w->vy[v] -= ky;
    w->vz[v] -= kz;
  }
  if (mat) {
    float m = -c->center[0], n = -c->center[1], o = -c->center[2];
    float tr[3] = {-kx, -ky, -kz};
    memset(mat, 0, sizeof(float) * 16);
    for (int r = 0; r < 3; r++) {
      mat[r * 4 + 0] = R[r * 3 + 0];
      mat[r * 4 + 1] = R[r * 3 + 1];
      mat[r * 4 + 2] = R[r * 3 + 2];
      mat[r * 4 + 3] = tr[r] + R[r * 3 + 0] * m + R[r * 3 + 1] * n + R[r * 3 + 2] * o;
    }
    mat[15] = 1.0f;
  }
}

/* Pixelize (STomoSh_INT3.cpp:285-364), default branch: triangles in index order. */
static void orc_pixelize(orc_ws *w, const orc_ctx *c, int32_t nTri, const int32_t *tri) {
  orc_clear(w);
  for (int32_t t = 0; t < nTri; t++) {
    int32_t i0 = tri[t * 3 + 0], i1 = tri[t * 3 + 1], i2 = tri[t * 3 + 2];
    w->cur_tri = t;
    float a0[3] = {w->vx[i0], w->vy[i0], w->vz[i0]};
    float a1[3] = {w->vx[i1], w->vy[i1], w->vz[i1]};
    float a2[3] = {w->vx[i2], w->vy[i2], w->vz[i2]};
    orc_trivoxel(w, c, a0, a1, a2, w->nz[i0], w->nz[i1], w->nz[i2]);
  }
}

static int orc_cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* createAlBePxls + matchAlBePairBriefly + createVoPixels + createShadowBriefly +
 * createVss_Explicit / createTCPixels + createVss_Implicit: the brief vslotPair
 * (STomoSh_INT3.cpp:1041-1062).  rows: [1 + ORC_ROWS][3] slot layout, rows beyond n zero.
 * Returns the new slot length. */
static int orc_vslot_pair(int16_t *slot, int n, float thr, int explicit_ss, int32_t *vo_out,
                          int32_t *vss_out, int32_t *flags) {
  int16_t *row = slot + 3; /* row p at row[p*3 + {0,1,2}] */
  /* sortSlotByZ (:398-446): descending z, ties by descending nZ.  Keys are distinct, so any
   * correct sort yields the same order; insertion sort on (z, nZ). */
  for (int i = 1; i < n; i++) {
    int16_t t = row[i * 3 + 0], z = row[i * 3 + 1], nz = row[i * 3 + 2];
    int j = i - 1;
    while (j >= 0 && (row[j * 3 + 1] < z || (row[j * 3 + 1] == z && row[j * 3 + 2] < nz))) {
      row[(j + 1) * 3 + 0] = row[j * 3 + 0];
      row[(j + 1) * 3 + 1] = row[j * 3 + 1];
      row[(j + 1) * 3 + 2] = row[j * 3 + 2];
      j--;
    }
    row[(j + 1) * 3 + 0] = t; row[(j + 1) * 3 + 1] = z; row[(j + 1) * 3 + 2] = nz;
  }
  /* splitAlBe (:493-513) */
  for (int p = 0; p < n; p++) {
    float f = (float)row[p * 3 + 2] * ORC_NF;
    if ((double)f > (double)ORC_MARGIN * 10.0) row[p * 3 + 0] |= ORC_AL;
    else if ((double)f < (double)ORC_MARGIN * -10.0) row[p * 3 + 0] |= ORC_BE;
  }
  /* matchAlBePairBriefly (:583-611) */
  if (n >= 2) {
    int started = 0;
    for (int p = 0; p < n; p++) {
      int16_t *ty = &row[p * 3 + 0];
      if (!started && *ty == ORC_AL) started = 1;
      else if (started && *ty == ORC_BE) started = 0;
      else if (*ty == ORC_AL || *ty == ORC_BE) *ty = 0;
    }
  }
  /* createVoPixels (:934-960) */
  int32_t al = 0, be = 0;
  for (int p = 0; p < n; p++) {
    if (row[p * 3 + 0] & ORC_AL) al += row[p * 3 + 1];
    else if (row[p * 3 + 0] & ORC_BE) be += row[p * 3 + 1];
  }
  *vo_out = (al > be) ? (al - be) : 0;
  /* createShadowBriefly (:749-812) */
  int ex = 0, im = 0;
  for (int p = 0; p < n; p++) {
    int16_t *ty = &row[p * 3 + 0];
    float fnz = (float)row[p * 3 + 2] / ORC_NF;
    if ((*ty & ORC_BE) && fnz < ORC_MARGIN) {
      if (!ex && fnz < thr) { *ty |= ORC_SSB; ex = 1; }
      if (!im && fnz > thr) { *ty |= ORC_NVB; im = 1; }
    } else if (*ty & ORC_AL) {
      if (ex) { *ty |= ORC_SSA; ex = 0; }
      if (im) { *ty |= ORC_NVA; im = 0; }
    }
  }
  if (ex) {
    if (n < ORC_MAXPIX) { row[n * 3 + 0] |= ORC_SSA; n++; } else *flags |= ORC_F_SENTINEL_DROP;
  }
  if (im) {
    if (n < ORC_MAXPIX) { row[n * 3 + 0] |= ORC_NVA; n++; } else *flags |= ORC_F_SENTINEL_DROP;
  }
  int32_t vss = 0;
  if (explicit_ss) { /* createVss_Explicit (:962-987) */
    int32_t sb = 0, sa = 0;
    for (int p = 0; p < n; p++) {
      if (row[p * 3 + 0] & ORC_SSB) sb += row[p * 3 + 1];
      else if (row[p * 3 + 0] & ORC_SSA) sa += row[p * 3 + 1];
    }
    vss = sb - sa;
  } else { /* createTCPixels (:734-747) + createVss_Implicit (:989-1025) */
    for (int p = 0; p < n; p++)
      if (row[p * 3 + 0] & ORC_AL) { row[p * 3 + 0] |= ORC_TC; break; }
    int32_t a = 0, b = 0, tc = 0, nvb = 0, nva = 0;
    for (int p = 0; p < n; p++) {
      int16_t ty = row[p * 3 + 0];
      int32_t z = row[p * 3 + 1];
      if (ty & ORC_AL) {
        a += z;
        if (ty & ORC_TC) tc += z;
        if (ty & ORC_NVA) nva += z;
      } else if (ty & ORC_BE) {
        b += z;
        if (ty & ORC_NVB) nvb += z;
      }
    }
    vss = -a + b + tc - nvb + nva;
  }
  *vss_out = vss;
  slot[0] = (int16_t)n;
  return n;
}

/* IsBedCandidate (STomoSh_INT3.cpp:1331-1373). */
static int orc_bed_candidate(orc_ws *w, const orc_ctx *c, int X, int Y) {
  int G = w->grid;
  int r = (int)c->bed_out;
  float md = 1e5f;
  int xa = (X - r > 0) ? X - r : 0, xb = (X + r < G) ? X + r : G;
  int ya = (Y - r > 0) ? Y - r : 0, yb = (Y + r < G) ? Y + r : G;
  for (int sx = xa; sx < xb; sx++)
    for (int sy = ya; sy < yb; sy++) {
      int16_t *slot = orc_slot(w, (int64_t)sx * G + sy);
      int n = slot[0];
      if (n <= 0) continue;
      int16_t z = slot[n * 3 + 1], ty = slot[n * 3 + 0];
      float dx = (float)X - (float)sx, dy = (float)Y - (float)sy;
      float d = sqrtf(dx * dx + dy * dy);
      if (z == 0 && (ty & (ORC_AL | ORC_BE | ORC_SSA | ORC_SS))) md = (md < d) ? md : d;
    }
  if (md > c->bed_out) return 0;
  if (c->bed_type == 1 || c->bed_type == 2) return 1;
  if (c->bed_type == 3 && md > c->bed_in) return 1;
  return 0;
}

/* GenerateBed (STomoSh_INT3.cpp:1375-1443): raster scan x outer, y inner, over the touched
 * bounding box grown by int(Bo)+1; each accepted slot gets a (z=1, nZ=0) bed row at once, which
 * later candidates see. */
static void orc_bed(orc_ws *w, const orc_ctx *c) {
  if (c->bed_type == 0 || w->ndirty == 0) return;
  int G = w->grid;
  int minX = G, minY = G, maxX = -1, maxY = -1;
  for (int64_t i = 0; i < w->ndirty; i++) {
    int sx = w->dirty[i] / G, sy = w->dirty[i] % G;
    if (sx < minX) minX = sx;
    if (sx > maxX) maxX = sx;
    if (sy < minY) minY = sy;
    if (sy > maxY) maxY = sy;
  }
  int m = (int)c->bed_out + 1;
  int bx0 = (minX - m > 0) ? minX - m : 0, bx1 = (maxX + m + 1 < G) ? maxX + m + 1 : G;
  int by0 = (minY - m > 0) ? minY - m : 0, by1 = (maxY + m + 1 < G) ? maxY + m + 1 : G;
  for (int x = bx0; x < bx1; x++)
    for (int y = by0; y < by1; y++) {
      int16_t *slot = orc_slot(w, (int64_t)x * G + y);
      int n = slot[0];
      int possible = 0;
      if (n == 0) possible = 1;
      else {
        int16_t z = slot[n * 3 + 1], ty = slot[n * 3 + 0];
        if (z == 0) {
          if (ty == ORC_NVA || ty == ORC_VO || ty == ORC_VSS) possible = 1;
          if (ty & (ORC_AL | ORC_BE | ORC_SSA | ORC_SS)) possible = 0;
          if (c->bed_type == 2 && (ty & (ORC_BE | ORC_SSA))) possible = 1;
        }
      }
      if (possible && orc_bed_candidate(w, c, x, y))
        orc_setbit(w, (float)x, (float)y, 1.0f, 0.0f, ORC_BED);
    }
}

/* VolToMass (STomoVolMassInfo.cpp:42-81) with pi = 3.141592. */
static void orc_vtm(const orc_ctx *c, orc_record *r) {
  float vo_clad = c->area * c->wall;
  float vo_core = r->fVo - vo_clad;
  float mo_clad = vo_clad * c->fclad * c->rho;
  float mo_core = vo_core * c->fcore * c->rho;
  r->Mo = mo_core + mo_clad;
  float mss = 0.0f;
  if (c->explicit_ss) {
    float vss = r->fVss;
    if (vss > ORC_MARGIN) {
      float radius = powf(vss / (4.f / 3.f * 3.141592f), 1.f / 3.f);
      float ass = (float)(4. * 3.141592 * (double)radius * (double)radius);
      float vss_clad = ass * c->wall;
      if (vss_clad > vss) vss_clad = vss;
      float vss_core = vss - vss_clad;
      float mss_clad = vss_clad * c->fclad * c->rho;
      float mss_core = vss_core * c->fcore * c->rho;
      mss = mss_core + mss_clad;
    }
  } else {
    mss = r->fVss * c->fss * c->css * c->rho;
  }
  float mbed = r->fVbed * c->wall * c->rho;
  mss += mbed;
  r->Mss = mss;
  r->Mtotal = r->Mo + mss;
}

void orc_vol_to_mass(const orc_opts *opt, orc_record *r) {
  orc_ctx c;
  orc_ctx_init(&c, opt, 0, NULL);
  orc_vtm(&c, r);
}

/* thread_func body (Tomo_Shell2026.cpp:99-117) for one orientation. */
static void orc_eval(orc_ws *w, const orc_ctx *c, float thr, int32_t nTri, const int32_t *tri,
                     int32_t nVtx, const float *vtx, const float *vnrm, const float ypr[3],
                     orc_record *r, int32_t *cnt_pre, int32_t *cnt_final, int32_t *svo,
                     int32_t *svss) {
  int G = w->grid;
  float R[9];
  memset(r, 0, sizeof(*r));
  w->flags = 0;
  orc_ypr_matrix(ypr[0], ypr[1], ypr[2], R);
  orc_rotate(w, c, R, nVtx, vtx, vnrm, r->mat4x4);
  orc_pixelize(w, c, nTri, tri);

  /* Pairing (STomoSh_INT3.cpp:366-395): ascending slot id, slots with more than one pixel. */
  qsort(w->dirty, (size_t)w->ndirty, sizeof(int32_t), orc_cmp_i32);
  if (svo) memset(svo, 0, sizeof(int32_t) * (size_t)G * G);
  if (svss) memset(svss, 0, sizeof(int32_t) * (size_t)G * G);
  if (cnt_pre) memset(cnt_pre, 0, sizeof(int32_t) * (size_t)G * G);
  int64_t abs_vo = 0, abs_vss = 0;
  float fvo = 0.0f, fvss = 0.0f;
  r->n_slots = w->ndirty;
  for (int64_t i = 0; i < w->ndirty; i++) {
    int32_t s = w->dirty[i];
    int16_t *slot = orc_slot(w, s);
    int n = slot[0];
    r->n_entries += n;
    r->chk_pre += (uint64_t)(s + 1) * (uint64_t)n;
    if (cnt_pre) cnt_pre[s] = n;
    if (n > 1) {
      int32_t vo = 0, vss = 0;
      n = orc_vslot_pair(slot, n, thr, c->explicit_ss, &vo, &vss, &w->flags);
      fvo += (float)vo;
      fvss += (float)vss;
      r->Vo += vo;
      r->Vss += vss;
      abs_vo += vo < 0 ? -(int64_t)vo : vo;
      abs_vss += vss < 0 ? -(int64_t)vss : vss;
      if (svo) svo[s] = vo;
      if (svss) svss[s] = vss;
    }
    /* the pre-fix reference kept per-slot sums in int16: flag slots whose rows could wrap */
    int32_t sabs = 0;
    for (int p = 0; p < n; p++) {
      int32_t z = slot[(p + 1) * 3 + 1];
      sabs += z < 0 ? -z : z;
    }
    if (sabs > 32767) w->flags |= ORC_F_INT16_WRAP;
  }
  orc_bed(w, c);

  /* Calculate (STomoSh_INT3.cpp:1211-1255): float running totals in ascending slot id. */
  qsort(w->dirty, (size_t)w->ndirty, sizeof(int32_t), orc_cmp_i32);
  float fva = 0.f, fvb = 0.f, fvnv = 0.f, fvtc = 0.f, fvbed = 0.f;
  int64_t abs_va = 0, abs_vb = 0, abs_vnv = 0, abs_vtc = 0;
  if (cnt_final) memset(cnt_final, 0, sizeof(int32_t) * (size_t)G * G);
  for (int64_t i = 0; i < w->ndirty; i++) {
    int32_t s = w->dirty[i];
    int16_t *slot = orc_slot(w, s);
    int n = slot[0];
    int32_t sa = 0, sb = 0, snvb = 0, snva = 0, stc = 0, sbed = 0;
    for (int p = 0; p < n; p++) {
      int16_t ty = slot[(p + 1) * 3 + 0];
      int32_t z = slot[(p + 1) * 3 + 1];
      if (ty & ORC_AL) sa += z;
      if (ty & ORC_BE) sb += z;
      if (ty & ORC_NVB) snvb += z;
      if (ty & ORC_NVA) snva += z;
      if (ty & ORC_TC) stc += z;
      if (ty & ORC_BED) { sbed += z; r->n_bed++; }
    }
    fva += (float)sa;
    fvb += (float)sb;
    fvnv += (float)(snvb - snva);
    fvtc += (float)stc;
    fvbed += (float)sbed * c->bed_thick;
    r->Va += sa; r->Vb += sb; r->Vnv += snvb - snva; r->Vtc += stc; r->Vbed += sbed;
    abs_va += sa < 0 ? -(int64_t)sa : sa;
    abs_vb += sb < 0 ? -(int64_t)sb : sb;
    abs_vnv += (snvb - snva) < 0 ? -(int64_t)(snvb - snva) : (snvb - snva);
    abs_vtc += stc < 0 ? -(int64_t)stc : stc;
    r->n_final += n;
    r->chk_final += (uint64_t)(s + 1) * (uint64_t)n;
    if (cnt_final) cnt_final[s] = n;
  }
  const int64_t lim = (int64_t)1 << 24;
  if (abs_vo > lim || abs_vss > lim || abs_va > lim || abs_vb > lim || abs_vnv > lim || abs_vtc > lim)
    w->flags |= ORC_F_FP32_INEXACT;
  r->fVo = fvo; r->fVss = fvss; r->fVa = fva; r->fVb = fvb; r->fVnv = fvnv; r->fVtc = fvtc;
  r->fVbed = fvbed;
  r->flags = w->flags;
  orc_vtm(c, r);
}

int orc_one(const orc_opts *opt, int32_t nTri, const int32_t *tri, int32_t nVtx,
            const float *vtx, const float *vnrm, const float ypr[3], orc_record *out,
            int32_t *cnt_pre, int32_t *cnt_final, int32_t *slot_vo, int32_t *slot_vss) {
  return orc_one_rows(opt, nTri, tri, nVtx, vtx, vnrm, ypr, out, cnt_pre, cnt_final, slot_vo, slot_vss,
                      NULL);
}

int orc_one_rows(const orc_opts *opt, int32_t nTri, const int32_t *tri, int32_t nVtx,
                 const float *vtx, const float *vnrm, const float ypr[3], orc_record *out,
                 int32_t *cnt_pre, int32_t *cnt_final, int32_t *slot_vo, int32_t *slot_vss,
                 uint64_t *rows) {
  if (!opt || !out || nTri < 0 || nVtx < 0 || (nTri > 0 && (!tri || !vtx || !vnrm))) return -1;
  orc_ctx c;
  orc_ctx_init(&c, opt, nVtx, vtx);
  orc_ws w;
  if (orc_ws_init(&w, c.grid, nVtx)) { orc_ws_free(&w); return -2; }
  float thr = (float)(-1.0 * sin((double)c.theta_c));
  orc_eval(&w, &c, thr, nTri, tri, nVtx, vtx, vnrm, ypr, out, cnt_pre, cnt_final, slot_vo, slot_vss);
  if (rows) { /* final slot contents: (type << 32) | ((z + 32768) << 16) | (nZ + 32768) */
    int64_t g2 = (int64_t)c.grid * c.grid;
    memset(rows, 0, sizeof(uint64_t) * 64 * (size_t)g2);
    for (int64_t sl = 0; sl < g2; sl++) {
      int16_t *slot = orc_slot(&w, sl);
      for (int p = 0; p < slot[0]; p++) {
        uint32_t key = ((uint32_t)(slot[(p + 1) * 3 + 1] + 32768) << 16) | (uint32_t)(slot[(p + 1) * 3 + 2] + 32768);
        rows[sl * 64 + p] = ((uint64_t)(uint16_t)slot[(p + 1) * 3 + 0] << 32) | key;
      }
    }
  }
  orc_ws_free(&w);
  return 0;
}

typedef struct {
  const orc_opts *opt;
  const orc_ctx *c;
  float thr;
  int32_t nTri, nVtx, nYPR;
  const int32_t *tri;
  const float *vtx, *vnrm, *ypr;
  orc_record *out;
  int32_t next; /* shared orientation counter */
  int32_t err;
} orc_job;

static void *orc_worker(void *arg) {
  orc_job *j = (orc_job *)arg;
  orc_ws w;
  if (orc_ws_init(&w, j->c->grid, j->nVtx)) {
    __atomic_store_n(&j->err, -2, __ATOMIC_RELAXED);
    orc_ws_free(&w);
    return NULL;
  }
  for (;;) {
    int32_t i = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
    if (i >= j->nYPR) break;
    orc_eval(&w, j->c, j->thr, j->nTri, j->tri, j->nVtx, j->vtx, j->vnrm, j->ypr + (size_t)i * 3,
             j->out + i, NULL, NULL, NULL, NULL);
  }
  orc_ws_free(&w);
  return NULL;
}

int orc_sweep(const orc_opts *opt, int32_t nTri, const int32_t *tri, int32_t nVtx,
              const float *vtx, const float *vnrm, int32_t nYPR, const float *ypr,
              orc_record *out, int32_t nthreads) {
  if (!opt || !out || nTri < 0 || nVtx < 0 || nYPR < 0 || (nYPR > 0 && !ypr) ||
      (nTri > 0 && (!tri || !vtx || !vnrm)))
    return -1;
  orc_ctx c;
  orc_ctx_init(&c, opt, nVtx, vtx);
  orc_job j;
  memset(&j, 0, sizeof(j));
  j.opt = opt; j.c = &c; j.thr = (float)(-1.0 * sin((double)c.theta_c));
  j.nTri = nTri; j.nVtx = nVtx; j.nYPR = nYPR; j.tri = tri; j.vtx = vtx; j.vnrm = vnrm;
  j.ypr = ypr; j.out = out;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nYPR && nYPR > 0) nthreads = nYPR;
  if (nthreads <= 1) {
    orc_worker(&j);
  } else {
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int k = 0; k < nthreads; k++) pthread_create(&th[k], NULL, orc_worker, &j);
    for (int k = 0; k < nthreads; k++) pthread_join(th[k], NULL);
    free(th);
  }
  return j.err;
}

int orc_pair_slot(int32_t n, const int16_t *z, const int16_t *nz, float theta_c, int32_t explicit_ss,
                  int32_t *vo, int32_t *vss, int16_t *z_out, int16_t *nz_out, int16_t *type_out) {
  if (n < 0 || n > ORC_MAXPIX) return -1;
  int16_t slot[(ORC_ROWS + 1) * 3];
  memset(slot, 0, sizeof(slot));
  slot[0] = (int16_t)n;
  for (int p = 0; p < n; p++) { slot[(p + 1) * 3 + 1] = z[p]; slot[(p + 1) * 3 + 2] = nz[p]; }
  int32_t flags = 0;
  *vo = 0; *vss = 0;
  if (n > 1) n = orc_vslot_pair(slot, n, (float)(-1.0 * sin((double)theta_c)), explicit_ss, vo, vss, &flags);
  for (int p = 0; p < n; p++) {
    if (z_out) z_out[p] = slot[(p + 1) * 3 + 1];
    if (nz_out) nz_out[p] = slot[(p + 1) * 3 + 2];
    if (type_out) type_out[p] = slot[(p + 1) * 3 + 0];
  }
  return n;
}

int orc_bary_point(const float p[2], const float a[2], const float b[2], const float c[2], float uvw[3]) {
  float a3[3] = {a[0], a[1], 0.0f}, b3[3] = {b[0], b[1], 0.0f}, c3[3] = {c[0], c[1], 0.0f};
  float u = -1.0f, v = -1.0f, w = -1.0f;
  int ok = orc_bary(p[0], p[1], a3, b3, c3, &u, &v, &w);
  uvw[0] = u; uvw[1] = v; uvw[2] = w;
  return ok;
}

int orc_bary_parts(const float p[2], const float a[2], const float b[2], const float c[2], float part[4]) {
  float a3[3] = {a[0], a[1], 0.0f}, b3[3] = {b[0], b[1], 0.0f}, c3[3] = {c[0], c[1], 0.0f};
  float u, v, w;
  return orc_bary_full(p[0], p[1], a3, b3, c3, &u, &v, &w, part);
}

int64_t orc_insert_log(const orc_opts *opt, int32_t nTri, const int32_t *tri, int32_t nVtx,
                       const float *vtx, const float *vnrm, const float ypr[3], int64_t cap,
                       int32_t *out) {
  if (!opt || nTri < 0 || nVtx < 0 || (nTri > 0 && (!tri || !vtx || !vnrm)) || cap < 0 || (cap > 0 && !out))
    return -1;
  orc_ctx c;
  orc_ctx_init(&c, opt, nVtx, vtx);
  orc_ws w;
  if (orc_ws_init(&w, c.grid, nVtx)) { orc_ws_free(&w); return -2; }
  float R[9], mat[16];
  orc_ypr_matrix(ypr[0], ypr[1], ypr[2], R);
  orc_rotate(&w, &c, R, nVtx, vtx, vnrm, mat);
  w.log = out;
  w.caplog = cap;
  w.nlog = 0;
  orc_pixelize(&w, &c, nTri, tri);
  int64_t n = w.nlog;
  orc_ws_free(&w);
  return n;
}

int32_t orc_find_first_optimal(int32_t n, const float *mss) {
  if (n <= 1) return 0;
  int32_t best = 0;
  float mv = 1e5f;
  for (int32_t i = 0; i < n; i++)
    if (mss[i] < mv) { mv = mss[i]; best = i; }
  return best;
}
