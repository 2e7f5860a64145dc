"""Seeded synthetic workloads for the support-tomography sweep.

Meshes and orientation grids for the benchmark configurations:
  * icosphere(subdiv=5, r=50 mm): 20,480 triangles, 1 mm voxels, 128 x 128 grid,
    yaw 0..350 x pitch 0..180 in 10 degree steps (684 orientations);
  * displaced lat-long sphere: 69,120 triangles (240 longitudes x 145 latitude bands),
    base radius 60 mm perturbed by real spherical harmonics l = 1..8 with coefficients
    N(0, 4 mm)/l, 0.5 mm voxels, 256 x 256 grid, yaw 0..359 x pitch 0..60 in 1 degree steps
    (21,960 orientations).
Meshes are expressed in voxel units (pre-scaled by 1/voxel), because the reference uses vertex
coordinates directly as voxel indices (STomoVoxelSpaceInfo.cpp:97-105) with dVoxel pinned to 1.
"""
import math

import numpy as np
from scipy import special

from .mesh import TomoMesh, mesh_area, vertex_normals


# ---------------------------------------------------------------- options (Float32Info / Int32Info)
def default_f32(area, theta_deg=60.0, bed_inner=0.0, bed_outer=2.0, bed_thick=0.3 + 0.27 + 2 * 0.2):
    """Float32Info[12] in the order of tomo_dll_interface.py:73-76 with SlicingOptions defaults
    (tomo_slicingOptions.py:41-58)."""
    return np.array([1.0, theta_deg * np.pi / 180.0, area, 0.8, 0.15, 1.0, 0.2, 1.0, 0.001121,
                     bed_inner, bed_outer, bed_thick], dtype=np.float32)


BED_NONE, BED_BRIM, BED_RAFT, BED_SKIRT = 0, 1, 2, 3


def default_i32(n_tri, n_vtx, n_ypr, grid=256, explicit_ss=True, shell=False, bed_type=BED_RAFT,
                shell_um=500):
    """Int32Info[11] in the order of tomo_dll_interface.py:78-82."""
    return np.array([0, int(explicit_ss), int(shell), grid, n_tri, n_vtx, n_ypr, 0, 0, bed_type,
                     shell_um], dtype=np.int32)


# ---------------------------------------------------------------- orientation grids
def ypr_grid(yaw_deg, pitch_deg):
    """Rows pitch-major, yaw fastest, roll 0 (tomo_dll_interface.py:38-43 ordering).
    Angles are float32(deg * pi / 180) computed in float64."""
    yaw = np.asarray(yaw_deg, np.float64) * np.pi / 180.0
    pit = np.asarray(pitch_deg, np.float64) * np.pi / 180.0
    yy, pp = np.meshgrid(yaw, pit)
    return np.column_stack([yy.ravel(), pp.ravel(), np.zeros(yy.size)]).astype(np.float32)


def reference_ypr_grid(theta_step_deg):
    """The drop-in grid of TomoDLLInterface.__init__ (tomo_dll_interface.py:37-43): yaw and pitch
    both linspace(0, 2 pi, 360/step + 1, float32), roll 0."""
    n = int(360 / theta_step_deg) + 1
    r = np.linspace(0.0, 2 * np.pi, num=n, endpoint=True, dtype=np.float32)
    xx, yy, zz = np.meshgrid(r, r, np.zeros(1))
    return np.column_stack([xx.ravel(), yy.ravel(), zz.ravel()]).astype(np.float32)


# ---------------------------------------------------------------- meshes
def icosphere(subdiv, radius):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e = np.sort(e, axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.ravel()
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v * radius, f


def real_sph_harm(l, m, polar, azim):
    """Real orthonormal spherical harmonic Y_lm."""
    y = special.sph_harm_y(l, abs(m), polar, azim)
    if m == 0:
        return y.real
    s = math.sqrt(2.0) * (-1) ** m
    return s * (y.real if m > 0 else y.imag)


def displaced_latlong_sphere(n_lon=240, n_bands=145, radius=60.0, lmax=8, sigma=4.0, seed=1):
    """Closed lat-long sphere with 2*n_lon*(n_bands-1) triangles, radius perturbed by real SH."""
    rng = np.random.default_rng(seed)
    polar_rings = np.arange(1, n_bands) * np.pi / n_bands
    azim = np.arange(n_lon) * 2 * np.pi / n_lon
    P, A = np.meshgrid(polar_rings, azim, indexing="ij")
    polar = np.concatenate([[0.0], P.ravel(), [np.pi]])
    az = np.concatenate([[0.0], A.ravel(), [0.0]])
    r = np.full(polar.shape, radius, np.float64)
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            a = rng.normal(0.0, sigma) / l
            r += a * real_sph_harm(l, m, polar, az)
    v = np.stack([r * np.sin(polar) * np.cos(az), r * np.sin(polar) * np.sin(az), r * np.cos(polar)], 1)
    nr = n_bands - 1
    top, bot = 0, 1 + nr * n_lon

    def ring(k, j):
        return 1 + k * n_lon + (j % n_lon)

    faces = []
    for j in range(n_lon):
        faces.append([top, ring(0, j), ring(0, j + 1)])
    for k in range(nr - 1):
        for j in range(n_lon):
            a, b = ring(k, j), ring(k, j + 1)
            c, d = ring(k + 1, j), ring(k + 1, j + 1)
            faces.append([a, c, d])
            faces.append([a, d, b])
    for j in range(n_lon):
        faces.append([bot, ring(nr - 1, j + 1), ring(nr - 1, j)])
    return v, np.array(faces, np.int64)


def fit_to_grid(v, grid, bed_outer=2.0, margin=2.0):
    """Uniform scale (<= 1) about the centroid so the mesh footprint fits the grid for every
    orientation: the placed mesh spans [Bo, Bo + extent] in x and y (STomoSh_Base.cpp:103-109),
    and the extent is at most the mesh diameter about its centre."""
    c = 0.5 * (v.min(0) + v.max(0))
    rmax = np.linalg.norm(v - c, axis=1).max()
    limit = grid - bed_outer - margin
    s = min(1.0, limit / (2.0 * rmax))
    return c + (v - c) * s, s


class Workload:
    """A mesh in voxel units plus its options and orientation grid."""

    def __init__(self, name, vertices, faces, grid, ypr, voxel_mm, scale_note="", bed_type=BED_RAFT,
                 explicit_ss=True, shell=False, shell_um=500):
        self.name = name
        self.vtx = np.ascontiguousarray(vertices, np.float32)
        self.tri = np.ascontiguousarray(faces, np.int32)
        self.vnrm = np.ascontiguousarray(vertex_normals(vertices, faces), np.float32)
        self.grid = int(grid)
        self.ypr = np.ascontiguousarray(ypr, np.float32)
        self.voxel_mm = voxel_mm
        self.scale_note = scale_note
        self.area = float(mesh_area(vertices, faces))
        self.bed_type = bed_type
        self.explicit_ss = explicit_ss
        self.shell = shell
        self.shell_um = shell_um

    @property
    def f32(self):
        return default_f32(self.area)

    @property
    def i32(self):
        return default_i32(len(self.tri), len(self.vtx), len(self.ypr), grid=self.grid,
                           explicit_ss=self.explicit_ss, shell=self.shell, bed_type=self.bed_type,
                           shell_um=self.shell_um)

    def mesh(self):
        return TomoMesh(self.vtx.astype(np.float64), self.tri, self.vnrm, self.name)


def icosphere_workload(bed_type=BED_RAFT):
    """20,480-triangle icosphere, r = 50 mm, 1 mm voxels, 128^2 grid, 684 orientations."""
    v, f = icosphere(5, 50.0)
    ypr = ypr_grid(np.arange(0, 360, 10), np.arange(0, 181, 10))
    return Workload("icosphere20480_r50_h1_g128_10deg", v, f, 128, ypr, 1.0, bed_type=bed_type)


def manikin_workload(bed_type=BED_RAFT, seed=7, yaw_step=1, pitch_max=60):
    """~87k-triangle open thin-shell manikin (inner and outer sheets), 1 mm voxels, 512^2 grid,
    shell mode, yaw 0..359 x pitch 0..60 in 1 degree steps."""
    v, f = manikin_shell(seed=seed)
    v, s = fit_to_grid(v, 512)
    note = "" if s == 1.0 else f"scaled by {s:.4f} to fit the 512 grid"
    ypr = ypr_grid(np.arange(0, 360, yaw_step), np.arange(0, pitch_max + 1, 1))
    return Workload(f"manikin{len(f)}_shell_h1_g512_1deg", v, f, 512, ypr, 1.0, note, bed_type=bed_type,
                    shell=True, shell_um=500)


def displaced_sphere_workload(bed_type=BED_RAFT, seed=1, yaw_step=1, pitch_max=60):
    """69,120-triangle displaced sphere, 0.5 mm voxels, 256^2 grid, yaw 0..359 x pitch 0..60."""
    v, f = displaced_latlong_sphere(seed=seed)
    v = v / 0.5
    v, s = fit_to_grid(v, 256)
    note = "" if s == 1.0 else f"scaled by {s:.4f} to fit the 256 grid"
    ypr = ypr_grid(np.arange(0, 360, yaw_step), np.arange(0, pitch_max + 1, 1))
    return Workload("dsphere69120_h0.5_g256_1deg", v, f, 256, ypr, 0.5, note, bed_type=bed_type)


# ---------------------------------------------------------------- thin-shell manikin
def _sd_ellipsoid(p, c, semi):
    """Approximate signed distance to an axis-aligned ellipsoid (gradient-normalised bound)."""
    q = (p - c) / semi
    k0 = np.linalg.norm(q, axis=-1)
    k1 = np.linalg.norm(q / semi, axis=-1)
    return k0 * (k0 - 1.0) / np.maximum(k1, 1e-12)


def _sd_capsule(p, a, b, r):
    pa, ba = p - a, b - a
    h = np.clip((pa @ ba) / (ba @ ba), 0.0, 1.0)
    return np.linalg.norm(pa - h[..., None] * ba, axis=-1) - r


def _sd_sphere(p, c, r):
    return np.linalg.norm(p - c, axis=-1) - r


def manikin_parts(seed=7):
    """Seeded manikin primitives (mm): torso ellipsoid with full axes 170 x 120 x 300, head sphere
    r = 90, two arm and two leg capsules with radii drawn uniformly in [35, 50]."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(35.0, 50.0, size=4)
    j = rng.normal(0.0, 4.0, size=(4, 2))
    torso = ("ellipsoid", np.array([0.0, 0.0, 0.0]), np.array([85.0, 60.0, 150.0]))
    head = ("sphere", np.array([0.0, 0.0, 210.0]), 90.0)
    arms = [("capsule", np.array([s * 95.0, 0.0, 95.0]), np.array([s * (112.0 + j[i, 0]), j[i, 1], -55.0]), r[i])
            for i, s in enumerate((-1.0, 1.0))]
    legs = [("capsule", np.array([s * 42.0, 0.0, -110.0]), np.array([s * (50.0 + j[2 + i, 0]), j[2 + i, 1], -140.0]),
             r[2 + i]) for i, s in enumerate((-1.0, 1.0))]
    return [torso, head] + arms + legs


def manikin_sdf(p, parts):
    d = None
    for part in parts:
        if part[0] == "ellipsoid":
            e = _sd_ellipsoid(p, part[1], part[2])
        elif part[0] == "sphere":
            e = _sd_sphere(p, part[1], part[2])
        else:
            e = _sd_capsule(p, part[1], part[2], part[3])
        d = e if d is None else np.minimum(d, e)
    return d


# Freudenthal split of a cube into 6 tetrahedra sharing the diagonal (0,0,0)-(1,1,1); neighbouring
# cubes split their shared faces along the same diagonal, so the extracted surface is closed.
_CUBE = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)])  # corner c = 4i + 2j + k
_TETS = []
for _perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
    _a = np.zeros(3, int)
    _t = [0]
    for _ax in _perm[:2]:
        _a[_ax] = 1
        _t.append(int(4 * _a[0] + 2 * _a[1] + _a[2]))
    _t.append(7)
    _TETS.append(_t)
_TETS = np.array(_TETS)


def marching_tetrahedra(F, origin, h, level=0.0):
    """Closed triangle mesh of the surface F = level over a regular grid (F sampled at
    origin + h * (i, j, k)).  Faces are oriented outward, towards increasing F."""
    nx, ny, nz = F.shape
    gid = lambda i, j, k: (i * ny + j) * nz + k  # noqa: E731
    ii, jj, kk = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    Fc = np.stack([F[ii + c[0], jj + c[1], kk + c[2]] for c in _CUBE], axis=1) - level
    Fc = np.where(Fc == 0.0, 1e-6 * h, Fc)  # no vertex exactly on a grid point: keeps faces non-degenerate
    keep = (Fc.min(1) < 0) & (Fc.max(1) > 0)
    ii, jj, kk, Fc = ii[keep], jj[keep], kk[keep], Fc[keep]
    corner_id = np.stack([gid(ii + c[0], jj + c[1], kk + c[2]) for c in _CUBE], axis=1)
    corner_pos = np.stack([np.stack([ii + c[0], jj + c[1], kk + c[2]], 1) for c in _CUBE], axis=1)
    tris, keys, ts, pa, pb, ins = [], [], [], [], [], []
    for tet in _TETS:
        v = Fc[:, tet]                      # (n, 4)
        ids = corner_id[:, tet]
        pos = corner_pos[:, tet].astype(np.float64)
        inside = v < 0
        cnt = inside.sum(1)
        for want in (1, 2, 3):
            sel = np.nonzero(cnt == want)[0]
            if sel.size == 0:
                continue
            vin = inside[sel]
            # order corners: inside ones first (stable), so the crossing edges are known
            order = np.argsort(~vin, axis=1, kind="stable")
            vv = np.take_along_axis(v[sel], order, 1)
            ii4 = np.take_along_axis(ids[sel], order, 1)
            pp = np.take_along_axis(pos[sel], order[:, :, None], 1)
            if want == 1:
                edges = [[(0, 1), (0, 2), (0, 3)]]
            elif want == 3:
                edges = [[(0, 3), (1, 3), (2, 3)]]
            else:
                edges = [[(0, 2), (0, 3), (1, 3)], [(0, 2), (1, 3), (1, 2)]]
            for tri_edges in edges:
                corners = []
                for (a, b) in tri_edges:
                    fa, fb = vv[:, a], vv[:, b]
                    t = fa / (fa - fb)
                    ka = np.minimum(ii4[:, a], ii4[:, b])
                    kb = np.maximum(ii4[:, a], ii4[:, b])
                    corners.append((ka, kb, pp[:, a] + t[:, None] * (pp[:, b] - pp[:, a])))
                keys.append(np.stack([np.stack([c[0], c[1]], 1) for c in corners], 1))  # (m, 3, 2)
                ts.append(np.stack([c[2] for c in corners], 1))                             # (m, 3, 3)
                ins.append(pp[:, 0])                                                        # an inside corner
    keys = np.concatenate(keys)
    pts = np.concatenate(ts)
    pin = np.concatenate(ins)
    flat = keys.reshape(-1, 2)
    uniq, inv = np.unique(flat, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    verts = np.zeros((len(uniq), 3))
    verts[inv] = pts.reshape(-1, 3)
    faces = inv.reshape(-1, 3)
    # orient outward: the normal points away from the tetrahedron's inside corner
    a, b, c = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    n = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", n, (a + b + c) / 3.0 - pin) < 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    return origin + h * verts, faces


def _compact(v, f):
    used = np.unique(f)
    remap = -np.ones(len(v), np.int64)
    remap[used] = np.arange(len(used))
    return v[used], remap[f]


def manikin_shell(seed=7, spacing=None, wall=2.0, open_mm=15.0, target_tris=87000):
    """Open thin-shell manikin (mm): the outer surface of the primitive union (marching
    tetrahedra), plus the same surface offset inward by `wall` with reversed faces, both opened
    by removing every face below `open_mm` above the lowest point (open leg ends).  The grid
    spacing is the given one, else the one whose triangle count is closest to target_tris."""
    parts = manikin_parts(seed)

    def build(hh):
        lo = np.array([-175.0, -100.0, -200.0]) - 2 * hh
        hi = np.array([175.0, 100.0, 305.0]) + 2 * hh
        n = np.ceil((hi - lo) / hh).astype(int) + 1
        g = [lo[d] + hh * np.arange(n[d]) for d in range(3)]
        P = np.stack(np.meshgrid(*g, indexing="ij"), -1)
        F = manikin_sdf(P, parts)
        vo, fo = marching_tetrahedra(F, lo, hh, 0.0)
        vi, fi = marching_tetrahedra(F, lo, hh, -wall)
        fi = fi[:, [0, 2, 1]] + len(vo)
        v = np.concatenate([vo, vi])
        f = np.concatenate([fo, fi])
        zc = v[f].mean(1)[:, 2]
        f = f[zc > v[:, 2].min() + open_mm]
        return _compact(v, f)

    if spacing is not None:
        return build(spacing)
    best = None
    for hh in (9.0, 8.5, 8.0, 7.6, 7.3, 7.0):
        v, f = build(hh)
        if best is None or abs(len(f) - target_tris) < abs(len(best[1]) - target_tris):
            best = (v, f, hh)
    return best[0], best[1]


# ---------------------------------------------------------------- batch of uneven meshes
def displaced_sphere_n(n_tri_target, radius, seed):
    """Displaced lat-long sphere with about n_tri_target triangles (same construction as the
    69,120-triangle one: longitudes : bands ~ 5 : 3, SH l <= 8, sigma = 4 mm x radius / 60)."""
    nr = max(2, int(round(math.sqrt(n_tri_target / (2 * 1.655)))))
    n_lon = max(3, int(round(n_tri_target / (2 * nr))))
    return displaced_latlong_sphere(n_lon, nr + 1, radius, 8, 4.0 * radius / 60.0, seed)


def cut_by_planes(v, f, n_parts=6, lo=2000, hi=30000, seed=5):
    """Split a mesh into n_parts open parts by seeded planes: repeatedly cut the largest part by
    a random plane through its triangle-centroid median (triangles go whole to one side).
    Retries until every part has lo..hi triangles."""
    rng = np.random.default_rng(seed)
    cen_all = v[f].mean(1)
    for _ in range(500):
        parts = [np.arange(len(f))]
        while len(parts) < n_parts:
            k = int(np.argmax([len(p) for p in parts]))
            idx = parts.pop(k)
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            proj = cen_all[idx] @ n
            d = np.median(proj) + rng.normal(0.0, 0.2) * proj.std()
            side = proj > d
            parts += [idx[side], idx[~side]]
        if all(lo <= len(p) <= hi for p in parts):
            out = []
            for p in sorted(parts, key=len, reverse=True):
                pf = f[p]
                used, inv = np.unique(pf, return_inverse=True)
                out.append((v[used], inv.reshape(-1, 3)))
            return out
    raise RuntimeError("no plane cut met the part-size bounds")


def member_workload(name, v_mm, f, shell, ypr=None):
    """Scale a mesh (mm) to a voxel size of 0.5..2 mm chosen from its size (diameter / 200),
    and pick the smallest grid of 128 / 256 / 512 that holds it in every orientation."""
    c = 0.5 * (v_mm.min(0) + v_mm.max(0))
    diam = 2.0 * np.linalg.norm(v_mm - c, axis=1).max()
    voxel = float(np.clip(diam / 200.0, 0.5, 2.0))
    ext = diam / voxel + 4.0
    grid = 128 if ext <= 124 else (256 if ext <= 252 else 512)
    v, s = fit_to_grid(v_mm / voxel, grid)
    note = "" if s == 1.0 else f"scaled by {s:.4f} to fit the {grid} grid"
    if ypr is None:
        ypr = ypr_grid(np.arange(0, 360, 1), np.arange(0, 61, 1))
    return Workload(name, v, f, grid, ypr, voxel, note, shell=shell)


def batch70_workloads(seed=3, ypr=None):
    """70 meshes of uneven size: the thin-shell manikin cut by seeded planes into 6 open parts of
    2k-30k triangles (shell mode), and 64 closed displaced spheres with triangle counts
    log-uniform in 1k-100k and radii uniform in 20-80 mm.  Each has its own voxel size and grid
    (member_workload) and, by default, the 1 degree yaw x 0..60 pitch grid (21,960)."""
    rng = np.random.default_rng(seed)
    out = []
    v, f = manikin_shell(seed=7)
    for i, (pv, pf) in enumerate(cut_by_planes(v, f, 6, 2000, 30000, seed=seed + 2)):
        out.append(member_workload(f"manikin_part{i}_{len(pf)}", pv, pf, shell=True, ypr=ypr))
    for k in range(64):
        n = int(round(math.exp(rng.uniform(math.log(1000.0), math.log(100000.0)))))
        radius = float(rng.uniform(20.0, 80.0))
        sv, sf = displaced_sphere_n(n, radius, seed=100 + k)
        out.append(member_workload(f"dsphere{k}_{len(sf)}", sv, sf, shell=False, ypr=ypr))
    return out
