"""Small seeded meshes for parity tests (edge cases the reference's behaviour defines)."""
import numpy as np

from tomo_shell2026_amd import synth
from tomo_shell2026_amd.mesh import vertex_normals


def box(lo, hi):
    """Closed axis-aligned box with outward faces."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [1, 2, 6], [1, 6, 5], [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int64)
    return v, f


def stacked_planes(n_layers=70, size=6.0, dz=1.7):
    """n horizontal squares stacked in z with alternating facing: saturates the 63-row slots."""
    vs, fs = [], []
    for k in range(n_layers):
        z = 1.3 + k * dz
        b = len(vs)
        vs += [[0.0, 0.0, z], [size, 0.0, z], [size, size, z], [0.0, size, z]]
        if k % 2 == 0:
            fs += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        else:
            fs += [[b, b + 2, b + 1], [b, b + 3, b + 2]]
    return np.array(vs, np.float64), np.array(fs, np.int64)


def tiny_triangles(n=40, seed=3):
    """Triangles far smaller than a voxel: every one takes the centroid fallback."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(5, 25, size=(n, 3))
    d = rng.normal(0, 0.05, size=(n, 3, 3))
    v = (c[:, None, :] + d).reshape(-1, 3)
    f = np.arange(3 * n).reshape(n, 3)
    return v, f


def normals_for(v, f):
    return vertex_normals(v, f).astype(np.float32)


def make(v, f, grid, ypr, bed=synth.BED_NONE, explicit_ss=True, shell=False, shell_um=500,
         theta_deg=60.0, bed_outer=2.0, bed_inner=0.0):
    v = np.asarray(v, np.float32)
    f = np.asarray(f, np.int32)
    n = normals_for(v, f)
    from tomo_shell2026_amd.mesh import mesh_area
    f32 = synth.default_f32(mesh_area(v, f), theta_deg=theta_deg, bed_inner=bed_inner,
                            bed_outer=bed_outer)
    ypr = np.asarray(ypr, np.float32).reshape(-1, 3)
    i32 = synth.default_i32(len(f), len(v), len(ypr), grid=grid, explicit_ss=explicit_ss,
                            shell=shell, bed_type=bed, shell_um=shell_um)
    return dict(f32=f32, i32=i32, grid=grid, tri=f, vtx=v, vnrm=n, ypr=ypr)


def coarse_ypr(step=30, pitch_max=180):
    return synth.ypr_grid(np.arange(0, 360, step), np.arange(0, pitch_max + 1, step))


INT_FIELDS = ("Va", "Vb", "Vo", "Vss", "Vnv", "Vtc", "Vbed", "n_entries", "n_slots", "n_final",
              "chk_pre", "chk_final", "n_bed")
FLOAT_FIELDS = ("fVa", "fVb", "fVtc", "fVnv", "fVss", "fVo", "fVbed", "Mo", "Mss", "Mtotal")


def compare_records(gpu, ref, label=""):
    """Bit-exact comparison of GPU and oracle records; returns a list of mismatch strings."""
    bad = []
    for k in INT_FIELDS:
        d = np.nonzero(gpu[k] != ref[k])[0]
        if d.size:
            i = d[0]
            bad.append(f"{label}{k}: {d.size} differ, first #{i}: gpu {gpu[k][i]} ref {ref[k][i]}")
    un = np.nonzero(gpu["flags"] & 64)[0]  # UNRESOLVED: the library could not keep an entry
    if un.size:
        bad.append(f"{label}unresolved entries in {un.size} orientations, first #{un[0]}")
    fl = np.nonzero(gpu["flags"] != ref["flags"])[0]
    if fl.size:
        i = fl[0]
        bad.append(f"{label}flags: {fl.size} differ, first #{i}: gpu {gpu['flags'][i]} ref {ref['flags'][i]}")
    for k in FLOAT_FIELDS:
        g = gpu[k].view(np.uint32)
        r = ref[k].view(np.uint32)
        d = np.nonzero(g != r)[0]
        if d.size:
            i = d[0]
            bad.append(f"{label}{k}: {d.size} differ, first #{i}: gpu {gpu[k][i]!r} ref {ref[k][i]!r}")
    m = np.nonzero(np.any(gpu["mat4x4"].view(np.uint32) != ref["mat4x4"].view(np.uint32), axis=1))[0]
    if m.size:
        bad.append(f"{label}mat4x4: {m.size} differ, first #{m[0]}")
    return bad


# ---- the 63-row cap inside one triangle ------------------------------------------------------
# Stacked squares fill every slot of a 50 x 50 region with 62 distinct rows (31 layers and their
# shell sub-rows in shell mode), then one sliver triangle puts the rows of several of its pixels
# into one of those slots: the interpolated point of a near-degenerate triangle leaves the pixel
# it was tested at.  SetBit keeps the sliver's row of the first such pixel in x-major order
# (STomoSh_INT3.cpp:1163-1186, STomoVoxelSpaceInfo.cpp:200-220); a rank of 2 * triangle + sub
# alone ties these rows.  Vertex values (float32) and normals come from a seeded search; the CPU
# tests check with the oracle's insertion log that each case really separates the two orders.
TIE_SLIVERS = {
    "solid_a": dict(shell=False, slot=(40, 29), nrm=(0.9566758871078491, 0.29115498065948486),
                    tri=(15.899605751037598, 16.96215057373047, 25.31871795654297,
                         49.05427932739258, 34.119590759277344, 42.59214401245117,
                         40.45609664916992, 29.670085906982422, 37.162994384765625)),
    "solid_b": dict(shell=False, slot=(29, 31), nrm=(0.7995659708976746, 0.600578248500824),
                    tri=(16.11545181274414, 19.215417861938477, 60.51253128051758,
                         42.303863525390625, 44.159305572509766, 52.198184967041016,
                         33.63581085205078, 35.903202056884766, 53.88571548461914)),
    "shell": dict(shell=True, slot=(17, 19), nrm=(0.9145917296409607, 0.4043785333633423),
                  tri=(17.74104881286621, 13.93664836883545, 76.2462387084961,
                       15.261141777038574, 35.62145233154297, 52.66484069824219,
                       16.142107009887695, 27.918020248413086, 60.374210357666016)),
}


def _square_layers(n, x0, x1, y0, y1, z0, dz, period=None):
    """n horizontal squares [x0, x1] x [y0, y1], layer k at z0 + dz * (k mod period), facing +z
    for even k mod period and -z for odd (vertices, faces, vertex normals)."""
    vs, fs, ns = [], [], []
    period = period or n
    for k in range(n):
        j = k % period
        z = z0 + j * dz
        b = len(vs)
        vs += [[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]]
        up = j % 2 == 0
        fs += [[b, b + 1, b + 2], [b, b + 2, b + 3]] if up else [[b, b + 2, b + 1], [b, b + 3, b + 2]]
        ns += [[0.0, 0.0, 1.0 if up else -1.0]] * 4
    return vs, fs, ns


def tie_at_cap(name, extra=0):
    """One TIE_SLIVERS case as a make() dict on a 64 grid at the identity orientation.  extra > 0
    appends that many small squares over the case's slot after the sliver, repeating the first
    layers' rows: the slot then holds over 512 (or 8,192) entries but no new key."""
    c = TIE_SLIVERS[name]
    shell = c["shell"]
    nl, dz = (31, 5.0) if shell else (62, 1.3)
    vs, fs, ns = _square_layers(nl, 2.0, 52.0, 2.0, 52.0, 1.0, dz)
    b = len(vs)
    t = c["tri"]
    vs += [list(t[0:3]), list(t[3:6]), list(t[6:9])]
    fs += [[b, b + 1, b + 2]]
    ns += [[0.0, c["nrm"][0], c["nrm"][1]]] * 3
    if extra:
        sx, sy = c["slot"]
        v2, f2, n2 = _square_layers(extra, sx + 0.2, sx + 0.9, sy + 0.3, sy + 0.7, 1.0, dz, period=nl)
        b = len(vs)
        vs += v2
        fs += [[q + b for q in tri_] for tri_ in f2]
        ns += n2
    case = make(np.array(vs, np.float32), np.array(fs, np.int32), 64, [[0.0, 0.0, 0.0]], shell=shell,
                shell_um=1500)
    case["vnrm"] = np.array(ns, np.float32)
    return case


def kept_by_order(log, slot):
    """The keys SetBit keeps in one slot from an insertion log: the first 63 distinct keys in call
    order, and the first 63 when ordered by (2 * triangle + sub, key) instead."""
    e = log[log["slot"] == slot]

    def first63(keys):
        seen = []
        for k in keys:
            k = int(k)
            if k not in seen:
                seen.append(k)
        return seen[:63], len(seen)

    exact, n_distinct = first63(e["key"])
    rank = 2 * e["tri"].astype(np.int64) + (e["seq"] & 1)
    by_rank, _ = first63(e["key"][np.lexsort((e["key"], rank))])
    return set(exact), set(by_rank), n_distinct, len(e)


# ---- the reference's validation primitives (cfms_tomo/shell_test/validate_primitives.py) ------
def slice_plane(v, f, normal, cap):
    """The part of a closed triangle mesh on the side x . n >= 0 of a plane through the origin,
    with the triangles that cross the plane clipped (one Sutherland-Hodgman pass each, vertices
    on the plane kept, cut points shared between neighbours).  With cap the cut is closed by a
    fan of triangles around its centroid, wound to face -n: for a convex solid the cut is one
    convex loop.  Stands in for trimesh's slice_plane(..., cap=...) as
    validate_primitives.py:25-30 uses it."""
    v = np.asarray(v, np.float64)
    f = np.asarray(f, np.int64)
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    d = v @ n
    d[np.abs(d) < 1e-12] = 0.0
    verts = [tuple(p) for p in v]
    cut_of = {}

    def cut(i, j):
        key = (min(i, j), max(i, j))
        if key not in cut_of:
            t = d[i] / (d[i] - d[j])
            verts.append(tuple(v[i] + t * (v[j] - v[i])))
            cut_of[key] = len(verts) - 1
        return cut_of[key]

    faces = []
    for tri in f:
        poly = []
        for k in range(3):
            i, j = int(tri[k]), int(tri[(k + 1) % 3])
            if d[i] >= 0:
                poly.append(i)
            if (d[i] > 0 and d[j] < 0) or (d[i] < 0 and d[j] > 0):
                poly.append(cut(i, j))
        for k in range(1, len(poly) - 1):
            faces.append([poly[0], poly[k], poly[k + 1]])
    V = np.array(verts)
    F_ = np.array(faces, np.int64)
    area = 0.5 * np.linalg.norm(np.cross(V[F_[:, 1]] - V[F_[:, 0]], V[F_[:, 2]] - V[F_[:, 0]]), axis=1)
    F_ = F_[area > 1e-12]
    if cap:
        e = np.concatenate([F_[:, [0, 1]], F_[:, [1, 2]], F_[:, [2, 0]]])
        und = {}
        for a, b in e:
            k = (min(a, b), max(a, b))
            und[k] = und.get(k, 0) + 1
        border = [(a, b) for a, b in e if und[(min(a, b), max(a, b))] == 1]
        ring = sorted({q for edge in border for q in edge})
        c = V[ring].mean(0)
        V = np.vstack([V, c])
        ci = len(V) - 1
        F_ = np.vstack([F_, np.array([[b, a, ci] for a, b in border], np.int64)])
    used = np.unique(F_)
    remap = -np.ones(len(V), np.int64)
    remap[used] = np.arange(len(used))
    return V[used], remap[F_]


def validation_primitives():
    """The ten cases of validate_primitives.py:149-179: a 20 mm cube and an r = 10 icosphere
    (3 subdivisions), each closed, upper half open (shell mode), upper half capped, slant-cut open
    (shell mode) and slant-capped, the slant plane's normal (1, 0, 0.7).  (name, v, f, shell)."""
    cv, cf = box((-10, -10, -10), (10, 10, 10))
    sv, sf = synth.icosphere(3, 10.0)
    out = []
    for base, (v, f) in (("cube", (cv, cf)), ("sphere", (sv, sf))):
        out.append((f"{base}_closed", v, f, False))
        for cut_name, nrm in (("upper", (0.0, 0.0, 1.0)), ("slant", (1.0, 0.0, 0.7))):
            ov, of = slice_plane(v, f, nrm, cap=False)
            out.append((f"{base}_{cut_name}_open", ov, of, True))
            kv, kf = slice_plane(v, f, nrm, cap=True)
            out.append((f"{base}_{cut_name}_capped", kv, kf, False))
    return out


def validation_ypr(step_deg=30):
    """run_tomo's grid (validate_primitives.py:38-50): yaw = pitch = linspace(0, 2 pi, 360/step + 1)
    in float32, roll 0, meshgrid order (yaw fastest)."""
    return synth.reference_ypr_grid(step_deg)
