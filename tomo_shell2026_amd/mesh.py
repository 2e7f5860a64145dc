"""Trimesh-free mesh front end.

The reference's run_dll reads vertices, vertex normals, faces, face normals, area,
watertightness and the convex hull from a trimesh object
(cfms_tomo/tomo_dll_interface.py:59-70).  trimesh is not a dependency here: this
module computes the same quantities with numpy so the drop-in path runs without it.
Vertex normals are the angle-weighted sum of incident face normals, normalised (trimesh's
default weighting); results are not pinned against trimesh itself.
"""
import numpy as np


class TomoMesh:
    """Minimal stand-in for the trimesh attributes run_dll reads."""

    def __init__(self, vertices, faces, vertex_normals=None, name="mesh"):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64)
        self.faces = np.ascontiguousarray(faces, dtype=np.int64)
        self.name = name
        self._vn = None if vertex_normals is None else np.asarray(vertex_normals, np.float64)

    @property
    def face_normals(self):
        return face_normals(self.vertices, self.faces)

    @property
    def vertex_normals(self):
        if self._vn is None:
            self._vn = vertex_normals(self.vertices, self.faces)
        return self._vn

    @property
    def area(self):
        return float(mesh_area(self.vertices, self.faces))

    @property
    def is_watertight(self):
        return is_watertight(self.faces)

    @property
    def bounds(self):
        return np.stack([self.vertices.min(0), self.vertices.max(0)])

    def copy(self):
        return TomoMesh(self.vertices.copy(), self.faces.copy(),
                        None if self._vn is None else self._vn.copy(), self.name)

    def apply_transform(self, m4):
        m4 = np.asarray(m4, np.float64)
        v = self.vertices @ m4[:3, :3].T + m4[:3, 3]
        n = None if self._vn is None else self._vn @ m4[:3, :3].T
        self.vertices = v
        self._vn = n
        return self


def face_normals(v, f):
    v = np.asarray(v, np.float64)
    e1 = v[f[:, 1]] - v[f[:, 0]]
    e2 = v[f[:, 2]] - v[f[:, 0]]
    n = np.cross(e1, e2)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    return n / np.where(ln > 0, ln, 1.0)


def mesh_area(v, f):
    v = np.asarray(v, np.float64)
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    return 0.5 * np.linalg.norm(n, axis=1).sum()


def face_angles(v, f):
    v = np.asarray(v, np.float64)
    out = np.zeros(f.shape, np.float64)
    for k in range(3):
        a = v[f[:, k]]
        b = v[f[:, (k + 1) % 3]] - a
        c = v[f[:, (k + 2) % 3]] - a
        nb = np.linalg.norm(b, axis=1)
        nc = np.linalg.norm(c, axis=1)
        cosang = np.einsum("ij,ij->i", b, c) / np.maximum(nb * nc, 1e-300)
        out[:, k] = np.arccos(np.clip(cosang, -1.0, 1.0))
    return out


def vertex_normals(v, f):
    fn = face_normals(v, f)
    ang = face_angles(v, f)
    acc = np.zeros((len(v), 3), np.float64)
    for k in range(3):
        np.add.at(acc, f[:, k], fn * ang[:, k:k + 1])
    ln = np.linalg.norm(acc, axis=1, keepdims=True)
    return acc / np.where(ln > 0, ln, 1.0)


def is_watertight(f):
    """Every undirected edge shared by exactly two faces with opposite orientation."""
    f = np.asarray(f, np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    und = np.sort(e, axis=1)
    key = und[:, 0] * (int(f.max()) + 1) + und[:, 1]
    _, cnt = np.unique(key, return_counts=True)
    if not np.all(cnt == 2):
        return False
    dkey = e[:, 0] * (int(f.max()) + 1) + e[:, 1]
    return np.unique(dkey).size == dkey.size


def load_mesh(filename):
    """Read a triangle mesh from an STL (binary or ASCII) or OBJ file without trimesh, merging
    vertices with equal coordinates (rounded to 8 decimals, as trimesh merges STL corners).
    Polygons of OBJ files are split into fans.  Returns a TomoMesh."""
    path = str(filename)
    with open(path, "rb") as fh:
        data = fh.read()
    low = path.lower()
    if low.endswith(".obj"):
        vs, fs = [], []
        for line in data.decode(errors="replace").splitlines():
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                vs.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = [int(p.split("/")[0]) for p in parts[1:]]
                idx = [i - 1 if i > 0 else len(vs) + i for i in idx]
                fs += [[idx[0], idx[k], idx[k + 1]] for k in range(1, len(idx) - 1)]
        return TomoMesh(np.array(vs, np.float64).reshape(-1, 3), np.array(fs, np.int64).reshape(-1, 3),
                        name=path)
    if not low.endswith(".stl"):
        raise ValueError(f"unsupported mesh file (STL or OBJ expected): {path}")
    corners = None
    if len(data) >= 84:
        n = int(np.frombuffer(data, "<u4", 1, 80)[0])
        if len(data) == 84 + 50 * n:
            rec = np.frombuffer(data, np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]), n, 84)
            corners = rec["v"].astype(np.float64).reshape(-1, 3)
    if corners is None:
        pts = [[float(x) for x in line.split()[1:4]] for line in data.decode(errors="replace").splitlines()
               if line.strip().startswith("vertex")]
        corners = np.array(pts, np.float64).reshape(-1, 3)
    uniq, inv = np.unique(np.round(corners, 8), axis=0, return_inverse=True)
    return TomoMesh(uniq, inv.reshape(-1, 3).astype(np.int64), name=path)


def save_stl(filename, vertices, faces):
    """Write a binary STL (the format validate_primitives.py:17-20 exports its cases in)."""
    v = np.asarray(vertices, np.float64)
    f = np.asarray(faces, np.int64)
    rec = np.zeros(len(f), np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    rec["n"] = face_normals(v, f).astype(np.float32)
    rec["v"] = v[f].astype(np.float32)
    with open(filename, "wb") as fh:
        fh.write(b"\0" * 80)
        fh.write(np.array([len(f)], "<u4").tobytes())
        fh.write(rec.tobytes())
