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
