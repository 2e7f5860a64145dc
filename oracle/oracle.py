"""ctypes wrapper of the CPU oracle (oracle/tomo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  The oracle restates the reference's INT3
CPU engine (cfms_tomo/Tomo_Shell_vc_src/cpu_src/STomoSh_INT3.cpp and friends); see the C
source for per-function citations.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

RECORD_DTYPE = np.dtype(
    [("Va", "<i8"), ("Vb", "<i8"), ("Vo", "<i8"), ("Vss", "<i8"), ("Vnv", "<i8"),
     ("Vtc", "<i8"), ("Vbed", "<i8"), ("n_entries", "<i8"), ("n_slots", "<i8"),
     ("n_final", "<i8"), ("chk_pre", "<u8"), ("chk_final", "<u8"), ("flags", "<i4"),
     ("n_bed", "<i4"), ("fVa", "<f4"), ("fVb", "<f4"), ("fVtc", "<f4"), ("fVnv", "<f4"),
     ("fVss", "<f4"), ("fVo", "<f4"), ("fVbed", "<f4"), ("Mo", "<f4"), ("Mss", "<f4"),
     ("Mtotal", "<f4"), ("mat4x4", "<f4", (16,))], align=True)
assert RECORD_DTYPE.itemsize == 208

F_FP32_INEXACT, F_INT16_WRAP, F_FOOTPRINT_OOB, F_CAP_DROP, F_SENTINEL_DROP = 1, 2, 4, 8, 16


class _Opts(ct.Structure):
    _fields_ = [("f32", ct.c_float * 12), ("i32", ct.c_int32 * 11), ("grid", ct.c_int32)]


_lib = None


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        P = ct.c_void_p
        L.orc_ypr_matrix.argtypes = (ct.c_float, ct.c_float, ct.c_float, P)
        L.orc_theta_cuts.argtypes = (ct.c_float, P, P)
        L.orc_sweep.argtypes = (P, ct.c_int32, P, ct.c_int32, P, P, ct.c_int32, P, P, ct.c_int32)
        L.orc_sweep.restype = ct.c_int
        L.orc_one.argtypes = (P, ct.c_int32, P, ct.c_int32, P, P, P, P, P, P, P, P)
        L.orc_one.restype = ct.c_int
        L.orc_one_rows.argtypes = (P, ct.c_int32, P, ct.c_int32, P, P, P, P, P, P, P, P, P)
        L.orc_one_rows.restype = ct.c_int
        L.orc_pair_slot.argtypes = (ct.c_int32, P, P, ct.c_float, ct.c_int32, P, P, P, P, P)
        L.orc_pair_slot.restype = ct.c_int
        L.orc_vol_to_mass.argtypes = (P, P)
        L.orc_bary_point.argtypes = (P, P, P, P, P)
        L.orc_bary_point.restype = ct.c_int
        L.orc_bary_parts.argtypes = (P, P, P, P, P)
        L.orc_bary_parts.restype = ct.c_int
        L.orc_insert_log.argtypes = (P, ct.c_int32, P, ct.c_int32, P, P, P, ct.c_int64, P)
        L.orc_insert_log.restype = ct.c_int64
        L.orc_find_first_optimal.argtypes = (ct.c_int32, P)
        L.orc_find_first_optimal.restype = ct.c_int32
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p) if a is not None else None


def make_opts(f32, i32, grid):
    o = _Opts()
    f = np.asarray(f32, dtype=np.float32)
    i = np.asarray(i32, dtype=np.int32)
    assert f.size == 12 and i.size == 11
    for k in range(12):
        o.f32[k] = float(f[k])
    for k in range(11):
        o.i32[k] = int(i[k])
    o.grid = int(grid)
    return o


def ypr_matrix(yaw, pitch, roll):
    R = np.zeros(9, dtype=np.float32)
    lib().orc_ypr_matrix(ct.c_float(yaw), ct.c_float(pitch), ct.c_float(roll), _ptr(R))
    return R


def theta_cuts(theta_c):
    a = np.zeros(1, np.int32)
    b = np.zeros(1, np.int32)
    lib().orc_theta_cuts(ct.c_float(theta_c), _ptr(a), _ptr(b))
    return int(a[0]), int(b[0])


def _mesh_arrays(tri, vtx, vnrm):
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    vtx = np.ascontiguousarray(vtx, dtype=np.float32)
    vnrm = np.ascontiguousarray(vnrm, dtype=np.float32)
    return tri, vtx, vnrm


def sweep(f32, i32, grid, tri, vtx, vnrm, ypr, nthreads=1):
    tri, vtx, vnrm = _mesh_arrays(tri, vtx, vnrm)
    ypr = np.ascontiguousarray(ypr, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(ypr.shape[0], dtype=RECORD_DTYPE)
    o = make_opts(f32, i32, grid)
    rc = lib().orc_sweep(ct.byref(o), tri.shape[0], _ptr(tri), vtx.shape[0], _ptr(vtx), _ptr(vnrm),
                         ypr.shape[0], _ptr(ypr), _ptr(out), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"orc_sweep failed: {rc}")
    return out


def one(f32, i32, grid, tri, vtx, vnrm, ypr3):
    """One orientation with per-slot dumps: (record, cnt_pre, cnt_final, slot_vo, slot_vss)."""
    tri, vtx, vnrm = _mesh_arrays(tri, vtx, vnrm)
    y = np.ascontiguousarray(ypr3, dtype=np.float32).reshape(3)
    out = np.zeros(1, dtype=RECORD_DTYPE)
    arrs = [np.zeros(grid * grid, np.int32) for _ in range(4)]
    o = make_opts(f32, i32, grid)
    rc = lib().orc_one(ct.byref(o), tri.shape[0], _ptr(tri), vtx.shape[0], _ptr(vtx), _ptr(vnrm),
                       _ptr(y), _ptr(out), *[_ptr(a) for a in arrs])
    if rc != 0:
        raise RuntimeError(f"orc_one failed: {rc}")
    return (out[0],) + tuple(arrs)


def one_rows(f32, i32, grid, tri, vtx, vnrm, ypr3):
    """One orientation: (record, cnt_final, rows[grid*grid*64] uint64)."""
    tri, vtx, vnrm = _mesh_arrays(tri, vtx, vnrm)
    y = np.ascontiguousarray(ypr3, dtype=np.float32).reshape(3)
    out = np.zeros(1, dtype=RECORD_DTYPE)
    fin = np.zeros(grid * grid, np.int32)
    rows = np.zeros(grid * grid * 64, np.uint64)
    o = make_opts(f32, i32, grid)
    rc = lib().orc_one_rows(ct.byref(o), tri.shape[0], _ptr(tri), vtx.shape[0], _ptr(vtx), _ptr(vnrm),
                            _ptr(y), _ptr(out), None, _ptr(fin), None, None, _ptr(rows))
    if rc != 0:
        raise RuntimeError(f"orc_one_rows failed: {rc}")
    return out[0], fin, rows


def pair_slot(z, nz, theta_c, explicit_ss=True):
    """Pair one hand-made slot; returns (vo, vss, z_sorted, nz_sorted, types)."""
    z = np.ascontiguousarray(z, dtype=np.int16)
    nz = np.ascontiguousarray(nz, dtype=np.int16)
    n = z.size
    vo = np.zeros(1, np.int32)
    vss = np.zeros(1, np.int32)
    zo = np.zeros(64, np.int16)
    nzo = np.zeros(64, np.int16)
    to = np.zeros(64, np.int16)
    m = lib().orc_pair_slot(n, _ptr(z), _ptr(nz), ct.c_float(theta_c), int(bool(explicit_ss)),
                            _ptr(vo), _ptr(vss), _ptr(zo), _ptr(nzo), _ptr(to))
    if m < 0:
        raise ValueError("bad slot")
    return int(vo[0]), int(vss[0]), zo[:m].copy(), nzo[:m].copy(), to[:m].copy()


def bary_point(p, a, b, c):
    """(accepted, u, v, w) of _getBaryCoord for one pixel centre."""
    arr = [np.ascontiguousarray(x, np.float32) for x in (p, a, b, c)]
    uvw = np.zeros(3, np.float32)
    ok = lib().orc_bary_point(*[_ptr(x) for x in arr], _ptr(uvw))
    return bool(ok), uvw


def bary_parts(p, a, b, c):
    """(accepted, den, v, w, u) of _getBaryCoord for one pixel centre (v, w, u = 0 if |den| <= eps)."""
    arr = [np.ascontiguousarray(x, np.float32) for x in (p, a, b, c)]
    part = np.zeros(4, np.float32)
    ok = lib().orc_bary_parts(*[_ptr(x) for x in arr], _ptr(part))
    return bool(ok), part[0], part[1], part[2], part[3]


INSERT_LOG_DTYPE = np.dtype([("tri", "<i4"), ("seq", "<i4"), ("slot", "<i4"), ("key", "<u4")])


def insert_log(f32, i32, grid, tri, vtx, vnrm, ypr3, cap=1 << 22):
    """Every SetBit call of Pixelize for one orientation, in call order: (tri, seq, slot, key)."""
    tri, vtx, vnrm = _mesh_arrays(tri, vtx, vnrm)
    y = np.ascontiguousarray(ypr3, dtype=np.float32).reshape(3)
    out = np.zeros(cap, dtype=INSERT_LOG_DTYPE)
    o = make_opts(f32, i32, grid)
    n = lib().orc_insert_log(ct.byref(o), tri.shape[0], _ptr(tri), vtx.shape[0], _ptr(vtx), _ptr(vnrm),
                             _ptr(y), cap, _ptr(out))
    if n < 0 or n > cap:
        raise RuntimeError(f"orc_insert_log failed or overflowed: {n}")
    return out[:n].copy()


def find_first_optimal(mss):
    m = np.ascontiguousarray(mss, dtype=np.float32)
    return int(lib().orc_find_first_optimal(m.size, _ptr(m)))
