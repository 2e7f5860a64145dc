"""ctypes binding of libtomo_hip.so (declared in include/tomo_hip.h).

The library is the only compute path: if it is missing or cannot be loaded this module raises,
there is no CPU fallback.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtomo_hip.so")

TOMO_OK = 0
F_FP32_INEXACT, F_INT16_WRAP, F_FOOTPRINT_OOB, F_CAP_DROP = 1, 2, 4, 8
F_SENTINEL_DROP, F_UNRESOLVED = 16, 64

RECORD_DTYPE = np.dtype(
    [("Va", "<i8"), ("Vb", "<i8"), ("Vo", "<i8"), ("Vss", "<i8"), ("Vnv", "<i8"),
     ("Vtc", "<i8"), ("Vbed", "<i8"), ("n_entries", "<i8"), ("n_slots", "<i8"),
     ("n_final", "<i8"), ("chk_pre", "<u8"), ("chk_final", "<u8"), ("flags", "<i4"),
     ("n_bed", "<i4"), ("fVa", "<f4"), ("fVb", "<f4"), ("fVtc", "<f4"), ("fVnv", "<f4"),
     ("fVss", "<f4"), ("fVo", "<f4"), ("fVbed", "<f4"), ("Mo", "<f4"), ("Mss", "<f4"),
     ("Mtotal", "<f4"), ("mat4x4", "<f4", (16,))], align=True)
assert RECORD_DTYPE.itemsize == 208

# Every symbol include/tomo_hip.h declares.
EXPORTS = (
    "tomo_create", "tomo_set_mesh", "tomo_set_orientations", "tomo_run", "tomo_run_static", "tomo_sync",
    "tomo_get_records", "tomo_sweep", "tomo_debug_slots", "tomo_debug_rows", "tomo_set_launch", "tomo_set_tuning",
    "tomo_last_kernel_ms", "tomo_get_info", "tomo_set_phase_profile", "tomo_phase_profile", "tomo_destroy", "tomo_last_error",
    "tomo_find_first_optimal", "TomoSh_HIP", "getMss", "getMo", "getVtc", "getVolMassInfo",
    "getMat4x4", "getnData2i", "getpData2i", "OnDestroy",
    "tomo_cvss_create", "tomo_cvss_forward", "tomo_cvss_backward", "tomo_cvss_destroy",
    "tomo_batch_create", "tomo_batch_add", "tomo_batch_tune", "tomo_batch_get_tuning", "tomo_batch_set_tuning",
    "tomo_batch_set_mode", "tomo_batch_get_plan", "tomo_batch_capture",
    "tomo_batch_launch", "tomo_batch_sync", "tomo_batch_last_ms", "tomo_batch_size", "tomo_batch_records",
    "tomo_batch_destroy",
)

_lib = None


class TomoError(RuntimeError):
    pass


def load():
    """Load libtomo_hip.so (build it first with `python -m tomo_shell2026_amd.build`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TomoError(f"{LIB_PATH} is not built; run python -m tomo_shell2026_amd.build")
    # PyTorch-ROCm ships its own HIP and HSA runtimes.  A process that uses both must load
    # torch's first: the library's libhsa-runtime64.so.1 then binds to the one already loaded,
    # while the other order leaves torch's HIP on a foreign HSA runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ct.CDLL(LIB_PATH)
    P, I32, F = ct.c_void_p, ct.c_int32, ct.c_float
    sig = {
        "tomo_create": ((P, P, I32, P), ct.c_int),
        "tomo_set_mesh": ((P, I32, P, I32, P, P), ct.c_int),
        "tomo_set_orientations": ((P, I32, P), ct.c_int),
        "tomo_run": ((P, I32, I32), ct.c_int),
        "tomo_run_static": ((P, I32, I32), ct.c_int),
        "tomo_sync": ((P,), ct.c_int),
        "tomo_get_records": ((P, I32, I32, P), ct.c_int),
        "tomo_sweep": ((P, I32, P, P), ct.c_int),
        "tomo_debug_slots": ((P, P, P, P, P, P, P), ct.c_int),
        "tomo_debug_rows": ((P, P, P, P, P), ct.c_int),
        "tomo_set_launch": ((P, I32), ct.c_int),
        "tomo_set_tuning": ((P, I32, I32), ct.c_int),
        "tomo_last_kernel_ms": ((P,), F),
        "tomo_get_info": ((P, P, P, P), ct.c_int),
        "tomo_set_phase_profile": ((P, I32), ct.c_int),
        "tomo_phase_profile": ((P, P), ct.c_int),
        "tomo_destroy": ((P,), ct.c_int),
        "tomo_last_error": ((), ct.c_char_p),
        "tomo_find_first_optimal": ((I32, P), I32),
        "TomoSh_HIP": ((P,) * 10, I32),
        "getMss": ((), ct.POINTER(ct.c_float)),
        "getMo": ((), ct.POINTER(ct.c_float)),
        "getVtc": ((), ct.POINTER(ct.c_float)),
        "getVolMassInfo": ((), ct.POINTER(ct.c_float)),
        "getMat4x4": ((), ct.POINTER(ct.c_float)),
        "getnData2i": ((ct.c_int16,), I32),
        "getpData2i": ((ct.c_int16,), ct.POINTER(ct.c_int16)),
        "OnDestroy": ((), None),
        "tomo_cvss_create": ((P, I32, P, I32, P, F, F, P), ct.c_int),
        "tomo_cvss_forward": ((P, I32, P, P, P, P), ct.c_int),
        "tomo_cvss_backward": ((P, I32, P, P, P, P, P), ct.c_int),
        "tomo_cvss_destroy": ((P,), ct.c_int),
        "tomo_batch_create": ((P,), ct.c_int),
        "tomo_batch_add": ((P, P, P, I32, I32, P, I32, P, P, I32, P, P), ct.c_int),
        "tomo_batch_tune": ((P, I32), ct.c_int),
        "tomo_batch_get_tuning": ((P, I32, P, P, P, P), ct.c_int),
        "tomo_batch_set_tuning": ((P, I32, I32, I32, I32), ct.c_int),
        "tomo_batch_set_mode": ((P, I32), ct.c_int),
        "tomo_batch_get_plan": ((P, P, P), ct.c_int),
        "tomo_batch_capture": ((P,), ct.c_int),
        "tomo_batch_launch": ((P,), ct.c_int),
        "tomo_batch_sync": ((P,), ct.c_int),
        "tomo_batch_last_ms": ((P,), F),
        "tomo_batch_size": ((P,), I32),
        "tomo_batch_records": ((P, I32, P), ct.c_int),
        "tomo_batch_destroy": ((P,), ct.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def last_error():
    return load().tomo_last_error().decode(errors="replace")


def check(rc, what):
    if rc != TOMO_OK:
        raise TomoError(f"{what} failed ({rc}): {last_error()}")


def ptr(a):
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


class Sweep:
    """One sweep context (a tomo_handle): options, a mesh, an orientation set."""

    def __init__(self, f32, i32, grid=0):
        L = load()
        self._f32 = np.ascontiguousarray(f32, np.float32)
        self._i32 = np.ascontiguousarray(i32, np.int32)
        if self._f32.size != 12 or self._i32.size != 11:
            raise ValueError("Float32Info must have 12 entries and Int32Info 11")
        h = ct.c_void_p()
        check(L.tomo_create(ptr(self._f32), ptr(self._i32), int(grid), ct.byref(h)), "tomo_create")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            load().tomo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_mesh(self, tri, vtx, vnrm):
        tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        vtx = np.ascontiguousarray(vtx, np.float32).reshape(-1, 3)
        vnrm = np.ascontiguousarray(vnrm, np.float32).reshape(-1, 3)
        if vtx.shape != vnrm.shape:
            raise ValueError("vertex and normal arrays differ in shape")
        check(load().tomo_set_mesh(self._h, tri.shape[0], ptr(tri), vtx.shape[0], ptr(vtx), ptr(vnrm)),
              "tomo_set_mesh")

    def set_orientations(self, ypr):
        ypr = np.ascontiguousarray(ypr, np.float32).reshape(-1, 3)
        check(load().tomo_set_orientations(self._h, ypr.shape[0], ptr(ypr)), "tomo_set_orientations")
        self.n = ypr.shape[0]

    def set_launch(self, n_groups):
        check(load().tomo_set_launch(self._h, int(n_groups)), "tomo_set_launch")

    TUNE_BATCH, TUNE_TILE_Y, TUNE_EXACT_FSUM, TUNE_INLINE = 0, 1, 2, 3

    def set_tuning(self, knob, value):
        check(load().tomo_set_tuning(self._h, int(knob), int(value)), "tomo_set_tuning")

    def run(self, begin=0, end=None):
        end = self.n if end is None else end
        check(load().tomo_run(self._h, int(begin), int(end)), "tomo_run")

    def run_static(self, begin=0, end=None):
        """tomo_run with the slot tables in the library's static device array (allocates nothing)."""
        end = self.n if end is None else end
        check(load().tomo_run_static(self._h, int(begin), int(end)), "tomo_run_static")

    def sync(self):
        check(load().tomo_sync(self._h), "tomo_sync")

    def records(self, begin=0, end=None):
        end = self.n if end is None else end
        out = np.zeros(end - begin, RECORD_DTYPE)
        check(load().tomo_get_records(self._h, int(begin), int(end), ptr(out)), "tomo_get_records")
        return out

    def sweep(self, ypr):
        self.set_orientations(ypr)
        self.run()
        return self.records()

    def kernel_ms(self):
        return float(load().tomo_last_kernel_ms(self._h))

    def info(self):
        g = ct.c_int32()
        n = ct.c_int32()
        b = ct.c_int64()
        check(load().tomo_get_info(self._h, ct.byref(g), ct.byref(n), ct.byref(b)), "tomo_get_info")
        return {"grid": g.value, "batch": n.value, "scratch_bytes": b.value}

    PHASES = ("aabb", "raster", "compact", "pair", "bed", "fsum")

    def set_phase_profile(self, enable=True):
        check(load().tomo_set_phase_profile(self._h, int(bool(enable))), "tomo_set_phase_profile")

    def phase_profile(self):
        out = np.zeros(8, np.float64)
        check(load().tomo_phase_profile(self._h, ptr(out)), "tomo_phase_profile")
        return {k: float(out[i]) for i, k in enumerate(self.PHASES)}

    def debug_rows(self, ypr3):
        y = np.ascontiguousarray(ypr3, np.float32).reshape(3)
        g2 = self.info()["grid"] ** 2
        rec = np.zeros(1, RECORD_DTYPE)
        rows = np.zeros(g2 * 64, np.uint64)
        nrows = np.zeros(g2, np.int32)
        check(load().tomo_debug_rows(self._h, ptr(y), ptr(rec), ptr(rows), ptr(nrows)), "tomo_debug_rows")
        return rec[0], rows, nrows

    def debug_slots(self, ypr3):
        y = np.ascontiguousarray(ypr3, np.float32).reshape(3)
        info = self.info()
        g2 = info["grid"] ** 2
        rec = np.zeros(1, RECORD_DTYPE)
        arrs = [np.zeros(g2, np.int32) for _ in range(4)]
        check(load().tomo_debug_slots(self._h, ptr(y), ptr(rec), *[ptr(a) for a in arrs]),
              "tomo_debug_slots")
        return (rec[0],) + tuple(arrs)


def decode_rows(rows, nrows):
    """Split tomo_debug_rows output into per-row (slot, type, z, nZ) arrays."""
    g2 = nrows.size
    r = rows.reshape(g2, 64)
    idx = np.arange(64)[None, :] < nrows[:, None]
    slot = np.nonzero(idx)[0]
    v = r[idx]
    typ = (v >> np.uint64(32)).astype(np.int64)
    key = (v & np.uint64(0xffffffff)).astype(np.int64)
    return slot, typ, (key >> 16) - 32768, (key & 0xffff) - 32768


def find_first_optimal(mss):
    m = np.ascontiguousarray(mss, np.float32)
    return int(load().tomo_find_first_optimal(m.size, ptr(m)))


class Batch:
    """Many meshes swept by one captured graph launch (tomo_batch_*)."""

    def __init__(self):
        h = ct.c_void_p()
        check(load().tomo_batch_create(ct.byref(h)), "tomo_batch_create")
        self._h = h
        self._n = []
        self._keys = []

    def add(self, f32, i32, grid, tri, vtx, vnrm, ypr):
        f = np.ascontiguousarray(f32, np.float32)
        i = np.ascontiguousarray(i32, np.int32)
        t = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        v = np.ascontiguousarray(vtx, np.float32).reshape(-1, 3)
        n = np.ascontiguousarray(vnrm, np.float32).reshape(-1, 3)
        y = np.ascontiguousarray(ypr, np.float32).reshape(-1, 3)
        m = ct.c_int32()
        check(load().tomo_batch_add(self._h, ptr(f), ptr(i), int(grid), t.shape[0], ptr(t), v.shape[0], ptr(v),
                                    ptr(n), y.shape[0], ptr(y), ct.byref(m)), "tomo_batch_add")
        self._n.append(y.shape[0])
        # the tuner's group (grid, ceil(log2 nTri)) and the options that change the work per slot
        lg = int(np.ceil(np.log2(max(t.shape[0], 1))))
        self._keys.append(f"grid{int(grid)}_lg{lg}_shell{int(i[2])}_ex{int(i[1])}_bed{int(i[9])}")
        return m.value

    def tune(self, sample=4096, cache=None):
        """Pick every member's launch shape (tomo_batch_tune).  With `cache` (a JSON file path),
        shapes found there for every member's group are applied without timing anything, and a
        fresh tuning is written back for the next process."""
        import json
        import os
        if cache and os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
            if all(k in known for k in self._keys):
                for m, k in enumerate(self._keys):
                    bs, ty, inl = known[k]
                    self.set_tuning(m, batch=min(int(bs), self._n[m]), tile_y=int(ty), inline=int(inl))
                return
        check(load().tomo_batch_tune(self._h, int(sample)), "tomo_batch_tune")
        if cache:
            known = {}
            if os.path.exists(cache):
                with open(cache) as fh:
                    known = json.load(fh)
            fresh = {}
            for m, k in enumerate(self._keys):  # a group shares one shape; batch clipped per member
                t = self.tuning(m)
                bs = max(t["batch"], fresh[k][0]) if k in fresh else t["batch"]
                fresh[k] = [bs, t["tile_y"], t["inline"]]
            known.update(fresh)
            tmp = cache + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(known, fh, indent=1, sort_keys=True)
            os.replace(tmp, cache)

    def tuning(self, member):
        b, t, i = ct.c_int32(), ct.c_int32(), ct.c_int32()
        us = ct.c_float()
        check(load().tomo_batch_get_tuning(self._h, int(member), ct.byref(b), ct.byref(t), ct.byref(i),
                                           ct.byref(us)), "tomo_batch_get_tuning")
        return {"batch": b.value, "tile_y": t.value, "inline": i.value,
                "us_per_orientation": round(us.value, 4)}

    def set_tuning(self, member, batch=0, tile_y=0, inline=16):
        check(load().tomo_batch_set_tuning(self._h, int(member), int(batch), int(tile_y), int(inline)),
              "tomo_batch_set_tuning")

    QUEUE, CHAIN = 0, 1  # TOMO_BATCH_QUEUE, TOMO_BATCH_CHAIN

    def set_mode(self, mode):
        """Batch.QUEUE: persistent work-queue launches per round of units (the default);
        Batch.CHAIN: each member's own chain of stage launches."""
        check(load().tomo_batch_set_mode(self._h, int(mode)), "tomo_batch_set_mode")

    def plan(self):
        """(rounds, units) of the last persistent capture."""
        r, u = ct.c_int32(), ct.c_int32()
        check(load().tomo_batch_get_plan(self._h, ct.byref(r), ct.byref(u)), "tomo_batch_get_plan")
        return r.value, u.value

    def capture(self):
        check(load().tomo_batch_capture(self._h), "tomo_batch_capture")

    def launch(self):
        check(load().tomo_batch_launch(self._h), "tomo_batch_launch")

    def sync(self):
        check(load().tomo_batch_sync(self._h), "tomo_batch_sync")

    def last_ms(self):
        return float(load().tomo_batch_last_ms(self._h))

    def records(self, member):
        out = np.zeros(self._n[member], RECORD_DTYPE)
        check(load().tomo_batch_records(self._h, int(member), ptr(out)), "tomo_batch_records")
        return out

    def __len__(self):
        return len(self._n)

    def close(self):
        if self._h:
            load().tomo_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
