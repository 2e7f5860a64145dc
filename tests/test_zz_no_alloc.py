"""tomo_run_static, the entry point that allocates nothing, against tomo_run through the C ABI.

Two calls in a row on the library's one static array: the 69,120-triangle sphere on a 256 grid
(many batches of about 20 orientations), then a smaller input, the 20,480-triangle icosphere on
a 128 grid with implicit support and a brim bed, which reuses the array over the first call's
contents.  Each must equal tomo_run's records bit for bit.  A grid whose single slot table
exceeds the array is refused with TOMO_E_INVALID_ARG."""
import ctypes as ct

import numpy as np
import pytest

from tests import _cases as C
from tomo_shell2026_amd import _lib, synth

pytestmark = pytest.mark.gpu


def _both(w, ypr):
    """Records of tomo_run and of tomo_run_static for the same handle and orientations."""
    L = _lib.load()
    with _lib.Sweep(w.f32, w.i32, w.grid) as s:
        s.set_mesh(w.tri, w.vtx, w.vnrm)
        s.set_orientations(ypr)
        s.run()
        ref = s.records()
        rc = L.tomo_run_static(s._h, 0, len(ypr))
        assert rc == 0, _lib.last_error()
        got = s.records()
        batch = s.info()["batch"]
    return got, ref, batch


def test_static_entry_matches_tomo_run():
    w = synth.displaced_sphere_workload()
    ypr = w.ypr[::73]
    print(f"[gpu-case] tomo_run_static call 1: {w.name}, grid {w.grid}, {len(ypr)} orientations", flush=True)
    got, ref, batch = _both(w, ypr)
    assert 1 <= batch < len(ypr)  # several batches through the array
    bad = C.compare_records(got, ref, "static vs tomo_run (1): ")
    assert not bad, "\n".join(bad)
    w2 = synth.icosphere_workload(bed_type=synth.BED_BRIM)
    w2.explicit_ss = False
    ypr2 = w2.ypr[::7]
    assert w2.grid < w.grid and len(ypr2) < len(ypr)
    print(f"[gpu-case] tomo_run_static call 2: {w2.name}, grid {w2.grid}, {len(ypr2)} orientations", flush=True)
    got2, ref2, _ = _both(w2, ypr2)
    bad = C.compare_records(got2, ref2, "static vs tomo_run (2): ")
    assert not bad, "\n".join(bad)
    assert (got2["Vtc"] > 0).all()  # implicit support really ran


def test_static_entry_refuses_oversized_tables():
    L = _lib.load()
    v, f = C.box((0, 0, 0), (10, 10, 10))
    case = C.make(v, f, 2048, [[0.0, 0.0, 0.0]])
    print("[gpu-case] tomo_run_static on a 2048 grid (refused)", flush=True)
    with _lib.Sweep(case["f32"], case["i32"], 2048) as s:
        s.set_mesh(case["tri"], case["vtx"], case["vnrm"])
        s.set_orientations(case["ypr"])
        assert L.tomo_run_static(s._h, 0, 1) == -1
        assert b"exceeds the static array" in L.tomo_last_error()
