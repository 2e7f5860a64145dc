"""CPU tests of product code that needs no GPU: the per-slot pairing logic of the kernels built
for the host and checked against the oracle, the C-ABI exports, and the host package."""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

from tomo_shell2026_amd import _lib, synth
from tomo_shell2026_amd.mesh import is_watertight, vertex_normals

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "_build")
THETA60 = np.float32(60.0 * np.pi / 180.0)


@pytest.fixture(scope="module")
def harness():
    os.makedirs(BUILD, exist_ok=True)
    so = os.path.join(BUILD, "libslot_harness.so")
    src = os.path.join(ROOT, "tests", "host_slot_harness.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", so, src],
                   check=True)
    L = ct.CDLL(so)
    for name in ("harness_pair2", "harness_pair4", "harness_pair8", "harness_pair16", "harness_pair63"):
        fn = getattr(L, name)
        fn.argtypes = (ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p)
        fn.restype = ct.c_int
    L.harness_make_key.argtypes = (ct.c_int, ct.c_int)
    L.harness_make_key.restype = ct.c_uint32
    L.harness_pair_two_mismatches.argtypes = (ct.c_int, ct.c_int)
    L.harness_pair_two_mismatches.restype = ct.c_int
    return L


def hpair(L, keys, n_max, cs, cn, explicit, bed_type=0, bed_r=2):
    keys = np.ascontiguousarray(keys, np.uint32)
    out = np.zeros(14 + n_max, np.int32)
    fn = {2: L.harness_pair2, 4: L.harness_pair4, 8: L.harness_pair8, 16: L.harness_pair16}.get(n_max, L.harness_pair63)
    rc = fn(keys.ctypes.data_as(ct.c_void_p), keys.size, cs, cn, int(explicit), bed_type, bed_r,
            out.ctypes.data_as(ct.c_void_p))
    assert rc == 0
    return out


def key(z, nz):
    return ((z + 32768) << 16) | (nz + 32768)


@pytest.mark.parametrize("n", [2, 4, 8, 16])
def test_sort_network_zero_one(harness, n):
    # 0-1 principle: the 1-, 5-, 19- and 80-comparator networks sort every binary input of length n
    lo, hi = key(3, 5), key(9, -5)
    for m in range(1 << n):
        ks = [hi if (m >> i) & 1 else lo for i in range(n)]
        out = hpair(harness, ks, n, -867, -866, True)
        srt = out[14:14 + n].astype(np.int64) & 0xffffffff
        assert list(srt) == sorted(srt, reverse=True)


@pytest.mark.parametrize("theta_deg", [30.0, 45.0, 60.0, 75.0, 89.0])
def test_two_entry_pairing_matches_network(harness, oracle_mod, theta_deg):
    # the kernels' straight-line pairing of slots with at most two entries against the 2-key
    # network, on every key pair of a pool straddling each decision, all counts, modes and beds
    cs, cn = oracle_mod.theta_cuts(np.float32(theta_deg * np.pi / 180.0))
    assert harness.harness_pair_two_mismatches(cs, cn) == 0


def test_known_answer_slot_product(harness):
    ks = [key(z, n) for z, n in zip([3, 1, 4, 2, 5, 9], [-999, -999, 999, 999, -999, 999])]
    for nmax in (8, 63):
        out = hpair(harness, ks, nmax, -867, -866, True)
        assert (out[0], out[1]) == (6, 3)
        out = hpair(harness, ks, nmax, -867, -866, False)
        assert (out[0], out[1]) == (6, 3)


@pytest.mark.parametrize("nmax", [2, 4, 8, 16, 63])
@pytest.mark.parametrize("explicit", [True, False])
def test_pairing_matches_oracle_random(harness, oracle_mod, nmax, explicit):
    rng = np.random.default_rng(11 + nmax + int(explicit))
    cs, cn = oracle_mod.theta_cuts(THETA60)
    nz_pool = np.array([-1000, -999, -900, -868, -867, -866, -865, -500, -1, 0, 1, 2, 500, 866, 999, 1000])
    for trial in range(1500):
        c = int(rng.integers(0, nmax + 1))
        z = rng.integers(0, 12, size=c)
        nz = rng.choice(nz_pool, size=c)
        ks = np.array([key(int(a), int(b)) for a, b in zip(z, nz)], np.int64)
        out = hpair(harness, ks, nmax, cs, cn, explicit)
        uk = np.unique(ks)
        dz = (uk >> 16) - 32768
        dnz = (uk & 0xffff) - 32768
        vo, vss, zo, nzo, ty = oracle_mod.pair_slot(dz, dnz, THETA60, explicit)
        ty = ty.astype(np.int64)
        zo = zo.astype(np.int64)
        assert out[6] == uk.size
        assert out[7] == len(zo), (trial, ks)
        assert (out[0], out[1]) == (vo, vss), (trial, ks)
        if uk.size > 1:
            assert out[2] == zo[(ty & 1) != 0].sum()
            assert out[3] == zo[(ty & 2) != 0].sum()
            assert out[4] == zo[(ty & 512) != 0].sum() - zo[(ty & 1024) != 0].sum()
            assert out[5] == zo[(ty & 256) != 0].sum()
            assert out[9] == zo[-1] and out[10] == ty[-1]
        assert out[8] == np.abs(zo).sum()


def test_exports_match_header():
    L = _lib.load()
    hdr = open(os.path.join(ROOT, "include", "tomo_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", " ", hdr, flags=re.S)
    hdr = re.sub(r"//[^\n]*", " ", hdr)
    hdr = re.sub(r"#[^\n]*", " ", hdr)
    declared = set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", hdr))
    assert set(_lib.EXPORTS) == declared
    for name in _lib.EXPORTS:
        assert hasattr(L, name), name


def test_create_reports_errors_without_exiting():
    L = _lib.load()
    h = ct.c_void_p()
    assert L.tomo_create(None, None, 0, ct.byref(h)) == -1
    assert b"null" in L.tomo_last_error()
    f32 = synth.default_f32(1.0)
    i32 = synth.default_i32(1, 3, 1, bed_type=7)
    assert L.tomo_create(_lib.ptr(f32), _lib.ptr(i32), 64, ct.byref(h)) == -1


def test_mesh_front_end():
    v, f = synth.icosphere(2, 10.0)
    assert len(f) == 320 and is_watertight(f)
    n = vertex_normals(v, f)
    radial = v / np.linalg.norm(v, axis=1, keepdims=True)
    assert np.all(np.einsum("ij,ij->i", n, radial) > 0.99)
    w = synth.displaced_sphere_workload()
    assert w.tri.shape == (69120, 3) and w.vtx.shape == (34562, 3)
    assert is_watertight(w.tri) and len(w.ypr) == 21960


def _bary_cases(rng, n_tri):
    """Random margined triangles: ordinary, slivers, tiny, huge, near-degenerate, vertices on
    pixel centres, far from the origin; each with its pixel box grown by one pixel per side."""
    out = []
    for k in range(n_tri):
        kind = k % 7
        c = rng.uniform(5, 200, 2)
        if kind == 5:  # vertices on pixel centres: edges through centres, u, v, w exactly 0 or 1
            t = np.floor(c) + 0.5 + rng.integers(-6, 7, (3, 2))
        elif kind == 6:  # far from the origin (coarse float spacing), ordinary size
            t = c + 1900.0 + rng.normal(0, 2.0, (3, 2))
        elif kind == 0:
            t = c + rng.normal(0, 1.2, (3, 2))
        elif kind == 1:  # sliver
            d = rng.normal(0, 1, 2)
            d /= np.linalg.norm(d)
            t = np.stack([c, c + d * rng.uniform(2, 40), c + d * rng.uniform(2, 40) + rng.normal(0, 0.02, 2)])
        elif kind == 2:  # tiny
            t = c + rng.normal(0, 0.05, (3, 2))
        elif kind == 3:  # huge
            t = c + rng.normal(0, 30, (3, 2))
        else:  # near-degenerate area ~ threshold
            t = np.stack([c, c + [1.0, 0.0], c + [rng.uniform(-1, 2), rng.uniform(0.01, 0.05)]])
        t = (t.astype(np.float32) + np.float32(0.001)).astype(np.float32)
        x0, y0 = np.floor(t.min(0)).astype(int) - 1
        x1, y1 = np.floor(t.max(0)).astype(int) + 1
        out.append((t, int(x0), int(x1), int(y0), int(y1)))
    return out


def test_raster_coverage_matches_oracle(harness, oracle_mod):
    """The rasteriser's exact coverage test equals _getBaryCoord bit for bit, and the affine
    prefilter never rejects a pixel the exact test accepts, over whole pixel boxes."""
    harness.harness_bary.argtypes = (ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float, ct.c_void_p)
    harness.harness_bary.restype = ct.c_int
    rng = np.random.default_rng(5)
    n_acc = n_cand = n_pix = 0
    h2 = np.float32(0.5)
    for k, (t, x0, x1, y0, y1) in enumerate(_bary_cases(rng, 700)):
        tri6 = np.ascontiguousarray(t.ravel(), np.float32)
        npx = (x1 - x0 + 1) * (y1 - y0 + 1)
        out = np.zeros((npx, 5), np.uint32)
        assert harness.harness_bary(tri6.ctypes.data_as(ct.c_void_p), x0, x1, y0, y1, h2,
                                    out.ctypes.data_as(ct.c_void_p)) == npx
        xs = np.arange(x0, x1 + 1, dtype=np.float32) + np.float32(0.001) + h2
        ys = np.arange(y0, y1 + 1, dtype=np.float32) + np.float32(0.001) + h2
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        pix = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
        step = max(1, npx // 300)
        for i in range(0, npx, step):  # oracle check on a subsample (python loop)
            ok, uvw = oracle_mod.bary_point(pix[i], t[0], t[1], t[2])
            assert bool(out[i, 1]) == ok
            if ok:
                assert tuple(out[i, 2:5]) == tuple(uvw.view(np.uint32)), (t, pix[i])
        assert np.all(out[out[:, 1] == 1, 0] == 1), "prefilter rejected an accepted pixel"
        if k % 7 != 1:  # 0.02-pixel-wide slivers: float noise in the numerators exceeds their width
            n_acc += int(out[:, 1].sum())
            n_cand += int(out[:, 0].sum())
            n_pix += npx
    assert n_acc > 1000
    # the prefilter passes few pixels the exact test rejects
    assert n_cand < 1.05 * n_acc + 0.01 * n_pix


def test_accept_box_holds_every_accepted_pixel(harness):
    """The rasteriser walks only the part of a triangle's pixel box that accept_range / clip_span
    leave: every pixel the exact test accepts must lie inside it (ordinary, sliver, tiny, huge,
    near-degenerate triangles, vertices on pixel centres, far from the origin), and for ordinary
    small triangles the walk must actually shrink."""
    harness.harness_bary.argtypes = (ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float, ct.c_void_p)
    harness.harness_bary.restype = ct.c_int
    harness.harness_accept_box.argtypes = (ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_float,
                                           ct.c_void_p)
    harness.harness_accept_box.restype = ct.c_int
    rng = np.random.default_rng(11)
    full = walked = 0
    for h2 in (np.float32(0.5), np.float32(0.25), np.float32(0.0)):
        for k, (t, x0, x1, y0, y1) in enumerate(_bary_cases(rng, 4200)):
            tri6 = np.ascontiguousarray(t.ravel(), np.float32)
            npx = (x1 - x0 + 1) * (y1 - y0 + 1)
            out = np.zeros((npx, 5), np.uint32)
            harness.harness_bary(tri6.ctypes.data_as(ct.c_void_p), x0, x1, y0, y1, h2, out.ctypes.data_as(ct.c_void_p))
            box = np.zeros(5, np.int32)
            harness.harness_accept_box(tri6.ctypes.data_as(ct.c_void_p), x0, x1, y0, y1, h2,
                                       box.ctypes.data_as(ct.c_void_p))
            acc = out[:, 1].reshape(x1 - x0 + 1, y1 - y0 + 1).astype(bool)
            xs, ys = np.nonzero(acc)
            if xs.size:
                assert box[0] <= x0 + xs.min() and x0 + xs.max() <= box[1], (k, t, box)
                assert box[2] <= y0 + ys.min() and y0 + ys.max() <= box[3], (k, t, box)
            assert x0 <= box[0] and box[1] <= x1 and y0 <= box[2] and box[3] <= y1
            if k % 7 in (0, 2):  # ordinary and tiny triangles
                full += npx
                walked += max(0, box[1] - box[0] + 1) * max(0, box[3] - box[2] + 1)
    assert walked < 0.5 * full


def test_manikin_generator():
    """Thin-shell manikin: ~87k triangles, two closed sheets before the opening cut, an open
    shell after it, deterministic per seed, and a footprint inside the 512 grid."""
    from tomo_shell2026_amd import synth
    from tomo_shell2026_amd.mesh import is_watertight
    v, f = synth.manikin_shell(seed=7)
    assert abs(len(f) - 87000) / 87000 < 0.03
    v2, f2 = synth.manikin_shell(seed=7)
    np.testing.assert_array_equal(f, f2)
    np.testing.assert_array_equal(v, v2)
    assert not is_watertight(f)
    vc, fc = synth.manikin_shell(seed=7, spacing=8.0, open_mm=-1e9)
    e = np.sort(np.concatenate([fc[:, [0, 1]], fc[:, [1, 2]], fc[:, [2, 0]]]), 1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()
    c = 0.5 * (v.min(0) + v.max(0))
    assert 2 * np.linalg.norm(v - c, axis=1).max() <= 512 - 4
    # both sheets present: outward normals of the outer sheet, inward of the inner one
    from tomo_shell2026_amd.mesh import face_normals
    n = face_normals(v, f)
    cen = v[f].mean(1)
    radial = cen - np.array([0.0, 0.0, cen[:, 2].mean()])
    out = np.einsum("ij,ij->i", n, radial) > 0
    assert 0.35 < out.mean() < 0.65


def test_marching_tetrahedra_sphere():
    """A sphere SDF gives a closed mesh whose enclosed volume and area approach 4/3 pi r^3 and
    4 pi r^2 as the grid refines, with outward faces."""
    from tomo_shell2026_amd import synth
    h = 1.0
    g = np.arange(-14, 15) * h
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    F = np.linalg.norm(P, axis=-1) - 10.0
    v, f = synth.marching_tetrahedra(F, np.array([-14.0, -14.0, -14.0]), h)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    assert abs(vol / (4 / 3 * np.pi * 1000) - 1) < 0.02
    assert abs(area / (4 * np.pi * 100) - 1) < 0.03
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()


def test_mesh_front_end_closed_forms():
    """Angle-weighted vertex normals, area and watertightness (the quantities run_dll reads,
    tomo_dll_interface.py:59-70) against closed forms.  The box's faces are split into two
    triangles with diagonals that meet some corners twice and others once: at every corner the
    three faces contribute 90 degrees each, so the angle-weighted normal is the corner's diagonal
    (+-1, +-1, +-1)/sqrt(3); weighting by area (or by triangle count) would not give it at the
    corners whose faces are split unevenly, which the test also checks."""
    from tests import _cases as C
    from tomo_shell2026_amd.mesh import face_normals, mesh_area
    v, f = C.box((0.0, 0.0, 0.0), (2.0, 3.0, 4.0))
    assert mesh_area(v, f) == pytest.approx(2 * (2 * 3 + 2 * 4 + 3 * 4), abs=1e-12)
    n = vertex_normals(v, f)
    ctr = v.mean(0)
    diag = np.sign(v - ctr) / np.sqrt(3.0)
    np.testing.assert_allclose(n, diag, atol=1e-12)
    fn = face_normals(v, f)
    area = 0.5 * np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1)
    by_area = np.zeros_like(v)
    for k in range(3):
        np.add.at(by_area, f[:, k], fn * area[:, None])
    by_area /= np.linalg.norm(by_area, axis=1, keepdims=True)
    assert np.abs(by_area - diag).max() > 0.05  # the fixture separates the weightings
    assert is_watertight(f)
    flipped = f.copy()
    flipped[3] = flipped[3, ::-1]
    assert not is_watertight(flipped)
    assert not is_watertight(f[1:])
