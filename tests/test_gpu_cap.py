"""GPU parity of the 63-row cap where it is hardest: the 63rd and 64th distinct rows of a slot
come from two pixels of one triangle (pairing re-walks that triangle to order them), slots of
more than 512 and 8,192 entries (the LDS-sorted and the pass-based selection of
pair_huge_kernel), an overflow pool that fills up (the orientation is swept again with larger
buffers), and float totals summed in slot order.  Every case prints its name first."""
import numpy as np
import pytest

from tests import _cases as C
from tomo_shell2026_amd import _lib, synth

pytestmark = pytest.mark.gpu


def announce(name):
    print(f"[gpu-case] {name}", flush=True)


def sweep_both(O, case, inline=0):
    with _lib.Sweep(case["f32"], case["i32"], case["grid"]) as s:
        s.set_mesh(case["tri"], case["vtx"], case["vnrm"])
        if inline:
            s.set_tuning(s.TUNE_INLINE, inline)
        g = s.sweep(case["ypr"])
    r = O.sweep(case["f32"], case["i32"], case["grid"], case["tri"], case["vtx"], case["vnrm"], case["ypr"],
                nthreads=8)
    return g, r


@pytest.mark.parametrize("extra", [0, 600, 8300])
@pytest.mark.parametrize("name", sorted(C.TIE_SLIVERS))
def test_cap_boundary_inside_one_triangle(oracle_mod, name, extra):
    """The fixture's slot keeps 62 rows of stacked squares and the sliver's first row in pixel
    order; extra squares after the sliver raise the slot to 83 .. 33,266 entries without new
    keys.  Rows of every slot after pairing, per-slot counts and the records must equal the
    oracle's."""
    case = C.tie_at_cap(name, extra)
    G = case["grid"]
    announce(f"cap boundary inside one triangle: {name}, {extra} extra squares, {len(case['tri'])} triangles")
    y = np.zeros(3, np.float32)
    with _lib.Sweep(case["f32"], case["i32"], G) as s:
        s.set_mesh(case["tri"], case["vtx"], case["vnrm"])
        rec, rows, nrows = s.debug_rows(y)
        drec, pre, fin, vo, vss = s.debug_slots(y)
    rrec, rfin, rrows = oracle_mod.one_rows(case["f32"], case["i32"], G, case["tri"], case["vtx"], case["vnrm"], y)
    _, rpre, rfin2, rvo, rvss = oracle_mod.one(case["f32"], case["i32"], G, case["tri"], case["vtx"], case["vnrm"], y)
    sx, sy = C.TIE_SLIVERS[name]["slot"]
    assert rfin[sx * G + sy] == 63
    np.testing.assert_array_equal(nrows, rfin)
    np.testing.assert_array_equal(rows, rrows)
    for a, b in ((pre, rpre), (fin, rfin2), (vo, rvo), (vss, rvss)):
        np.testing.assert_array_equal(a, b)
    bad = C.compare_records(np.array([rec]), np.array([rrec]))
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("name", ["solid_a", "shell"])
def test_cap_boundary_sweep(oracle_mod, name):
    """The same fixtures over a coarse orientation grid (the tie occurs at the identity; other
    orientations exercise the large-slot paths on tilted stacks)."""
    case = C.tie_at_cap(name, 600)
    case["ypr"] = C.coarse_ypr(60, 60)
    announce(f"cap fixture {name} over {len(case['ypr'])} orientations")
    g, r = sweep_both(oracle_mod, case)
    bad = C.compare_records(g, r)
    assert not bad, "\n".join(bad)


def test_pool_overflow_is_swept_again(oracle_mod):
    """200 stacked 60 x 60 squares on a 64 grid with 8 inline entries: about 690,000 entries
    overflow the 131,072-entry pool, so every orientation is swept again with larger pools.  The
    records must be exact (no UNRESOLVED flag) and equal the oracle's."""
    v, f = C.stacked_planes(n_layers=200, size=60.0, dz=1.3)
    case = C.make(v, f, 64, synth.ypr_grid([0, 180], [0, 180]), bed=synth.BED_RAFT)
    announce("pool overflow: 200 stacked squares, inline 8")
    g, r = sweep_both(oracle_mod, case, inline=8)
    assert (r["flags"] & 8).all()
    bad = C.compare_records(g, r)
    assert not bad, "\n".join(bad)


def test_float_totals_staircase(oracle_mod):
    """120 open boxes of heights 60 .. 112 on 4-voxel strips of a 512 grid: totals near 2^24.4,
    so fVo, fVa and the masses come from the ordered float replay; they must equal the oracle's
    bit for bit (the oracle's order is pinned by test_float_totals_run_in_ascending_slot_order)."""
    from tests.test_oracle_pins import build, sheet
    H = [60 + (7 * i) % 53 for i in range(120)]
    parts = []
    for i, h in enumerate(H):
        t = sheet(2 + 4 * i, 6 + 4 * i, 2, 482, float(h), up=True)
        b = sheet(2 + 4 * i, 6 + 4 * i, 2, 482, 0.0, up=False)
        parts += [(t[0], t[1], [0, 0, 1]), (b[0], b[1], [0, 0, -1])]
    v, f, n = build(parts)
    case = C.make(v, f, 512, synth.ypr_grid([0, 90, 180], [0, 30]))
    case["vnrm"] = n
    announce("float totals staircase on a 512 grid")
    g, r = sweep_both(oracle_mod, case)
    assert (r["flags"][0] & 1)
    bad = C.compare_records(g, r)
    assert not bad, "\n".join(bad)
