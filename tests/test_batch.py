"""Batched sweeps (many meshes, one captured graph launch): the generator of the 70-mesh
workload (CPU) and, on the GPU, bit-exact agreement of the graph launch with per-mesh sweeps and
with the oracle, including members of different grids and inline capacities in sequence."""
import numpy as np
import pytest

from tests import _cases as C
from tomo_shell2026_amd import synth


def test_batch70_generator():
    ws = synth.batch70_workloads()
    assert len(ws) == 70
    parts = [w for w in ws if w.name.startswith("manikin_part")]
    spheres = [w for w in ws if w.name.startswith("dsphere")]
    assert len(parts) == 6 and len(spheres) == 64
    assert all(2000 <= len(w.tri) <= 30000 for w in parts) and all(w.shell for w in parts)
    n = np.array([len(w.tri) for w in spheres])
    assert n.min() >= 900 and n.max() <= 110000 and not any(w.shell for w in spheres)
    assert all(0.5 <= w.voxel_mm <= 2.0 for w in ws)
    assert all(len(w.ypr) == 21960 for w in ws)
    # every member fits its grid in every orientation (footprint within grid - 2 Bo)
    for w in ws:
        c = 0.5 * (w.vtx.min(0) + w.vtx.max(0))
        assert 2 * np.linalg.norm(w.vtx - c, axis=1).max() <= w.grid - 4 + 1e-3


def _small_members():
    ypr = synth.ypr_grid(np.arange(0, 360, 15), np.arange(0, 61, 15))  # 24 x 5 = 120
    v, f = synth.icosphere(3, 30.0)
    a = synth.member_workload("ico1280", v, f, shell=False, ypr=ypr)
    sv, sf = synth.displaced_sphere_n(6000, 55.0, seed=101)
    b = synth.member_workload("dsphere6000", sv, sf, shell=False, ypr=ypr)
    mv, mf = synth.manikin_shell(seed=7)
    part = synth.cut_by_planes(mv, mf, 6, 2000, 30000, seed=5)[-1]
    c = synth.member_workload("manikin_part", part[0], part[1], shell=True, ypr=ypr)
    sv2, sf2 = synth.displaced_sphere_n(1500, 22.0, seed=102)
    d = synth.member_workload("dsphere1500", sv2, sf2, shell=False, ypr=ypr)
    return [a, b, c, d]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["queue", "chain"])
def test_batch_graph_matches_single_sweeps_and_oracle(oracle_mod, mode):
    """Both capture forms: persistent work-queue launches over rounds of units, and each member's
    own chain of launches."""
    from tomo_shell2026_amd import _lib
    ws = _small_members()
    print(f"[gpu-case] batch graph ({mode}), members:", [(w.name, w.grid, len(w.tri)) for w in ws], flush=True)
    assert len({w.grid for w in ws}) > 1  # layouts change between members
    singles = []
    for w in ws:
        with _lib.Sweep(w.f32, w.i32, w.grid) as s:
            s.set_mesh(w.tri, w.vtx, w.vnrm)
            singles.append(s.sweep(w.ypr))
    with _lib.Batch() as B:
        for w in ws:
            B.add(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr)
        B.set_mode(B.QUEUE if mode == "queue" else B.CHAIN)
        B.launch()
        B.sync()
        rounds, units = B.plan()
        if mode == "queue":
            assert 1 <= rounds <= units and units >= len(ws)
        else:
            assert (rounds, units) == (0, 0)
        first = [B.records(m) for m in range(len(ws))]
        B.launch()  # a second launch of the same graph gives the same results
        B.sync()
        second = [B.records(m) for m in range(len(ws))]
        B.tune(sample=120)
        tuned = [B.tuning(m) for m in range(len(ws))]
        print("[gpu-case] tuned shapes:", tuned, flush=True)
        B.launch()
        B.sync()
        third = [B.records(m) for m in range(len(ws))]
    for m, w in enumerate(ws):
        for got in (first[m], second[m], third[m]):
            bad = C.compare_records(got, singles[m], f"{w.name} graph vs single: ")
            assert not bad, "\n".join(bad)
        ref = oracle_mod.sweep(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr, nthreads=8)
        bad = C.compare_records(first[m], ref, f"{w.name} graph vs oracle: ")
        assert not bad, "\n".join(bad)


@pytest.mark.gpu
def test_tuned_shapes_cache(tmp_path):
    """Batch.tune(cache=...) writes the tuned shapes per group; a later batch of the same groups
    applies them without timing anything and sweeps to the same records."""
    from tomo_shell2026_amd import _lib
    ws = _small_members()[:2]
    path = str(tmp_path / "shapes.json")
    print("[gpu-case] tuned-shape cache", flush=True)
    with _lib.Batch() as B:
        for w in ws:
            B.add(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr)
        B.tune(sample=256, cache=path)
        first = [B.tuning(m) for m in range(len(ws))]
        B.launch()
        ref = [B.records(m) for m in range(len(ws))]
    assert all(t["us_per_orientation"] > 0 for t in first)
    with _lib.Batch() as B:
        for w in ws:
            B.add(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr)
        B.tune(sample=256, cache=path)
        again = [B.tuning(m) for m in range(len(ws))]
        B.launch()
        got = [B.records(m) for m in range(len(ws))]
    for a, b in zip(first, again):
        assert (a["batch"], a["tile_y"], a["inline"]) == (b["batch"], b["tile_y"], b["inline"])
        assert b["us_per_orientation"] < 0  # applied from the cache, not timed
    for m, w in enumerate(ws):
        bad = C.compare_records(got[m], ref[m], f"{w.name} cached shapes: ")
        assert not bad, "\n".join(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["queue", "chain"])
def test_members_reuse_larger_layouts_in_any_order(oracle_mod, mode):
    """Members share one scratch block and each lays its slot table out over the previous
    member's data, so every table a sweep reads must be written by that sweep first.  A 512-grid
    member with 24 inline entries and 256-tall tiles runs first, then a 256-grid member with 8
    inline entries, a 128-grid member with 16, and the 512-grid one again with 8: graph results
    must equal single sweeps with default shapes, and the oracle.  In the persistent form the
    units of a round lie side by side, and every round's tables lie over the previous round's."""
    from tomo_shell2026_amd import _lib
    ypr = synth.ypr_grid(np.arange(0, 360, 24), np.arange(0, 61, 20))  # 15 x 4
    big_v, big_f = synth.displaced_sphere_n(3000, 260.0, seed=11)
    a = synth.member_workload("dsphere_big", big_v, big_f, shell=False, ypr=ypr)
    bv, bf = synth.displaced_sphere_n(6000, 60.0, seed=12)
    b = synth.member_workload("dsphere_mid", bv, bf, shell=False, ypr=ypr)
    cv, cf = synth.icosphere(3, 25.0)
    c = synth.member_workload("ico", cv, cf, shell=False, ypr=ypr)
    ws = [a, b, c, a]
    shapes = [(len(ypr), 256, 24), (40, 32, 8), (7, 0, 16), (len(ypr), 64, 8)]
    assert [w.grid for w in ws] == [512, 256, 128, 512]
    print(f"[gpu-case] batch member order ({mode}): grids 512/256/128/512, inline 24/8/16/8", flush=True)
    singles = []
    for w in ws[:3]:
        with _lib.Sweep(w.f32, w.i32, w.grid) as s:
            s.set_mesh(w.tri, w.vtx, w.vnrm)
            singles.append(s.sweep(w.ypr))
    singles.append(singles[0])
    with _lib.Batch() as B:
        for w in ws:
            B.add(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr)
        B.set_mode(B.QUEUE if mode == "queue" else B.CHAIN)
        for m, (bs, ty, inl) in enumerate(shapes):
            B.set_tuning(m, batch=bs, tile_y=ty, inline=inl)
            got = B.tuning(m)
            assert (got["batch"], got["tile_y"], got["inline"]) == (bs, ty, inl)
        for _ in range(2):
            B.launch()
            B.sync()
            for m, w in enumerate(ws):
                bad = C.compare_records(B.records(m), singles[m], f"{w.name}#{m} graph vs single: ")
                assert not bad, "\n".join(bad)
    for m, w in enumerate(ws[:3]):
        ref = oracle_mod.sweep(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr, nthreads=8)
        bad = C.compare_records(singles[m], ref, f"{w.name} single vs oracle: ")
        assert not bad, "\n".join(bad)


@pytest.mark.gpu
def test_batch70_workload(oracle_mod):
    """The 70-mesh workload at its full size: 6 manikin parts and 64 displaced spheres, each on
    its own 1 degree yaw x 0..60 pitch grid (1,537,200 orientation evaluations), tuned, captured
    into one graph and launched.  Every member's 21,960 records must equal its own per-mesh
    sweep bit for bit, and every 499th orientation of every member must equal the oracle's.
    This is the per-part flow of cfms_tomo/__init__.py:26-30 done in one launch."""
    import time
    from tomo_shell2026_amd import _lib
    t0 = time.time()
    ws = synth.batch70_workloads()
    print(f"[gpu-case] batch70: {len(ws)} meshes generated in {time.time() - t0:.1f} s", flush=True)
    with _lib.Batch() as B:
        for w in ws:
            B.add(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr)
        B.tune(4096)
        print(f"[gpu-case] batch70: tuned at {time.time() - t0:.1f} s", flush=True)
        B.launch()
        B.sync()
        graph = [B.records(m) for m in range(len(ws))]
        rounds, units = B.plan()
    print(f"[gpu-case] batch70: graph launched at {time.time() - t0:.1f} s ({rounds} rounds of {units} units)",
          flush=True)
    assert 1 <= rounds < units
    for m, w in enumerate(ws):
        with _lib.Sweep(w.f32, w.i32, w.grid) as s:
            s.set_mesh(w.tri, w.vtx, w.vnrm)
            single = s.sweep(w.ypr)
        bad = C.compare_records(graph[m], single, f"{w.name} graph vs single: ")
        assert not bad, "\n".join(bad)
        assert not (graph[m]["flags"] & 64).any()
        idx = np.arange(0, len(w.ypr), 499)
        ref = oracle_mod.sweep(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr[idx], nthreads=8)
        bad = C.compare_records(graph[m][idx], ref, f"{w.name} graph vs oracle: ")
        assert not bad, "\n".join(bad)
    print(f"[gpu-case] batch70: compared at {time.time() - t0:.1f} s", flush=True)
