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
def test_batch_graph_matches_single_sweeps_and_oracle(oracle_mod):
    from tomo_shell2026_amd import _lib
    ws = _small_members()
    print("[gpu-case] batch graph, members:", [(w.name, w.grid, len(w.tri)) for w in ws], flush=True)
    assert len({w.grid for w in ws}) > 1  # layouts change between members
    singles = []
    for w in ws:
        with _lib.Sweep(w.f32, w.i32, w.grid) as s:
            s.set_mesh(w.tri, w.vtx, w.vnrm)
            singles.append(s.sweep(w.ypr))
    with _lib.Batch() as B:
        for w in ws:
            B.add(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr)
        B.launch()
        B.sync()
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
