"""The overlapped batch schedule against the one-stream schedule, through the C ABI.

With its side stream, a handle runs each batch's bed and ordered-sum kernels beside the next
batch's rasteriser, and each next batch's placement beside the current batch's pairing, with
two alternating header blocks.  Stage timing keeps every stage on the main stream, one batch
after another.  Small batches give many of them per sweep; both schedules must give the same
records bit for bit, and the oracle must agree on a sample.  The icosphere is swept with a raft
(bed kernel) and with a brim and implicit support structures, and a box in shell mode."""
import os

import numpy as np
import pytest

from tests import _cases as C
from tomo_shell2026_amd import _lib, synth

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


def _records(case, batch, one_stream):
    with _lib.Sweep(case["f32"], case["i32"], case["grid"]) as s:
        s.set_mesh(case["tri"], case["vtx"], case["vnrm"])
        s.set_orientations(case["ypr"])
        s.set_tuning(s.TUNE_BATCH, batch)
        s.set_phase_profile(one_stream)
        s.run()
        s.run()  # a second run over the same handle (header blocks reused)
        s.sync()
        return s.records()


def _case(w):
    return dict(f32=w.f32, i32=w.i32, grid=w.grid, tri=w.tri, vtx=w.vtx, vnrm=w.vnrm, ypr=w.ypr)


def _check(name, case, batch, oracle_mod, every):
    print(f"[gpu-case] schedule: {name}, {len(case['ypr'])} orientations in batches of {batch}", flush=True)
    over = _records(case, batch, one_stream=False)
    seq = _records(case, batch, one_stream=True)
    bad = C.compare_records(over, seq, f"{name} overlapped vs one stream: ")
    assert not bad, "\n".join(bad[:20])
    sub = dict(case, ypr=case["ypr"][::every])
    ref = oracle_mod.sweep(sub["f32"], sub["i32"], sub["grid"], sub["tri"], sub["vtx"], sub["vnrm"],
                           sub["ypr"], nthreads=NTHREADS)
    bad = C.compare_records(over[::every], ref, f"{name} vs oracle: ")
    assert not bad, "\n".join(bad[:20])


def test_icosphere_raft_many_batches(oracle_mod):
    w = synth.icosphere_workload(bed_type=synth.BED_RAFT)
    w.ypr = w.ypr[::5]
    case = _case(w)
    _check("icosphere raft", case, 16, oracle_mod, 3)


def test_icosphere_implicit_brim_many_batches(oracle_mod):
    w = synth.icosphere_workload(bed_type=synth.BED_BRIM)
    w.explicit_ss = False
    w.ypr = w.ypr[::7]
    case = _case(w)
    _check("icosphere implicit brim", case, 8, oracle_mod, 3)


def test_box_shell_many_batches(oracle_mod):
    v, f = C.box((0, 0, 0), (60, 40, 30))
    case = C.make(v, f, 128, synth.ypr_grid(np.arange(0, 360, 40), np.arange(0, 61, 30)),
                  bed=synth.BED_RAFT, shell=True)
    _check("box shell", case, 4, oracle_mod, 3)
