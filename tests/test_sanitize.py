"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer: the oracle
(oracle/tomo_oracle.c), the host build of the kernels' pairing networks and rasteriser coverage
helpers (tests/host_slot_harness.cpp), driven by tests/sanitize_driver.cpp.  The sanitized oracle
must run every case without a finding and give the same records as the ordinary build."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests import _cases as C
from tomo_shell2026_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1",
       "-ffp-contract=off", "-fno-fast-math"]


def _cases():
    out = []
    v, f = synth.icosphere(3, 20.0)
    ypr = C.coarse_ypr(60)
    out.append(C.make(v, f, 64, ypr, bed=synth.BED_RAFT))
    out.append(C.make(v, f, 64, ypr, bed=synth.BED_SKIRT, explicit_ss=False))
    bv, bf = C.box((0, 0, 0), (20, 20, 20))
    out.append(C.make(bv, bf, 64, ypr, shell=True, shell_um=1500, bed=synth.BED_BRIM))
    tv, tf = C.tiny_triangles()
    out.append(C.make(tv, tf, 64, ypr))
    sv, sf = C.stacked_planes()
    out.append(C.make(sv, sf, 160, C.coarse_ypr(90), bed=synth.BED_RAFT))
    ov, of = C.box((0, 0, 0), (40, 30, 10))
    out.append(C.make(ov, of, 32, C.coarse_ypr(90)))  # footprint outside the grid
    name, pv, pf, shell = C.validation_primitives()[3]
    out.append(C.make(pv, pf, 64, C.coarse_ypr(90), shell=shell, theta_deg=45.0))
    return out


def _write_cases(path, cases, oracle_mod):
    with open(path, "wb") as fh:
        for c in cases:
            tri = np.ascontiguousarray(c["tri"], np.int32)
            vtx = np.ascontiguousarray(c["vtx"], np.float32)
            vn = np.ascontiguousarray(c["vnrm"], np.float32)
            ypr = np.ascontiguousarray(c["ypr"], np.float32)
            fh.write(np.array([len(tri), len(vtx), len(ypr)], np.int32).tobytes())
            o = oracle_mod.make_opts(c["f32"], c["i32"], c["grid"])
            fh.write(bytes(o))
            for a in (tri, vtx, vn, ypr):
                fh.write(a.tobytes())


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_oracle_and_host_harness_under_asan_ubsan(tmp_path, oracle_mod):
    d = str(tmp_path)
    subprocess.run(["gcc", "-std=gnu11", "-pthread", *SAN, "-c", os.path.join(ROOT, "oracle", "tomo_oracle.c"),
                    "-o", os.path.join(d, "oracle.o")], check=True)
    subprocess.run(["g++", "-std=c++17", *SAN, "-c", os.path.join(ROOT, "tests", "host_slot_harness.cpp"),
                    "-o", os.path.join(d, "harness.o")], check=True)
    exe = os.path.join(d, "sanitize_driver")
    subprocess.run(["g++", "-std=c++17", *SAN, os.path.join(ROOT, "tests", "sanitize_driver.cpp"),
                    os.path.join(d, "oracle.o"), os.path.join(d, "harness.o"), "-o", exe, "-pthread", "-lm",
                    "-static-libasan", "-static-libubsan"], check=True)
    cases = _cases()
    cpath, opath = os.path.join(d, "cases.bin"), os.path.join(d, "records.bin")
    _write_cases(cpath, cases, oracle_mod)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, cpath, opath], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert f"cases {len(cases)}" in r.stdout and "pairs ok" in r.stdout and "coverage ok" in r.stdout
    got = np.fromfile(opath, dtype=oracle_mod.RECORD_DTYPE)
    want = np.concatenate([oracle_mod.sweep(c["f32"], c["i32"], c["grid"], c["tri"], c["vtx"], c["vnrm"],
                                            c["ypr"], nthreads=2) for c in cases])
    assert got.tobytes() == want.tobytes()
