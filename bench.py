"""Benchmark: orientations/s of the support-tomography sweep on one MI355X.

Workload: the 69,120-triangle displaced sphere (0.5 mm voxels, 256 x 256 slot grid) over the
yaw 0..359 x pitch 0..60 grid at 1 degree (21,960 orientations), explicit support, raft bed.
One step = one full sweep: kernel, record readback and host finishing (float totals, masses).
Inputs (mesh, rotation matrices) are resident in HBM before timing starts.

Prints one JSON line.  --gpus above 1 is not supported (the sweep runs on one GPU).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "orientations/s at 69k tris on a 1° grid; batched+refined vs plain at equal Vss"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n_vtx, n_tri, rec):
    """Bytes one orientation must move at minimum over the whole sweep: SoA fp32 position +
    normal per vertex, int32 indices, a write + read of an 8-byte entry per rasterised row, a
    write + read of a 4-byte counter per touched slot, and the 64-byte result record."""
    p_bar = float(rec["n_entries"].mean())
    s_bar = float(rec["n_slots"].mean())
    return 24.0 * n_vtx + 12.0 * n_tri + 16.0 * p_bar + 8.0 * s_bar + 64.0, p_bar, s_bar


def raster_bytes(n_vtx, n_tri, p_bar, s_bar):
    """Bytes the rasteriser kernel must move per orientation: int32 indices, the raw fp32
    position and normal of every vertex (it rotates them itself), an 8-byte slot entry written
    per rasterised row and a 4-byte slot counter read and written per touched slot."""
    return 12.0 * n_tri + 24.0 * n_vtx + 8.0 * p_bar + 8.0 * s_bar


def copy_calibration(gib=1):
    """A plain device-to-device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = gib << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 4, "launches": 3}), flush=True)


def summarize_pmc(fetch_dir, write_dir, cfetch_dir, cwrite_dir, kernel="raster_kernel"):
    """Per-launch FETCH_SIZE / WRITE_SIZE of `kernel` from two rocprofv3 --pmc runs, scaled by the
    same counters on the copy calibration (bytes read / written by a copy of known size)."""
    import csv
    import glob

    def per_launch(d, name, counter):
        vals = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if name in row["Kernel_Name"] and row["Counter_Name"] == counter:
                        vals.append(float(row["Counter_Value"]))
        return sum(vals) / len(vals), len(vals)

    fk, n1 = per_launch(fetch_dir, kernel, "FETCH_SIZE")
    wk, n2 = per_launch(write_dir, kernel, "WRITE_SIZE")
    cf, _ = per_launch(cfetch_dir, "copy", "FETCH_SIZE")
    cw, _ = per_launch(cwrite_dir, "copy", "WRITE_SIZE")
    copy_bytes = float(1 << 30)
    f_scale = copy_bytes / (cf * 1024.0)
    w_scale = copy_bytes / (cw * 1024.0)
    return {"kernel": kernel, "launches": [n1, n2], "fetch_kib_per_launch": fk, "write_kib_per_launch": wk,
            "copy_fetch_kib": cf, "copy_write_kib": cw, "fetch_scale": f_scale, "write_scale": w_scale,
            "bytes_per_launch": fk * 1024.0 * f_scale + wk * 1024.0 * w_scale}


def summarize_counters(lds_dir, wr_dir):
    """Per kernel, averaged over its launches: LDS bank-conflict cycles over conflict-free LDS
    cycles (SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT), the rate rocprof's
    LdsBankConflict metric reports), LDS and VALU instructions per wave, and the bytes the L2
    wrote to memory from its 64-byte and 32-byte write requests (TCC_EA0_WRREQ counts both,
    TCC_EA0_WRREQ_64B the 64-byte ones)."""
    import collections
    import csv
    import glob

    def load(d):
        out = collections.defaultdict(lambda: collections.defaultdict(list))
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    k = row["Kernel_Name"].split("(")[0].replace("tomo::", "")
                    out[k][row["Counter_Name"]].append(float(row["Counter_Value"]))
        return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in out.items()}

    lds, wr = load(lds_dir), load(wr_dir)
    summ = {}
    for k in sorted(set(lds) | set(wr)):
        if k.startswith("__amd"):
            continue
        a, b = lds.get(k, {}), wr.get(k, {})
        e = {}
        if a:
            bc, act, waves = a["SQ_LDS_BANK_CONFLICT"], a["SQ_LDS_IDX_ACTIVE"], max(a["SQ_WAVES"], 1.0)
            e.update({"lds_bank_conflict_cycles": bc, "lds_active_cycles": act,
                      "lds_bank_conflict_rate": round(bc / (act - bc), 4) if act > bc else None,
                      "lds_addr_conflict_cycles": a["SQ_LDS_ADDR_CONFLICT"],
                      "lds_insts_per_wave": round(a["SQ_INSTS_LDS"] / waves, 1),
                      "valu_insts_per_wave": round(a["SQ_INSTS_VALU"] / waves, 1), "waves": waves})
        if b:
            req, req64 = b["TCC_EA0_WRREQ_sum"], b["TCC_EA0_WRREQ_64B_sum"]
            e.update({"write_requests": req, "write_requests_64B": req64,
                      "write_bytes_per_launch": 64.0 * req64 + 32.0 * (req - req64)})
        summ[k] = e
    return summ


_T0 = time.perf_counter()


def progress(msg):
    """One progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    """Host CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, os.cpu_count() or 1, avail


def cpu_baseline(w, records, stride, runs=3):
    """The C oracle (orientation-parallel pthreads, as Tomo_Shell2026.cpp:155-214 fans out) on a
    bounded sample of the same sweep: the median of `runs` runs on all available cores (at most
    16, this GPU's share of the host), and one run on one thread over a 20x sparser sample."""
    from oracle import oracle as O
    O.build()
    model, nproc, avail = cpu_info()
    threads = max(1, min(16, avail))
    idx = np.arange(0, len(w.ypr), stride)
    times = []
    same = True
    for _ in range(runs):
        t0 = time.perf_counter()
        ref = O.sweep(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr[idx], nthreads=threads)
        times.append(time.perf_counter() - t0)
        same = same and bool(np.array_equal(ref["Vss"], records["Vss"][idx]) and
                             np.array_equal(ref["Vo"], records["Vo"][idx]) and
                             np.array_equal(ref["Mss"].view(np.uint32), records["Mss"][idx].view(np.uint32)))
    med = float(np.median(times))
    idx1 = np.arange(0, len(w.ypr), stride * 20)
    t0 = time.perf_counter()
    ref1 = O.sweep(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr[idx1], nthreads=1)
    dt1 = time.perf_counter() - t0
    same = same and bool(np.array_equal(ref1["Vss"], records["Vss"][idx1]))
    return {"value": round(len(idx) / med, 2), "unit": "orientations/s", "cores": threads, "kind": "port",
            "sample": f"one orientation in {stride} of the same sweep ({len(idx)} orientations), median of "
                      f"{runs} runs ({', '.join(f'{t:.2f}' for t in times)} s wall), oracle in C, "
                      f"orientation-parallel pthreads",
            "one_thread": {"value": round(len(idx1) / dt1, 2), "unit": "orientations/s",
                           "sample": f"one orientation in {stride * 20} ({len(idx1)}, {dt1:.1f} s wall)"},
            "cpu_model": model, "nproc": nproc, "cpus_available": avail,
            "matches_gpu_on_sample": same}


def bench_batch70(args):
    """70 meshes (6 thin-shell manikin parts, 64 displaced spheres of 1k-100k triangles), each on
    its own 1 degree x 0..60 grid: autotune, capture one graph of persistent work-queue launches,
    time `steps` launches; compare with the same graph in the chained form (each member's own
    launches) and with the sweeps launched mesh by mesh with the default shape and no graph."""
    from tomo_shell2026_amd import _lib, synth
    ws = synth.batch70_workloads()
    total = sum(len(w.ypr) for w in ws)
    B = _lib.Batch()
    for w in ws:
        B.add(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr)
    t0 = time.perf_counter()
    progress("batch70: tuning launch shapes")
    B.tune(4096)
    progress("batch70: capturing the graph")
    tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    B.capture()
    capture_s = time.perf_counter() - t0
    for _ in range(max(args.warmup, 0)):
        B.launch()
        B.sync()
    ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        B.launch()
        B.sync()
        ms.append(B.last_ms())
    dt = time.perf_counter() - t0
    rounds, units = B.plan()
    queue_recs = [B.records(m) for m in range(len(ws))]
    progress("batch70: the chained form")
    B.set_mode(B.CHAIN)
    B.capture()
    B.launch()
    B.sync()
    chain_ms = []
    for _ in range(args.steps):
        B.launch()
        B.sync()
        chain_ms.append(B.last_ms())
    chain_same = all(B.records(m).tobytes() == queue_recs[m].tobytes() for m in range(len(ws)))
    shapes = {}
    for m in range(len(ws)):
        t = B.tuning(m)
        k = (t["batch"], t["tile_y"], t["inline"])
        shapes[str(k)] = shapes.get(str(k), 0) + 1
    flags = {"cap_drop": 0, "fp32_inexact": 0, "footprint_oob": 0, "unresolved": 0}
    for m in range(len(ws)):
        r = queue_recs[m]
        flags["cap_drop"] += int(((r["flags"] & 8) != 0).sum())
        flags["fp32_inexact"] += int(((r["flags"] & 1) != 0).sum())
        flags["footprint_oob"] += int(((r["flags"] & 4) != 0).sum())
        flags["unresolved"] += int(((r["flags"] & 64) != 0).sum())
    # every 499th orientation of every member against the oracle
    from oracle import oracle as O
    O.build()
    sample_ok, sample_n = True, 0
    for m, w in enumerate(ws):
        r = queue_recs[m]
        idx = np.arange(0, len(w.ypr), 499)
        ref = O.sweep(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm, w.ypr[idx], nthreads=16)
        sample_n += len(idx)
        for k in ("Vo", "Vss", "Va", "Vb", "Vbed", "n_entries", "chk_final", "flags"):
            sample_ok = sample_ok and bool(np.array_equal(r[k][idx], ref[k]))
        sample_ok = sample_ok and bool(np.array_equal(r["Mss"][idx].view(np.uint32), ref["Mss"].view(np.uint32)))
    B.close()
    # the same sweeps, mesh by mesh, default launch shape, no graph
    plain_ms = 0.0
    for k, w in enumerate(ws):
        if k % 10 == 0:
            progress(f"batch70: mesh-by-mesh sweep {k} of {len(ws)}")
        with _lib.Sweep(w.f32, w.i32, w.grid) as s:
            s.set_mesh(w.tri, w.vtx, w.vnrm)
            s.set_orientations(w.ypr)
            s.run()
            s.sync()
            s.run()
            s.sync()
            plain_ms += s.kernel_ms()
    graph_ms = float(np.mean(ms))
    return {
        "metric": "orientation evaluations/s, 70 uneven meshes in one persistent graph launch",
        "value": round(total * args.steps / dt, 1), "unit": "orientations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32",
        "data": "synthetic",
        "config": {"workload": "batch70", "meshes": len(ws), "triangles_total": int(sum(len(w.tri) for w in ws)),
                   "orientations_total": total, "grids": sorted({w.grid for w in ws})},
        "graph_ms": round(graph_ms, 2), "rounds": rounds, "units": units,
        "chain_graph_ms": round(float(np.mean(chain_ms)), 2), "chain_matches_queue": chain_same,
        "queue_vs_chain_speedup": round(float(np.mean(chain_ms)) / graph_ms, 3),
        "plain_per_mesh_ms": round(plain_ms, 2),
        "graph_vs_plain_speedup": round(plain_ms / graph_ms, 3),
        "tune_s": round(tune_s, 2), "capture_s": round(capture_s, 2), "tuned_shapes": shapes, "flags": flags,
        "matches_oracle_on_sample": sample_ok, "oracle_sample": f"every 499th orientation of every member ({sample_n})",
    }


def refine_comparison(w):
    """Least voxel Vss on the dense 0.1 degree yaw x pitch grid (0..359.9 x 0..60) by the plain
    sweep, against the coarse 5 degree sweep + batched gradient refinement of the continuous
    support volume + voxel judging on the 0.1 degree grid."""
    from tomo_shell2026_amd.refine import OrientationSearch
    srch = OrientationSearch(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm)
    srch.refined()  # warm-up (kernel loads, torch initialisation)
    ref = srch.refined()
    progress("refine: dense 0.1 degree grid")
    dense = srch.dense(0.1)
    srch.close()
    return {"dense_grid_deg": 0.1, "dense": dense, "refined": ref,
            "vss_gap_rel": round((ref["vss"] - dense["vss"]) / dense["vss"], 6),
            "wall_speedup": round(dense["wall_s"] / ref["wall_s"], 1),
            "kernel_speedup": round(dense["kernel_ms"] / ref["kernel_ms"], 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0, help="orientations in flight (0 = auto)")
    ap.add_argument("--cpu-stride", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("dsphere", "manikin", "batch70"), default="dsphere",
                    help="dsphere: 69,120-triangle displaced sphere, 256 grid (the headline workload); "
                         "manikin: ~87k-triangle thin-shell manikin, 512 grid, shell mode; "
                         "batch70: 70 uneven meshes in one tuned, captured graph launch")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 70-mesh batched graph launch measurement")
    ap.add_argument("--no-refine", action="store_true",
                    help="skip the dense 0.1 degree grid vs coarse + gradient-refined search comparison")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy (for scaling rocprof FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="rocprofv3 --pmc output folders: write the traffic summary under profiles/ and exit")
    ap.add_argument("--summarize-counters", nargs=2, metavar=("LDS_PMC", "WRITE_PMC"),
                    help="rocprofv3 --pmc output folders of the LDS and write-request counter passes: write "
                         "profiles/sweep_dsphere69120_counters.json and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.summarize_counters:
        summ = summarize_counters(*args.summarize_counters)
        with open(os.path.join(ROOT, "profiles", "sweep_dsphere69120_counters.json"), "w") as fh:
            json.dump(summ, fh, indent=1)
        print(json.dumps(summ))
        return
    if args.summarize_pmc:
        summ = summarize_pmc(*args.summarize_pmc)
        with open(os.path.join(ROOT, "profiles", "raster_kernel_dsphere69120_traffic.json"), "w") as fh:
            json.dump(summ, fh, indent=1)
        print(json.dumps(summ))
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; the sweep runs on exactly one GPU "
              f"(orientation sharding across GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    from tomo_shell2026_amd import _lib, synth
    if args.workload == "batch70":
        print(json.dumps(bench_batch70(args)), flush=True)
        return
    w = synth.manikin_workload() if args.workload == "manikin" else synth.displaced_sphere_workload()
    n = len(w.ypr)
    s = _lib.Sweep(w.f32, w.i32, w.grid)
    s.set_mesh(w.tri, w.vtx, w.vnrm)
    s.set_orientations(w.ypr)
    if args.batch:
        s.set_launch(args.batch)
    rec = None
    for _ in range(max(args.warmup, 0)):
        s.run()
        rec = s.records()
    s.sync()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.run()
        rec = s.records()  # synchronises
        kernel_ms.append(s.kernel_ms())
    s.sync()
    dt = time.perf_counter() - t0
    info = s.info()
    value = n * args.steps / dt
    b_orient, p_bar, s_bar = algorithmic_bytes(len(w.vtx), len(w.tri), rec)
    k_ms = float(np.mean(kernel_ms))
    # per-stage times from HIP events between the kernels (an untimed pass)
    s.set_phase_profile(True)
    s.run()
    s.sync()
    stages = s.phase_profile()
    s.set_phase_profile(False)
    batch = info["batch"]
    launches = (n + batch - 1) // batch
    raster_ms_launch = stages["raster"] / launches
    rb = raster_bytes(len(w.vtx), len(w.tri), p_bar, s_bar)
    achieved = rb * batch / (raster_ms_launch * 1e-3) / 1e9
    traffic, counters = None, None
    tpath = os.path.join(ROOT, "profiles", "raster_kernel_dsphere69120_traffic.json")
    cpath = os.path.join(ROOT, "profiles", "sweep_dsphere69120_counters.json")
    if args.workload == "dsphere" and os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    if args.workload == "dsphere" and os.path.exists(cpath):
        with open(cpath) as fh:
            counters = json.load(fh).get("raster_kernel")
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "orientations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+i32",
        "data": "synthetic",
        "config": {"workload": w.name, "triangles": int(len(w.tri)), "vertices": int(len(w.vtx)),
                   "orientations": n, "grid": w.grid, "voxel_mm": w.voxel_mm, "bed": "raft",
                   "support": "explicit", "shell": bool(w.shell), "note": w.scale_note, "batch": info["batch"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "raster_kernel", "launch_orientations": batch,
                     "kernel_ms_per_launch": round(raster_ms_launch, 3),
                     "lds_bank_conflict_rate": None if counters is None else counters["lds_bank_conflict_rate"],
                     "write_bytes_from_requests": None if counters is None else counters["write_bytes_per_launch"],
                     "algorithmic_bytes_per_launch": round(rb * batch),
                     "algorithmic_bytes_per_orientation": round(rb),
                     "mean_entries": round(p_bar, 1), "mean_slots": round(s_bar, 1),
                     "stage_ms_per_sweep": {k: round(v, 3) for k, v in stages.items()},
                     "sweep_kernels_ms": round(k_ms, 3),
                     "sweep_bytes_per_orientation": round(b_orient),
                     "sweep_achieved_gbs": round(b_orient * n / (k_ms * 1e-3) / 1e9, 1)},
        "flags": {"cap_drop_orientations": int(((rec["flags"] & 8) != 0).sum()),
                  "fp32_inexact": int(((rec["flags"] & 1) != 0).sum()),
                  "footprint_oob": int(((rec["flags"] & 4) != 0).sum()),
                  "unresolved": int(((rec["flags"] & 64) != 0).sum())},
        "argmin": int(_lib.find_first_optimal(rec["Mss"])),
    }
    s.close()
    progress(f"sweep: {value:.0f} orientations/s, batch {info['batch']}")
    if not args.no_refine:
        out["refine"] = refine_comparison(w)
        progress(f"refine: {out['refine']['wall_speedup']}x less wall time")
        if args.workload == "dsphere":
            # the differentiation workload's mesh: the thin-shell manikin (512 grid, shell mode)
            mw = synth.manikin_workload()
            out["refine_manikin"] = dict(refine_comparison(mw), workload=mw.name)
            progress(f"refine manikin: {out['refine_manikin']['wall_speedup']}x less wall time, "
                     f"gap {out['refine_manikin']['vss_gap_rel']}")
    if not args.no_batch and args.workload == "dsphere":
        b = bench_batch70(argparse.Namespace(steps=2, warmup=1))
        progress(f"batch70: {b['value']:.0f} orientations/s")
        out["batch70"] = {k: b[k] for k in ("value", "unit", "config", "graph_ms", "rounds", "units",
                                            "chain_graph_ms", "chain_matches_queue", "queue_vs_chain_speedup",
                                            "plain_per_mesh_ms", "graph_vs_plain_speedup", "tune_s", "capture_s",
                                            "tuned_shapes", "flags", "matches_oracle_on_sample", "oracle_sample")}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, rec, args.cpu_stride)
        progress(f"cpu baseline: {out['cpu_baseline']['value']} orientations/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
