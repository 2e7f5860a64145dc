"""Orientation search: plain dense grid against coarse grid + gradient refinement.

Plain: the voxel sweep over a dense yaw x pitch grid; the answer is the orientation of least
voxel Vss (integer support volume, the quantity Pairing accumulates, STomoSh_INT3.cpp:366-395).

Batched + refined:
  1. voxel sweep of a coarse grid (default 5 degrees);
  2. the top_k coarse cells by voxel Vss start a batched gradient descent on the continuous
     support volume Vss_c (HIP forward/backward through torch.autograd), each start held to a
     box of +-coarse_step around its cell;
  3. the voxel sweep judges: every fine-grid orientation (default 0.1 degrees) within
     +-patch of each refined point and of each start cell is evaluated;
  4. polish: up to polish_rounds more patches around the polish_top best judged points, while
     they keep improving; the least voxel Vss wins.
Both searches report kernel time (HIP events around the voxel sweeps, CUDA events around the
gradient steps) and wall time.
"""
import math
import time

import numpy as np
import torch

from . import _lib
from .cvss import ContinuousVssMesh


def grid_deg(yaw_step, pitch_step, pitch_max=60.0, yaw_max=360.0):
    """(yaw, pitch) in degrees, pitch-major rows, yaw fastest; yaw in [0, yaw_max)."""
    ny = int(round(yaw_max / yaw_step))
    npch = int(round(pitch_max / pitch_step)) + 1
    yaw = np.arange(ny) * yaw_step
    pit = np.arange(npch) * pitch_step
    yy, pp = np.meshgrid(yaw, pit)
    return np.stack([yy.ravel(), pp.ravel()], 1)


def to_ypr(deg):
    """Degrees (yaw, pitch) -> float32 radians (yaw, pitch, 0) computed in float64."""
    d = np.asarray(deg, np.float64).reshape(-1, 2)
    return np.column_stack([d * (np.pi / 180.0), np.zeros(len(d))]).astype(np.float32)


class OrientationSearch:
    """Voxel sweep and continuous model of one mesh (arrays as for tomo_set_mesh)."""

    def __init__(self, f32, i32, grid, tri, vtx, vnrm, pitch_max=60.0):
        self.f32, self.i32, self.grid = f32, i32, grid
        self.sweep = _lib.Sweep(f32, i32, grid)
        self.sweep.set_mesh(tri, vtx, vnrm)
        self.cmesh = ContinuousVssMesh(f32, tri, vtx)
        self.pitch_max = float(pitch_max)
        self.voxel_evals = 0

    def close(self):
        self.sweep.close()
        self.cmesh.close()

    def voxel_vss(self, deg):
        """Integer voxel Vss for each (yaw, pitch) in degrees, and the sweep's kernel ms."""
        ypr = to_ypr(deg)
        self.sweep.set_orientations(ypr)
        self.sweep.run()
        self.sweep.sync()
        ms = self.sweep.kernel_ms()
        rec = self.sweep.records()
        self.voxel_evals += len(ypr)
        return rec["Vss"].astype(np.int64), ms

    # ---------------------------------------------------------------- plain dense grid
    def dense(self, step_deg=0.1):
        t0 = time.perf_counter()
        deg = grid_deg(step_deg, step_deg, self.pitch_max)
        vss, ms = self.voxel_vss(deg)
        i = int(np.argmin(vss))
        return {"best_deg": deg[i].tolist(), "vss": int(vss[i]), "evals": len(deg), "kernel_ms": ms,
                "wall_s": time.perf_counter() - t0}

    # ---------------------------------------------------------------- batched + refined
    def refined(self, coarse_step=5.0, top_k=16, gd_steps=40, lr_deg=0.5, patch_deg=0.5, fine_step=0.1,
                polish_rounds=4, polish_top=4):
        t0 = time.perf_counter()
        evals0 = self.voxel_evals
        kernel_ms = 0.0
        coarse = grid_deg(coarse_step, coarse_step, self.pitch_max)
        vss_c, ms = self.voxel_vss(coarse)
        kernel_ms += ms
        order = np.argsort(vss_c, kind="stable")[:top_k]
        starts = coarse[order]
        # batched gradient descent on Vss_c, each start boxed to +-coarse_step around its cell
        dev = torch.device("cuda")
        rad = math.pi / 180.0
        x0 = torch.tensor(starts * rad, dtype=torch.float32, device=dev)
        x = x0.clone().requires_grad_(True)
        lo = torch.stack([x0[:, 0] - coarse_step * rad, torch.clamp(x0[:, 1] - coarse_step * rad, min=0.0)], 1)
        hi = torch.stack([x0[:, 0] + coarse_step * rad,
                          torch.clamp(x0[:, 1] + coarse_step * rad, max=self.pitch_max * rad)], 1)
        opt = torch.optim.Adam([x], lr=lr_deg * rad)
        zero = torch.zeros((len(starts), 1), dtype=torch.float32, device=dev)
        best_x = x0.clone()
        best_v = torch.full((len(starts),), float("inf"), device=dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(gd_steps):
            opt.zero_grad(set_to_none=True)
            v = self.cmesh(torch.cat([x, zero], 1))
            with torch.no_grad():
                better = v < best_v
                best_v = torch.where(better, v, best_v)
                best_x = torch.where(better[:, None], x.detach(), best_x)
            v.sum().backward()
            opt.step()
            with torch.no_grad():
                x.copy_(torch.minimum(torch.maximum(x, lo), hi))
        with torch.no_grad():
            v = self.cmesh(torch.cat([x, zero], 1))
            better = v < best_v
            best_x = torch.where(better[:, None], x, best_x)
        e1.record()
        torch.cuda.synchronize()
        gd_ms = e0.elapsed_time(e1)
        kernel_ms += gd_ms
        refined = best_x.cpu().numpy().astype(np.float64) / rad
        # judge: fine-grid patches around the refined points and the start cells, then polish
        # rounds around the best judged points until no patch improves
        q = int(round(360.0 / fine_step))
        seen = {}

        def patch_points(centres):
            k = int(round(patch_deg / fine_step))
            offs = np.arange(-k, k + 1)
            oy, op = np.meshgrid(offs, offs)
            pts = []
            for cy, cp in centres:
                gy = int(round(cy / fine_step)) + oy.ravel()
                gp = int(round(cp / fine_step)) + op.ravel()
                pts.append(np.stack([gy % q, gp], 1))
            pts = np.unique(np.concatenate(pts), axis=0)
            ok = (pts[:, 1] >= 0) & (pts[:, 1] * fine_step <= self.pitch_max + 1e-9)
            return [tuple(p) for p in pts[ok] if tuple(p) not in seen]

        def judge(points):
            nonlocal kernel_ms
            if not points:
                return
            arr = np.array(points, np.float64) * fine_step
            vals, ms = self.voxel_vss(arr)
            kernel_ms += ms
            for p_, v_ in zip(points, vals):
                seen[p_] = int(v_)

        for p_, v_ in zip(coarse, vss_c):  # coarse points lie on the fine grid
            seen[(int(round(p_[0] / fine_step)) % q, int(round(p_[1] / fine_step)))] = int(v_)
        judge(patch_points(np.concatenate([refined, starts])))
        polished = 0
        for _ in range(polish_rounds):
            best_now = min(seen.values())
            top = sorted(seen.items(), key=lambda kv: kv[1])[:polish_top]
            pts = patch_points([(a * fine_step, b * fine_step) for (a, b), _ in top])
            if not pts:
                break
            judge(pts)
            polished += 1
            if min(seen.values()) >= best_now:
                break
        (bi, bp), bv = min(seen.items(), key=lambda kv: (kv[1], kv[0]))
        best = ([bi * fine_step, bp * fine_step], bv)
        n_judge = len(seen) - len(coarse)
        return {"best_deg": best[0], "vss": best[1], "evals": self.voxel_evals - evals0,
                "coarse_evals": len(coarse), "judge_evals": n_judge, "polish_rounds": polished,
                "gd_steps": gd_steps, "starts": len(starts),
                "gd_ms": gd_ms, "kernel_ms": kernel_ms, "wall_s": time.perf_counter() - t0}
