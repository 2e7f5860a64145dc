"""Continuous support volume Vss_c: the FP64 oracle pinned by closed forms and finite
differences (CPU), and the HIP forward/backward kernels checked against it (GPU).

Tolerances: the HIP kernels evaluate each term in float32 and sum in float64, so values agree
with the FP64 oracle to ~1e-5 relative; gradients are compared as vector norms per
orientation."""
import numpy as np
import pytest

from oracle import cvss_oracle as O
from tomo_shell2026_amd import synth

TH = np.pi / 3


def test_single_triangle_by_hand():
    # One face, normal pointing straight down at yaw = pitch = 0, area 50, centroid height 20/3
    # above the lowest vertex at z = 0:  Vss_c = A * 1 * h = 50 * 20/3.
    v = np.array([[0, 0, 10.0], [0, 10, 10.0], [10, 0, 10.0], [3, 3, 0.0]])
    f = np.array([[0, 1, 2]])
    # the lone vertex (3, 3, 0) is not part of the face but sets the build plate
    val = O.cvss(f, v, TH, [[0, 0, 0]])
    assert abs(val[0] - 50.0 * 10.0) < 1e-9
    # tilted 20 degrees about x (yaw): the face stays below the critical angle; the projected
    # area times height changes smoothly; compare with the hand formula
    y = np.deg2rad(20.0)
    r = np.array([0.0, np.sin(y), np.cos(y)])
    n = np.array([0.0, 0.0, -1.0])
    p = v - 0.5 * (v.min(0) + v.max(0))
    c = p[:3].mean(0)
    want = 50.0 * (-(r @ n)) * (r @ c - (p @ r).min())
    assert abs(O.cvss(f, v, TH, [[y, 0, 0]])[0] - want) < 1e-9


def test_sphere_closed_form():
    v, f = synth.icosphere(5, 50.0)
    got = O.cvss(f, v, TH, synth.ypr_grid([0, 45, 90], [0, 30]))
    want = O.sphere_closed_form(50.0, TH)
    assert np.all(np.abs(got / want - 1) < 0.02)   # triangulation error of the 20,480-face sphere
    # the error shrinks with refinement
    v4, f4 = synth.icosphere(4, 50.0)
    e4 = abs(O.cvss(f4, v4, TH, [[0, 0, 0]])[0] / want - 1)
    e5 = abs(O.cvss(f, v, TH, [[0, 0, 0]])[0] / want - 1)
    assert e5 < e4


def test_roll_invariance_and_fd():
    w = synth.displaced_sphere_workload()
    ypr = O.so3_uniform_ypr(12, 5).astype(np.float64)
    base = O.cvss(w.tri, w.vtx, TH, ypr, tau=0.1, temp=0.1)
    rolled = ypr.copy()
    rolled[:, 2] += 0.7
    np.testing.assert_array_equal(O.cvss(w.tri, w.vtx, TH, rolled, tau=0.1, temp=0.1), base)
    g = O.cvss_grad(w.tri, w.vtx, TH, ypr, tau=0.1, temp=0.1)
    fd = O.cvss_fd(w.tri, w.vtx, TH, ypr, h=1e-3, tau=0.1, temp=0.1)
    rel = np.linalg.norm(g - fd, axis=1) / np.linalg.norm(fd, axis=1)
    assert np.median(rel) < 1e-4 and rel.max() < 1e-3
    assert np.all(g[:, 2] == 0) and np.all(np.abs(fd[:, 2]) < 1e-6 * np.abs(fd[:, :2]).max())


def test_grad_detects_planted_mistake():
    """A gradient with the lowest-point term dropped (a plausible slip) fails the FD check."""
    w = synth.displaced_sphere_workload()
    ypr = O.so3_uniform_ypr(6, 9).astype(np.float64)
    fd = O.cvss_fd(w.tri, w.vtx, TH, ypr, h=1e-3, tau=0.1, temp=0.1)
    n, A, c, v = O.geometry(w.tri, w.vtx)
    r, dy, dp = O.build_dir(ypr)
    d = r @ n.T
    zmin, pv = O.low(r @ v.T, 0.1)
    h = r @ c.T - zmin[:, None]
    sig, dsig = O.step(-d - np.sin(TH), 0.1)
    w_ = sig * A * (-d)
    gr = (-A * (sig + (-d) * dsig) * h) @ n + w_ @ c   # missing "- sum(w) v*"
    bad = np.stack([(gr * dy).sum(1), (gr * dp).sum(1)], 1)
    rel = np.linalg.norm(bad - fd[:, :2], axis=1) / np.linalg.norm(fd[:, :2], axis=1)
    assert rel.min() > 1e-2


def test_so3_uniform_ypr_is_rotation_of_reference_convention():
    from oracle import oracle
    ypr = O.so3_uniform_ypr(64, 3)
    r, _, _ = O.build_dir(ypr)
    for k in range(0, 64, 7):
        R = oracle.ypr_matrix(*ypr[k]).reshape(3, 3)
        np.testing.assert_allclose(R[2], r[k], atol=2e-6)
    # build directions are uniform on the sphere: mean ~ 0
    r = O.build_dir(O.so3_uniform_ypr(20000, 4))[0]
    assert np.abs(r.mean(0)).max() < 0.03


# ---------------------------------------------------------------------------------- GPU
def _gpu_eval(mesh, ypr, grad=True):
    import torch
    y = torch.tensor(np.asarray(ypr, np.float32), device="cuda", requires_grad=grad)
    vss = mesh(y)
    g = None
    if grad:
        vss.sum().backward()
        g = y.grad.detach().cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()
    return vss.detach().cpu().numpy().astype(np.float64), g


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["manikin", "dsphere"])
def test_gpu_cvss_against_fp64(which):
    """4,096 seeded starts uniform on SO(3): HIP forward vs the FP64 oracle, HIP backward vs the
    FP64 gradient and vs FP64 central differences (h = 1e-3 rad) on a subset."""
    from tomo_shell2026_amd.cvss import ContinuousVssMesh, TAU, TEMP
    print(f"[gpu-case] cvss {which} 4096 SO(3) starts", flush=True)
    w = synth.manikin_workload() if which == "manikin" else synth.displaced_sphere_workload()
    m = ContinuousVssMesh(w.f32, w.tri, w.vtx)
    ypr = O.so3_uniform_ypr(4096, 11)
    val, grad = _gpu_eval(m, ypr)
    s32 = float(np.sin(np.float32(w.f32[1]), dtype=np.float32))
    ref = O.cvss(w.tri, w.vtx, w.f32[1], ypr, tau=TAU, temp=TEMP, s=s32)
    relv = np.abs(val - ref) / np.maximum(np.abs(ref), 1.0)
    print(f"value rel err: median {np.median(relv):.2e} max {relv.max():.2e}", flush=True)
    assert np.median(relv) < 1e-5 and relv.max() < 1e-3
    gref = O.cvss_grad(w.tri, w.vtx, w.f32[1], ypr, tau=TAU, temp=TEMP, s=s32)
    relg = np.linalg.norm(grad - gref, axis=1) / np.maximum(np.linalg.norm(gref, axis=1), 1e-6 * np.abs(ref).max())
    print(f"grad rel err vs FP64: median {np.median(relg):.2e} p99 {np.quantile(relg, 0.99):.2e}", flush=True)
    assert np.median(relg) < 1e-4 and np.quantile(relg, 0.99) < 1e-2
    assert np.all(grad[:, 2] == 0)
    sub = np.arange(0, 4096, 128)
    fd = O.cvss_fd(w.tri, w.vtx, w.f32[1], ypr[sub], h=1e-3, tau=TAU, temp=TEMP, s=s32)
    relf = np.linalg.norm(grad[sub] - fd, axis=1) / np.linalg.norm(fd, axis=1)
    print(f"grad rel err vs central FD: median {np.median(relf):.2e} max {relf.max():.2e}", flush=True)
    assert np.median(relf) < 1e-3


@pytest.mark.gpu
def test_gpu_cvss_hard_form_matches_oracle():
    """tau = temp = 0: the hard indicator and the lowest vertex (icosphere against its closed form
    and the FP64 oracle)."""
    from tomo_shell2026_amd.cvss import ContinuousVssMesh
    print("[gpu-case] cvss hard form icosphere", flush=True)
    v, f = synth.icosphere(5, 50.0)
    f32 = synth.default_f32(1.0)
    m = ContinuousVssMesh(f32, f, v, tau=0.0, temp=0.0)
    ypr = synth.ypr_grid(np.arange(0, 360, 30), np.arange(0, 181, 30))
    val, _ = _gpu_eval(m, ypr, grad=False)
    s32 = float(np.sin(np.float32(f32[1]), dtype=np.float32))
    ref = O.cvss(f, v, f32[1], ypr, s=s32)
    np.testing.assert_allclose(val, ref, rtol=2e-5)
    assert np.all(np.abs(val / O.sphere_closed_form(50.0, f32[1]) - 1) < 0.02)


@pytest.mark.gpu
def test_gpu_refined_search_against_tenth_degree_grid():
    """Coarse 5 degree voxel sweep + gradient refinement + 0.1 degree voxel judging must reach
    the plain dense 0.1 degree grid's least support volume (2,163,600 voxel evaluations) within
    0.5%, with at most a hundredth of its voxel evaluations, and beat the 1 degree grid."""
    from tomo_shell2026_amd.refine import OrientationSearch
    print("[gpu-case] refined search dsphere against the 0.1 degree grid", flush=True)
    w = synth.displaced_sphere_workload()
    s = OrientationSearch(w.f32, w.i32, w.grid, w.tri, w.vtx, w.vnrm)
    r = s.refined()
    d1 = s.dense(1.0)
    d01 = s.dense(0.1)
    s.close()
    print(r, d1, d01, flush=True)
    assert r["vss"] <= d1["vss"]
    assert (r["vss"] - d01["vss"]) <= 0.005 * d01["vss"]
    assert r["evals"] * 100 <= d01["evals"]
