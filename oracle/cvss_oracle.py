"""FP64 NumPy oracle for the continuous support volume and its gradient.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks, never by the product).

    Vss_c(yaw, pitch) = sum_t sigma(-r.n_t - s) A_t (-r.n_t) (r.c_t - zlow),   s = sin(theta_c)

r = (-sin pitch, cos pitch sin yaw, cos pitch cos yaw) is the third row of the reference's
rotation Rz(roll) Ry(pitch) Rx(yaw) (SMatrix4f.cpp:213-240; the sweep's ypr_matrix); n_t / A_t /
c_t are the unit face normal, area and centroid.  The per-triangle projected-area-times-height
precedent is STomoNV_CvxH::getCHVss (STomoNV_CvxH.cpp:97-174), which the reference never
compiles; this form measures heights from the build plate.

Two smoothings make the function continuously differentiable (both 0 = the hard form):
  * sigma, a C1 smoothstep of width tau centred on the critical angle (u = clamp(x / tau + 1/2),
    sigma = u^2 (3 - 2u)); with the hard step [x > 0] a flat triangle switches on as a whole
    and Vss_c jumps;
  * zlow, the soft minimum -T log sum_v exp(-r.v / T) of the vertex heights (T = temp); the
    hard minimum has a kink wherever the lowest vertex changes.
Everything here is float64.

Pins (tests/test_cvss.py): the closed form of a sphere, 2 pi rho^3 (1/6 - s^2/2 + s^3/3);
invariance under roll; central differences of cvss() against cvss_grad(); a single triangle
worked by hand.
"""
import numpy as np


def geometry(tri, vtx):
    """Face unit normals, areas, centroids and vertices in float64, relative to the AABB centre."""
    v = np.asarray(vtx, np.float64).reshape(-1, 3)
    t = np.asarray(tri, np.int64).reshape(-1, 3)
    c0 = 0.5 * (v.min(0) + v.max(0)) if len(v) else np.zeros(3)
    v = v - c0
    p0, p1, p2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    ln = np.linalg.norm(n, axis=1)
    unit = np.where(ln[:, None] > 0, n / np.where(ln > 0, ln, 1.0)[:, None], 0.0)
    return unit, 0.5 * ln, (p0 + p1 + p2) / 3.0, v


def build_dir(ypr):
    y = np.asarray(ypr, np.float64).reshape(-1, 3)
    sy, cy, sp, cp = np.sin(y[:, 0]), np.cos(y[:, 0]), np.sin(y[:, 1]), np.cos(y[:, 1])
    r = np.stack([-sp, cp * sy, cp * cy], 1)
    dr_dyaw = np.stack([np.zeros_like(sp), cp * cy, -cp * sy], 1)
    dr_dpitch = np.stack([-cp, -sp * sy, -sp * cy], 1)
    return r, dr_dyaw, dr_dpitch


def _chunks(n, size):
    for a in range(0, n, size):
        yield a, min(n, a + size)


def step(x, tau):
    """(sigma, d sigma / dx) of the smoothstep of width tau (tau = 0: hard step, derivative 0)."""
    if tau <= 0:
        return (x > 0).astype(np.float64), np.zeros_like(x)
    u = np.clip(x / tau + 0.5, 0.0, 1.0)
    return u * u * (3.0 - 2.0 * u), 6.0 * u * (1.0 - u) / tau


def low(zv, temp):
    """(zlow, weights over vertices whose weighted sum is d zlow / d z_v) per row of zv."""
    vi = zv.argmin(1)
    zmin = zv[np.arange(len(zv)), vi]
    if temp <= 0:
        p = np.zeros_like(zv)
        p[np.arange(len(zv)), vi] = 1.0
        return zmin, p
    e = np.exp(-(zv - zmin[:, None]) / temp)
    se = e.sum(1)
    return zmin - temp * np.log(se), e / se[:, None]


def cvss(tri, vtx, theta_c, ypr, tau=0.0, temp=0.0, s=None, chunk=64):
    """Vss_c for each row of ypr (float64).  s overrides sin(theta_c) (for matching a float32
    threshold exactly)."""
    n, A, c, v = geometry(tri, vtx)
    s = np.sin(np.float64(theta_c)) if s is None else np.float64(s)
    r, _, _ = build_dir(ypr)
    out = np.zeros(len(r))
    for a, b in _chunks(len(r), chunk):
        rr = r[a:b]
        d = rr @ n.T                                # (m, T)
        zmin, _ = low(rr @ v.T, temp)               # (m,)
        h = rr @ c.T - zmin[:, None]
        sig, _ = step(-d - s, tau)
        out[a:b] = (sig * A[None, :] * (-d) * h).sum(1)
    return out


def cvss_grad(tri, vtx, theta_c, ypr, tau=0.0, temp=0.0, s=None, chunk=64):
    """d Vss_c / d (yaw, pitch, roll) for each row of ypr (float64); with temp = 0 the lowest
    vertex is the argmin (first index on ties)."""
    n, A, c, v = geometry(tri, vtx)
    s = np.sin(np.float64(theta_c)) if s is None else np.float64(s)
    r, dy, dp = build_dir(ypr)
    g = np.zeros((len(r), 3))
    for a, b in _chunks(len(r), chunk):
        rr = r[a:b]
        d = rr @ n.T
        zmin, pv = low(rr @ v.T, temp)
        vstar = pv @ v                              # d zlow / d r
        h = rr @ c.T - zmin[:, None]
        sig, dsig = step(-d - s, tau)
        w = sig * A[None, :] * (-d)
        dw_dd = -A[None, :] * (sig + (-d) * dsig)
        # d/dr = sum_t [dw/dd n (r.c - zlow) + w (c - v*)]
        gr = (dw_dd * h) @ n + w @ c - w.sum(1)[:, None] * vstar
        g[a:b, 0] = (gr * dy[a:b]).sum(1)
        g[a:b, 1] = (gr * dp[a:b]).sum(1)
    return g


def cvss_fd(tri, vtx, theta_c, ypr, h=1e-3, tau=0.0, temp=0.0, s=None):
    """Central differences of cvss() in yaw, pitch and roll with step h (radians)."""
    y = np.asarray(ypr, np.float64).reshape(-1, 3)
    g = np.zeros_like(y)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        g[:, k] = (cvss(tri, vtx, theta_c, y + e, tau, temp, s) - cvss(tri, vtx, theta_c, y - e, tau, temp, s)) / (2 * h)
    return g


def sphere_closed_form(rho, theta_c):
    """Vss_c of a sphere of radius rho: 2 pi rho^3 (1/6 - s^2/2 + s^3/3), s = sin(theta_c)."""
    s = np.sin(theta_c)
    return 2 * np.pi * rho ** 3 * (1.0 / 6.0 - s * s / 2.0 + s ** 3 / 3.0)


def so3_uniform_ypr(n, seed):
    """n rotations uniform on SO(3) (normalised Gaussian quaternions), as (yaw, pitch, roll) with
    R = Rz(roll) Ry(pitch) Rx(yaw); float32 radians."""
    q = np.random.default_rng(seed).normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - w * z); R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y); R[:, 2, 1] = 2 * (y * z + w * x); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    pitch = np.arcsin(np.clip(-R[:, 2, 0], -1.0, 1.0))
    yaw = np.arctan2(R[:, 2, 1], R[:, 2, 2])
    roll = np.arctan2(R[:, 1, 0], R[:, 0, 0])
    return np.stack([yaw, pitch, roll], 1).astype(np.float32)
