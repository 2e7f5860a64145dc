"""Continuous support volume Vss_c(yaw, pitch) on the GPU, differentiable through
torch.autograd.

    Vss_c = sum over triangles t of sigma(-r.n_t - sin(theta_c)) A_t (-r.n_t) (r.c_t - zlow)

sigma is a C1 smoothstep of width tau around the critical angle and zlow a soft minimum
(temperature temp) of the vertex heights; with tau = temp = 0 it is the hard form
sum over r.n_t < -sin(theta_c) of A_t (-r.n_t) (r.c_t - min_v r.v).

r = (-sin pitch, cos pitch sin yaw, cos pitch cos yaw) is the build direction in mesh
coordinates: the third row of the reference's rotation Rz(roll) Ry(pitch) Rx(yaw)
(SMatrix4f.cpp:213-240), so roll does not enter.  n_t, A_t, c_t are the unit face normal, area
and centroid.  The per-triangle projected-area-times-height form follows getCHVss
(STomoNV_CvxH.cpp:97-174), with heights measured from the build plate (the lowest vertex).

Forward and backward are the hand-written HIP kernels of libtomo_hip.so (tomo_cvss_forward /
tomo_cvss_backward); PyTorch only supplies device memory, the stream and the autograd plumbing.
"""
import ctypes as ct

import numpy as np
import torch

from . import _lib

AUX_DOUBLES = 12
TAU, TEMP = 0.1, 0.1  # default smoothstep width (units of r.n) and softmin temperature (voxels)


class ContinuousVssMesh:
    """A mesh resident on the GPU for Vss_c evaluations (f32x12 supplies theta_c = f32x12[1]).
    tau: smoothstep width around the critical angle; temp: softmin temperature of the lowest
    point (both 0 = the hard, piecewise form)."""

    def __init__(self, f32, tri, vtx, tau=TAU, temp=TEMP):
        f = np.ascontiguousarray(f32, np.float32)
        t = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        v = np.ascontiguousarray(vtx, np.float32).reshape(-1, 3)
        h = ct.c_void_p()
        _lib.check(_lib.load().tomo_cvss_create(_lib.ptr(f), t.shape[0], _lib.ptr(t), v.shape[0], _lib.ptr(v),
                                                ct.c_float(tau), ct.c_float(temp), ct.byref(h)),
                   "tomo_cvss_create")
        self.tau, self.temp = float(tau), float(temp)
        self._h = h
        self.n_tri, self.n_vtx = t.shape[0], v.shape[0]

    def close(self):
        if self._h:
            _lib.load().tomo_cvss_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, ypr):
        """Vss_c for a (B, 3) float32 CUDA tensor of (yaw, pitch, roll) in radians."""
        return ContinuousVss.apply(ypr, self)

    # raw calls (tensors must be contiguous CUDA tensors on the current device)
    def forward_raw(self, ypr, vss, aux, stream):
        _lib.check(_lib.load().tomo_cvss_forward(self._h, ypr.shape[0], ct.c_void_p(ypr.data_ptr()),
                                                 ct.c_void_p(vss.data_ptr()), ct.c_void_p(aux.data_ptr()),
                                                 ct.c_void_p(stream)), "tomo_cvss_forward")

    def backward_raw(self, ypr, aux, grad_vss, grad_ypr, stream):
        _lib.check(_lib.load().tomo_cvss_backward(self._h, ypr.shape[0], ct.c_void_p(ypr.data_ptr()),
                                                  ct.c_void_p(aux.data_ptr()), ct.c_void_p(grad_vss.data_ptr()),
                                                  ct.c_void_p(grad_ypr.data_ptr()), ct.c_void_p(stream)),
                   "tomo_cvss_backward")


def _check_ypr(ypr):
    if not (isinstance(ypr, torch.Tensor) and ypr.is_cuda and ypr.dtype == torch.float32 and ypr.dim() == 2
            and ypr.shape[1] == 3):
        raise ValueError("ypr must be a (B, 3) float32 CUDA tensor")
    return ypr.contiguous()


class ContinuousVss(torch.autograd.Function):
    """vss = ContinuousVss.apply(ypr, mesh): forward and backward run the HIP kernels on torch's
    current stream."""

    @staticmethod
    def forward(ctx, ypr, mesh):
        ypr = _check_ypr(ypr.detach())
        n = ypr.shape[0]
        vss = torch.empty(n, dtype=torch.float32, device=ypr.device)
        aux = torch.empty((n, AUX_DOUBLES), dtype=torch.float64, device=ypr.device)
        stream = torch.cuda.current_stream(ypr.device).cuda_stream
        mesh.forward_raw(ypr, vss, aux, stream)
        ctx.save_for_backward(ypr, aux)
        ctx.mesh = mesh
        return vss

    @staticmethod
    def backward(ctx, grad_vss):
        ypr, aux = ctx.saved_tensors
        g = grad_vss.detach().to(torch.float32).contiguous()
        grad = torch.empty_like(ypr)
        stream = torch.cuda.current_stream(ypr.device).cuda_stream
        ctx.mesh.backward_raw(ypr, aux, g, grad, stream)
        return grad, None
