"""TomoDLLInterface: the drop-in replacement of cfms_tomo/tomo_dll_interface.py.

Same constructor and run_dll contract as the reference (tomo_dll_interface.py:55-135): the mesh
and options are packed into Float32Info[12] / Int32Info[11] and the ten-pointer entry point is
called; per-orientation Mss, the optimum and the placement matrix are read back through the
same getters.  The entry point is TomoSh_HIP of libtomo_hip.so (MI355X), which reproduces the
reference CPU engine's integer results; both bUseCUDA settings use it.
"""
import ctypes as ct
import os
from copy import copy, deepcopy
from typing import Optional

import numpy as np

from .. import _lib
from ..mesh import TomoMesh
from .tomo_io import (Cptr1d, Cptr1d_to_np, Cptr1i, Cptr1i_to_np, Cptr1iL, findOptimals, np_to_Cptr1d,
                      np_to_Cptr1iL, toRadian)
from .tomo_object import TomoObject
from .tomo_slicingOptions import TOMO_nPixelFormat, TOMO_PixelVarNames, SlicingOptions, TomoPixels, enumPixelType


class TomoDLLInterface:
    CppDLL = None
    slicing_option = None
    nYPR_Intervals = 1
    (yaw_range, pitch_range, roll_range) = ([], [], [])
    (xx, yy, zz, YPR) = ([], [], [], [])
    (Float32Info, Int32Info) = ([], [])
    (bUseExplicitSS, bUseClosedVolumeVoxel) = (True, False)

    def __init__(self, dll_fname, s_o: Optional[SlicingOptions], bUseCUDA, bShellMesh):
        # dll_fname names the reference's Windows DLL; the HIP library is loaded instead.
        self.CppDLL = _lib.load()
        self.CppFunction = self.CppDLL.TomoSh_HIP
        self.bUseCUDA = bUseCUDA
        self.bShellMesh = bShellMesh
        L = self.CppDLL
        L.getMo.restype = Cptr1d
        L.getMss.restype = Cptr1d
        L.getMat4x4.restype = Cptr1d
        L.getnData2i.argtypes = (ct.c_int16,)
        L.getnData2i.restype = ct.c_int32
        L.getpData2i.argtypes = (ct.c_int16,)
        L.getpData2i.restype = Cptr1i
        S_O = self.slicing_option = copy(s_o) if s_o is not None else SlicingOptions()
        self.nYPR_Intervals = int(360 / S_O.theta_YP) + 1
        self.yaw_range = np.linspace(toRadian(0), toRadian(360), num=self.nYPR_Intervals, endpoint=True,
                                     dtype=np.float32)
        self.pitch_range = np.linspace(toRadian(0), toRadian(360), num=self.nYPR_Intervals, endpoint=True,
                                       dtype=np.float32)
        self.roll_range = np.zeros(1)
        (self.xx, self.yy, self.zz) = np.meshgrid(self.yaw_range, self.pitch_range, self.roll_range)
        self.YPR = np.column_stack([self.xx.ravel(), self.yy.ravel(), self.zz.ravel()]).astype(np.float32)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def free(self):
        if self.CppDLL is not None:
            self.CppDLL.OnDestroy()

    @staticmethod
    def _mesh_of(cutobject):
        return cutobject.tmesh if hasattr(cutobject, "tmesh") else cutobject

    def pack(self, T, bVerbose):
        """Float32Info / Int32Info in the reference's order (tomo_dll_interface.py:73-82)."""
        S_O = self.slicing_option
        vtx = np.asarray(T.vertices).astype(np.float32, copy=True)
        tri = np.asarray(T.faces).astype(np.int32, copy=True)
        H = getattr(T, "convex_hull", None)
        n_htri = len(H.faces) if H is not None and hasattr(H, "faces") else 0
        n_hvtx = len(H.vertices) if H is not None and hasattr(H, "vertices") else 0
        self.Float32Info = np.array([
            S_O.dVoxel, S_O.theta_c, T.area, S_O.wall_thickness,
            S_O.Fcore, S_O.Fclad, S_O.Fss, S_O.Css,
            S_O.PLA_density, S_O.BedType[1], S_O.BedType[2], S_O.BedType[3]]).astype(np.float32)
        self.Int32Info = np.array([
            bVerbose, self.bUseExplicitSS, self.bShellMesh, S_O.nVoxel,
            tri.shape[0], vtx.shape[0], self.YPR.shape[0],
            n_htri, n_hvtx, S_O.BedType[0],
            int(round(S_O.ShellThickness * 1000.0))]).astype(np.int32)
        return vtx, tri

    def run_dll(self, cutobject0, bVerbose):
        if bVerbose:
            print("Tomo processing: ", getattr(cutobject0, "name", "mesh"))
        c_o = deepcopy(cutobject0)
        T = self._mesh_of(c_o)
        self.bShellMesh = self.slicing_option.bShellMesh and (not T.is_watertight)
        vtx, tri = self.pack(T, bVerbose)
        vtx_nrm = np.asarray(T.vertex_normals).astype(np.float32, copy=True)
        tri_nrm = np.asarray(T.face_normals).astype(np.float32, copy=True)
        empty_i = np.zeros(3, np.int32)
        empty_f = np.zeros(3, np.float32)
        opt = self.CppFunction(
            np_to_Cptr1d(self.Float32Info), np_to_Cptr1iL(self.Int32Info), np_to_Cptr1d(self.YPR),
            np_to_Cptr1iL(tri), np_to_Cptr1d(vtx), np_to_Cptr1d(vtx_nrm), np_to_Cptr1d(tri_nrm),
            np_to_Cptr1iL(empty_i), np_to_Cptr1d(empty_f), np_to_Cptr1d(empty_f))
        if opt < 0:
            raise _lib.TomoError(f"TomoSh_HIP failed ({opt}): {_lib.last_error()}")
        self.Cdll_opt_id = int(opt)
        Mss3D = np.array(Cptr1d_to_np(self.CppDLL.getMss(), self.YPR.shape[0])).astype(np.float32, copy=True)
        return self._tomo_object(c_o, T, Mss3D, bVerbose)

    def run_group(self, cutobjects, bVerbose):
        """run_dll for every part of a cut group, with all parts' sweeps in one captured graph
        launch (tomo_batch_*).  Per part, the optimum's render groups and placement matrix come
        from a one-orientation TomoSh_HIP call at that orientation, which gives what the full call
        gives for it (orientations are independent).  Returns one TomoObject per part."""
        work = []
        B = _lib.Batch()
        try:
            for c0 in cutobjects:
                c_o = deepcopy(c0)
                T = self._mesh_of(c_o)
                self.bShellMesh = self.slicing_option.bShellMesh and (not T.is_watertight)
                vtx, tri = self.pack(T, bVerbose)
                vtx_nrm = np.asarray(T.vertex_normals).astype(np.float32, copy=True)
                B.add(self.Float32Info, self.Int32Info, 0, tri, vtx, vtx_nrm, self.YPR)
                work.append((c_o, T, vtx, tri, vtx_nrm, self.Float32Info.copy(), self.Int32Info.copy()))
            B.launch()
            B.sync()
            mss = [B.records(m)["Mss"].astype(np.float32) for m in range(len(work))]
        finally:
            B.close()
        self.Mss3D_group = mss  # per part, in cutobjects order
        out = []
        empty_i = np.zeros(3, np.int32)
        empty_f = np.zeros(3, np.float32)
        for (c_o, T, vtx, tri, vtx_nrm, f32, i32), Mss3D in zip(work, mss):
            if bVerbose:
                print("Tomo processing: ", getattr(c_o, "name", "mesh"))
            opt = _lib.find_first_optimal(Mss3D)
            one = np.ascontiguousarray(self.YPR[opt:opt + 1])
            i1 = i32.copy()
            i1[6] = 1
            rc = self.CppFunction(
                np_to_Cptr1d(f32), np_to_Cptr1iL(i1), np_to_Cptr1d(one), np_to_Cptr1iL(tri), np_to_Cptr1d(vtx),
                np_to_Cptr1d(vtx_nrm), np_to_Cptr1d(np.zeros(3 * len(tri), np.float32)),
                np_to_Cptr1iL(empty_i), np_to_Cptr1d(empty_f), np_to_Cptr1d(empty_f))
            if rc < 0:
                raise _lib.TomoError(f"TomoSh_HIP failed ({rc}): {_lib.last_error()}")
            self.Float32Info, self.Int32Info = f32, i32
            self.Cdll_opt_id = int(opt)
            out.append(self._tomo_object(c_o, T, Mss3D, bVerbose))
        return out

    def _tomo_object(self, c_o, T, Mss3D, bVerbose):
        """TomoObject from per-orientation Mss and the optimum's getters (tomo_dll_interface.py:99-135)."""
        L = self.CppDLL
        t_o = TomoObject(c_o)
        if os.environ.get("TOMO_DEBUG_MSS", "0").lower() in ("1", "true", "yes", "on"):
            best_ids = np.argsort(Mss3D)[:8]
            print("TOMO_DEBUG_MSS", "min", float(np.min(Mss3D)), "max", float(np.max(Mss3D)),
                  "sum", float(np.sum(Mss3D)), "best_ids", best_ids.tolist(),
                  "best_vals", [float(Mss3D[i]) for i in best_ids])
        self.Mss3D = Mss3D
        (t_o.best_vminfo, _) = findOptimals(self.YPR, Mss3D, bVerbose)
        t_o.best_vminfo.set_max_length(T)
        for p_type in (enumPixelType.eptSS, enumPixelType.eptBed):
            p_name = TOMO_PixelVarNames[p_type]
            n_2i = L.getnData2i(int(p_type))
            if n_2i > 0:
                p_2i = L.getpData2i(int(p_type))
                pxls = Cptr1i_to_np(p_2i, n_2i * TOMO_nPixelFormat).reshape(n_2i, TOMO_nPixelFormat)
                pxls = np.array(pxls).astype(float)[:, 0:3]
                t_o.pxl_groups.append(TomoPixels(p_type, p_name, pxls))
        t_o.mat4x4 = np.array(Cptr1d_to_np(L.getMat4x4(), 4 * 4)).astype(np.float32, copy=True).reshape(4, 4)
        return t_o
