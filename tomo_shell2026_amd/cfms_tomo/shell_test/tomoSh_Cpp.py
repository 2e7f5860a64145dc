"""tomoSh_Cpp: the validation harness's direct caller of the library, as
cfms_tomo/shell_test/tomoSh_Cpp.py defines it (used by validate_primitives.py:38-50 and
TSE_TomoSh1.py:74): load a mesh file, centre it, pack Float32Info / Int32Info in the same order
(:54-67), call the ten-pointer entry point and read Mo, Mss, the optimum's VolMassInfo and, with
bVerbose, the pixel groups (:81-111).

Every entry-point name the reference accepts (TomoSh_INT3, TomoSh_CUDA) runs TomoSh_HIP of
libtomo_hip.so.  Meshes load without trimesh (mesh.load_mesh: STL, OBJ); a mesh object with
.vertices and .faces, or a (vertices, faces) pair, may be passed in place of the file name.
The legacy TomoSh_TMPxl engine is not part of the library: asking for it raises ValueError.
"""
import ctypes as ct

import numpy as np

from ... import _lib
from ...mesh import TomoMesh, face_normals, load_mesh
from ..tomo_io import Cptr1d, Cptr1d_to_np, Cptr1i, Cptr1i_to_np, np_to_Cptr1d, np_to_Cptr1iL, toRadian
from ..tomo_slicingOptions import TOMO_nPixelFormat, enumBedType, enumPixelType

g_PixelEnums = ("eptAl", "eptBe", "eptSSB", "eptSSA", "eptSS", "eptBed", "eptVo", "eptVss", "eptTC",
                "eptNVB", "eptNVA")
g_PixelVarNames = ("al_pxls", "be_pxls", "SSB_pxls", "SSA_pxls", "SS_pxls", "Bed_pxls", "Vo_pxls",
                   "Vss_pxls", "TC_pxls", "NVB_pxls", "NVA_pxls")
ENTRY_POINTS = ("TomoSh_INT3", "TomoSh_CUDA", "TomoSh_HIP")


def LoadInputMesh(source):
    """The tuple LoadInputMesh returns (tomoSh_io_trimesh.py:530-558), vertices centred on their
    bounding box: (mesh, tri, vtx, vtx_nrm, tri_nrm, tri_area, tri_center, chull_tri, chull_vtx,
    chull_trinrm).  The hull arrays are empty: the INT3 engine does not read them."""
    if isinstance(source, tuple):
        mesh = TomoMesh(source[0], source[1])
    elif hasattr(source, "vertices") and hasattr(source, "faces"):
        mesh = TomoMesh(np.asarray(source.vertices), np.asarray(source.faces))
    else:
        mesh = load_mesh(source)
    lo, hi = mesh.vertices.min(0), mesh.vertices.max(0)
    mesh.vertices = mesh.vertices - (lo + hi) * 0.5
    vtx = np.asarray(mesh.vertices).astype(np.float32)
    tri = np.asarray(mesh.faces).astype(np.int32)
    tri_nrm = face_normals(mesh.vertices, mesh.faces).astype(np.float32)
    vtx_nrm = np.asarray(mesh.vertex_normals).astype(np.float32)
    e = np.cross(vtx[tri[:, 1]] - vtx[tri[:, 0]], vtx[tri[:, 2]] - vtx[tri[:, 0]])
    tri_area = (0.5 * np.linalg.norm(e, axis=1)).astype(np.float32)
    tri_center = np.average(vtx[tri], axis=1).astype(np.float32)
    return (mesh, tri, vtx, vtx_nrm, tri_nrm, tri_area, tri_center, np.zeros((0, 3), np.int32),
            np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))


class tomoSh_Cpp:
    wall_thickness = 1.0
    PLA_density = 0.001121
    Fclad = 1.0
    Fcore = 0.15
    Fss = 0.2
    Css = 1.0
    bVerbose = False
    bUseExplicitSS = False
    bUseClosedVolumeVoxel = False
    bShellMesh = False
    ShellThickness = 0.5
    dVoxel = 1.0
    nVoxel = 256
    theta_c = toRadian(60.0)
    BedType = (enumBedType.ebtRaft, 0.0, 0.0, 0.0)

    def __init__(self, filename, nYPR_Intervals, yaw_range, pitch_range, roll_range, bVerbose):
        self.FileName = filename
        (self.mesh0, self.tri, self.vtx, self.vtx_nrm, self.tri_nrm, self.tri_area, self.tri_center,
         self.chull_tri, self.chull_vtx, self.chull_trinrm) = LoadInputMesh(filename)
        self.xx, self.yy, self.zz = np.meshgrid(yaw_range, pitch_range, roll_range)
        self.YPR = np.column_stack([self.xx.ravel(), self.yy.ravel(), self.zz.ravel()]).astype(np.float32)
        self.nYPR_Intervals = nYPR_Intervals
        self.bVerbose = bVerbose
        ext = self.vtx.max(0).astype(int) - self.vtx.min(0).astype(int)
        self.AABB2D = (0, 0, int(ext.max()), int(ext.max()))

    def Run(self, cpp_function_name):
        if cpp_function_name not in ENTRY_POINTS:
            raise ValueError(f"{cpp_function_name} is not provided by libtomo_hip.so "
                             f"(entry points: {', '.join(ENTRY_POINTS)})")
        self.Float32Info = np.array([
            self.dVoxel, self.theta_c, self.mesh0.area, self.wall_thickness,
            self.Fcore, self.Fclad, self.Fss, self.Css,
            self.PLA_density, self.BedType[1], self.BedType[2], self.BedType[3]]).astype(np.float32)
        self.Int32Info = np.array([
            self.bVerbose, self.bUseExplicitSS, self.bShellMesh, self.nVoxel,
            self.tri.shape[0], self.vtx.shape[0], self.YPR.shape[0],
            self.chull_tri.shape[0], self.chull_vtx.shape[0], self.BedType[0],
            int(round(self.ShellThickness * 1000.0))]).astype(np.int32)
        L = self.CppDLL = _lib.load()
        self.Cdll_opt_id = L.TomoSh_HIP(
            np_to_Cptr1d(self.Float32Info), np_to_Cptr1iL(self.Int32Info), np_to_Cptr1d(self.YPR),
            np_to_Cptr1iL(self.tri), np_to_Cptr1d(self.vtx), np_to_Cptr1d(self.vtx_nrm),
            np_to_Cptr1d(self.tri_nrm), np_to_Cptr1iL(np.ascontiguousarray(self.chull_tri)),
            np_to_Cptr1d(np.ascontiguousarray(self.chull_vtx)),
            np_to_Cptr1d(np.ascontiguousarray(self.chull_trinrm)))
        if self.Cdll_opt_id < 0:
            raise _lib.TomoError(f"TomoSh_HIP failed ({self.Cdll_opt_id}): {_lib.last_error()}")
        n = self.YPR.shape[0]
        get = lambda fn: ct.cast(fn(), Cptr1d)  # noqa: E731
        self.Mo3D = np.array(Cptr1d_to_np(get(L.getMo), n)).astype(np.float32)
        self.Mss3D = np.array(Cptr1d_to_np(get(L.getMss), n)).astype(np.float32)
        self.Mtotal3D = self.Mo3D + self.Mss3D
        self.vm_info = np.array(Cptr1d_to_np(get(L.getVolMassInfo), 21)).astype(np.float32)
        self.mat4x4 = np.array(Cptr1d_to_np(get(L.getMat4x4), 16)).astype(np.float32).reshape(4, 4)
        if self.bVerbose:
            for p_type, p_name in zip(g_PixelEnums, g_PixelVarNames):
                k = int(getattr(enumPixelType, p_type))
                n_2i = L.getnData2i(k)
                if n_2i == 0:
                    setattr(self, p_name, np.array([[0, 0, 0, 0, 0, 0]]))
                else:
                    p = ct.cast(L.getpData2i(k), Cptr1i)
                    setattr(self, p_name, Cptr1i_to_np(p, n_2i * TOMO_nPixelFormat).reshape(n_2i, TOMO_nPixelFormat))
        L.OnDestroy()
