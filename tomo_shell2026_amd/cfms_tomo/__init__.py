"""Drop-in cfms_tomo API (mirrors cfms_tomo/__init__.py of the reference) over libtomo_hip.so."""
from .tomo_dll_interface import TomoDLLInterface
from .tomo_object import TomoObject
from .tomo_result import TomoResult
from .tomo_slicingOptions import SlicingOptions, enumBedType, enumPixelType

g_bVerbose = True
g_bUseCUDA = False
g_bShellMesh = True
g_tomo_theta_YP = 10


class TomoManager:
    """Manage the tomography library and per-part results (cfms_tomo/__init__.py:17-42)."""

    def __init__(self, s_o: SlicingOptions):
        self.slicing_option = s_o
        self.CppDLL = TomoDLLInterface("libtomo_hip.so", s_o, s_o.bUseCUDA, s_o.bShellMesh)

    def get_best_orientation(self, C_G, batched=True):
        """C_G: an object with .name and .cutobjects (each with .tmesh and .name).  batched: all
        parts' sweeps in one graph launch (run_group); otherwise one run_dll per part as the
        reference does (cfms_tomo/__init__.py:26-30).  Both give the same TomoObjects."""
        self.name = f"{getattr(C_G, 'name', 'group')}_tomo"
        self.input_gr = C_G
        if batched:
            for c_o, t_o in zip(C_G.cutobjects, self.CppDLL.run_group(C_G.cutobjects, self.slicing_option.bVerbose)):
                c_o.tomo_object = t_o
            return
        for c_o in C_G.cutobjects:
            c_o.tomo_object = self.CppDLL.run_dll(c_o, self.slicing_option.bVerbose)

    def get_shells(self, offset_size):
        C_G = self.input_gr
        unrotated = C_G.get_shells(offset_size)
        rotated = C_G.get_shells(offset_size)
        unrotated.name = f"{self.name}_unrotated"
        rotated.name = f"{self.name}_rotated"
        for c_o in rotated.cutobjects:
            c_o.transform(c_o.tomo_object.mat4x4)
        return unrotated, rotated


__all__ = ["TomoManager", "TomoDLLInterface", "TomoObject", "TomoResult", "SlicingOptions", "enumBedType",
           "enumPixelType"]
