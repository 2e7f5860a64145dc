"""TomoObject (mirrors cfms_tomo/tomo_object.py of the reference, minus the polyscope
rendering hook that depends on the mesh-cutting package)."""
from copy import copy

import numpy as np

from .tomo_result import TomoResult


class TomoObject:
    def __init__(self, cutobject, pxl_groups=None, best=None, colorspan=None, mat4x4=None):
        self.raw_cutobject = cutobject
        self.pxl_groups = []
        if pxl_groups:
            for pxls in pxl_groups:
                self.pxl_groups.append(pxls.copy())
        self.best_vminfo = copy(best) if best else TomoResult()
        self.colorspan = copy(colorspan) if colorspan else None
        self.mat4x4 = mat4x4.copy() if mat4x4 is not None else None

    def translate_by(self, vec):
        for p_g in self.pxl_groups:
            p_g.pxls = np.add(np.array(p_g.pxls), vec)

    def copy(self):
        return TomoObject(self.raw_cutobject, self.pxl_groups, self.best_vminfo, self.colorspan, self.mat4x4)
