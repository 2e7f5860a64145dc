"""TomoResult (mirrors cfms_tomo/tomo_result.py of the reference)."""
from collections import namedtuple

import numpy as np

from .tomo_io import FStr

VolMassData = namedtuple("VolMassData", "yaw pitch roll Mo Mss Mtotal")


class TomoResult:
    vmData = VolMassData(0, 0, 0, 0, 0, 0)

    def __init__(self, yprm6f=None):
        if yprm6f is not None:
            self.vmData = VolMassData(yaw=yprm6f[0], pitch=yprm6f[1], roll=yprm6f[2],
                                      Mo=yprm6f[3], Mss=yprm6f[4], Mtotal=yprm6f[5])
        self.mat4x4 = None

    def get_ypr_in_degree(self):
        return (self.vmData.yaw * 180 / np.pi, self.vmData.pitch * 180 / np.pi, self.vmData.roll * 180 / np.pi)

    def get_ypr_in_radian(self):
        return (self.vmData.yaw, self.vmData.pitch, self.vmData.roll)

    def set_max_length(self, tmesh):
        """Longest edge of an oriented bounding box.  With a trimesh object its
        bounding_box_oriented is used (tomo_result.py:26-27); otherwise the box in the mesh's
        principal axes (not the minimum-volume box)."""
        obb = getattr(tmesh, "bounding_box_oriented", None)
        if obb is not None:
            self.BeamLength = np.max(obb.extents, axis=0)
            return
        v = np.asarray(tmesh.vertices, np.float64)
        c = v - v.mean(0)
        _, _, vt = np.linalg.svd(c, full_matrices=False)
        p = c @ vt.T
        self.BeamLength = float(np.max(p.max(0) - p.min(0)))

    def get_max_length(self):
        return self.BeamLength

    def GetMss(self):
        return self.vmData[4]

    def print(self):
        print(FStr(self.vmData))
