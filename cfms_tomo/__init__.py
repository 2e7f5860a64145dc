"""The reference's module path for the drop-in API: `from cfms_tomo import TomoManager`,
`cfms_tomo.tomo_object`, `cfms_tomo.tomo_slicingOptions`, `cfms_tomo.tomo_result`,
`cfms_tomo.tomo_io`, `cfms_tomo.tomo_dll_interface` and `cfms_tomo.shell_test.tomoSh_Cpp`
(cfms_tomo/__init__.py, tomo_*.py, shell_test/tomoSh_Cpp.py of the reference) resolve to the
modules of tomo_shell2026_amd.cfms_tomo, which run on libtomo_hip.so."""
from tomo_shell2026_amd.cfms_tomo import *  # noqa: F401,F403
from tomo_shell2026_amd.cfms_tomo import (TomoDLLInterface, TomoManager, TomoObject, TomoResult,  # noqa: F401
                                          SlicingOptions, enumBedType, enumPixelType)
