"""cfms_tomo.tomo_dll_interface of the reference: the drop-in module tomo_shell2026_amd.cfms_tomo.tomo_dll_interface."""
import sys

from tomo_shell2026_amd.cfms_tomo import tomo_dll_interface as _module

sys.modules[__name__] = _module
