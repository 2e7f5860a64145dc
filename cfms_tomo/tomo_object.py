"""cfms_tomo.tomo_object of the reference: the drop-in module tomo_shell2026_amd.cfms_tomo.tomo_object."""
import sys

from tomo_shell2026_amd.cfms_tomo import tomo_object as _module

sys.modules[__name__] = _module
