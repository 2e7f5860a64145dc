"""cfms_tomo.tomo_result of the reference: the drop-in module tomo_shell2026_amd.cfms_tomo.tomo_result."""
import sys

from tomo_shell2026_amd.cfms_tomo import tomo_result as _module

sys.modules[__name__] = _module
