"""cfms_tomo.tomo_io of the reference: the drop-in module tomo_shell2026_amd.cfms_tomo.tomo_io."""
import sys

from tomo_shell2026_amd.cfms_tomo import tomo_io as _module

sys.modules[__name__] = _module
