"""cfms_tomo.tomo_slicingOptions of the reference: the drop-in module tomo_shell2026_amd.cfms_tomo.tomo_slicingOptions."""
import sys

from tomo_shell2026_amd.cfms_tomo import tomo_slicingOptions as _module

sys.modules[__name__] = _module
