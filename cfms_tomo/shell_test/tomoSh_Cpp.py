"""cfms_tomo.shell_test.tomoSh_Cpp of the reference: the drop-in module
tomo_shell2026_amd.cfms_tomo.shell_test.tomoSh_Cpp."""
import sys

from tomo_shell2026_amd.cfms_tomo.shell_test import tomoSh_Cpp as _module

sys.modules[__name__] = _module
