"""cfms_tomo.shell_test of the reference: the validation-harness caller."""
