"""The validation-harness caller of the reference (cfms_tomo/shell_test/)."""
