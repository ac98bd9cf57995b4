"""CPU parity checker for the HIP CVL backend (test infrastructure only)."""
