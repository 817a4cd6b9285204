"""CPU parity oracle (test infrastructure only)."""
