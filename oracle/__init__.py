"""CPU oracle for the dense-contraction path (test infrastructure only)."""
