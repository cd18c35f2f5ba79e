"""CPU oracle for the gammaray hot path (test infrastructure only)."""
