"""CPU oracle for the GROUP BY hot path (test infrastructure only)."""
