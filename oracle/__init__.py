"""CPU oracle for the LLG hot path (test infrastructure only)."""
