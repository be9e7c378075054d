"""CPU oracle for the cut-evaluation / local-search path (test infrastructure only)."""
