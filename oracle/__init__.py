"""CPU oracle of the P3T hot path (test infrastructure only)."""
