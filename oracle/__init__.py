"""CPU oracle of the SVO ray pipeline (test infrastructure only)."""
