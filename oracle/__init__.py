"""CPU oracle of the Padenti hot path (test infrastructure only; see oracle.c)."""
