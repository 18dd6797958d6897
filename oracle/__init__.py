"""CPU oracle of the step path (test infrastructure)."""
