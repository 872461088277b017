"""CPU oracle of the laser-polio daily step (test infrastructure only)."""
