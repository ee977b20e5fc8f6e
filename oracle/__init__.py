"""CPU oracle of the reference projector's method (test infrastructure only)."""
