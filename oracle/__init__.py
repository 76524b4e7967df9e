"""CPU oracle of the explicit Moelans phase-field step (test infrastructure only)."""
