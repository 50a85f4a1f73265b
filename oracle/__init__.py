"""CPU oracle of the reference render path (test infrastructure only)."""
