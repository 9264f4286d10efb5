"""CPU oracle of the orthonormalization path (test infrastructure only)."""
