"""FP64 CPU oracle of the reference's algorithm (test infrastructure only)."""
