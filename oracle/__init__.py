"""CPU oracle of the sparse-DNN inference path (test infrastructure only)."""
