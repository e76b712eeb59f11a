"""CPU oracle of the funlbm timestep (test infrastructure only)."""
