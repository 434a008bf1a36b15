"""CPU oracle of the Monte Carlo road-evaluation path (test infrastructure)."""
