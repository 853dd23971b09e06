"""CPU oracle of the SolverFT step: test infrastructure, never the product."""
