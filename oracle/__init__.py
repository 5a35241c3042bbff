"""CPU oracle of the stochastic splatting renderer: test infrastructure only (see oracle.py)."""
