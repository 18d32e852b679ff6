"""CPU oracle of the SPH step (test infrastructure only; see oracle.py)."""
