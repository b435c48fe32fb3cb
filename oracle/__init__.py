"""FP64 CPU oracle of the LUT simulation chain (test infrastructure only; see oracle.py)."""
