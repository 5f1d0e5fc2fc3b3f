"""CPU oracle (FP64 NumPy) of the co-kriging interpolator -- test infrastructure only."""
