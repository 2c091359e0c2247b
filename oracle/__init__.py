"""CPU FP64 oracle for the 2-D RBC step (test infrastructure only; see oracle.py)."""
