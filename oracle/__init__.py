"""CPU oracle for the K-lane PHAST path — test infrastructure only (see oracle.py)."""
