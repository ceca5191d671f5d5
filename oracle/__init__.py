"""CPU oracle for the LCS hot path -- test infrastructure only (see oracle.py)."""
