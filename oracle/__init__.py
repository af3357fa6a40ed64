"""CPU oracle of the MKF hot path: test infrastructure only (see oracle.py)."""
