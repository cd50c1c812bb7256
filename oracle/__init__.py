"""TEST INFRASTRUCTURE: the CPU parity oracle (see oracle.py)."""
