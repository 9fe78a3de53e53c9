"""Test infrastructure only: the CPU oracle (see oracle/oracle.py)."""
