"""Synthetic meshes and the benchmark workloads (airfoil finite-volume loops,
P1 finite-element assembly) used by bench.py, the tests and the demos."""
