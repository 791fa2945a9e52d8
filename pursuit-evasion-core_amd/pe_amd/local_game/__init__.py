"""Finite-body contact game (8D semi-Lagrangian min-max backup)."""
