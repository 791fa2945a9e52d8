"""Reversible-car free-space game (5D relative HJI)."""
