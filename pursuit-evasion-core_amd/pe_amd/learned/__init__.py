"""Learned value surrogate (BF16-MFMA MLP) for warm starts and control ordering."""
