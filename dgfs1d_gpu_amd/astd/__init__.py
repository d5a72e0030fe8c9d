"""Asymptotic-preserving (IMEX) single-species solver path (reference dgfs1D/astd)."""
