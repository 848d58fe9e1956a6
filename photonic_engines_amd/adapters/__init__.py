"""Adapters of the SLOS engine to other front ends (the strong-simulation backend face)."""
