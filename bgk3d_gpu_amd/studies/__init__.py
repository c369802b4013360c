"""Measurement studies of the BASELINE configurations (run on one GPU)."""
