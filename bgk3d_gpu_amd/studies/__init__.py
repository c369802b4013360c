"""Measurement studies: storage precision on a Poiseuille channel, flow past a sphere (one GPU)."""
