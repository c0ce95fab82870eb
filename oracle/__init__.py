"""CPU oracle of the 3DGS training step (test infrastructure: checker and CPU baseline only)."""
