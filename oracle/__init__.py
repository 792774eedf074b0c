"""CPU FP64 oracle of sph07's step loop -- test infrastructure only."""
