"""sph07_amd -- MI355X-native per-timestep particle loop of the sph07 SPH solver."""
from .deck import Deck, Material  # noqa: F401
