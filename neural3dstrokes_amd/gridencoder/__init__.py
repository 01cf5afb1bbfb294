"""Drop-in for the reference's ``source.gridencoder`` package."""
from .grid import GridEncoder, grid_encode, _grid_encode  # noqa: F401
