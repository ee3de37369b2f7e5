"""fresco.distances: the distance-method enum (distance_enum.py:4-6) taken by
InversePower and WCAPotential."""
from fresco_amd.distances import Distance  # noqa: F401

__all__ = ["Distance"]
