"""Distance methods (mirror of fresco/distances/distance_enum.py)."""
from enum import Enum, unique


@unique
class Distance(Enum):
    CARTESIAN = 1
    PERIODIC = 2


__all__ = ["Distance"]
