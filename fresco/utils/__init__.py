"""fresco.utils: the structure-factor report of the reference's utils
(export_structure, utils.py:16-52; radial_profile, utils.py:138-152) on the
GPU.  The reference's other helpers (lattices, drawing, pickling, logging,
point-pattern statistics) are not part of this package."""
from fresco_amd.utils import export_structure, radial_profile  # noqa: F401

__all__ = ["export_structure", "radial_profile"]
