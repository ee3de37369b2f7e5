"""The import path of martiniani-lab/FReSCo (``fresco.potentials``,
``fresco.optimize``, ``fresco.utils``, ``fresco.distances``) over the
MI355X implementation in ``fresco_amd``, so that scripts and notebooks
written for the reference (``from fresco.potentials import NUwU``,
``from fresco.optimize import LBFGS_CPP``,
``from fresco.utils import export_structure``) run unchanged.

Only the structure-factor path and what drives it are provided; the
reference's plotting, lattice, file and point-statistics helpers are not.
"""
from fresco_amd import __version__  # noqa: F401
