"""fresco_amd: MI355X-native structure-factor energy/gradient path.

Drop-in for the potentials (fresco.potentials.NUwU, ...) and optimizers
(fresco.optimize.LBFGS_CPP, ModifiedFireCPP) of martiniani-lab/FReSCo, with
the arithmetic in hand-written gfx950 HIP kernels behind a C ABI
(include/fresco_hip.h, libfresco_hip.so).
"""
__version__ = "0.1.0"
