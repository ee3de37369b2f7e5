"""CPU oracle for the structure-factor energy/gradient path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from this package, and
only as the checker (or as the timed CPU baseline).  The product path in
``fresco_amd`` never imports it and fails loudly when its HIP library is absent.

Contents
--------
sf_direct    exact FP64 direct sums (NUDFT) restating the four potentials of
             source/fresco/{nuwu,nuwnu,uwu,uwnu}.hpp, plus the coordinate map,
             mode grid and weight conventions of those files.
optim        line-for-line FP64 restatement of source/lbfgs.cpp and
             source/fresco/modified_fire.hpp (the optimizers that drive the path).
pair         brute-force FP64 restatement of the inverse-power / WCA pair
             energies of source/fresco/{inversepower_potential,wca}.hpp.
nufft_cpu    C/OpenMP exponential-of-semicircle NUFFT (type 1 / type 2) used as
             the CPU restatement of the reference path at sizes where a direct
             sum is too slow (the CPU baseline of the benchmark).
_ref/        (built, git-ignored) the reference's own optimizer and pair-potential
             sources compiled by ``oracle/Makefile``.

Parity pinning of each function is listed in the module docstrings.
"""
