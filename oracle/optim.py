"""FP64 restatement of the reference optimizers that drive the path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

LBFGS            source/lbfgs.cpp:9-257, source/fresco/lbfgs.hpp, driver
                 source/fresco/optimizer.hpp:55-241
ModifiedFire     source/fresco/modified_fire.hpp:44-230, source/modified_fire.cpp

`potential` is any object with get_energy_gradient(x) -> (E, grad).

Pinning: tests/test_oracle_optim.py runs these against the reference's own
lbfgs.cpp / modified_fire.cpp compiled from its sources (oracle/_ref) on the
same callback potentials and compares every iterate, plus the analytic
minimum of a harmonic potential (cpp_tests/source/test_base_potential.cpp).
"""
from __future__ import annotations

import math

import numpy as np


class LBFGS:
    def __init__(self, x0, potential, tol=1e-4, M=4):
        self.pot = potential
        self.tol = tol
        self.maxstep = 0.1
        self.maxiter = 1000
        self.iter = 0
        self.nfev = 0
        self.x = np.array(x0, dtype=np.float64).copy()
        self.f = 0.0
        self.g = np.zeros_like(self.x)
        self.rms = 1e10
        self.initialized = False
        # lbfgs.cpp:9-31
        self.M = M
        self.max_f_rise = 1e-4
        self.use_relative_f = False
        self.rho = np.zeros(M)
        self.H0 = 0.1
        self.k = 0
        self.alpha = np.zeros(M)
        n = len(self.x)
        self.inv_sqrt_size = 1.0 / math.sqrt(n)
        self.s = np.zeros((M, n))
        self.y = np.zeros((M, n))

    # optimizer.hpp:208-231
    def compute_func_gradient(self, x):
        self.nfev += 1
        e, g = self.pot.get_energy_gradient(x)
        return float(e), np.array(g, dtype=np.float64)

    def initialize_func_gradient(self):
        self.f, self.g = self.compute_func_gradient(self.x)
        self.rms = math.sqrt(float(np.dot(self.g, self.g)) / len(self.x))
        self.initialized = True

    def set_func_gradient(self, f, grad):
        self.f = float(f)
        self.g = np.array(grad, dtype=np.float64)
        self.rms = math.sqrt(float(np.dot(self.g, self.g)) / len(self.g))
        self.initialized = True

    def stop_criterion_satisfied(self):
        if not self.initialized:
            self.initialize_func_gradient()
        return self.rms <= self.tol

    def run(self, niter=None):
        if niter is None:
            niter = self.maxiter - self.iter
        if not self.initialized:
            self.initialize_func_gradient()
        for _ in range(niter):
            if self.stop_criterion_satisfied():
                break
            self.one_iteration()

    # lbfgs.cpp:36-64
    def one_iteration(self):
        if not self.initialized:
            self.initialize_func_gradient()
        xold = self.x.copy()
        gold = self.g.copy()
        step = self.compute_lbfgs_step()
        self.stepnorm = self.backtracking_linesearch(step, xold)
        self.update_memory(xold, gold, self.x, self.g)
        self.iter += 1

    # lbfgs.cpp:66-105
    def update_memory(self, xold, gold, xnew, gnew):
        kl = self.k % self.M
        self.y[kl] = gnew - gold
        self.s[kl] = xnew - xold
        ys = float(np.dot(self.y[kl], self.s[kl]))
        yy = float(np.dot(self.y[kl], self.y[kl]))
        if ys == 0.0:
            ys = 1.0
        self.rho[kl] = 1.0 / ys
        if yy == 0.0:
            yy = 1.0
        self.H0 = ys / yy
        self.k += 1

    # lbfgs.cpp:107-170
    def compute_lbfgs_step(self):
        if self.k == 0:
            gnorm = math.sqrt(float(np.dot(self.g, self.g)))
            if gnorm > 1.0:
                gnorm = 1.0 / gnorm
            return (-gnorm * self.H0) * self.g
        step = self.g.copy()
        jmin = max(0, self.k - self.M)
        jmax = self.k
        self.alpha[:] = 0.0
        for j in range(jmax - 1, jmin - 1, -1):
            i = j % self.M
            a = self.rho[i] * float(np.dot(self.s[i], step))
            step -= a * self.y[i]
            self.alpha[i] = a
        step *= -self.H0
        for j in range(jmin, jmax):
            i = j % self.M
            beta = -self.rho[i] * float(np.dot(self.y[i], step))
            ab = self.alpha[i] - beta
            step -= self.s[i] * ab
        return step

    # lbfgs.cpp:172-245
    def backtracking_linesearch(self, step, xold):
        if float(np.dot(step, self.g)) > 0.0:
            step = -step
        factor = 1.0
        stepnorm = math.sqrt(float(np.dot(step, step)))
        if factor * stepnorm > self.maxstep:
            factor = self.maxstep / stepnorm
        fnew = self.f
        for _ in range(10):
            self.x = xold + factor * step
            fnew, self.g = self.compute_func_gradient(self.x)
            df = fnew - self.f
            if self.use_relative_f:
                absf = abs(self.f) if self.f != 0 else 1e-100
                df /= absf
            if df < self.max_f_rise:
                break
            factor *= 0.5
        self.f = fnew
        self.rms = math.sqrt(float(np.dot(self.g, self.g))) * self.inv_sqrt_size
        return stepnorm * factor

    # lbfgs.cpp:247-257
    def reset(self, x0):
        if len(x0) != len(self.x):
            raise ValueError("The number of degrees of freedom (x0.size()) cannot change when calling reset()")
        self.k = 0
        self.iter = 0
        self.nfev = 0
        self.x = np.array(x0, dtype=np.float64).copy()
        self.initialize_func_gradient()


class ModifiedFire:
    """modified_fire.cpp:43-118 and modified_fire.hpp:84-230."""

    def __init__(self, x0, potential, dtstart=0.1, dtmax=1.0, maxstep=0.5, Nmin=5, finc=1.1, fdec=0.5, fa=0.99,
                 astart=0.1, tol=1e-4, stepback=True):
        self.pot = potential
        self.tol = tol
        self.maxiter = 1000
        self.iter = 0
        self.nfev = 0
        self.x = np.array(x0, dtype=np.float64).copy()
        self.f = 0.0
        self.g = np.zeros_like(self.x)
        self.rms = 1e10
        self.initialized = False
        self.dtstart = dtstart
        self.dt = dtstart
        self.dtmax = dtmax
        self.maxstep = maxstep
        self.Nmin = Nmin
        self.finc = finc
        self.fdec = fdec
        self.fa = fa
        self.astart = astart
        self.a = astart
        self.fold = self.f
        self.ifnorm = 0.0
        self.vnorm = 0.0
        self.v = np.zeros_like(self.x)
        self.xold = self.x.copy()
        self.gold = self.g.copy()
        self.fire_iter = 0
        self.N = len(self.x)
        self.stepback = stepback

    def initialize_func_gradient(self):
        self.nfev += 1
        e, g = self.pot.get_energy_gradient(self.x)
        self.f = float(e)
        self.g = np.array(g, dtype=np.float64)
        self.fold = self.f
        self.v = -self.g * self.dt
        gnorm = float(np.dot(self.g, self.g))
        self.vnorm = math.sqrt(float(np.dot(self.v, self.v)))
        self.ifnorm = 1.0 / math.sqrt(gnorm)
        self.rms = 1.0 / (self.ifnorm * math.sqrt(self.N))
        self.initialized = True

    def stop_criterion_satisfied(self):
        if not self.initialized:
            self.initialize_func_gradient()
        return self.rms <= self.tol

    def run(self, niter=None):
        if niter is None:
            niter = self.maxiter - self.iter
        if not self.initialized:
            self.initialize_func_gradient()
        for _ in range(niter):
            if self.stop_criterion_satisfied():
                break
            self.one_iteration()

    def _forward_euler(self):
        self.v = self.v - self.dt * self.g
        dx = self.dt * self.v
        self.gold = self.g.copy()
        normdx = math.sqrt(float(np.dot(dx, dx)))
        factor = 1.0
        if normdx > self.maxstep:
            factor = self.maxstep / normdx
        self.x = self.x + dx * factor
        e, g = self.pot.get_energy_gradient(self.x)
        self.f = float(e)
        self.g = np.array(g, dtype=np.float64)

    def one_iteration(self):
        self.nfev += 1
        self.iter += 1
        self.fire_iter += 1
        self.fold = self.f
        self.xold = self.x.copy()
        self.v = (1.0 - self.a) * self.v - self.a * self.g * self.ifnorm * self.vnorm
        self._forward_euler()
        P = -float(np.dot(self.v, self.g))
        if P > 0:
            if self.fire_iter > self.Nmin:
                self.dt = min(self.dt * self.finc, self.dtmax)
                self.a *= self.fa
            gnorm = float(np.dot(self.g, self.g))
            self.vnorm = math.sqrt(float(np.dot(self.v, self.v)))
            self.ifnorm = 1.0 / math.sqrt(gnorm)
            self.rms = 1.0 / (self.ifnorm * math.sqrt(self.N))
        else:
            self.dt *= self.fdec
            self.a = self.astart
            self.fire_iter = 0
            self.v[:] = 0.0
            if self.stepback:
                self.f = self.fold
                self.x = self.xold.copy()
                self.g = self.gold.copy()
