"""Fluid models -- drop-in for taichi_lbm/FluidModels.py of the reference.

The kernels evaluate nu_app = m * shear_rate^(n-1); a Newtonian fluid is the
case (m, n) = (nu, 1), which gives nu exactly (pow(x, 0) = 1 for every x).
"""
from __future__ import annotations

import numpy as np


class Fluid:
    def apparent_viscosity(self, shear_rate, rho):
        raise NotImplementedError

    def visc_params(self):
        """(m, n - 1) as the fp32 values the kernels use."""
        raise NotImplementedError


class NewtonianFluid(Fluid):
    """FluidModels.py:9-16."""

    def __init__(self, nu):
        self.nu = nu

    def apparent_viscosity(self, shear_rate, rho):
        return np.full_like(np.asarray(shear_rate, np.float32), np.float32(self.nu))

    def visc_params(self):
        return float(np.float32(self.nu)), 0.0


class PowerLawFluid(Fluid):
    """FluidModels.py:20-29: nu = m * shear_rate**(n - 1)."""

    def __init__(self, m, n):
        self.m = m
        self.n = n

    def apparent_viscosity(self, shear_rate, rho):
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            return np.float32(self.m) * np.power(np.asarray(shear_rate, np.float32), np.float32(self.n - 1.0))

    def visc_params(self):
        return float(np.float32(self.m)), float(np.float32(self.n - 1.0))
