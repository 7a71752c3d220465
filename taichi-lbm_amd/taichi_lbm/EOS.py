"""Pseudopotentials -- drop-in for taichi_lbm/EOS.py.

SC_psi (EOS.py:40-47): psi = rho0 (1 - exp(-rho / rho0)).
Pseudopotential (EOS.py:5-37) with a cubic equation of state p(rho, T):
psi = sqrt(|6/G (p(rho, T) - rho/3)|), T the node's temperature
(lb_field.T); vdW_psi (:50-58), RK_psi (:61-69), PR_psi (:72-83), CS_psi
(:86-95).  The HIP kernels evaluate the same expressions (psi kind and the
parameters travel in lbm_lattice); get_psi here is the host-side form for
initial fields.
"""
from __future__ import annotations

import numpy as np

from . import _hip as H

f32 = np.float32


class SC_psi:
    """psi = rho0 * (1 - exp(-rho / rho0)), EOS.py:40-47."""

    KIND = H.PSI_SC

    def __init__(self, rho0):
        self.rho0 = rho0

    def get_psi(self, dens, T=None):
        d = np.asarray(dens, np.float32)
        r0 = np.float32(self.rho0)
        return r0 * (np.float32(1.0) - np.exp(-d / r0))

    def fill(self, L, c):
        L.psi_kind[c] = H.PSI_SC
        L.psi_rho0[c] = float(f32(self.rho0))


class Pseudopotential:
    """EOS.py:5-37: params a, b, R (default 1), G (default 1), rho_cr,
    pressure_cr, temperature_cr (PR_psi also omega)."""

    KIND = None

    def __init__(self, params: dict):
        default_params = {'a': None, 'b': None, 'R': 1.0, 'G': 1.0}
        params = {**default_params, **params}
        self.a = params['a']
        self.b = params['b']
        self.R = params['R']
        self.G = params['G']
        self.rho_cr = params['rho_cr']
        self.pressure_cr = params['pressure_cr']
        self.temperature_cr = params['temperature_cr']
        self.omega = params.get('omega', 0.0)
        self.show_info()

    def show_info(self):
        print("=" * 20)
        print("EOS_pars")
        print(f"  a: {self.a}, b={self.b}")
        print(f"  rho_cr: {self.rho_cr}, p_cr={self.pressure_cr},T_cr={self.temperature_cr}")

    def EOS(self, dens, T):
        raise NotImplementedError

    def get_psi(self, dens, T):
        d = np.asarray(dens, np.float32)
        T = np.asarray(T, np.float32)
        p = self.EOS(d, T)
        return np.sqrt(np.abs(f32(6.0) / f32(self.G) * (p - d / f32(3.0)))).astype(np.float32)

    def fill(self, L, c):
        L.psi_kind[c] = self.KIND
        L.psi_rho0[c] = 0.0
        for q, v in enumerate((self.a, self.b, self.R, self.G, self.omega, self.temperature_cr or 1.0)):
            L.eos[c][q] = float(f32(v))


def _rep(d, T, R, b):
    return d * f32(R) * T / (f32(1.0) - f32(b) * d)


class vdW_psi(Pseudopotential):
    """van der Waals, EOS.py:50-58."""

    KIND = H.PSI_VDW

    def EOS(self, dens, T):
        return _rep(dens, T, self.R, self.b) - f32(self.a) * dens ** 2


class RK_psi(Pseudopotential):
    """Redlich-Kwong, EOS.py:61-69."""

    KIND = H.PSI_RK

    def EOS(self, dens, T):
        return _rep(dens, T, self.R, self.b) - f32(self.a) * dens ** 2 / np.sqrt(T) / (f32(1.0) + f32(self.b) * dens)


class PR_psi(Pseudopotential):
    """Peng-Robinson, EOS.py:72-83."""

    KIND = H.PSI_PR

    def __init__(self, params: dict):
        super().__init__(params)
        self.omega = params['omega']

    def EOS(self, dens, T):
        w = f32(self.omega)
        k = f32(1.0) + (f32(0.37464) + f32(1.54226) * w - f32(0.269922) * w * w) * \
            (f32(1.0) - np.sqrt(T / f32(self.temperature_cr)))
        bd = f32(self.b) * dens
        return _rep(dens, T, self.R, self.b) - f32(self.a) * (k * k) * dens ** 2 / (f32(1.0) + f32(2.0) * bd - bd * bd)


class CS_psi(Pseudopotential):
    """Carnahan-Starling, EOS.py:86-95."""

    KIND = H.PSI_CS

    def EOS(self, dens, T):
        x = f32(self.b) * dens / f32(4.0)
        return dens * f32(self.R) * T * (f32(1.0) + x + x * x - x ** 3) / (f32(1.0) - x) ** 3 - f32(self.a) * dens ** 2
