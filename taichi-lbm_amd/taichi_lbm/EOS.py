"""Pseudopotentials -- drop-in for taichi_lbm/EOS.py (SC_psi only, EOS.py:40-47)."""
from __future__ import annotations

import numpy as np


class SC_psi:
    """psi = rho0 * (1 - exp(-rho / rho0)), EOS.py:40-47."""

    def __init__(self, rho0):
        self.rho0 = rho0

    def get_psi(self, dens, T=None):
        d = np.asarray(dens, np.float32)
        r0 = np.float32(self.rho0)
        return r0 * (np.float32(1.0) - np.exp(-d / r0))


def _eos_out_of_scope(name):
    def init(self, *a, **k):
        raise NotImplementedError(f"{name} (EOS.py) is not part of the hot path; use SC_psi")
    return type(name, (), {"__init__": init})


vdW_psi = _eos_out_of_scope("vdW_psi")
RK_psi = _eos_out_of_scope("RK_psi")
PR_psi = _eos_out_of_scope("PR_psi")
CS_psi = _eos_out_of_scope("CS_psi")
