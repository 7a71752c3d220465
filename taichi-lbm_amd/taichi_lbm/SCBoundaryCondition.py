"""psi halo boundaries -- drop-in for taichi_lbm/SCBoundaryCondition.py
(periodic ones, :126-160)."""
from __future__ import annotations

import ctypes

from . import _hip as H
from ._device import stream
from .lbm import gfields, lattice_struct


class _PeriodicPsi:
    WHICH = 0

    def apply(self, lb_field):
        sc = lb_field.sc_field
        sc.psi_halo()
        sc._halo_written()
        H.check(H.lib().lbm_g_sc_periodic_psi(ctypes.byref(lattice_struct(lb_field, sc=sc)),
                                              ctypes.byref(gfields(lb_field, sc=sc)), self.WHICH, stream()),
                type(self).__name__)


class PeriodicPsi_LR(_PeriodicPsi):
    """SCBoundaryCondition.py:126-134: halo columns x = -1, NX."""
    WHICH = 1


class PeriodicPsi_BT(_PeriodicPsi):
    """SCBoundaryCondition.py:136-144: halo rows y = -1, NY (after LR, so corners wrap)."""
    WHICH = 2


class BoundaryEngineSC:
    """SCBoundaryCondition.py:149-160."""

    def __init__(self):
        self.boundary_conditions = {}

    def add_boundary_condition(self, name, bc):
        self.boundary_conditions[name] = bc

    def apply_boundary_conditions(self, lb_field):
        for bc in self.boundary_conditions.values():
            bc.apply(lb_field)

    def is_periodic(self):
        """True when the list is PeriodicPsi_LR then PeriodicPsi_BT: a periodic wrap."""
        kinds = [type(b) for b in self.boundary_conditions.values()]
        return kinds == [PeriodicPsi_LR, PeriodicPsi_BT]


def _psi_bc_out_of_scope(name):
    def init(self, *a, **k):
        raise NotImplementedError(f"{name} (SCBoundaryCondition.py:9-123) is not supported; psi boundaries "
                                  "are periodic (PeriodicPsi_LR, PeriodicPsi_BT)")
    return type(name, (), {"__init__": init})


OpenNeumannPsi = _psi_bc_out_of_scope("OpenNeumannPsi")
OpenConvective1orderPsi = _psi_bc_out_of_scope("OpenConvective1orderPsi")
OpenConvective2orderPsi = _psi_bc_out_of_scope("OpenConvective2orderPsi")
OpenExtrapolationPsi = _psi_bc_out_of_scope("OpenExtrapolationPsi")
FixedPsi = _psi_bc_out_of_scope("FixedPsi")
