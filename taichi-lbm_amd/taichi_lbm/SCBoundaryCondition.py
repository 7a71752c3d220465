"""psi halo boundaries -- drop-in for taichi_lbm/SCBoundaryCondition.py.

The periodic pair (PeriodicPsi_LR then PeriodicPsi_BT, :126-144) is what the
fused Shan-Chen step reads (a periodic wrap of the psi planes).  The other
psi boundaries (:9-123) act on groups of the psi halo grid, nodes (i, j) in
[-1, NX] x [-1, NY] from a BoundaryClassifier over that grid (shape (NX+2,
NY+2), offset (-1, -1)); they run as granular kernels after update_psi.  The
fused path also takes FixedPsi on nodes of the lattice that are outside the
fluid group (solid walls with a fixed psi: their psi never changes, so the
value is written into the psi planes once).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _hip as H
from ._device import stream
from .BoundaryCondition import BoundaryCondition, PlaneBoundary
from .lbm import gfields, group_device, lattice_struct


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


def _psi_apply(bc, lb_field, kind, direction=0, value=0.0):
    sc = lb_field.sc_field
    NX, NY = lb_field.NX, lb_field.NY
    nodes = np.asarray(bc.group.nodes, np.int64).reshape(-1, 2)
    if len(nodes):
        reach = {H.PSI_BC_FIXED: 0, H.PSI_BC_NEUMANN: 1, H.PSI_BC_CONVECTIVE1: 1, H.PSI_BC_CONVECTIVE2: 2,
                 H.PSI_BC_EXTRAPOLATION: 2}[kind]
        c = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1]])[direction if reach else 0]
        for r in range(reach + 1):
            I, J = nodes[:, 0] - r * c[0], nodes[:, 1] - r * c[1]
            if I.min() < -1 or I.max() > NX or J.min() < -1 or J.max() > NY:
                raise H.LBMError(f"{type(bc).__name__}: reads psi outside the halo grid")
        if kind in (H.PSI_BC_CONVECTIVE1, H.PSI_BC_CONVECTIVE2):
            I, J = nodes[:, 0] - c[0], nodes[:, 1] - c[1]
            if I.min() < 0 or I.max() >= NX or J.min() < 0 or J.max() >= NY:
                raise H.LBMError(f"{type(bc).__name__}: reads the velocity outside the lattice")
    sc.psi_halo()
    sc._halo_written()
    g = group_device(bc.group)
    H.check(H.lib().lbm_g_sc_psi_boundary(ctypes.byref(lattice_struct(lb_field, sc=sc)),
                                          ctypes.byref(gfields(lb_field, sc=sc)), kind, int(direction),
                                          float(np.float32(value)), g[0], g[1], stream()), type(bc).__name__)


class OpenBoundaryPsi(PlaneBoundary):
    """SCBoundaryCondition.py:9-16."""
    KIND = 0

    def apply(self, lb_field):
        lb_field.materialize()
        _psi_apply(self, lb_field, self.KIND, self.direction)


class OpenNeumannPsi(OpenBoundaryPsi):
    """:19-33: psi_c(x) := psi_c(x - c_d)."""
    KIND = H.PSI_BC_NEUMANN


class OpenConvective1orderPsi(OpenBoundaryPsi):
    """:36-58: lambda = |mean u(x - c_d)| over the group;
    psi(x) := (psi(x) + lambda psi(x - c_d)) / (1 + lambda)."""
    KIND = H.PSI_BC_CONVECTIVE1


class OpenConvective2orderPsi(OpenBoundaryPsi):
    """:61-89: psi(x) := (psi(x) + 2 lambda psi(x - c_d) - lambda/2 psi(x - 2 c_d)) / (1 + 3 lambda / 2)."""
    KIND = H.PSI_BC_CONVECTIVE2


class OpenExtrapolationPsi(OpenBoundaryPsi):
    """:92-107: psi(x) := 2 psi(x - c_d) - psi(x - 2 c_d)."""
    KIND = H.PSI_BC_EXTRAPOLATION


class FixedPsi(BoundaryCondition):
    """:110-123: psi_0(x) := psi_value (component 0 only)."""

    def __init__(self, spec, psi_value):
        super().__init__(spec)
        self.psi_value = psi_value

    def apply(self, lb_field):
        _psi_apply(self, lb_field, H.PSI_BC_FIXED, 0, self.psi_value)


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

    def fused_plan(self, lb_field, fluid_nodes):
        """What the fused Shan-Chen step can do with this list: the periodic
        pair, followed by FixedPsi entries over lattice nodes outside the fluid
        group and off the lattice edges (their psi is constant, written once).
        Returns [(nodes, value)]; raises NotImplementedError otherwise."""
        bcs = list(self.boundary_conditions.values())
        if [type(b) for b in bcs[:2]] != [PeriodicPsi_LR, PeriodicPsi_BT]:
            raise NotImplementedError("the fused Shan-Chen path needs the psi boundaries PeriodicPsi_LR then "
                                      "PeriodicPsi_BT")
        NX, NY = lb_field.NX, lb_field.NY
        active = np.zeros((NX, NY), bool)
        fn = np.asarray(fluid_nodes).reshape(-1, 2)
        active[fn[:, 0], fn[:, 1]] = True
        plan = []
        for b in bcs[2:]:
            nodes = np.asarray(b.group.nodes, np.int64).reshape(-1, 2) if isinstance(b, FixedPsi) else None
            ok = nodes is not None and (len(nodes) == 0 or (
                nodes[:, 0].min() > 0 and nodes[:, 0].max() < NX - 1 and nodes[:, 1].min() > 0 and
                nodes[:, 1].max() < NY - 1 and not active[nodes[:, 0], nodes[:, 1]].any()))
            if not ok:
                raise NotImplementedError(f"psi boundary {type(b).__name__} has no fused form (FixedPsi over "
                                          "non-fluid lattice nodes off the edges only): use the granular path")
            plan.append((nodes, float(np.float32(b.psi_value))))
        return plan
