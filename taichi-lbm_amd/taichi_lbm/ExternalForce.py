"""Shan-Chen pseudopotential forces -- drop-in for taichi_lbm/ExternalForce.py.

``psi_field`` keeps the reference's shape (NX+2, NY+2, nc) and its offset
indexing (index -1 is the lower halo, ExternalForce.py:38): integer indices
given to ``psi_field[...]`` are shifted by +1, ``to_numpy()`` returns the
whole halo array.  The fused path keeps psi in two plain (NX*NY) planes and
reads the halo as a periodic wrap, which is what PeriodicPsi_LR/BT produce.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _hip as H
from ._device import DeviceField, device, stream, torch
from .lbm import gfields, group_device, lattice_struct

f32 = np.float32


class _OffsetPsiField(DeviceField):
    """psi_field with the reference's offset (-1, -1, 0)."""

    def __init__(self, sc):
        super().__init__(sc._halo_view, name="psi_field", after_write=sc._halo_written)

    @staticmethod
    def _shift(key):
        key = key if isinstance(key, tuple) else (key,)
        out = []
        for ax, k in enumerate(key):
            if ax < 2 and isinstance(k, (int, np.integer)):
                out.append(int(k) + 1)
            elif ax < 2 and isinstance(k, np.ndarray) and k.dtype != bool:
                out.append(k + 1)
            else:
                out.append(k)
        return tuple(out)

    def __getitem__(self, key):
        return super().__getitem__(self._shift(key))

    def __setitem__(self, key, val):
        super().__setitem__(self._shift(key), val)


class ShanChenModel:
    """ExternalForce.py:33-52."""

    SC_KIND = H.SC_NONE

    def __init__(self, lb_field, group):
        self.lb_field = lb_field
        self.group = group
        self.NX, self.NY, self.nc = lb_field.NX, lb_field.NY, lb_field.nc
        w = lb_field.weights
        self.wc = np.stack([w * lb_field.c[:, 0].astype(f32), w * lb_field.c[:, 1].astype(f32)], 1)
        self._halo = None          # (nc, (NX+2)*(NY+2)) device, reference layout
        self._planes = None        # (2, nc, ps) device, fused-path ping-pong
        self._halo_valid = True    # halo holds the current psi
        self._planes_valid = False
        self._planes_idx = 0       # plane set holding the current psi
        self.psi_field = _OffsetPsiField(self)

    # ---- storage ----------------------------------------------------------
    def psi_halo(self):
        if self._halo is None:
            t = torch()
            self._halo = t.zeros((self.nc, (self.NX + 2) * (self.NY + 2)), dtype=t.float32, device=device())
        if not self._halo_valid:
            self._halo_from_planes(self._planes_idx)
        return self._halo

    def _halo_view(self):
        h = self.psi_halo()
        return h.reshape(self.nc, self.NX + 2, self.NY + 2).permute(1, 2, 0)

    def _halo_written(self):
        self._planes_valid = False

    def _halo_from_planes(self, idx):
        """Halo interior <- plane set idx, then the periodic ring (PeriodicPsi_LR, _BT)."""
        lbf = self.lb_field
        h = self._halo.reshape(self.nc, self.NX + 2, self.NY + 2)
        h[:, 1:-1, 1:-1] = self._planes[idx, :, :lbf.N].reshape(self.nc, self.NX, self.NY)
        self._halo_valid = True
        H.check(H.lib().lbm_g_sc_periodic_psi(ctypes.byref(lattice_struct(lbf, sc=self)),
                                              ctypes.byref(gfields(lbf, sc=self)), 3, stream()), "psi halo")

    def planes_for_run(self):
        """(psi ping-pong tensor, index of the current set) for a fused run."""
        lbf = self.lb_field
        t = torch()
        if self._planes is None:
            self._planes = t.zeros((2, self.nc, lbf.ps), dtype=t.float32, device=device())
        if not self._planes_valid:
            h = self.psi_halo().reshape(self.nc, self.NX + 2, self.NY + 2)[:, 1:-1, 1:-1]
            for q in range(2):
                self._planes[q, :, :lbf.N] = h.reshape(self.nc, -1)
            self._planes_valid = True
        return self._planes

    def keep_previous(self, written):
        """Before a one-step fused run that writes its psi into plane set
        `written`: the pressure observable of that step reads the psi of the
        step before from the other set, so if `written` holds it (the newest
        psi), it is copied there first."""
        if self._planes is not None and self._planes_idx == written:
            self._planes[1 - written].copy_(self._planes[written])

    def after_run(self, idx):
        """The fused run left the newest psi in plane set idx."""
        self._planes_idx = idx
        self._halo_valid = False

    # ---- reference API ----------------------------------------------------
    def _granular(self, fn, lb_field, what):
        lb_field.materialize()
        g = group_device(self.group)
        H.check(fn(ctypes.byref(lattice_struct(lb_field, sc=self)), ctypes.byref(gfields(lb_field, sc=self)),
                   g[0], g[1], stream()), what)

    def update_psi(self, lb_field):
        self.psi_halo()
        self._halo_written()
        self._granular(H.lib().lbm_g_sc_update_psi, lb_field, "update_psi")

    def cau_SCforce(self, lb_field):
        self._granular(H.lib().lbm_g_sc_force, lb_field, "cau_SCforce")


class ShanChenForceC1(ShanChenModel):
    """ExternalForce.py:55-91: single-component force on component 0."""

    SC_KIND = H.SC_C1

    def __init__(self, params: dict):
        default_params = {'lb_field': None, 'g_coh': None, 'group': None, 'psi': None}
        params = {**default_params, **params}
        super().__init__(params['lb_field'], params['group'])
        self.g_coh = params['g_coh']
        self.psi = params['psi']

    def fill_lattice(self, L):
        L.shan_chen = H.SC_C1
        L.g_coh[0] = float(f32(self.g_coh))
        for c in range(self.nc):
            self.psi.fill(L, c)


class ShanChenForceC2(ShanChenModel):
    """ExternalForce.py:66-116: two-component cohesion; a neighbour x + c_k for
    which fluid_strategy(x2, y2) is false contributes the adhesion w_k c_k
    gadh[c1] instead (without psi(x + c_k), commented out at :114).
    fluid_strategy is evaluated once on the host over the halo grid
    (x2, y2 in [-1, NX] x [-1, NY]), vectorised (integer index arrays) or
    node by node."""

    SC_KIND = H.SC_C2

    def __init__(self, params: dict):
        default_params = {'lb_field': None, 'g_coh': None, 'gadh': None, 'group': None, 'psi_list': None,
                          'fluid_strategy': None}
        params = {**default_params, **params}
        super().__init__(params['lb_field'], params['group'])
        if self.nc != 2:
            raise ValueError("ShanChenForceC2 needs a two-component LBField")
        self.g_coh = np.asarray(params['g_coh'], f32)
        self.g_adh = params['gadh']
        self.psi1, self.psi2 = params['psi_list'][0], params['psi_list'][1]
        self.fluid_strategy = params['fluid_strategy']
        self.fluid_map = None          # (NX+2, NY+2) bool over the halo grid, offset -1
        self._fluid_dev = None
        if self.fluid_strategy is not None:
            from .BoundaryCondition import evaluate_predicate
            fm = evaluate_predicate(lambda I, J: self.fluid_strategy(I - 1, J - 1), self.NX + 2, self.NY + 2)
            if not fm.all():
                self.fluid_map = fm

    @property
    def adhesion(self):
        return self.fluid_map is not None

    def sc_fluid_device(self):
        """The fluid map as uint8 on the device (lbm_gfields.sc_fluid)."""
        if self.fluid_map is None:
            return None
        if self._fluid_dev is None:
            self._fluid_dev = torch().as_tensor(np.ascontiguousarray(self.fluid_map.astype(np.uint8)).reshape(-1),
                                                device=device())
        return self._fluid_dev

    def solid_codes(self):
        """Node-code bits (LBM_CODE_SC_SOLID) of the lattice's nodes for the
        fused kernel, which wraps the halo periodically: refused unless the
        halo's flags are the periodic wrap of the lattice's."""
        if self.fluid_map is None:
            return None
        inner = self.fluid_map[1:-1, 1:-1]
        if not np.array_equal(np.pad(inner, 1, mode="wrap"), self.fluid_map):
            raise NotImplementedError("the fused Shan-Chen path reads the psi halo as a periodic wrap: fluid_strategy "
                                      "on the halo must equal its value at the wrapped node (use the granular "
                                      "cau_SCforce otherwise)")
        return np.where(inner.reshape(-1), 0, H.CODE_SC_SOLID).astype(np.uint16)

    def fill_lattice(self, L):
        L.shan_chen = H.SC_C2
        for a in range(2):
            for b in range(2):
                L.g_coh[a * H.MAX_NC + b] = float(self.g_coh[a, b])
        self.psi1.fill(L, 0)
        self.psi2.fill(L, 1)
        gadh = [0.0, 0.0] if self.g_adh is None else list(self.g_adh)
        L.g_adh[0], L.g_adh[1] = (float(f32(v)) for v in gadh[:2])
        L.sc_adhesion = 1 if self.fluid_map is not None else 0
