"""LBField and MacroscopicEngine -- drop-in for taichi_lbm/lbm.py of the reference.

Fields live in HIP device memory in the plane layout of include/lbm_abi.h and
are exposed with the reference's Taichi shapes (lbm.py:22-52): ``f``/``f2``
(NX, NY, nc, 9), ``rho`` (NX, NY, nc), ``vel`` (NX, NY, 2), ``mask`` (NX, NY)
and so on, through ``to_numpy`` / ``from_numpy`` / ``fill`` / indexing.
The setup-time fields ``mask`` and ``neighbor`` stay on the host.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _hip as H
from ._codes import C as _C, neighbor_table, neighbor_valid
from ._device import DeviceField, HostField, _Scalar, device, plane_stride, planes_view, stream, torch


def _f32(x) -> float:
    return float(np.float32(x))


class LBField:
    """lbm.py:6-142.  ``LBField(name, NX, NY, num_components)``."""

    def __init__(self, name, NX, NY, num_components):
        self.name = name
        self.NX = int(NX)
        self.NY = int(NY)
        if not 1 <= int(num_components) <= H.MAX_NC:
            raise ValueError(f"num_components must be 1..{H.MAX_NC}")
        self.num_components = _Scalar(int(num_components))
        self.nc = int(num_components)
        self.N = self.NX * self.NY
        self.ps = plane_stride(self.N)
        self.NPOP = 9
        # lbm.py:44-45
        self.weights = np.array([4, 1, 1, 1, 1, 1 / 4, 1 / 4, 1 / 4, 1 / 4], np.float32) / np.float32(9.0)
        self.c = _C.copy()
        self.gravity_force = (0.0, 0.0)
        self.time = _Scalar(0)
        self.sc_field = None
        # unit conversion, lbm.py:57-71
        self.Ct = self.Cl = self.C_rho = self.Cu = self.C_pressure = 1.0
        self.C_force = self.C_torque = self.C_temperature = self.Cnu = self.C_T = 1.0
        self.shear_viscosity_LB = np.zeros(self.nc, np.float32)
        self.bulk_viscosity_LB = np.zeros(self.nc, np.float32)

        self.mask = HostField(np.ones((self.NX, self.NY), np.int32), on_change=self._geometry_changed)
        self._neighbor = None            # HostField, built on first access
        self._neighbor_from_mask = False  # neighbor_classify ran and the table is mask-derived
        self._neighbor_dev = None
        self._st = None          # device state, allocated on first device use
        self._round_step = 0     # steps taken by the fused runs (seeds of 16-bit stochastic rounding)
        self._finisher = None    # pending boundary pass of the last fused run
        self._pop_owner = None   # a Stepper whose 16-bit populations are the newest state (None: FP32)
        self._pressure_refresh = None
        self.geometry_version = 0

        nc, NX_, NY_ = self.nc, self.NX, self.NY
        pv = planes_view
        self.f = DeviceField(lambda: pv(self._pop(f=True), (nc, 9), NX_, NY_), self.materialize,
                             self._after_f_write, name="f")
        self.f2 = DeviceField(lambda: pv(self._pop(f=False), (nc, 9), NX_, NY_), self.materialize,
                              self._pops_written, name="f2")
        self.rho = DeviceField(lambda: pv(self._mac()[:nc], (nc,), NX_, NY_), self.materialize, name="rho")
        self.total_rho = DeviceField(lambda: pv(self._mac()[nc:nc + 1], (), NX_, NY_), self.materialize,
                                     name="total_rho")
        self.vel = DeviceField(lambda: pv(self._mac()[nc + 1:nc + 3], (2,), NX_, NY_), self.materialize,
                               name="vel")
        self.body_force = DeviceField(lambda: pv(self.aux("body_force"), (nc, 2), NX_, NY_), name="body_force")
        self.SCforce = DeviceField(lambda: pv(self.aux("SCforce"), (nc, 2), NX_, NY_), name="SCforce")
        self.shear_rate = DeviceField(lambda: pv(self.aux("shear_rate"), (nc,), NX_, NY_), name="shear_rate")
        self.pressure = DeviceField(lambda: pv(self.aux("pressure"), (nc,), NX_, NY_), self._refresh_pressure,
                                    name="pressure")
        self.total_pressure = DeviceField(lambda: pv(self.aux("total_pressure"), (), NX_, NY_),
                                          self._refresh_pressure, name="total_pressure")
        self.T = DeviceField(lambda: pv(self.aux("T"), (), NX_, NY_), name="T")

        print("=" * 20)
        print("init LBField")
        print(f"  NX: {self.NX}")
        print(f"  NY: {self.NY}")
        print(f"  num_components: {self.nc}")

    # ---- reference API ---------------------------------------------------
    def init_conversion(self, params: dict):
        """lbm.py:77-104."""
        default_params = {'Cl': None, 'Ct': None, 'C_rho': None, 'C_pressure': None, 'C_temperature': None}
        params = {**default_params, **params}
        self.Ct = params['Ct']
        self.Cl = params['Cl']
        self.Cu = self.Cl / self.Ct
        self.Cnu = self.Cl ** 2 / self.Ct
        self.C_rho = params['C_rho']
        self.C_pressure = params['C_pressure']
        self.C_temperature = params['C_temperature']
        self.C_force = self.Cl ** 3 * self.C_rho / self.Ct ** 2
        self.C_torque = self.C_force * self.Cl
        print("=" * 20)
        print("init conversion")
        print(f"  Cl : {self.Cl}")
        print(f"  Ct : {self.Ct}")
        print(f"  C_rho : {self.C_rho}")

    def set_gravity(self, g):
        """lbm.py:106-107.  Read at every step (the reference freezes it at first compile)."""
        self.gravity_force = tuple(float(x) for x in g)

    def get_Cforce(self):
        return self.C_force

    def get_Ctorque(self):
        return self.C_torque

    def get_Cu(self):
        return self.Cu

    def init_LBM(self, collsion, group):
        """lbm.py:119-125: f = f2 = feq(vel, rho) on the group."""
        self.materialize()
        g = group_device(group)
        H.check(H.lib().lbm_g_init_lbm(ctypes.byref(lattice_struct(self)), ctypes.byref(gfields(self)),
                                       g[0], g[1], stream()), "init_LBM")
        self._round_step = 0
        self._pops_written()
        print("init LBM")

    def neighbor_classify(self):
        """lbm.py:127-142 (non-periodic, solid neighbours give -1).  The table is
        a function of the mask; it is materialised on first access."""
        self._neighbor = None
        self._neighbor_from_mask = True
        self._neighbor_mask = self.mask.arr.copy()   # later mask edits do not touch the table
        self._neighbor_dev = None
        self.geometry_version += 1

    @property
    def neighbor(self):
        """(NX, NY, 9, 2) int32 neighbour table (lbm.py:50)."""
        if self._neighbor is None:
            arr = neighbor_table(self._neighbor_mask) if self._neighbor_from_mask else \
                np.zeros((self.NX, self.NY, 9, 2), np.int32)
            self._neighbor = HostField(arr, on_change=self._table_edited)
        return self._neighbor

    def _table_edited(self):
        self._neighbor_from_mask = False
        self._neighbor_dev = None
        self.geometry_version += 1

    def neighbor_valid(self):
        """(NX, NY, 9) bool map of the table's valid entries (what bounce-back tests)."""
        if self._neighbor is None and self._neighbor_from_mask:
            return neighbor_valid(self._neighbor_mask)
        return self.neighbor.arr[..., 0] != -1

    # ---- device state ----------------------------------------------------
    def _geometry_changed(self):
        self.geometry_version += 1

    def state(self):
        if self._st is None:
            t = torch()
            dev = device()
            nc, ps = self.nc, self.ps
            self._st = {
                "pop": t.zeros((2, nc * 9, ps), dtype=t.float32, device=dev),
                "mac": t.zeros((2, nc + 3, ps), dtype=t.float32, device=dev),
                "f2_idx": 0,          # buffer holding f2; the other holds f when f_valid
                "f_valid": True,
                "aux": {},
            }
        return self._st

    def aux(self, name):
        st = self.state()
        if name not in st["aux"]:
            t = torch()
            planes = {"body_force": 2 * self.nc, "SCforce": 2 * self.nc, "shear_rate": self.nc,
                      "pressure": self.nc, "total_pressure": 1, "T": 1}[name]
            st["aux"][name] = t.zeros((planes, self.ps), dtype=t.float32, device=device())
        return st["aux"][name]

    def _pop(self, f: bool):
        st = self.state()
        return st["pop"][1 - st["f2_idx"] if f else st["f2_idx"]]

    def _mac(self):
        st = self.state()
        return st["mac"][st["f2_idx"]]

    def materialize(self):
        """Runs the pending boundary pass of the last fused run, so that f, rho,
        vel and total_rho hold what the reference's fields hold after lbm_solve()."""
        st = self.state()
        if not st["f_valid"]:
            fin = self._finisher
            if fin is None:
                raise RuntimeError("populations are post-collision but no boundary pass is registered")
            self._refresh_pressure()   # pressure sees rho before the boundary overrides
            fin()
            st["f_valid"] = True

    def _after_f_write(self):
        self._pops_written()

    def _pops_written(self):
        """The FP32 populations were written directly (field writes, granular
        collide / boundary kernels, init_LBM): they are the newest state, and a
        Stepper with 16-bit storage encodes them again before its next run."""
        self._pop_owner = None

    def _refresh_pressure(self):
        if self._pressure_refresh is not None:
            fn, self._pressure_refresh = self._pressure_refresh, None
            fn()

    def neighbor_device(self):
        if self._neighbor_dev is None:
            t = torch()
            self._neighbor_dev = t.as_tensor(np.ascontiguousarray(self.neighbor.arr), device=device())
        return self._neighbor_dev


# ---------------------------------------------------------------------------
# structs for the C-ABI
# ---------------------------------------------------------------------------
def group_device(group):
    """(device pointer, count) of a MaskAndGroup's node list."""
    t = group.device_nodes()
    return (H.ptr(t) if t is not None else None), int(group.count[None])


def lattice_struct(lbf: LBField, collision=None, sc=None, storage=H.STORE_F32, tau_mode=H.TAU_CONST,
                   neem=()) -> H.Lattice:
    L = H.Lattice()
    L.nx, L.ny, L.nc = lbf.NX, lbf.NY, lbf.nc
    L.plane_stride = lbf.ps
    L.round_step = lbf._round_step & 0xFFFFFFFF
    L.storage = storage
    L.tau_mode = tau_mode
    L.tau_min, L.tau_max = _f32(0.5000001), _f32(4.0)
    L.gx, L.gy = (_f32(v) for v in lbf.gravity_force)
    L.visc_m, L.visc_nm1 = 0.0, 0.0
    if collision is not None:
        collision.fill_lattice(L)
    if sc is not None:
        sc.fill_lattice(L)
    L.n_neem = len(neem)
    return L


def gfields(lbf: LBField, collision=None, sc=None, neighbor=False) -> H.GFields:
    st = lbf.state()
    g = H.GFields()
    nc = lbf.nc
    g.f = H.ptr(lbf._pop(True))
    g.f2 = H.ptr(lbf._pop(False))
    mac = lbf._mac()
    g.rho = H.ptr(mac[0])
    g.total_rho = H.ptr(mac[nc])
    g.vel = H.ptr(mac[nc + 1])
    g.body_force = H.ptr(lbf.aux("body_force"))
    g.sc_force = H.ptr(lbf.aux("SCforce"))
    g.shear_rate = H.ptr(lbf.aux("shear_rate"))
    g.pressure = H.ptr(lbf.aux("pressure"))
    g.total_pressure = H.ptr(lbf.aux("total_pressure"))
    g.temperature = H.ptr(lbf.aux("T"))
    if collision is not None:
        g.tau = H.ptr(collision.state_tensor())
    if sc is not None:
        g.psi = H.ptr(sc.psi_halo())
        fm = sc.sc_fluid_device() if hasattr(sc, "sc_fluid_device") else None
        g.sc_fluid = H.ptr(fm) if fm is not None else None
    if neighbor:
        g.neighbor = H.ptr(lbf.neighbor_device())
    return g


# ---------------------------------------------------------------------------
class MacroscopicEngine:
    """lbm.py:145-237: density / pressure / force_density / velocity over a group."""

    def __init__(self, group):
        self.group = group

    def time_updata(self, lb_field):
        """lbm.py:150-151."""
        lb_field.time[None] += 1

    def _call(self, fn, lb_field, what, sc=None):
        lb_field.materialize()
        g = group_device(self.group)
        H.check(fn(ctypes.byref(lattice_struct(lb_field, sc=sc)), ctypes.byref(gfields(lb_field, sc=sc)),
                   g[0], g[1], stream()), what)

    def density(self, lb_field):
        """lbm.py:153-164."""
        self._call(H.lib().lbm_g_density, lb_field, "density")

    def pressure(self, lb_field):
        """lbm.py:205-212 -> pressure0 / pressure1 / pressure2."""
        lb_field._pressure_refresh = None
        self._call(H.lib().lbm_g_pressure, lb_field, "pressure", sc=lb_field.sc_field)

    def force_density(self, lb_field):
        """lbm.py:215-220."""
        self._call(H.lib().lbm_g_force_density, lb_field, "force_density")

    def velocity(self, lb_field):
        """lbm.py:223-237."""
        self._call(H.lib().lbm_g_velocity, lb_field, "velocity")
