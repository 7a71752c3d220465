"""BGK / TRT / MRT collision -- drop-in for taichi_lbm/CollisionEngine.py.

Each collision owns one per-node relaxation state plane per component, which
holds what the reference stores: BGK ``tau``, TRT ``tau_sym`` and MRT
``diag[7]`` (= 1/tau).  TRT ``tau_antisym`` and the MRT ``diag`` vector are
fixed functions of that state (CollisionEngine.py:155, 164, 301-318) and are
exposed as derived read-only fields.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _hip as H
from ._device import DeviceField, device, plane_stride, planes_view, stream, torch
from .FluidModels import NewtonianFluid
from .lbm import gfields, group_device, lattice_struct

f32 = np.float32


def _clamp(t, lo, hi):
    """Collision.tau_max_min, CollisionEngine.py:28-34."""
    if t < lo:
        return f32(lo)
    if t > hi:
        return f32(hi)
    return f32(t)


class Collision:
    """CollisionEngine.py:6-34."""

    KIND = None

    def __init__(self, group, fluid_model):
        self.group = group
        self.fluid_model = fluid_model if fluid_model is not None else NewtonianFluid(0.0)
        self.tau_min = 0.5000001
        self.tau_max = 4.0
        self._state = None
        self._uniform = True   # state still equals the initTau value on the group

    # host versions of the @ti.func helpers (CollisionEngine.py:14-25)
    @staticmethod
    def f_eq(c, w, vel, rho):
        c = np.asarray(c, f32)
        vel = np.asarray(vel, f32)
        eu = c @ vel
        uv = vel @ vel
        return np.asarray(w, f32) * f32(rho) * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uv)

    @staticmethod
    def f_force(c, w, F, vel):
        c = np.asarray(c, f32)
        cF = c @ np.asarray(F, f32)
        cu = c @ np.asarray(vel, f32)
        uF = np.asarray(vel, f32) @ np.asarray(F, f32)
        return np.asarray(w, f32) * (3 * cF + 9 * cF * cu - 3 * uF)

    def _setup(self, params):
        self.NX, self.NY, self.nc = int(params['NX']), int(params['NY']), int(params['num_components'])
        self.N = self.NX * self.NY
        self.ps = plane_stride(self.N)
        self.nu = np.array([f32(v) for v in params['shearviscosity']], f32)

    # ---- device state ---------------------------------------------------
    def state_tensor(self):
        if self._state is None:
            t = torch()
            self._state = t.zeros((self.nc, self.ps), dtype=t.float32, device=device())
            self._write_init_state()
        return self._state

    def _group_mask(self):
        m = np.zeros((self.NX, self.NY), bool)
        if self.group is not None and len(self.group.nodes):
            m[self.group.nodes[:, 0], self.group.nodes[:, 1]] = True
        return m

    def _write_init_state(self):
        t = torch()
        vals = self.init_state_values()
        host = np.zeros((self.nc, self.ps), f32)
        gm = self._group_mask().reshape(-1)
        for c in range(self.nc):
            host[c, :self.N] = np.where(gm, vals[c], self.off_group_value())
        self._state.copy_(t.as_tensor(host))
        self._uniform = True

    def off_group_value(self):
        return f32(0.0)

    def initTau(self):
        if self._state is not None:
            self._write_init_state()
        else:
            self._uniform = True

    def _state_field(self, name):
        return DeviceField(lambda: planes_view(self.state_tensor(), (self.nc,), self.NX, self.NY),
                           after_write=self._state_written, name=name)

    def _state_written(self):
        self._uniform = False

    def fill_lattice(self, L):
        L.collision = self.KIND
        L.tau_min, L.tau_max = float(f32(self.tau_min)), float(f32(self.tau_max))
        L.visc_m, L.visc_nm1 = self.fluid_model.visc_params()

    # ---- granular kernels -----------------------------------------------
    def _granular(self, fn, lb_field, what):
        lb_field.materialize()
        g = group_device(self.group)
        H.check(fn(ctypes.byref(lattice_struct(lb_field, collision=self)),
                   ctypes.byref(gfields(lb_field, collision=self)), g[0], g[1], stream()), what)

    def apply(self, lb_field):
        """Collide f -> f2 (BGK :96-110, TRT :223-245, MRT :424-451)."""
        self._granular(H.lib().lbm_g_collide, lb_field, f"{type(self).__name__}.apply")
        lb_field._pops_written()

    def shear_rate(self, lb_field):
        """BGK :80-92, TRT :166-178, MRT :332-377."""
        self._granular(H.lib().lbm_g_shear_rate, lb_field, f"{type(self).__name__}.shear_rate")

    def relaxation_time(self, lb_field):
        """BGK :72-78, TRT :157-164, MRT :322-329."""
        self._uniform = False
        self._granular(H.lib().lbm_g_relaxation_time, lb_field, f"{type(self).__name__}.relaxation_time")


class BGKCollision(Collision):
    """CollisionEngine.py:36-110."""

    KIND = H.LBM_BGK

    def __init__(self, params: dict):
        default_params = {'group': None, 'fluid_model': None, 'NX': None, 'NY': None,
                          'num_components': None, 'shearviscosity': None}
        params = {**default_params, **params}
        super().__init__(params['group'], params['fluid_model'])
        self._setup(params)
        self.tau = self._state_field("tau")
        self.initTau()

    def init_state_values(self):
        """initTau, CollisionEngine.py:65-70: tau = 3 nu + 0.5 (no clamp)."""
        return [self.nu[c] * f32(3.0) + f32(0.5) for c in range(self.nc)]

    def fill_lattice(self, L):
        super().fill_lattice(L)
        for c, v in enumerate(self.init_state_values()):
            L.tau[c] = float(v)


class TRTCollision(Collision):
    """CollisionEngine.py:114-245."""

    KIND = H.LBM_TRT

    def __init__(self, params: dict):
        default_params = {'group': None, 'fluid_model': None, 'NX': None, 'NY': None,
                          'num_components': None, 'shearviscosity': None, 'magic': None}
        params = {**default_params, **params}
        super().__init__(params['group'], params['fluid_model'])
        self._setup(params)
        self.magic = np.array([f32(v) for v in params['magic']], f32)
        self.tau_sym = self._state_field("tau_sym")
        self.tau_antisym = DeviceField(self._antisym_view, readonly=True, name="tau_antisym")
        self.initTau()

    def init_state_values(self):
        """initTau, CollisionEngine.py:149-155."""
        return [self.nu[c] * f32(3.0) + f32(0.5) for c in range(self.nc)]

    def _antisym_view(self):
        t = torch()
        s = planes_view(self.state_tensor(), (self.nc,), self.NX, self.NY)
        mag = t.as_tensor(self.magic, device=s.device)
        gm = t.as_tensor(self._group_mask(), device=s.device)[..., None]
        return t.where(gm, mag / (s - 0.5) + 0.5, t.zeros_like(s))

    def fill_lattice(self, L):
        super().fill_lattice(L)
        for c, v in enumerate(self.init_state_values()):
            L.tau[c] = float(v)
            L.tau_anti[c] = float(self.magic[c] / (v - f32(0.5)) + f32(0.5))
            L.magic[c] = float(self.magic[c])


class MRTCollision(Collision):
    """CollisionEngine.py:247-451 (M of :273-282; exact inverse M^T D^-1)."""

    KIND = H.LBM_MRT
    M = np.array([
        [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],
        [-4.0, -1.0, -1.0, -1.0, -1.0, 2.0, 2.0, 2.0, 2.0],
        [4.0, -2.0, -2.0, -2.0, -2.0, 1.0, 1.0, 1.0, 1.0],
        [0.0, 1.0, 0.0, -1.0, 0.0, 1.0, -1.0, -1.0, 1.0],
        [0.0, -2.0, 0.0, 2.0, 0.0, 1.0, -1.0, -1.0, 1.0],
        [0.0, 0.0, 1.0, 0.0, -1.0, 1.0, 1.0, -1.0, -1.0],
        [0.0, 0.0, -2.0, 0.0, 2.0, 1.0, 1.0, -1.0, -1.0],
        [0.0, 1.0, -1.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0],
        [0.0, 0.0, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0, -1.0]], f32)

    def __init__(self, params: dict):
        default_params = {'group': None, 'fluid_model': None, 'NX': None, 'NY': None,
                          'num_components': None, 'shearviscosity': None, 'bulkviscosity': None}
        params = {**default_params, **params}
        super().__init__(params['group'], params['fluid_model'])
        self._setup(params)
        self.nu_bulk = np.array([f32(v) for v in params['bulkviscosity']], f32)
        self.M_inverse = (self.M.T.astype(np.float64) /
                          np.array([9.0, 36.0, 36.0, 6.0, 12.0, 6.0, 12.0, 4.0, 4.0])).astype(f32)
        self.diag = DeviceField(self._diag_view, readonly=True, name="diag")
        self.initTau()

    def _rates(self, c):
        tau = _clamp(self.nu[c] * f32(3.0) + f32(0.5), self.tau_min, self.tau_max)
        tb = _clamp(self.nu_bulk[c] * f32(3.0) + f32(0.5), self.tau_min, self.tau_max)
        return f32(1.0) / tau, f32(1.0) / tb

    def init_state_values(self):
        """initTau, CollisionEngine.py:301-318: diag[7] = diag[8] = 1/clamp(3 nu + 0.5)."""
        return [self._rates(c)[0] for c in range(self.nc)]

    def off_group_value(self):
        return f32(1.0)   # diag.fill(1) (CollisionEngine.py:303-304)

    def _diag_view(self):
        t = torch()
        s = planes_view(self.state_tensor(), (self.nc,), self.NX, self.NY)
        d = t.ones((self.NX, self.NY, self.nc, 9), dtype=t.float32, device=s.device)
        gm = t.as_tensor(self._group_mask(), device=s.device)
        for c in range(self.nc):
            sb = float(self._rates(c)[1])
            row = t.tensor([0.0, sb, 1.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0], device=s.device)
            blk = row.expand(self.NX, self.NY, 9).clone()
            blk[..., 7] = s[..., c]
            blk[..., 8] = s[..., c]
            d[:, :, c, :] = t.where(gm[..., None], blk, d[:, :, c, :])
        return d

    def fill_lattice(self, L):
        super().fill_lattice(L)
        for c in range(self.nc):
            s, sb = self._rates(c)
            L.s_shear[c] = float(s)
            L.s_bulk[c] = float(sb)


class HuangMRTCollision(MRTCollision):
    """CollisionEngine.py:453-515: MRT with the source moments of m_Q
    (:466-478) added to the collision as S Q (:503-507):
    Q1 = 3 (k1 + 2 k2) |F|^2 / (G psi^2), Q2 = -Q1,
    Q7 = k1 (Fx^2 - Fy^2) / (G psi^2), Q8 = k1 Fx Fy / (G psi^2),
    with k2 = epslion / (-8) - k1 (:462), G = g_coh[c][c] (:463), F the
    component's body force and psi its Shan-Chen psi of the step.

    The reference class cannot run: its constructor passes
    (num_components, group) to MRTCollision's params-dict constructor
    (:455) and apply indexes the per-node rates diag[i, j, c] as
    diag[c, ii] (:507).  Here the constructor takes MRTCollision's params
    plus 'k1', 'epslion' (one value per component) and 'g_coh' (the
    coupling matrix), and S is the node's MRT rates.  apply(lb_field,
    sc_field) needs the Shan-Chen field's psi; the fused path takes it
    from the Shan-Chen force of the run."""

    def __init__(self, params: dict):
        for key in ("k1", "epslion", "g_coh"):
            if params.get(key) is None:
                raise ValueError(f"HuangMRTCollision needs '{key}'")
        super().__init__(params)
        self.k1 = np.array([f32(v) for v in params["k1"]], f32)
        eps = np.array([f32(v) for v in params["epslion"]], f32)
        self.epslion = [float(v) for v in eps]
        self.k2 = (eps / f32(-8.0) - self.k1).astype(f32)
        g = np.asarray(params["g_coh"], f32).reshape(self.nc, self.nc) if np.ndim(params["g_coh"]) \
            else np.full((self.nc, self.nc), f32(params["g_coh"]), f32)
        self.g_coh = np.array([g[c, c] for c in range(self.nc)], f32)
        if np.any(self.g_coh == 0):
            raise ValueError("HuangMRTCollision divides by G psi^2: g_coh[c][c] must be non-zero")
        if len(self.k1) != self.nc or len(eps) != self.nc:
            raise ValueError("k1 and epslion need one value per component")

    def fill_lattice(self, L):
        super().fill_lattice(L)
        L.huang = 1
        for c in range(self.nc):
            L.huang_k1[c] = float(self.k1[c])
            L.huang_k2[c] = float(self.k2[c])
            L.huang_g[c] = float(self.g_coh[c])

    def apply(self, lb_field, sc_field=None):
        """Collide f -> f2 with the Huang source (CollisionEngine.py:480-515)."""
        sc = sc_field if sc_field is not None else lb_field.sc_field
        if sc is None:
            raise ValueError("HuangMRTCollision.apply needs the Shan-Chen field (its psi)")
        lb_field.materialize()
        g = group_device(self.group)
        H.check(H.lib().lbm_g_collide(ctypes.byref(lattice_struct(lb_field, collision=self, sc=sc)),
                                      ctypes.byref(gfields(lb_field, collision=self, sc=sc)), g[0], g[1],
                                      stream()), "HuangMRTCollision.apply")
        lb_field._pops_written()
