"""Immersed-boundary particles -- drop-in for taichi_lbm/ibm.py of the reference.

Circular particles (``Sphere``) carry a ring of Lagrangian markers
(``IBMNodes``); ``IBDEMCouplerEngine`` places the markers of the particles
that lie on the lattice, and ``IBEngine`` couples them to the fluid by
multi-direct forcing (interpolate the lattice velocity at the markers, force
2 (u_desired - u), spread it back onto ``vel`` and ``body_force``).

The struct fields live in HIP device memory as arrays per member (a vector
member is two planes, x then y) and are exposed like Taichi struct fields:
``field.member.to_numpy()`` gives the reference's shape (``(Max_Number, 2)``
for ``Sphere.force``), ``field[m].member = value`` writes one record,
``field.to_numpy()`` gives a dict of members.  Every kernel runs on the GPU
(csrc/lbm_ibm.hip); the host only checks sizes.

IB coupling acts on the granular fields: call ``MultiDirectForcingLoop``
after ``MacroscopicEngine.velocity`` and before the collision's ``apply``
(the collision reads the corrected ``vel`` and ``body_force``).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _hip as H
from ._device import _Scalar, device, stream, torch
from .lbm import gfields, lattice_struct


class _Member:
    """One member of a struct field, with the reference's shape (*shape, width)."""

    def __init__(self, owner, name):
        self._o, self.name = owner, name

    def _t(self):
        return self._o._data[self.name]

    def view(self):
        t = self._t()
        w = self._o._width[self.name]
        shape = self._o.shape
        if w == 1:
            return t.view(shape)
        nd = len(shape)                      # (w, cap) planes -> (*shape, w) view
        return t.view(w, *shape).permute(*range(1, nd + 1), 0)

    @property
    def shape(self):
        return tuple(self.view().shape)

    def to_numpy(self):
        return self.view().detach().cpu().numpy().copy()

    def from_numpy(self, a):
        t = torch()
        v = self.view()
        v.copy_(t.as_tensor(np.asarray(a), dtype=v.dtype).reshape(v.shape))

    def fill(self, val):
        self.view().fill_(val)

    def __getitem__(self, key):
        v = self.view()[key]
        return v.item() if v.dim() == 0 else v.detach().cpu().numpy()

    def __setitem__(self, key, val):
        t = torch()
        v = self.view()
        v[key] = t.as_tensor(np.asarray(val), dtype=v.dtype, device=v.device)


class _Record:
    """``field[idx]``: attribute access to one record of a struct field."""

    def __init__(self, owner, idx):
        object.__setattr__(self, "_o", owner)
        object.__setattr__(self, "_idx", idx)

    def __getattr__(self, name):
        if name not in self._o._width:
            raise AttributeError(name)
        return getattr(self._o, name)[self._idx]

    def __setattr__(self, name, value):
        if name not in self._o._width:
            raise AttributeError(name)
        getattr(self._o, name)[self._idx] = value


class StructField:
    """A Taichi-style struct field in device memory: members of the schema as
    arrays over the field's cells (vectors as two planes)."""

    def __init__(self, schema: dict, shape):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        cap = int(np.prod(self.shape))
        t = torch()
        dev = device()
        self._width = {}
        self._data = {}
        for name, (dtype, width) in schema.items():
            dt = t.int32 if dtype is int else t.float32
            self._width[name] = width
            self._data[name] = t.zeros((width, cap) if width > 1 else (cap,), dtype=dt, device=dev)
        self._members = {n: _Member(self, n) for n in schema}

    def __getattr__(self, name):
        members = self.__dict__.get("_members", {})
        if name in members:
            return members[name]
        raise AttributeError(name)

    def __getitem__(self, idx):
        return _Record(self, idx)

    def to_numpy(self):
        return {n: m.to_numpy() for n, m in self._members.items()}

    def ptr(self, name):
        return H.ptr(self._data[name])


class IBMNodes:
    """ibm.py:8-14: one Lagrangian marker."""
    schema = {"pos": (float, 2), "vel": (float, 2), "vel_desired": (float, 2), "force": (float, 2),
              "force_temp": (float, 2), "mask_insiderDomain": (int, 1)}

    @classmethod
    def field(cls, shape):
        return StructField(cls.schema, shape)


class Sphere:
    """ibm.py:17-26: one circular particle."""
    schema = {"id": (int, 1), "pos": (float, 2), "vel": (float, 2), "force": (float, 2), "torque": (float, 1),
              "radius": (float, 1), "angle": (float, 1), "NumIB": (int, 1)}

    @classmethod
    def field(cls, shape):
        return StructField(cls.schema, shape)


class _DeviceScalar:
    """0-d int field in device memory (x[None] reads it back)."""

    def __init__(self):
        self.t = torch().zeros(1, dtype=torch().int32, device=device())

    def __getitem__(self, key):
        return int(self.t.item())

    def __setitem__(self, key, v):
        self.t.fill_(int(v))


class _VectorField:
    """ti.Vector.field(2, int, shape=(n,)) in device memory."""

    def __init__(self, n):
        self.t = torch().zeros((n, 2), dtype=torch().int32, device=device())

    @property
    def shape(self):
        return (self.t.shape[0],)

    def to_numpy(self):
        return self.t.cpu().numpy().copy()

    def from_numpy(self, a):
        self.t.copy_(torch().as_tensor(np.asarray(a, np.int32)))

    def fill(self, v):
        self.t[:] = torch().as_tensor(np.asarray(v, np.int32), device=self.t.device)

    def __getitem__(self, key):
        return self.t[key].cpu().numpy()


def num_ib(radius, cof) -> int:
    """ibm.py:61: round(pi * 2 * radius / NUMIB_cof) in FP32 (halves up)."""
    x = np.float32(2.0 * math.pi) * np.float32(radius) / np.float32(cof)
    if not (x > -1.0 and x < 2147483520.0):
        return 0
    return int(np.floor(x + np.float32(0.5)))


class SphereIB:
    """ibm.py:29-83: ``SphereIB(Max_Number, Max_NumIB)``."""

    def __init__(self, Max_Number, Max_NumIB):
        self.Max_Number = int(Max_Number)
        self.Max_NumIB = int(Max_NumIB)
        if self.Max_Number < 1 or self.Max_NumIB < 1:
            raise ValueError("Max_Number and Max_NumIB must be positive")
        self.num = _Scalar(0)
        self.Sphere = Sphere.field(shape=(self.Max_Number,))
        self.NUMIB_cof = 1
        self.IBNodes = IBMNodes.field(shape=(self.Max_Number, self.Max_NumIB))
        self.IBNodeInsiderGroup = _VectorField(self.Max_Number * self.Max_NumIB)
        self.IBInsiderCount = _DeviceScalar()
        n_work = H.lib().lbm_ib_work_ints(self.Max_Number, self.Max_NumIB)
        if n_work <= 0:
            raise H.LBMError("bad particle capacity")
        self._work = torch().zeros(int(n_work), dtype=torch().int32, device=device())

    def abi(self) -> H.IBParticles:
        p = H.IBParticles()
        p.max_number, p.max_num_ib = self.Max_Number, self.Max_NumIB
        n = int(self.num[None])
        if not 0 <= n <= self.Max_Number:
            raise H.LBMError(f"num = {n} outside [0, Max_Number = {self.Max_Number}]")
        p.num = n
        p.numib_cof = float(np.float32(self.NUMIB_cof))
        S, N = self.Sphere, self.IBNodes
        p.pos, p.vel, p.force = S.ptr("pos"), S.ptr("vel"), S.ptr("force")
        p.torque, p.radius, p.angle, p.num_ib = S.ptr("torque"), S.ptr("radius"), S.ptr("angle"), S.ptr("NumIB")
        p.node_pos, p.node_vel = N.ptr("pos"), N.ptr("vel")
        p.node_vel_desired, p.node_force, p.node_force_temp = N.ptr("vel_desired"), N.ptr("force"), N.ptr("force_temp")
        p.group = H.ptr(self.IBNodeInsiderGroup.t)
        p.count = H.ptr(self.IBInsiderCount.t)
        p.work = H.ptr(self._work)
        return p

    def init_Balls(self, num, id_np, pos_np, vel_np, force_np, torque_np, radius_np, angle_np):
        """ibm.py:43-55 (one upload per member)."""
        num = int(num)
        if not 0 <= num <= self.Max_Number:
            raise ValueError(f"num = {num} outside [0, Max_Number = {self.Max_Number}]")
        self.num[None] = num
        S = self.Sphere
        for name, arr in (("id", id_np), ("pos", pos_np), ("vel", vel_np), ("force", force_np),
                          ("torque", torque_np), ("radius", radius_np), ("angle", angle_np)):
            a = np.asarray(arr)[:num]
            getattr(S, name)[:num] = a
        print(f"init {num} balls")

    def init_IB_nodes(self):
        """ibm.py:57-67.  A sphere needing more than Max_NumIB markers is refused
        (the reference writes past its node field)."""
        num = int(self.num[None])
        radius = self.Sphere.radius.to_numpy()[:num]
        counts = [num_ib(r, self.NUMIB_cof) for r in radius]
        for m, n in enumerate(counts):
            if n > self.Max_NumIB:
                raise H.LBMError(f"sphere {m} (radius {radius[m]}) needs {n} IB nodes, more than "
                                 f"Max_NumIB = {self.Max_NumIB}")
        H.check(H.lib().lbm_ib_init_nodes(ctypes.byref(self.abi()), stream()), "init_IB_nodes")
        print(f"init {num} balls with {sum(max(n, 0) for n in counts)} IB nodes")

    def computer_IB_nodes_pos(self, m, n):
        """ibm.py:69-78, evaluated on the host from the device record (FP32)."""
        f = np.float32
        r = f(self.Sphere.radius[m])
        N = int(self.Sphere.NumIB[m])
        ang = f(2.0 * math.pi) * f(n) / f(N)
        xref, yref = r * f(np.cos(ang)), r * f(np.sin(ang))
        th = f(self.Sphere.angle[m])
        c, s = f(np.cos(th)), f(np.sin(th))
        px, py = (f(v) for v in self.Sphere.pos[m])
        return np.array([px + xref * c - yref * s, py + xref * s + yref * c], np.float32)

    def reset_force(self):
        """ibm.py:80-83."""
        H.check(H.lib().lbm_ib_reset_force(ctypes.byref(self.abi()), stream()), "reset_force")


class IBDEMCouplerEngine:
    """ibm.py:86-161."""

    def __init__(self, lb_field):
        self.NX = lb_field.NX
        self.NY = lb_field.NY

    def CheckPos(self, pos, diameter):
        """ibm.py:92-108 (host form of the kernel's test)."""
        f = np.float32
        x, y, d = f(pos[0]), f(pos[1]), f(diameter)
        x_min, x_max, y_min, y_max = x - d, x + d, y - d, y + d
        nx, ny = self.NX - 1, self.NY - 1
        if x_min >= 0 and x_max <= nx and y_min >= 0 and y_max <= ny:
            return 1
        if x_max < 0 or x_min > nx or y_max < 0 or y_min > ny:
            return 0
        return 2

    def InDomain(self, pos):
        """ibm.py:111-116."""
        return int(pos[0] > 0 and pos[0] < self.NX - 1 and pos[1] > 0 and pos[1] < self.NY - 1)

    def _lattice(self):
        L = H.Lattice()
        L.nx, L.ny, L.nc = self.NX, self.NY, 1
        return L

    def SphereIBUpdate(self, sphere_field):
        """ibm.py:120-141."""
        H.check(H.lib().lbm_ib_update(ctypes.byref(self._lattice()), ctypes.byref(sphere_field.abi()), stream()),
                "SphereIBUpdate")

    def SphereInfoUpdata(self, sphere_field, pos_np, vel_np, angle_np):
        """ibm.py:143-150."""
        n = int(sphere_field.num[None])
        S = sphere_field.Sphere
        S.pos[:n] = np.asarray(pos_np)[:n]
        S.vel[:n] = np.asarray(vel_np)[:n]
        S.angle[:n] = np.asarray(angle_np)[:n]
        self.SphereIBUpdate(sphere_field)

    def GetSphereForce(self, sphere_field):
        """ibm.py:152-155."""
        force_data = np.array(sphere_field.Sphere.force.to_numpy(), dtype=np.float64)
        return force_data[:sphere_field.num[None], :]

    def GetSphereTorque(self, sphere_field):
        """ibm.py:157-161."""
        torque_data = np.array(sphere_field.Sphere.torque.to_numpy(), dtype=np.float64)
        return torque_data[:sphere_field.num[None]]


class IBEngine:
    """ibm.py:166-253: multi-direct forcing with ``MDFIteration`` iterations."""

    def __init__(self, MDFIteration):
        self.MDFIteration = int(MDFIteration)

    def ComputeParticleForce(self, sphere_field):
        """ibm.py:171-185: per-sphere sums over its markers in the group (the
        reference resets the sums inside the marker loop, a race; here a sphere
        with markers gets their full sum)."""
        H.check(H.lib().lbm_ib_particle_force(ctypes.byref(sphere_field.abi()), stream()), "ComputeParticleForce")

    @staticmethod
    def BottomAndTop(pos):
        """ibm.py:187-194."""
        return np.array([math.floor(pos[0]), math.floor(pos[1]), math.ceil(pos[0]), math.ceil(pos[1])], np.int64)

    @staticmethod
    def weight(pos, iX, iY):
        """ibm.py:196-198."""
        return (1.0 - abs(pos[0] - iX)) * (1.0 - abs(pos[1] - iY))

    @staticmethod
    def _lattice_args(lb_filed):
        lb_filed.materialize()
        return ctypes.byref(lattice_struct(lb_filed)), ctypes.byref(gfields(lb_filed))

    def InterpolateBodyVelocities(self, sphere_field, lb_filed):
        """ibm.py:201-223."""
        L, G = self._lattice_args(lb_filed)
        H.check(H.lib().lbm_ib_interpolate(L, G, ctypes.byref(sphere_field.abi()), stream()),
                "InterpolateBodyVelocities")

    def SpreadBodyForces(self, sphere_field, lb_filed):
        """ibm.py:225-244."""
        L, G = self._lattice_args(lb_filed)
        H.check(H.lib().lbm_ib_spread(L, G, ctypes.byref(sphere_field.abi()), stream()), "SpreadBodyForces")

    def MultiDirectForcingLoop(self, sphere_field, lb_filed):
        """ibm.py:248-253: reset_force, then MDFIteration x (interpolate, spread)."""
        L, G = self._lattice_args(lb_filed)
        H.check(H.lib().lbm_ib_mdf(L, G, ctypes.byref(sphere_field.abi()), self.MDFIteration, stream()),
                "MultiDirectForcingLoop")
