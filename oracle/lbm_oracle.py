"""CPU oracle for the D2Q9 hot path of taichi-lbm -- TEST INFRASTRUCTURE ONLY.

This module restates, in NumPy, the per-time-step semantics of the reference
(Chiuzzmm/taichi-lbm) kernel by kernel, in the reference's own (unfused) order
and with the reference's data layout (fields indexed ``[i, j, component, k]``,
``i`` = x slowest).  It exists only to check the HIP implementation: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import it, and only as the checker.  The product path never calls it.

Every function runs in the dtype of the state it is given (float32 to mirror the
reference's ``float`` = f32 fields, float64 as the high-precision oracle).  All
arithmetic is written in the reference's operation order; Taichi may contract to
FMA and reorder, so bit-identity with Taichi is not claimed anywhere.

Pinning (see tests/test_oracle_*.py): every function is checked against closed
forms or invariants -- equilibrium moments, the exact MRT transform identities,
MRT/TRT <-> BGK equivalences, the rest-state fixed point, mass and momentum
budgets, Poiseuille (Newtonian and power-law) profiles, Taylor-Green decay,
NEEM/Zou-He identities, the Shan-Chen stencil on uniform/linear psi (one and
two components: F_c1 = -psi_c1 G[c1,c2] grad psi_c2 / 3, and the cross forces
summing to zero over a periodic lattice), the reference's own tabulated
Shan-Chen coexistence densities (cases/Bubble or two-phase separation.py:48-71;
pressure1 is equal at the tabulated gas and liquid densities), pressure0 as
the rest equilibrium's momentum flux, the pressure2 uniform-state form and its
single-component limit, the nc = 2 force split (shares sum to g; the total
momentum gains N g per step) and barycentric velocity, and the table stream
against a plain shift.  Nothing from the reference
could be executed here (Taichi is not installed), so no function is pinned by a
reference run: parity unpinned for Taichi's own FP32 rounding trajectory
(fast-math contraction, summation order) and for its ti.random() stream.
tests/test_planted_mistakes.py plants plausible slips here and checks that a
pin fails for each; tests/golden/*.npz freeze this module's outputs on eleven
lattices (written by tests/golden/make_golden.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

# ---------------------------------------------------------------------------
# D2Q9 constants -- lbm.py:43-45, BoundaryCondition.py:242, :63-72
# ---------------------------------------------------------------------------
C = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
              [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int64)
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int64)
UNKNOWN = {1: [3, 6, 7], 3: [1, 5, 8], 4: [2, 5, 6], 2: [4, 7, 8]}


def weights(dtype):
    """lbm.py:44 -- a ``float`` vector (4,1,1,1,1,1/4,...) divided by 9.0 in ``dtype``."""
    dt = np.dtype(dtype)
    base = np.array([4, 1, 1, 1, 1, 1 / 4, 1 / 4, 1 / 4, 1 / 4], dtype=dt)
    return base / dt.type(9.0)


# MRT transform, CollisionEngine.py:273-282.
M = np.array([
    [1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],
    [-4.0, -1.0, -1.0, -1.0, -1.0, 2.0, 2.0, 2.0, 2.0],
    [4.0, -2.0, -2.0, -2.0, -2.0, 1.0, 1.0, 1.0, 1.0],
    [0.0, 1.0, 0.0, -1.0, 0.0, 1.0, -1.0, -1.0, 1.0],
    [0.0, -2.0, 0.0, 2.0, 0.0, 1.0, -1.0, -1.0, 1.0],
    [0.0, 0.0, 1.0, 0.0, -1.0, 1.0, 1.0, -1.0, -1.0],
    [0.0, 0.0, -2.0, 0.0, 2.0, 1.0, 1.0, -1.0, -1.0],
    [0.0, 1.0, -1.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0],
    [0.0, 0.0, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0, -1.0]])
# The reference inverts M numerically (CollisionEngine.py:283); the rows of M are
# orthogonal, so the exact inverse is M^T D^-1 with D = diag(M M^T).
M_DIAG = np.array([9.0, 36.0, 36.0, 6.0, 12.0, 6.0, 12.0, 4.0, 4.0])
M_INV = M.T / M_DIAG[None, :]


def _cf(dt):
    return C.astype(dt)


# ---------------------------------------------------------------------------
# Per-node formulas (vectorised over a batch of nodes), CollisionEngine.py:14-34
# ---------------------------------------------------------------------------
def f_eq(vel, rho, dt):
    """Collision.f_eq, CollisionEngine.py:14-18.  vel (m,2), rho (m,) -> (m,9)."""
    cf = _cf(dt)
    w = weights(dt)
    vx = vel[:, 0:1]
    vy = vel[:, 1:2]
    eu = cf[None, :, 0] * vx + cf[None, :, 1] * vy
    uv = vx * vx + vy * vy
    return w[None, :] * rho[:, None] * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uv)


def f_force(F, vel, dt):
    """Collision.f_force (Guo term), CollisionEngine.py:20-25."""
    cf = _cf(dt)
    w = weights(dt)
    cF = cf[None, :, 0] * F[:, 0:1] + cf[None, :, 1] * F[:, 1:2]
    cu = cf[None, :, 0] * vel[:, 0:1] + cf[None, :, 1] * vel[:, 1:2]
    uF = vel[:, 0:1] * F[:, 0:1] + vel[:, 1:2] * F[:, 1:2]
    return w[None, :] * (3 * cF + 9 * cF * cu - 3 * uF)


def tau_max_min(tau, tau_min, tau_max):
    """Collision.tau_max_min, CollisionEngine.py:28-34 (NaN passes through)."""
    tau = np.where(tau < tau_min, tau_min, tau)
    return np.where(tau > tau_max, tau_max, tau).astype(tau.dtype)


# ---------------------------------------------------------------------------
# State (LBField, lbm.py:6-75)
# ---------------------------------------------------------------------------
@dataclass
class Lattice:
    NX: int
    NY: int
    nc: int
    dtype: type = np.float32
    gravity: tuple = (0.0, 0.0)
    time: int = 0

    def __post_init__(self):
        dt = np.dtype(self.dtype)
        NX, NY, nc = self.NX, self.NY, self.nc
        self.f = np.zeros((NX, NY, nc, 9), dt)
        self.f2 = np.zeros((NX, NY, nc, 9), dt)
        self.SCforce = np.zeros((NX, NY, nc, 2), dt)
        self.rho = np.zeros((NX, NY, nc), dt)
        self.pressure = np.zeros((NX, NY, nc), dt)
        self.body_force = np.zeros((NX, NY, nc, 2), dt)
        self.shear_rate = np.zeros((NX, NY, nc), dt)
        self.vel = np.zeros((NX, NY, 2), dt)
        self.total_rho = np.zeros((NX, NY), dt)
        self.total_pressure = np.zeros((NX, NY), dt)
        self.T = np.zeros((NX, NY), dt)
        self.neighbor = np.zeros((NX, NY, 9, 2), np.int32)
        self.mask = np.ones((NX, NY), np.int32)          # lbm.py:74 mask.fill(1)
        self.psi = None                                   # (NX+2, NY+2, nc), offset -1
        self.g = np.array(self.gravity, dt)

    @property
    def dt(self):
        return np.dtype(self.dtype)

    def copy(self):
        import copy
        return copy.deepcopy(self)


def group_ij(group):
    g = np.asarray(group, dtype=np.int64).reshape(-1, 2)
    return g[:, 0], g[:, 1]


# ---------------------------------------------------------------------------
# Geometry / classification -- BoundaryCondition.py:22-39, :464-475, lbm.py:127-142
# ---------------------------------------------------------------------------
def mask_rectangle_identify(lat: Lattice, xmin, xmax, ymin, ymax):
    """BoundaryEngine.Mask_rectangle_identify, BoundaryCondition.py:464-469 (strict, f32 compare)."""
    I, J = np.meshgrid(np.arange(lat.NX), np.arange(lat.NY), indexing="ij")
    f32 = np.float32
    hit = (I < f32(xmax)) & (I > f32(xmin)) & (J < f32(ymax)) & (J > f32(ymin))
    lat.mask[hit] = -1


def mask_circle_identify(lat: Lattice, x, y, r):
    """BoundaryEngine.Mask_cricle_identify, BoundaryCondition.py:471-475.

    The predicate is evaluated in f32 as Taichi does for ``float`` arguments:
    ``(ix-x)**2 + (iy-y)**2 < r**2``.
    """
    f32 = np.float32
    I, J = np.meshgrid(np.arange(lat.NX), np.arange(lat.NY), indexing="ij")
    dx = I.astype(f32) - f32(x)
    dy = J.astype(f32) - f32(y)
    hit = dx * dx + dy * dy < f32(r) * f32(r)
    lat.mask[hit] = -1


def classify_scalar(NX, NY, pred: Callable[[int, int], bool]):
    """BoundaryClassifier._classify_nodes, BoundaryCondition.py:33-39, as a scalar loop.

    Returns the group as an (m, 2) int array in (i, j) lexicographic order; the
    reference compacts with atomic_add so its order is arbitrary and groups are
    compared as sets.
    """
    out = [(i, j) for i in range(NX) for j in range(NY) if pred(i, j)]
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def neighbor_classify(lat: Lattice):
    """LBField.neighbor_classify, lbm.py:127-142 (non-periodic; solid nodes untouched)."""
    NX, NY = lat.NX, lat.NY
    for i in range(NX):
        for j in range(NY):
            if lat.mask[i, j] == -1:
                continue
            for k in range(9):
                i2, j2 = i - C[k, 0], j - C[k, 1]
                if 0 <= i2 < NX and 0 <= j2 < NY and lat.mask[i2, j2] != -1:
                    lat.neighbor[i, j, k] = (i2, j2)
                else:
                    lat.neighbor[i, j, k] = (-1, -1)


# ---------------------------------------------------------------------------
# Collision models -- CollisionEngine.py:36-451
# ---------------------------------------------------------------------------
TAU_MIN = 0.5000001   # CollisionEngine.py:10
TAU_MAX = 4.0         # CollisionEngine.py:11


@dataclass
class FluidModel:
    """NewtonianFluid / PowerLawFluid, FluidModels.py:9-29."""
    kind: str = "newtonian"
    nu: float = 0.0
    m: float = 0.0
    n: float = 1.0

    def apparent_viscosity(self, shear_rate, rho):
        dt = shear_rate.dtype
        if self.kind == "newtonian":
            return np.full_like(shear_rate, dt.type(self.nu))
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            return (dt.type(self.m) * np.power(shear_rate, dt.type(self.n - 1.0))).astype(dt)


@dataclass
class Collision:
    kind: str                      # "bgk" | "trt" | "mrt"
    group: np.ndarray
    fluid_model: FluidModel
    nu: Sequence[float]
    magic: Optional[Sequence[float]] = None
    nu_bulk: Optional[Sequence[float]] = None
    tau_min: float = TAU_MIN
    tau_max: float = TAU_MAX
    # HuangMRTCollision (CollisionEngine.py:453-515): per component k1,
    # epslion and G (the coupling of the component with itself)
    huang: Optional[dict] = None
    # state
    tau: Optional[np.ndarray] = None          # BGK tau / TRT tau_sym
    tau_antisym: Optional[np.ndarray] = None  # TRT
    diag: Optional[np.ndarray] = None         # MRT (NX,NY,nc,9)

    def init_tau(self, lat: Lattice):
        """initTau: BGK CollisionEngine.py:65-70, TRT :149-155, MRT :301-318."""
        dt = lat.dt
        I, J = group_ij(self.group)
        nc = len(self.nu)
        if self.kind == "bgk":
            self.tau = np.zeros((lat.NX, lat.NY, nc), dt)
            for c in range(nc):
                self.tau[I, J, c] = dt.type(self.nu[c]) * 3.0 + 0.5
        elif self.kind == "trt":
            self.tau = np.zeros((lat.NX, lat.NY, nc), dt)
            self.tau_antisym = np.zeros((lat.NX, lat.NY, nc), dt)
            for c in range(nc):
                ts = dt.type(self.nu[c]) * dt.type(3.0) + dt.type(0.5)
                self.tau[I, J, c] = ts
                self.tau_antisym[I, J, c] = dt.type(self.magic[c]) / (ts - dt.type(0.5)) + dt.type(0.5)
        elif self.kind == "mrt":
            self.diag = np.ones((lat.NX, lat.NY, nc, 9), dt)
            for c in range(nc):
                tau = tau_max_min(np.array(dt.type(self.nu[c]) * 3.0 + 0.5, dt),
                                  dt.type(self.tau_min), dt.type(self.tau_max))
                tb = tau_max_min(np.array(dt.type(self.nu_bulk[c]) * 3.0 + 0.5, dt),
                                 dt.type(self.tau_min), dt.type(self.tau_max))
                self.diag[I, J, c, 7] = dt.type(1.0) / tau
                self.diag[I, J, c, 8] = dt.type(1.0) / tau
                self.diag[I, J, c, 1] = dt.type(1.0) / tb
                self.diag[I, J, c, 0] = 0.0
                self.diag[I, J, c, 3] = 0.0
                self.diag[I, J, c, 5] = 0.0
        else:
            raise ValueError(self.kind)

    # -- shear rate / relaxation time --------------------------------------
    def shear_rate(self, lat: Lattice):
        """BGK :80-92, TRT :166-178, MRT :332-377."""
        dt = lat.dt
        I, J = group_ij(self.group)
        for c in range(lat.nc):
            rho = lat.rho[I, J, c]
            fneq = lat.f[I, J, c, :] - f_eq(lat.vel[I, J], rho, dt)
            if self.kind in ("bgk", "trt"):
                cof = dt.type(-1.5) / rho / self.tau[I, J, c]
                src = fneq
            else:
                mneq = mf(fneq)
                smneq = self.diag[I, J, c, :] * mneq
                src = minv_apply(smneq, dt)
                cof = dt.type(-1.5) / rho
            Sxx = (src[:, 1] + src[:, 3] + src[:, 5] + src[:, 6] + src[:, 7] + src[:, 8]) * cof
            Syy = (src[:, 2] + src[:, 4] + src[:, 5] + src[:, 6] + src[:, 7] + src[:, 8]) * cof
            Sxy = (src[:, 5] - src[:, 6] + src[:, 7] - src[:, 8]) * cof
            lat.shear_rate[I, J, c] = np.sqrt(2 * (Sxx ** 2 + Syy ** 2 + 2 * Sxy ** 2))

    def relaxation_time(self, lat: Lattice):
        """BGK :72-78, TRT :157-164, MRT :322-329."""
        dt = lat.dt
        I, J = group_ij(self.group)
        for c in range(lat.nc):
            nu = self.fluid_model.apparent_viscosity(lat.shear_rate[I, J, c], lat.rho[I, J, c])
            with np.errstate(over="ignore", invalid="ignore"):
                tau = nu * dt.type(3.0) + dt.type(0.5)
            tau = tau_max_min(tau, dt.type(self.tau_min), dt.type(self.tau_max))
            if self.kind == "bgk":
                self.tau[I, J, c] = tau
            elif self.kind == "trt":
                self.tau[I, J, c] = tau
                self.tau_antisym[I, J, c] = dt.type(self.magic[c]) / (tau - dt.type(0.5)) + dt.type(0.5)
            else:
                inv = dt.type(1.0) / tau
                self.diag[I, J, c, 7] = inv
                self.diag[I, J, c, 8] = inv

    # -- collide ------------------------------------------------------------
    def apply(self, lat: Lattice):
        dt = lat.dt
        I, J = group_ij(self.group)
        vel = lat.vel[I, J]
        for c in range(lat.nc):
            f = lat.f[I, J, c, :]
            rho = lat.rho[I, J, c]
            F = lat.body_force[I, J, c, :]
            if self.kind == "bgk":          # CollisionEngine.py:96-110
                omega = (dt.type(1.0) / self.tau[I, J, c])[:, None]
                feq = f_eq(vel, rho, dt)
                fF = f_force(F, vel, dt)
                coll = -omega * (f - feq)
                fterm = (1.0 - 0.5 * omega) * fF
                lat.f2[I, J, c, :] = f + coll + fterm
            elif self.kind == "trt":        # CollisionEngine.py:223-245
                ws = (dt.type(1.0) / self.tau[I, J, c])[:, None]
                wa = (dt.type(1.0) / self.tau_antisym[I, J, c])[:, None]
                feq = f_eq(vel, rho, dt)
                fF = f_force(F, vel, dt)
                fs, es, Fs = trt_sym(f), trt_sym(feq), trt_sym(fF)
                fa, ea, Fa = f - fs, feq - es, fF - Fs
                coll = -ws * (fs - es) - wa * (fa - ea)
                fterm = (1.0 - 0.5 * ws) * Fs + (1.0 - 0.5 * wa) * Fa
                lat.f2[I, J, c, :] = f + coll + fterm
            else:                           # CollisionEngine.py:424-451
                m = mf(f)
                meq = m_eq(vel, rho, dt)
                mF = m_force(F, vel, dt)
                S = self.diag[I, J, c, :]
                a = S * (m - meq)
                b = (1.0 - S / 2.0) * mF
                m2 = m - a + b
                if self.huang is not None:   # HuangMRTCollision.apply, :503-507
                    h = self.huang
                    k1 = dt.type(h["k1"][c])
                    k2 = dt.type(h["epslion"][c]) / dt.type(-8.0) - k1
                    Q = m_q(F, lat.psi[I + 1, J + 1, c], k1, k2, dt.type(h["g"][c]), dt)
                    m2 = m2 + S * Q
                lat.f2[I, J, c, :] = minv_apply(m2, dt)


def trt_sym(f):
    """TRTCollision.f_sym / feq_sym / force_sym, CollisionEngine.py:181-221."""
    s = np.empty_like(f)
    s[:, 0] = f[:, 0]
    s[:, 1] = (f[:, 1] + f[:, 3]) / 2
    s[:, 2] = (f[:, 2] + f[:, 4]) / 2
    s[:, 3] = s[:, 1]
    s[:, 4] = s[:, 2]
    s[:, 5] = (f[:, 5] + f[:, 7]) / 2
    s[:, 6] = (f[:, 6] + f[:, 8]) / 2
    s[:, 7] = s[:, 5]
    s[:, 8] = s[:, 6]
    return s


def mf(f):
    """MRTCollision.mf, CollisionEngine.py:379-385: m = M f (M has small-integer entries)."""
    return f @ M.astype(f.dtype).T


def minv_apply(m, dt):
    """f = M^-1 m with the exact inverse M^T D^-1 cast to ``dt`` (CollisionEngine.py:447-449)."""
    return m @ M_INV.astype(dt).T


def m_eq(vel, rho, dt):
    """MRTCollision.m_eq, CollisionEngine.py:388-406."""
    jx = vel[:, 0]
    jy = vel[:, 1]
    meq = np.empty((vel.shape[0], 9), dt)
    meq[:, 0] = 1.0
    meq[:, 1] = -2.0 + 3.0 * (jx ** 2 + jy ** 2)
    meq[:, 2] = 1.0 - 3.0 * (jx ** 2 + jy ** 2)
    meq[:, 3] = jx
    meq[:, 4] = -jx
    meq[:, 5] = jy
    meq[:, 6] = -jy
    meq[:, 7] = jx ** 2 - jy ** 2
    meq[:, 8] = jx * jy
    return meq * rho[:, None]


def m_q(F, psi, k1, k2, g, dt):
    """HuangMRTCollision.m_Q, CollisionEngine.py:466-478."""
    Q = np.zeros((F.shape[0], 9), dt)
    F2 = F[:, 0] * F[:, 0] + F[:, 1] * F[:, 1]
    a = 3 * (k1 + 2 * k2) * F2
    b = g * psi ** 2
    Q[:, 1] = a / b
    Q[:, 2] = -Q[:, 1]
    Q[:, 7] = k1 * (F[:, 0] ** 2 - F[:, 1] ** 2) / b
    Q[:, 8] = k1 * F[:, 0] * F[:, 1] / b
    return Q


def m_force(F, vel, dt):
    """MRTCollision.m_force, CollisionEngine.py:408-422."""
    mF = np.zeros((vel.shape[0], 9), dt)
    Fu = F[:, 0] * vel[:, 0] + F[:, 1] * vel[:, 1]
    mF[:, 1] = 6 * Fu
    mF[:, 2] = -6 * Fu
    mF[:, 3] = F[:, 0]
    mF[:, 4] = -F[:, 0]
    mF[:, 5] = F[:, 1]
    mF[:, 6] = -F[:, 1]
    mF[:, 7] = 2 * (F[:, 0] * vel[:, 0] - F[:, 1] * vel[:, 1])
    mF[:, 8] = F[:, 1] * vel[:, 0] + F[:, 0] * vel[:, 1]
    return mF


# ---------------------------------------------------------------------------
# MacroscopicEngine -- lbm.py:146-237; LBField.init_LBM lbm.py:119-125
# ---------------------------------------------------------------------------
def init_lbm(lat: Lattice, group):
    I, J = group_ij(group)
    for c in range(lat.nc):
        fe = f_eq(lat.vel[I, J], lat.rho[I, J, c], lat.dt)
        lat.f[I, J, c, :] = fe
        lat.f2[I, J, c, :] = fe


def density(lat: Lattice, group):
    """lbm.py:153-164: fill 0, then rho_c = sum_k f (k in order), total += rho_c."""
    lat.total_rho.fill(0)
    lat.rho.fill(0)
    I, J = group_ij(group)
    tot = np.zeros(len(I), lat.dt)
    for c in range(lat.nc):
        f = lat.f[I, J, c, :]
        s = np.zeros(len(I), lat.dt)
        for k in range(9):
            s = s + f[:, k]
        lat.rho[I, J, c] = s
        tot = tot + s
    lat.total_rho[I, J] = tot


def pressure(lat: Lattice, group, sc=None):
    """MacroscopicEngine.pressure -> pressure0/1/2, lbm.py:166-212."""
    I, J = group_ij(group)
    lat.pressure.fill(0)
    dt = lat.dt
    if sc is None:
        lat.pressure[I, J, 0] = lat.pressure[I, J, 0] + lat.rho[I, J, 0] / 3.0
        lat.total_pressure[I, J] = lat.pressure[I, J, 0]
    elif lat.nc == 1:
        psi = lat.psi[I + 1, J + 1, 0]
        lat.pressure[I, J, 0] = lat.pressure[I, J, 0] + (lat.rho[I, J, 0] + 0.5 * dt.type(sc.g_coh) * psi ** 2) / 3.0
        lat.total_pressure[I, J] = lat.pressure[I, J, 0]
    else:
        G = np.asarray(sc.g_coh, dt)
        for c1 in range(lat.nc):
            psi = lat.psi[I + 1, J + 1, c1]
            lat.pressure[I, J, c1] = lat.pressure[I, J, c1] + (lat.rho[I, J, c1] + 0.5 * G[c1, c1] * psi ** 2) / 3.0
        p_ij = np.zeros(len(I), dt)
        p_ii = np.zeros(len(I), dt)
        for c1 in range(lat.nc):
            p_ii = p_ii + lat.pressure[I, J, c1]
            for c2 in range(c1 + 1, lat.nc):
                p_ij = p_ij + G[c1, c2] * lat.psi[I + 1, J + 1, c1] * lat.psi[I + 1, J + 1, c2]
        lat.total_pressure[I, J] = p_ii + p_ij / 6.0


def force_density(lat: Lattice, group):
    """lbm.py:215-220: F_c = SCforce_c + rho_c/rho_tot * g."""
    I, J = group_ij(group)
    for c in range(lat.nc):
        frac = (lat.rho[I, J, c] / lat.total_rho[I, J])[:, None]
        lat.body_force[I, J, c, :] = lat.SCforce[I, J, c, :] + frac * lat.g[None, :]


def velocity(lat: Lattice, group):
    """lbm.py:223-237: barycentric velocity with half-force shift."""
    lat.vel.fill(0)
    I, J = group_ij(group)
    dt = lat.dt
    cf = _cf(dt)
    acc = np.zeros((len(I), 2), dt)
    for c in range(lat.nc):
        f = lat.f[I, J, c, :]
        vx = np.zeros(len(I), dt)
        vy = np.zeros(len(I), dt)
        for k in range(9):
            vx = vx + f[:, k] * cf[k, 0]
            vy = vy + f[:, k] * cf[k, 1]
        vt = np.stack([vx, vy], axis=1) + 0.5 * lat.body_force[I, J, c, :]
        acc = acc + vt
    lat.vel[I, J] = acc / lat.total_rho[I, J][:, None]


# ---------------------------------------------------------------------------
# Boundary conditions -- BoundaryCondition.py:104-259, :419-456
# ---------------------------------------------------------------------------
def rhow(direction, vel, f):
    """NEEMMethod.rhow, BoundaryCondition.py:118-129.  vel (2,) or (m,2); f (m,9)."""
    vel = np.broadcast_to(vel, (f.shape[0], 2))
    if direction == 2:
        return 1. / (1. + vel[:, 1]) * (f[:, 0] + f[:, 1] + f[:, 3] + 2 * (f[:, 2] + f[:, 5] + f[:, 6]))
    if direction == 4:
        return 1. / (1. - vel[:, 1]) * (f[:, 0] + f[:, 1] + f[:, 3] + 2 * (f[:, 4] + f[:, 7] + f[:, 8]))
    if direction == 1:
        return 1. / (1. + vel[:, 0]) * (f[:, 0] + f[:, 2] + f[:, 4] + 2 * (f[:, 1] + f[:, 5] + f[:, 8]))
    if direction == 3:
        return 1. / (1. - vel[:, 0]) * (f[:, 0] + f[:, 2] + f[:, 4] + 2 * (f[:, 3] + f[:, 6] + f[:, 7]))
    return np.zeros(f.shape[0], f.dtype)


def neem(vel1, rho1, vel2, rho2, fnb, dt):
    """NEEMMethod.NEEM, BoundaryCondition.py:112-116: feq(u1,rho1) + f2 - feq(u2,rho2)."""
    return f_eq(vel1, rho1, dt) + fnb - f_eq(vel2, rho2, dt)


@dataclass
class BC:
    """One entry of BoundaryEngine's ordered dict (BoundaryCondition.py:459-483)."""
    kind: str            # fluid | periodic | inside | bounceback | velocity_neem | pressure_neem |
                         # velocity_bb | open_neumann | velocity_eq | pressure_abc
    group: np.ndarray
    direction: Optional[int] = None
    value: object = None  # velocity (2,) or rho scalar

    def apply(self, lat: Lattice):
        I, J = group_ij(self.group)
        dt = lat.dt
        if len(I) == 0 or self.kind == "fluid":
            return
        if self.kind == "periodic":          # :424-432
            for k in range(9):
                x2 = (I - C[k, 0] + lat.NX) % lat.NX
                y2 = (J - C[k, 1] + lat.NY) % lat.NY
                lat.f[I, J, :, k] = lat.f2[x2, y2, :, k]
        elif self.kind == "inside":          # :440-449 -- no -1 guard: poison with NaN
            for k in range(9):
                nb = lat.neighbor[I, J, k]
                ok = nb[:, 0] != -1
                vals = np.full((len(I), lat.nc), np.nan, dt)
                vals[ok] = lat.f2[nb[ok, 0], nb[ok, 1], :, k]
                lat.f[I, J, :, k] = vals
        elif self.kind == "bounceback":      # :244-259
            for k in range(9):
                nb = lat.neighbor[I, J, k]
                ok = nb[:, 0] != -1
                vals = lat.f2[I, J, :, OPP[k]].copy()
                vals[ok] = lat.f2[nb[ok, 0], nb[ok, 1], :, k]
                lat.f[I, J, :, k] = vals
        elif self.kind == "velocity_neem":   # :154-166
            d = self.direction
            uw = np.asarray(self.value, dt)
            I2, J2 = I - C[d, 0], J - C[d, 1]
            lat.vel[I, J] = uw
            for c in range(lat.nc):
                lat.rho[I, J, c] = rhow(d, uw, lat.f[I, J, c, :])
            for c in range(lat.nc):
                lat.f[I, J, c, :] = neem(lat.vel[I, J], lat.rho[I, J, c], lat.vel[I2, J2],
                                         lat.rho[I2, J2, c], lat.f[I2, J2, c, :], dt)
        elif self.kind == "pressure_neem":   # :215-226
            d = self.direction
            I2, J2 = I - C[d, 0], J - C[d, 1]
            lat.vel[I, J] = lat.vel[I2, J2]
            lat.total_rho[I, J] = dt.type(self.value)
            unk = np.array(UNKNOWN[d])
            for c in range(lat.nc):
                full = neem(lat.vel[I, J], lat.rho[I, J, c], lat.vel[I2, J2], lat.rho[I2, J2, c],
                            lat.f[I2, J2, c, :], dt)
                lat.f[I[:, None], J[:, None], c, unk[None, :]] = full[:, unk]
        elif self.kind == "velocity_eq":     # VelocityEquilibrium :132-145: f = feq(u_w, rho_w), all 9
            d = self.direction
            uw = np.asarray(self.value, dt)
            lat.vel[I, J] = uw
            for c in range(lat.nc):
                lat.rho[I, J, c] = rhow(d, uw, lat.f[I, J, c, :])
            for c in range(lat.nc):
                lat.f[I, J, c, :] = f_eq(lat.vel[I, J], lat.rho[I, J, c], dt)
        elif self.kind == "pressure_abc":    # PressureABC :177-205 (anti-bounce-back on the unknowns)
            d = self.direction
            I2, J2 = I - C[d, 0], J - C[d, 1]
            uw = 1.5 * lat.vel[I, J] - 0.5 * lat.vel[I2, J2]
            lat.vel[I, J] = uw
            lat.total_rho[I, J] = dt.type(self.value)
            w = weights(dt)
            cf = _cf(dt)
            cu = uw @ cf.T                                    # (m, 9)
            uw2 = (uw * uw).sum(axis=1)[:, None]
            unk = np.array(UNKNOWN[d])
            for c in range(lat.nc):
                ftemp = -lat.f[I, J, c, :] + 2 * w[None, :] * dt.type(self.value) * (1.0 + 4.5 * cu * cu - 1.5 * uw2)
                lat.f[I[:, None], J[:, None], c, unk[None, :]] = ftemp[:, unk]
        elif self.kind == "velocity_bb":     # :85-102 (additive on the unknowns)
            d = self.direction
            uw = np.asarray(self.value, dt)
            w = weights(dt)
            for c in range(lat.nc):
                rho_w = lat.rho[I, J, c]
                for k in UNKNOWN[d]:
                    cv = dt.type(C[k, 0]) * uw[0] + dt.type(C[k, 1]) * uw[1]
                    lat.f[I, J, c, k] = lat.f[I, J, c, k] + 6 * w[k] * rho_w * cv
        elif self.kind == "open_extrapolation":   # :399-413, density / vel not recomputed (:412-413)
            d = self.direction
            I2, J2 = I - C[d, 0], J - C[d, 1]
            I3, J3 = I2 - C[d, 0], J2 - C[d, 1]
            for c in range(lat.nc):
                for k in UNKNOWN[d]:
                    lat.f[I, J, c, k] = 2 * lat.f[I2, J2, c, k] - lat.f[I3, J3, c, k]
        elif self.kind in ("open_neumann", "open_convective1"):   # :321-334 / :343-359, density/vel :274-298
            d = self.direction
            I2, J2 = I - C[d, 0], J - C[d, 1]
            if self.kind == "open_neumann":
                for c in range(lat.nc):
                    for k in UNKNOWN[d]:
                        lat.f[I, J, c, k] = lat.f[I2, J2, c, k]
            else:
                U = lat.vel[I2, J2]
                lam = np.sqrt(U[:, 0] * U[:, 0] + U[:, 1] * U[:, 1])     # lambdda = 1. * |u(x - c_d)|
                for c in range(lat.nc):
                    for k in UNKNOWN[d]:
                        lat.f[I, J, c, k] = (lat.f[I, J, c, k] + lam * lat.f[I2, J2, c, k]) / (1.0 + lam)
            tot = np.zeros(len(I), dt)
            for c in range(lat.nc):
                r = np.zeros(len(I), dt)
                for k in range(9):
                    r = r + lat.f[I, J, c, k]
                lat.rho[I, J, c] = r
                tot = tot + r
            lat.total_rho[I, J] = tot
            cf = _cf(dt)
            acc = np.zeros((len(I), 2), dt)
            for c in range(lat.nc):
                vx = np.zeros(len(I), dt)
                vy = np.zeros(len(I), dt)
                for k in range(9):
                    vx = vx + lat.f[I, J, c, k] * cf[k, 0]
                    vy = vy + lat.f[I, J, c, k] * cf[k, 1]
                acc = acc + np.stack([vx, vy], 1) + 0.5 * lat.body_force[I, J, c, :]
            lat.vel[I, J] = acc / tot[:, None]
        else:
            raise ValueError(self.kind)


def apply_bcs(lat: Lattice, bcs: List[BC]):
    """BoundaryEngine.apply_boundary_conditions, BoundaryCondition.py:481-483 (insertion order)."""
    for bc in bcs:
        bc.apply(lat)


# ---------------------------------------------------------------------------
# Shan-Chen -- ExternalForce.py:33-117, EOS.py:40-47, SCBoundaryCondition.py:126-160
# ---------------------------------------------------------------------------
def sc_psi(rho, rho0):
    """SC_psi.get_psi, EOS.py:40-47."""
    dt = rho.dtype
    return dt.type(rho0) * (1.0 - np.exp(-rho / dt.type(rho0)))


@dataclass
class EOS:
    """Pseudopotential.get_psi with a cubic equation of state, EOS.py:5-95:
    psi = sqrt(|6/G (p_EOS(rho, T) - rho/3)|).  kind: "vdw" (:50-58), "rk"
    (:61-69), "pr" (:72-83, omega, T_cr), "cs" (:86-95)."""
    kind: str
    a: float
    b: float
    R: float = 1.0
    G: float = 1.0
    omega: float = 0.0
    T_cr: float = 1.0

    def pressure(self, rho, T):
        dt = rho.dtype.type
        a, b, R = dt(self.a), dt(self.b), dt(self.R)
        if self.kind == "vdw":
            return rho * R * T / (1.0 - b * rho) - a * rho ** 2.0
        if self.kind == "rk":
            return rho * R * T / (1.0 - b * rho) - a * rho ** 2.0 / np.sqrt(T) / (1.0 + b * rho)
        if self.kind == "pr":
            w = dt(self.omega)
            alpha = (1 + (0.37464 + 1.54226 * w - 0.269922 * w ** 2) * (1 - np.sqrt(T / dt(self.T_cr)))) ** 2
            return rho * R * T / (1.0 - b * rho) - a * alpha * rho ** 2 / (1.0 + 2.0 * b * rho - (b * rho) ** 2)
        if self.kind == "cs":
            x = b * rho / 4.0
            return rho * R * T * (1.0 + x + x ** 2 - x ** 3) / (1 - x) ** 3 - a * rho ** 2
        raise ValueError(self.kind)

    def get_psi(self, rho, T):
        dt = rho.dtype.type
        return np.sqrt(np.abs(dt(6.0) / dt(self.G) * (self.pressure(rho, T) - rho / 3.0)))


@dataclass
class ShanChen:
    kind: str                        # "c1" | "c2"
    group: np.ndarray
    g_coh: object                    # scalar (c1) or (nc,nc) matrix (c2)
    rho0: Sequence[float] = (1.0,)   # SC_psi rho0 per psi object (c1: one; c2: one per component)
    g_adh: Optional[Sequence[float]] = None
    fluid: Optional[np.ndarray] = None   # c2: fluid_strategy(x, y) over the halo grid (NX+2, NY+2),
                                         # offset -1; None: fluid everywhere
    eos: Optional[Sequence[Optional[EOS]]] = None   # per psi object: a cubic EOS in place of SC_psi

    def _psi(self, q, rho, T):
        e = self.eos[q] if self.eos is not None else None
        return sc_psi(rho, self.rho0[q]) if e is None else e.get_psi(rho, T)

    def update_psi(self, lat: Lattice):
        """C1 ExternalForce.py:42-47; C2 :86-92 (psi objects: EOS.py)."""
        I, J = group_ij(self.group)
        T = lat.T[I, J]
        if self.kind == "c1":
            for c in range(lat.nc):
                lat.psi[I + 1, J + 1, c] = self._psi(0, lat.rho[I, J, c], T)
        else:
            lat.psi[I + 1, J + 1, 0] = self._psi(0, lat.rho[I, J, 0], T)
            lat.psi[I + 1, J + 1, 1] = self._psi(1, lat.rho[I, J, 1], T)

    def cau_scforce(self, lat: Lattice):
        """C1 ExternalForce.py:78-91; C2 :124-145 (fluid_strategy == everywhere fluid)."""
        dt = lat.dt
        lat.SCforce.fill(0)
        I, J = group_ij(self.group)
        w = weights(dt)
        cf = _cf(dt)
        wc = w[:, None] * cf
        if self.kind == "c1":
            F = np.zeros((len(I), 2), dt)
            for k in range(1, 9):
                F = F + wc[k][None, :] * lat.psi[I + 1 + C[k, 0], J + 1 + C[k, 1], 0][:, None]
            F = F * (-dt.type(self.g_coh) * lat.psi[I + 1, J + 1, 0])[:, None]
            lat.SCforce[I, J, 0, :] = lat.SCforce[I, J, 0, :] + F
        else:
            # :96-117: a neighbour for which fluid_strategy is false contributes
            # the adhesion w_k c_k gadh[c1] (without psi(x + c_k): the factor is
            # commented out at :114) in place of the cohesion terms
            G = np.asarray(self.g_coh, dt)
            gadh = np.zeros(lat.nc, dt) if self.g_adh is None else np.asarray(self.g_adh, dt)
            for c1 in range(lat.nc):
                Fc = np.zeros((len(I), 2), dt)
                Fa = np.zeros((len(I), 2), dt)
                for k in range(1, 9):
                    I2, J2 = I + 1 + C[k, 0], J + 1 + C[k, 1]
                    fl = np.ones(len(I), bool) if self.fluid is None else self.fluid[I2, J2]
                    for c2 in range(lat.nc):
                        if c2 != c1:
                            add = wc[k][None, :] * lat.psi[I2, J2, c2][:, None] * G[c1, c2]
                            Fc = np.where(fl[:, None], Fc + add, Fc)
                    Fa = np.where(fl[:, None], Fa, Fa + wc[k][None, :] * gadh[c1])
                lat.SCforce[I, J, c1, :] = (-lat.psi[I + 1, J + 1, c1])[:, None] * (Fc + Fa)


@dataclass
class PsiBC:
    """One entry of BoundaryEngineSC (SCBoundaryCondition.py:149-160) other than
    the periodic pair.  nodes: (m, 2) halo-grid coordinates (i, j) in
    [-1, NX] x [-1, NY]; direction d: the interior neighbour is x - c_d.
    kind: "neumann" (OpenNeumannPsi :19-33), "convective1" (:36-58),
    "convective2" (:61-89), "extrapolation" (:92-107), "fixed" (FixedPsi
    :110-123, component 0 only)."""
    kind: str
    nodes: np.ndarray
    direction: Optional[int] = None
    value: float = 0.0

    def apply(self, lat: Lattice):
        p = lat.psi
        I = np.asarray(self.nodes, np.int64)[:, 0] + 1
        J = np.asarray(self.nodes, np.int64)[:, 1] + 1
        dt = p.dtype.type
        if self.kind == "fixed":
            p[I, J, 0] = dt(self.value)
            return
        d = self.direction
        I2, J2 = I - C[d, 0], J - C[d, 1]
        if self.kind == "neumann":
            p[I, J, :] = p[I2, J2, :]
        elif self.kind == "extrapolation":
            I3, J3 = I2 - C[d, 0], J2 - C[d, 1]
            p[I, J, :] = 2 * p[I2, J2, :] - p[I3, J3, :]
        elif self.kind in ("convective1", "convective2"):
            # U = mean velocity of the interior neighbours (lattice indices), lambda = |U|
            U = lat.vel[I2 - 1, J2 - 1].sum(axis=0) / dt(len(I))
            lam = np.sqrt(U @ U)
            if self.kind == "convective1":
                p[I, J, :] = (p[I, J, :] + lam * p[I2, J2, :]) / (1.0 + lam)
            else:
                I3, J3 = I - 2 * C[d, 0], J - 2 * C[d, 1]
                p[I, J, :] = (p[I, J, :] + 2 * lam * p[I2, J2, :] - 0.5 * lam * p[I3, J3, :]) / (1.0 + 1.5 * lam)
        else:
            raise ValueError(self.kind)


def periodic_psi(lat: Lattice):
    """PeriodicPsi_LR then PeriodicPsi_BT, SCBoundaryCondition.py:126-144."""
    p = lat.psi
    NX, NY = lat.NX, lat.NY
    p[0, :, :] = p[NX, :, :]        # psi[-1, y] = psi[NX-1, y]
    p[NX + 1, :, :] = p[1, :, :]    # psi[NX, y] = psi[0, y]
    p[:, 0, :] = p[:, NY, :]        # psi[x, -1] = psi[x, NY-1]
    p[:, NY + 1, :] = p[:, 1, :]    # psi[x, NY] = psi[x, 0]


# ---------------------------------------------------------------------------
# One lbm_solve() step in the reference's call order (cases/*.py lbm_solve)
# ---------------------------------------------------------------------------
@dataclass
class Solver:
    lat: Lattice
    group: np.ndarray
    collision: Collision
    bcs: List[BC]
    power_law: bool = False
    sc: Optional[ShanChen] = None
    advance_time: bool = False
    psi_bcs: Optional[List] = None   # BoundaryEngineSC list: "periodic" (LR then BT) and PsiBC entries

    def step(self):
        lat = self.lat
        density(lat, self.group)
        # cases/Bubble or two-phase separation.py:199-204: pressure runs before
        # update_psi, so it sees the psi of the previous step.
        pressure(lat, self.group, self.sc)
        if self.sc is not None:
            self.sc.update_psi(lat)
            for bc in (self.psi_bcs if self.psi_bcs is not None else ["periodic"]):
                if bc == "periodic":
                    periodic_psi(lat)
                else:
                    bc.apply(lat)
            self.sc.cau_scforce(lat)
        force_density(lat, self.group)
        velocity(lat, self.group)
        if self.power_law:
            self.collision.shear_rate(lat)
            self.collision.relaxation_time(lat)
        self.collision.apply(lat)
        apply_bcs(lat, self.bcs)
        if self.advance_time:
            lat.time += 1

    def run(self, n):
        for _ in range(n):
            self.step()
