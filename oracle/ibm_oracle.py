"""CPU oracle for the immersed-boundary coupling of taichi-lbm -- TEST INFRASTRUCTURE ONLY.

Restates, in NumPy, the sphere / IB-node kernels of the reference's
``taichi_lbm/ibm.py`` (SphereIB, IBDEMCouplerEngine, IBEngine) with the
reference's field layout: sphere arrays indexed ``[m]``, IB-node arrays
``[m, n]``, lattice fields ``vel[i, j, axis]``, ``rho[i, j, c]``,
``total_rho[i, j]``, ``body_force[i, j, c, axis]``.  Only ``tests/`` may import
it, as the checker of the HIP kernels; the product path never calls it.

Arithmetic runs in the dtype of the arrays given (float32 mirrors the
reference's ``float`` fields) in the reference's operation order.  The
reference accumulates spread forces with parallel atomic adds, so its sums
have no fixed order; here they run sequentially in group order (np.add.at).

Pins (tests/test_ibm_oracle.py): marker positions on the circle (radius,
spacing, rotation), NumIB = round(2 pi r / cof), CheckPos / InDomain on their
boundary cases, bilinear interpolation exact on linear velocity fields,
weights summing to one, spreading conserving the total force, spreading the
transpose of interpolation (sum F . I(u) = sum u . S(F)), the closed form of
one multi-direct-forcing iteration for an isolated marker, and the force /
torque sums for uniform and tangential node forces;
tests/test_planted_mistakes.py plants slips here (transposed stencil,
dropped factor 2, torque sign, no half on the velocity correction) and checks
that a pin fails for each.  Parity unpinned: Taichi's rounding of ti.round at
exact halves (2 pi r / cof is never a half for the radii tested) and the
order of its atomic sums.

Deviation, stated: ComputeParticleForce (ibm.py:171-185) resets the sphere's
force and torque inside the loop over IB nodes; run in parallel that is a
race, run serially it keeps only the last node's share.  Here (and on the
GPU) every sphere with at least one node in the group gets the sum over its
nodes; spheres without one keep their previous force and torque, as in the
reference.
"""
from __future__ import annotations

import math

import numpy as np

# ibm.py:221: ForceCorrection = 2 * (vel_desired - vel)
FORCE_GAIN = 2.0
# ibm.py:241: r2 = total_rho * 2.0 (the velocity takes half the spread force)
VEL_SHARE = 2.0


def torque_arm(fx, fy, rx, ry):
    """ibm.py:183: torque += fx * ry - fy * rx."""
    return fx * ry - fy * rx


class Spheres:
    """SphereIB's state (ibm.py:29-41): Sphere.field(Max_Number) and
    IBMNodes.field(Max_Number, Max_NumIB), as plain arrays."""

    def __init__(self, max_number: int, max_num_ib: int, dtype=np.float32):
        dt = np.dtype(dtype)
        self.dtype = dt
        self.max_number, self.max_num_ib = int(max_number), int(max_num_ib)
        M, K = self.max_number, self.max_num_ib
        self.num = 0
        self.numib_cof = 1
        self.id = np.zeros(M, np.int32)
        self.pos = np.zeros((M, 2), dt)
        self.vel = np.zeros((M, 2), dt)
        self.force = np.zeros((M, 2), dt)
        self.torque = np.zeros(M, dt)
        self.radius = np.zeros(M, dt)
        self.angle = np.zeros(M, dt)
        self.num_ib = np.zeros(M, np.int32)
        self.node_pos = np.zeros((M, K, 2), dt)
        self.node_vel = np.zeros((M, K, 2), dt)
        self.node_vel_desired = np.zeros((M, K, 2), dt)
        self.node_force = np.zeros((M, K, 2), dt)
        self.node_force_temp = np.zeros((M, K, 2), dt)
        self.group = np.zeros((M * K, 2), np.int32)
        self.count = 0

    def init_balls(self, num, id_np, pos_np, vel_np, force_np, torque_np, radius_np, angle_np):
        """SphereIB.init_Balls, ibm.py:43-55."""
        dt = self.dtype
        self.num = int(num)
        s = slice(0, self.num)
        self.id[s] = np.asarray(id_np)[:self.num]
        self.pos[s] = np.asarray(pos_np, dt)[:self.num]
        self.vel[s] = np.asarray(vel_np, dt)[:self.num]
        self.force[s] = np.asarray(force_np, dt)[:self.num]
        self.torque[s] = np.asarray(torque_np, dt)[:self.num]
        self.radius[s] = np.asarray(radius_np, dt)[:self.num]
        self.angle[s] = np.asarray(angle_np, dt)[:self.num]


def num_ib(radius, cof, dtype=np.float32):
    """ibm.py:61: int(round(pi * 2 * radius / cof)) in the field dtype."""
    dt = np.dtype(dtype).type
    x = dt(math.pi) * dt(2) * dt(radius) / dt(cof)
    return int(np.floor(x + dt(0.5)))


def _marker_angle(n, N, dt):
    """ibm.py:63 / :72: 2.0 * math.pi * n / NumIB."""
    return dt(2.0 * math.pi) * dt(n) / dt(N)


def init_ib_nodes(sp: Spheres):
    """SphereIB.init_IB_nodes, ibm.py:57-67 (no rotation: the sphere angle is
    not applied here)."""
    dt = sp.dtype.type
    for m in range(sp.num):
        N = num_ib(sp.radius[m], sp.numib_cof, sp.dtype)
        if N > sp.max_num_ib:
            raise ValueError(f"sphere {m} needs {N} IB nodes, more than Max_NumIB = {sp.max_num_ib}")
        sp.num_ib[m] = N
        for n in range(max(N, 0)):
            a = _marker_angle(n, N, dt)
            sp.node_pos[m, n, 0] = sp.pos[m, 0] + sp.radius[m] * dt(np.cos(a))
            sp.node_pos[m, n, 1] = sp.pos[m, 1] + sp.radius[m] * dt(np.sin(a))


def node_pos(sp: Spheres, m, n):
    """SphereIB.computer_IB_nodes_pos, ibm.py:69-78: the marker rotated by the
    sphere's angle about its centre."""
    dt = sp.dtype.type
    a = _marker_angle(n, sp.num_ib[m], dt)
    xref = sp.radius[m] * dt(np.cos(a))
    yref = sp.radius[m] * dt(np.sin(a))
    th = sp.angle[m]
    c, s = dt(np.cos(th)), dt(np.sin(th))
    x = sp.pos[m, 0] + xref * c - yref * s
    y = sp.pos[m, 1] + xref * s + yref * c
    return np.array([x, y], sp.dtype)


def check_pos(pos, diameter, NX, NY):
    """IBDEMCouplerEngine.CheckPos, ibm.py:92-108: 1 inside, 0 outside, 2 crossing."""
    x_min, x_max = pos[0] - diameter, pos[0] + diameter
    y_min, y_max = pos[1] - diameter, pos[1] + diameter
    nx, ny = NX - 1, NY - 1
    if x_min >= 0 and x_max <= nx and y_min >= 0 and y_max <= ny:
        return 1
    if x_max < 0 or x_min > nx or y_max < 0 or y_min > ny:
        return 0
    return 2


def in_domain(pos, NX, NY):
    """IBDEMCouplerEngine.InDomain, ibm.py:111-116 (strict)."""
    return int(pos[0] > 0 and pos[0] < NX - 1 and pos[1] > 0 and pos[1] < NY - 1)


def sphere_ib_update(sp: Spheres, NX, NY):
    """IBDEMCouplerEngine.SphereIBUpdate, ibm.py:120-141.  The group is built
    in (m, n) order (the reference's atomic compaction has no fixed order)."""
    dt = sp.dtype.type
    sp.group[:] = 0
    entries = []
    for m in range(sp.num):
        flag = check_pos(sp.pos[m], dt(2) * sp.radius[m], NX, NY)
        if flag == 0:
            continue
        for n in range(min(sp.num_ib[m], sp.max_num_ib)):
            sp.node_pos[m, n] = node_pos(sp, m, n)
            sp.node_vel_desired[m, n] = sp.vel[m]
            if flag == 1 or in_domain(sp.node_pos[m, n], NX, NY):
                entries.append((m, n))
    sp.count = len(entries)
    if entries:
        sp.group[:len(entries)] = np.array(entries, np.int32)


def reset_force(sp: Spheres):
    """SphereIB.reset_force, ibm.py:80-83 (every (m, n) of the capacity)."""
    sp.node_force[:] = 0


def stencil(pos):
    """IBEngine.BottomAndTop, ibm.py:187-194 -> the (iX, iY) of
    ti.ndrange((floor x, ceil x + 1), (floor y, ceil y + 1)), iX outer."""
    x0, y0 = int(np.floor(pos[0])), int(np.floor(pos[1]))
    x1, y1 = int(np.ceil(pos[0])), int(np.ceil(pos[1]))
    return [(ix, iy) for ix in range(x0, x1 + 1) for iy in range(y0, y1 + 1)]


def weight(pos, ix, iy):
    """IBEngine.weight, ibm.py:196-198."""
    dt = type(pos[0])
    return (dt(1.0) - abs(pos[0] - dt(ix))) * (dt(1.0) - abs(pos[1] - dt(iy)))


def _groups(sp):
    return [tuple(sp.group[g]) for g in range(sp.count)]


def interpolate(sp: Spheres, vel):
    """IBEngine.InterpolateBodyVelocities, ibm.py:201-223: node velocity by
    bilinear interpolation of the lattice velocity, force_temp = 2 (u_d - u),
    force += force_temp."""
    dt = sp.dtype.type
    for m, n in _groups(sp):
        p = sp.node_pos[m, n]
        v = np.zeros(2, sp.dtype)
        for ix, iy in stencil(p):
            w = weight(p, ix, iy)
            v = v + vel[ix, iy].astype(sp.dtype) * w
        sp.node_vel[m, n] = v
        fc = dt(FORCE_GAIN) * (sp.node_vel_desired[m, n] - v)
        sp.node_force_temp[m, n] = fc
        sp.node_force[m, n] = sp.node_force[m, n] + fc


def spread(sp: Spheres, vel, body_force, rho, total_rho):
    """IBEngine.SpreadBodyForces, ibm.py:225-244: at each stencil node,
    vel += w F / (2 rho_tot) and body_force_c += w F rho_c / rho_tot."""
    dt = sp.dtype.type
    nc = rho.shape[2]
    for m, n in _groups(sp):
        p = sp.node_pos[m, n]
        for ix, iy in stencil(p):
            w = weight(p, ix, iy)
            ft = sp.node_force_temp[m, n] * w
            r2 = total_rho[ix, iy] * dt(VEL_SHARE)
            vel[ix, iy] = vel[ix, iy] + ft / r2
            for c in range(nc):
                body_force[ix, iy, c] = body_force[ix, iy, c] + ft * rho[ix, iy, c] / total_rho[ix, iy]


def multi_direct_forcing(sp: Spheres, iterations, vel, body_force, rho, total_rho):
    """IBEngine.MultiDirectForcingLoop, ibm.py:248-253."""
    reset_force(sp)
    for _ in range(int(iterations)):
        interpolate(sp, vel)
        spread(sp, vel, body_force, rho, total_rho)


def particle_force(sp: Spheres):
    """IBEngine.ComputeParticleForce, ibm.py:171-185, summed over each
    sphere's group nodes (see the module docstring): force = sum F_n,
    torque = sum (F_x r_y - F_y r_x), r = node position - sphere centre."""
    dt = sp.dtype.type
    touched = {}
    for m, n in _groups(sp):
        fx, fy = sp.node_force[m, n]
        rx = sp.node_pos[m, n, 0] - sp.pos[m, 0]
        ry = sp.node_pos[m, n, 1] - sp.pos[m, 1]
        acc = touched.setdefault(m, [dt(0), dt(0), dt(0)])
        acc[0] = acc[0] + torque_arm(fx, fy, rx, ry)
        acc[1] = acc[1] + fx
        acc[2] = acc[2] + fy
    for m, (tq, fx, fy) in touched.items():
        sp.torque[m] = tq
        sp.force[m] = (fx, fy)
