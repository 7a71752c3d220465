"""Pins of the immersed-boundary oracle (oracle/ibm_oracle.py) against closed
forms and invariants of the reference's multi-direct forcing
(taichi_lbm/ibm.py): marker geometry, the marker selection, bilinear
interpolation, force spreading, the multi-direct-forcing iteration and the
particle force and torque.  None of them restates the oracle's formulas."""
import math

import numpy as np
import pytest

from oracle import ibm_oracle as I


def _spheres(pos, vel, radius, angle, NX, NY, max_num_ib=96, dtype=np.float64):
    n = len(pos)
    sp = I.Spheres(n + 1, max_num_ib, dtype)
    sp.init_balls(n, np.arange(n), pos, vel, np.zeros((n, 2)), np.zeros(n), radius, angle)
    I.init_ib_nodes(sp)
    return sp


def test_marker_ring_geometry():
    # NumIB = round(2 pi r) markers spaced 2 pi / NumIB on the circle of radius r;
    # the sphere's angle rotates every marker about the centre
    radii = [3.3, 5.0, 7.7, 10.0, 12.25]
    pos = [(30.0 + 3 * k, 40.0 - 2 * k) for k in range(len(radii))]
    angle = [0.0, 0.3, -1.1, 2.5, 0.75]
    sp = _spheres(pos, np.zeros((5, 2)), radii, angle, 100, 100)
    for m, r in enumerate(radii):
        N = int(sp.num_ib[m])
        assert N == round(2 * math.pi * r)
        p0 = sp.node_pos[m, :N] - np.asarray(pos[m])              # init_IB_nodes: no rotation
        assert np.allclose(np.hypot(p0[:, 0], p0[:, 1]), r, rtol=1e-12)
        ang = np.unwrap(np.arctan2(p0[:, 1], p0[:, 0]))
        assert np.allclose(np.diff(ang), 2 * math.pi / N, atol=1e-12)
        assert abs(ang[0]) < 1e-12                                # marker 0 on the +x axis
        th = angle[m]
        R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
        for n in (0, 1, N // 3, N - 1):
            got = I.node_pos(sp, m, n) - np.asarray(pos[m])
            assert np.allclose(got, R @ p0[n], atol=1e-12)


def test_marker_count_in_float32_and_capacity():
    assert I.num_ib(10.0, 1, np.float32) == 63          # 62.83
    assert I.num_ib(10.0, 2, np.float32) == 31          # 31.42
    assert I.num_ib(0.05, 1, np.float32) == 0           # 0.31
    sp = I.Spheres(2, 10, np.float32)
    sp.init_balls(1, [0], [(20.0, 20.0)], [(0.0, 0.0)], [(0.0, 0.0)], [0.0], [5.0], [0.0])
    with pytest.raises(ValueError):
        I.init_ib_nodes(sp)                              # 31 markers > Max_NumIB = 10


def test_check_pos_and_in_domain_boundaries():
    NX, NY = 50, 40
    # inside: the box [x - d, x + d] x [y - d, y + d] within [0, NX-1] x [0, NY-1], edges included
    assert I.check_pos((10.0, 10.0), 10.0, NX, NY) == 1
    assert I.check_pos((39.0, 29.0), 10.0, NX, NY) == 1
    assert I.check_pos((9.5, 10.0), 10.0, NX, NY) == 2
    assert I.check_pos((39.5, 20.0), 10.0, NX, NY) == 2
    # outside: the box entirely beyond one side
    assert I.check_pos((-10.5, 20.0), 10.0, NX, NY) == 0
    assert I.check_pos((-10.0, 20.0), 10.0, NX, NY) == 2        # x_max = 0 touches
    assert I.check_pos((20.0, 49.5), 10.0, NX, NY) == 0
    assert I.check_pos((20.0, 49.0), 10.0, NX, NY) == 2
    # InDomain is strict on all four sides
    assert I.in_domain((0.5, 0.5), NX, NY) == 1
    assert I.in_domain((0.0, 5.0), NX, NY) == 0
    assert I.in_domain((49.0, 5.0), NX, NY) == 0
    assert I.in_domain((48.999, 38.999), NX, NY) == 1
    assert I.in_domain((5.0, 39.0), NX, NY) == 0


def test_sphere_update_selects_markers():
    NX, NY = 60, 40
    # sphere 0 inside, sphere 1 crossing the left edge (centre x = 2, r = 5), sphere 2 outside
    pos = [(30.0, 20.0), (2.0, 20.0), (-30.0, 20.0)]
    vel = [(0.01, -0.02), (0.03, 0.0), (0.5, 0.5)]
    sp = _spheres(pos, vel, [6.0, 5.0, 4.0], [0.0, 0.0, 0.0], NX, NY)
    before = sp.node_pos[2].copy()
    I.sphere_ib_update(sp, NX, NY)
    N0, N1 = int(sp.num_ib[0]), int(sp.num_ib[1])
    # crossing sphere: markers with 2 + 5 cos(theta) > 0, i.e. cos(theta) > -0.4 (no marker within 1e-3 of it)
    c = np.cos(2 * math.pi * np.arange(N1) / N1)
    assert np.min(np.abs(c + 0.4)) > 1e-3
    inside1 = np.nonzero(c > -0.4)[0]
    want = [(0, n) for n in range(N0)] + [(1, int(n)) for n in inside1]
    assert sp.count == len(want)
    assert [tuple(g) for g in sp.group[:sp.count]] == want
    assert not sp.group[sp.count:].any()
    assert np.array_equal(sp.node_pos[2], before)              # outside: markers not moved
    assert np.allclose(sp.node_vel_desired[0, :N0], vel[0]) and np.allclose(sp.node_vel_desired[1, :N1], vel[1])
    assert not sp.node_vel_desired[2].any()


def _lattice(NX, NY, nc, rng, uniform_rho=False):
    if uniform_rho:
        rho = np.full((NX, NY, nc), 1.0 / nc)
    else:
        rho = rng.uniform(0.4, 0.9, (NX, NY, nc))
    return rho, rho.sum(axis=2)


def _markers(sp_pos, NX, NY, rng, k=40, integer=6):
    """A crossing-free sphere whose ring gives k markers; some markers set to integer positions."""
    pos = rng.uniform(4.0, min(NX, NY) - 5.0, (k, 2))
    pos[:integer] = np.round(pos[:integer])
    return pos


def _set_markers(sp, pos):
    k = len(pos)
    sp.node_pos[0, :k] = pos
    sp.count = k
    sp.group[:k] = np.stack([np.zeros(k, np.int32), np.arange(k, dtype=np.int32)], axis=1)


def test_interpolation_reproduces_linear_fields():
    rng = np.random.default_rng(3)
    NX, NY = 30, 22
    sp = I.Spheres(1, 64, np.float64)
    pos = _markers(None, NX, NY, rng)
    _set_markers(sp, pos)
    i, j = np.meshgrid(np.arange(NX), np.arange(NY), indexing="ij")
    a = (0.01, 0.002, -0.0035)      # ux = a0 + a1 i + a2 j (different slopes in i and j)
    b = (-0.02, -0.001, 0.0045)
    vel = np.stack([a[0] + a[1] * i + a[2] * j, b[0] + b[1] * i + b[2] * j], axis=-1)
    sp.node_vel_desired[0, :len(pos)] = 0.0
    I.interpolate(sp, vel)
    exact = np.stack([a[0] + a[1] * pos[:, 0] + a[2] * pos[:, 1], b[0] + b[1] * pos[:, 0] + b[2] * pos[:, 1]], 1)
    assert np.allclose(sp.node_vel[0, :len(pos)], exact, atol=1e-14)
    # force correction 2 (u_d - u) with u_d = 0
    assert np.allclose(sp.node_force_temp[0, :len(pos)], -2 * exact, atol=1e-14)
    assert np.allclose(sp.node_force[0, :len(pos)], -2 * exact, atol=1e-14)
    # the weights of a stencil sum to one
    for p in pos:
        assert abs(sum(I.weight(p, ix, iy) for ix, iy in I.stencil(p)) - 1.0) < 1e-14


def test_spread_conserves_force_and_is_the_transpose_of_interpolation():
    rng = np.random.default_rng(4)
    NX, NY, nc = 28, 24, 2
    rho, tot = _lattice(NX, NY, nc, rng)
    sp = I.Spheres(1, 64, np.float64)
    pos = _markers(None, NX, NY, rng)
    k = len(pos)
    _set_markers(sp, pos)
    F = rng.normal(0, 1e-3, (k, 2))
    sp.node_force_temp[0, :k] = F
    vel = np.zeros((NX, NY, 2))
    bf = np.zeros((NX, NY, nc, 2))
    I.spread(sp, vel, bf, rho, tot)
    # total force: the components' body-force shares and 2 rho_tot du each add up to sum F
    assert np.allclose(bf.sum(axis=(0, 1, 2)), F.sum(0), atol=1e-15)
    assert np.allclose((2 * tot[..., None] * vel).sum(axis=(0, 1)), F.sum(0), atol=1e-15)
    # each component receives its density share at every node
    for c in range(nc):
        assert np.allclose(bf[:, :, c, :], (rho[..., c] / tot)[..., None] * bf.sum(axis=2), rtol=1e-12, atol=1e-20)
    # adjoint: sum_m F_m . I(u)_m = sum_x u(x) . S(F)(x), S(F) = 2 rho_tot du
    u = rng.normal(0, 1e-2, (NX, NY, 2))
    sp.node_vel_desired[0, :k] = 0.0
    I.interpolate(sp, u)
    lhs = float((F * sp.node_vel[0, :k]).sum())
    rhs = float((u * 2 * tot[..., None] * vel).sum())
    assert abs(lhs - rhs) < 1e-15 * max(1.0, abs(lhs)) + 1e-18


@pytest.mark.parametrize("x0,y0", [(7.25, 9.6), (8.0, 11.0)])
def test_multi_direct_forcing_closed_form(x0, y0):
    # one marker in a uniform flow u0 with rho_tot = 1: after k iterations the
    # interpolated error is (1 - s)^k (u0 - u_d), s = sum of squared weights,
    # and the marker force is 2 (u_d - u0) sum_{j<k} (1 - s)^j
    NX, NY, nc, K = 20, 20, 1, 3
    rng = np.random.default_rng(5)
    rho, tot = _lattice(NX, NY, nc, rng, uniform_rho=True)
    u0, ud = np.array([0.03, -0.01]), np.array([0.0, 0.02])
    vel = np.broadcast_to(u0, (NX, NY, 2)).copy()
    bf = np.zeros((NX, NY, nc, 2))
    sp = I.Spheres(1, 4, np.float64)
    _set_markers(sp, np.array([[x0, y0]]))
    sp.node_vel_desired[0, 0] = ud
    p = np.array([x0, y0])
    s = sum(I.weight(p, ix, iy) ** 2 for ix, iy in I.stencil(p))
    I.multi_direct_forcing(sp, K, vel, bf, rho, tot)
    u_last = ud - (1 - s) ** (K - 1) * (ud - u0)              # what the K-th interpolation saw
    assert np.allclose(sp.node_vel[0, 0], u_last, atol=1e-15)
    Fk = 2 * (ud - u0) * sum((1 - s) ** j for j in range(K))
    assert np.allclose(sp.node_force[0, 0], Fk, atol=1e-15)
    assert np.allclose(bf.sum(axis=(0, 1, 2)), Fk, atol=1e-15)
    if float(x0).is_integer() and float(y0).is_integer():
        assert s == 1.0 and np.allclose(vel[int(x0), int(y0)], ud, atol=1e-16)


def test_particle_force_and_torque():
    NX, NY = 80, 80
    sp = _spheres([(30.0, 30.0), (50.0, 50.0), (60.0, 20.0)], np.zeros((3, 2)), [6.0, 4.5, 3.0],
                  [0.2, 0.0, 0.0], NX, NY)
    I.sphere_ib_update(sp, NX, NY)
    N0, N1 = int(sp.num_ib[0]), int(sp.num_ib[1])
    r0 = sp.node_pos[0, :N0] - sp.pos[0]
    r1 = sp.node_pos[1, :N1] - sp.pos[1]
    sp.node_force[0, :N0] = (1e-3, -2e-3)                     # uniform: no torque
    k = 5e-4
    sp.node_force[1, :N1] = np.stack([-k * r1[:, 1], k * r1[:, 0]], 1)   # tangential
    sp.force[2], sp.torque[2] = (7.0, 8.0), 9.0
    # sphere 2 with no markers in the group keeps its values
    keep = sp.group[:sp.count].copy()
    sp.group[:] = 0
    sel = keep[keep[:, 0] != 2]
    sp.group[:len(sel)] = sel
    sp.count = len(sel)
    I.particle_force(sp)
    assert np.allclose(sp.force[0], (N0 * 1e-3, -N0 * 2e-3), rtol=1e-12)
    assert abs(sp.torque[0]) < 1e-15
    assert np.allclose(sp.force[1], 0.0, atol=1e-15)
    # torque = sum (F_x r_y - F_y r_x) = -k sum |r|^2 = -k N r^2 (the reference's sign)
    assert np.isclose(sp.torque[1], -k * N1 * 4.5 ** 2, rtol=1e-12)
    assert np.allclose(sp.force[2], (7.0, 8.0)) and sp.torque[2] == 9.0
