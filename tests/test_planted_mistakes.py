"""Mistakes planted in the oracle, on the CPU, must each fail an independent
pin: a closed form, special value or invariant of tests/test_oracle_kat.py
or the C++ restatement (oracle/cpu_ref.cpp, written separately from the
NumPy oracle).  Where a planted change
moves the fields of a golden case beyond the FP32 parity bounds, the
committed vectors and the C++ restatement must flag it too; the sign of the
Guo cross term does not (9 (c.F)(c.u) with F = 2e-5 and u = 4e-4 changes f
by about 1e-8 per step, under the FP32 bound of 4e-6), so only its moment
pin catches it.

Each mutation is a slip a restatement of the reference could plausibly make:
a dropped term of the equilibrium (CollisionEngine.py:14-18), a sign of the
Guo forcing term (CollisionEngine.py:20-25), the diagonal opposites of the
bounce map (BoundaryCondition.py:242), an untransposed MRT inverse
(CollisionEngine.py:283-287), the other common pseudopotential (EOS.py:40-47).
"""
from pathlib import Path

import numpy as np
import pytest

import test_oracle_kat as kat
from oracle import lbm_oracle as O
from tolerances import check

def _feq_without_quadratic(vel, rho, dt):
    cf = O.C.astype(dt)
    w = O.weights(dt)
    eu = cf[None, :, 0] * vel[:, 0:1] + cf[None, :, 1] * vel[:, 1:2]
    uv = vel[:, 0:1] ** 2 + vel[:, 1:2] ** 2
    return w[None, :] * rho[:, None] * (1 + 3 * eu - 1.5 * uv)


def _guo_with_flipped_cross_term(F, vel, dt):
    cf = O.C.astype(dt)
    w = O.weights(dt)
    cF = cf[None, :, 0] * F[:, 0:1] + cf[None, :, 1] * F[:, 1:2]
    cu = cf[None, :, 0] * vel[:, 0:1] + cf[None, :, 1] * vel[:, 1:2]
    uF = vel[:, 0:1] * F[:, 0:1] + vel[:, 1:2] * F[:, 1:2]
    return w[None, :] * (3 * cF - 9 * cF * cu - 3 * uF)


def _psi_exponential_form(rho, rho0):
    return rho0 * np.exp(-rho0 / rho)


# name -> (attribute of the oracle module, planted value, KAT pins, golden case
#          that exercises it, golden vectors resolve it, C++ restatement resolves it)
MUTATIONS = {
    "feq_quadratic_term_dropped": ("f_eq", _feq_without_quadratic, [kat.test_feq_moments],
                                   "cylinder_bgk_48x24", True, True),
    "guo_cross_term_sign": ("f_force", _guo_with_flipped_cross_term, [kat.test_guo_force_moments],
                            "porous_powerlaw_trt_48x40", False, False),
    "diagonal_opposites_swapped": ("OPP", np.array([0, 3, 4, 1, 2, 8, 7, 6, 5], dtype=np.int64),
                                   [kat.test_constants], "cylinder_bgk_48x24", True, True),
    "mrt_inverse_not_transposed": ("M_INV", O.M / O.M_DIAG[:, None],
                                   [kat.test_mrt_transform_identities, kat.test_mrt_equals_bgk_at_unit_rates],
                                   "cylinder_mrt_48x24", True, True),
    "psi_exponential_form": ("sc_psi", _psi_exponential_form, [kat.test_sc_psi_and_periodic_halo],
                             "shan_chen_c1_32x32", True, False),   # no Shan-Chen in the C++ restatement
}


def _fails(fn):
    try:
        fn()
    except AssertionError:
        return True
    return False


def _golden(name):
    z = np.load(Path(__file__).resolve().parent / "golden" / f"{name}.npz")
    return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(name, monkeypatch):
    attr, value, pins, case_name, by_golden, by_cpp = MUTATIONS[name]
    from golden.make_golden import build, run_oracle
    case, steps = build(case_name)
    # the unmutated oracle agrees with the golden vectors (control)
    fx = _golden(case_name)
    good = run_oracle(case, steps, np.float32, params32=False)
    check("f", good.f, fx["out32_f"], log=False)

    monkeypatch.setattr(O, attr, value)
    caught = [p.__name__ for p in pins if _fails(p)]
    print(f"  {name}: closed-form pins that fail: {caught}")
    assert caught, f"{name}: no closed-form pin fails"

    case, steps = build(case_name)
    bad = run_oracle(case, steps, np.float32, params32=False)
    assert _fails(lambda: check("f", bad.f, fx["out32_f"], log=False)) == by_golden, name
    if case["lbf"].nc == 1 and case.get("sc_spec") is None:
        from oracle import cpu_ref
        cl = cpu_ref.from_case(case, nthreads=2)
        cl.run(steps)
        assert _fails(lambda: check("f", bad.f, cl.f, log=False)) == by_cpp, name


# ---------------------------------------------------------------------------
# Slips in the two-component Shan-Chen force (ExternalForce.py:96-117), the
# pressure observables (lbm.py:166-212), the nc = 2 force split and
# barycentric velocity (lbm.py:215-237) and the table stream
# (BoundaryCondition.py:435-449).  Each is planted as a wrapper around the
# oracle's own function and must fail one of the pins listed beside it.
# ---------------------------------------------------------------------------
_orig_pressure = O.pressure
_orig_force = O.ShanChen.cau_scforce
_orig_force_density = O.force_density
_orig_bc_apply = O.BC.apply


def _pressure_cross_term_unscaled(lat, group, sc=None):
    _orig_pressure(lat, group, sc)
    if sc is not None and lat.nc == 2:       # p_ij in place of p_ij / 6
        I, J = O.group_ij(group)
        G = np.asarray(sc.g_coh, lat.dt)
        lat.total_pressure[I, J] += 5 * G[0, 1] * lat.psi[I + 1, J + 1, 0] * lat.psi[I + 1, J + 1, 1] / 6


def _pressure1_without_half(lat, group, sc=None):
    _orig_pressure(lat, group, sc)
    if sc is not None and lat.nc == 1:       # (rho + G psi^2)/3 in place of (rho + G psi^2 / 2)/3
        I, J = O.group_ij(group)
        lat.pressure[I, J, 0] += 0.5 * float(sc.g_coh) * lat.psi[I + 1, J + 1, 0] ** 2 / 3
        lat.total_pressure[I, J] = lat.pressure[I, J, 0]


def _force_self_coupling(self, lat):
    if self.kind != "c2":
        return _orig_force(self, lat)
    G = np.asarray(self.g_coh)
    keep = self.g_coh
    self.g_coh = np.array([[G[0, 0], G[0, 0]], [G[1, 1], G[1, 1]]])   # G[c1, c1] where G[c1, c2] belongs
    try:
        _orig_force(self, lat)
    finally:
        self.g_coh = keep


def _force_sign_flipped(self, lat):
    _orig_force(self, lat)
    if self.kind == "c2":
        lat.SCforce *= -1                    # +psi_c1(x) in place of -psi_c1(x)


def _force_split_by_total_density(lat, group):
    I, J = O.group_ij(group)
    for c in range(lat.nc):                  # rho_tot / rho_tot in place of rho_c / rho_tot
        lat.body_force[I, J, c, :] = lat.SCforce[I, J, c, :] + lat.g[None, :]


def _velocity_of_first_component(lat, group):
    lat.vel.fill(0)
    I, J = O.group_ij(group)
    cf = O.C.astype(lat.dt)
    j = lat.f[I, J, 0, :] @ cf + 0.5 * lat.body_force[I, J, 0, :]   # component 0 over its own density
    lat.vel[I, J] = j / lat.rho[I, J, 0][:, None]


def _inside_wraps_periodically(self, lat):
    if self.kind != "inside":
        return _orig_bc_apply(self, lat)
    keep = self.kind
    self.kind = "periodic"                  # pulls across the edges and from solids in place of the table
    try:
        _orig_bc_apply(self, lat)
    finally:
        self.kind = keep


def _convective_with_own_velocity(self, lat):
    # l = |u(x)| of the outlet node in place of its interior neighbour's
    if self.kind != "open_convective1":
        return _orig_bc_apply(self, lat)
    I, J = O.group_ij(self.group)
    d = self.direction
    keep = lat.vel[I - O.C[d, 0], J - O.C[d, 1]].copy()
    lat.vel[I - O.C[d, 0], J - O.C[d, 1]] = lat.vel[I, J]
    try:
        _orig_bc_apply(self, lat)
    finally:
        lat.vel[I - O.C[d, 0], J - O.C[d, 1]] = keep


def _extrapolation_from_own_node(self, lat):
    # 2 f(x - c_d) - f(x) in place of 2 f(x - c_d) - f(x - 2 c_d)
    if self.kind != "open_extrapolation":
        return _orig_bc_apply(self, lat)
    I, J = O.group_ij(self.group)
    d = self.direction
    I2, J2 = I - O.C[d, 0], J - O.C[d, 1]
    for c in range(lat.nc):
        for k in O.UNKNOWN[d]:
            lat.f[I, J, c, k] = 2 * lat.f[I2, J2, c, k] - lat.f[I, J, c, k]


# name -> (object, attribute, planted value, pins that must fail)
MUTATIONS_SC = {
    "convective_outlet_own_velocity": (O.BC, "apply", _convective_with_own_velocity,
                                       [kat.test_open_convective1_relaxes_unknowns_towards_the_neighbour]),
    "extrapolation_from_own_node": (O.BC, "apply", _extrapolation_from_own_node,
                                    [kat.test_open_extrapolation_continues_linear_profiles]),
    "pressure2_cross_term_unscaled": (O, "pressure", _pressure_cross_term_unscaled, [kat.test_pressure2_uniform_states]),
    "pressure1_half_dropped": (O, "pressure", _pressure1_without_half,
                               [kat.test_pressure1_equal_at_reference_coexistence_densities]),
    "sc_c2_self_coupling": (O.ShanChen, "cau_scforce", _force_self_coupling, [kat.test_sc_c2_linear_psi_closed_form]),
    "sc_c2_sign_flipped": (O.ShanChen, "cau_scforce", _force_sign_flipped, [kat.test_sc_c2_linear_psi_closed_form]),
    "force_split_by_total_density": (O, "force_density", _force_split_by_total_density,
                                     [kat.test_two_component_force_split, kat.test_two_component_momentum_budget_and_mass]),
    "velocity_of_first_component": (O, "velocity", _velocity_of_first_component,
                                    [kat.test_two_component_barycentric_velocity]),
    "inside_wraps_periodically": (O.BC, "apply", _inside_wraps_periodically, [kat.test_inside_boundary_table_stream]),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS_SC))
def test_planted_multicomponent_mistake_is_caught(name, monkeypatch):
    obj, attr, value, pins = MUTATIONS_SC[name]
    for p in pins:                           # control: the unmutated oracle passes every pin
        p()
    monkeypatch.setattr(obj, attr, value)
    failed = [p.__name__ for p in pins if _fails(p)]
    print(f"  {name}: pins that fail: {failed}")
    assert failed == [p.__name__ for p in pins], f"{name}: a pin does not catch it"


# HuangMRTCollision's source moments (oracle m_q)
def _mq_q2_sign(F, psi, k1, k2, g, dt):
    Q = _ORIG_M_Q(F, psi, k1, k2, g, dt)
    Q[:, 2] = Q[:, 1]
    return Q


def _mq_psi_not_squared(F, psi, k1, k2, g, dt):
    # G psi in place of G psi^2 (psi > 0 here: sqrt(psi)^2 = psi)
    return _ORIG_M_Q(F, np.sqrt(psi), k1, k2, g, dt)


def _mq_k2_sign(F, psi, k1, k2, g, dt):
    # k2 = epslion / 8 - k1 instead of epslion / -8 - k1
    eps = (k2 + k1) * dt.type(-8.0)
    return _ORIG_M_Q(F, psi, k1, eps / dt.type(8.0) - k1, g, dt)


def _mq_q8_factor_two(F, psi, k1, k2, g, dt):
    Q = _ORIG_M_Q(F, psi, k1, k2, g, dt)
    Q[:, 8] = 2 * Q[:, 8]
    return Q


_ORIG_M_Q = O.m_q
MUTATIONS_HUANG = {
    "q2_sign": (O, "m_q", _mq_q2_sign, [kat.test_huang_source_closed_form]),
    "psi_not_squared": (O, "m_q", _mq_psi_not_squared, [kat.test_huang_source_closed_form,
                                                        kat.test_huang_source_scales_with_inverse_g_psi_squared]),
    "k2_sign": (O, "m_q", _mq_k2_sign, [kat.test_huang_source_closed_form]),
    "q8_factor_two": (O, "m_q", _mq_q8_factor_two, [kat.test_huang_source_closed_form]),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS_HUANG))
def test_planted_huang_mistake_is_caught(name, monkeypatch):
    obj, attr, value, pins = MUTATIONS_HUANG[name]
    for p in pins:
        p()
    monkeypatch.setattr(obj, attr, value)
    failed = [p.__name__ for p in pins if _fails(p)]
    print(f"  {name}: pins that fail: {failed}")
    assert failed == [p.__name__ for p in pins], f"{name}: a pin does not catch it"


# Immersed-boundary coupling (oracle/ibm_oracle.py, the reference's ibm.py)
import test_ibm_oracle as ibk  # noqa: E402
from oracle import ibm_oracle as IBO  # noqa: E402


def _weight_transposed(pos, ix, iy):
    dt = type(pos[0])
    return (dt(1.0) - abs(pos[1] - dt(ix))) * (dt(1.0) - abs(pos[0] - dt(iy)))


def _torque_sign_flipped(fx, fy, rx, ry):
    return rx * fy - ry * fx


MUTATIONS_IBM = {
    "stencil_weight_transposed": ("weight", _weight_transposed,
                                  [ibk.test_interpolation_reproduces_linear_fields]),
    "force_gain_dropped": ("FORCE_GAIN", 1.0, [ibk.test_interpolation_reproduces_linear_fields]),
    "velocity_share_not_halved": ("VEL_SHARE", 1.0,
                                  [ibk.test_spread_conserves_force_and_is_the_transpose_of_interpolation]),
    "torque_sign_flipped": ("torque_arm", _torque_sign_flipped, [ibk.test_particle_force_and_torque]),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS_IBM))
def test_planted_ibm_mistake_is_caught(name, monkeypatch):
    attr, value, pins = MUTATIONS_IBM[name]
    for p in pins:
        p()
    monkeypatch.setattr(IBO, attr, value)
    failed = [p.__name__ for p in pins if _fails(p)]
    print(f"  {name}: pins that fail: {failed}")
    assert failed == [p.__name__ for p in pins], f"{name}: a pin does not catch it"
    # the multi-direct-forcing closed form catches every slip of the coupling itself
    if attr != "torque_arm":
        assert _fails(lambda: ibk.test_multi_direct_forcing_closed_form(7.25, 9.6)), name
