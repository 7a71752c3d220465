"""The reference's table-stream boundary lists (cases/contraction flow.py:96-149,
cases/Power Law Fluid.py:64-117) as built by taichi_lbm.cases, on the host:
the groups equal scalar restatements of the reference's predicates (with
the out-of-range mask reads made explicit), the lists compile for the fused
path, and a list that leaves a -1 table entry to the unguarded table stream
is refused (the reference would read outside its field there)."""
import contextlib
import io

import numpy as np
import pytest

from oracle import lbm_oracle as O
from taichi_lbm import cases
from taichi_lbm._codes import UnsupportedBoundaryList, compile_codes

C = O.C


def _quiet(fn, *a, **k):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


def _ref_inside(mask, i, j):
    # cases/contraction flow.py:96-104 (all reads in range: 0 < ix2 < NX-1, 0 < iy2 < NY-1)
    NX, NY = mask.shape
    if mask[i, j] != 1:
        return False
    for k in range(9):
        ix2, iy2 = i - C[k, 0], j - C[k, 1]
        if 0 < ix2 < NX - 1 and 0 < iy2 < NY - 1 and mask[ix2, iy2] == 1:
            return True
    return False


def _ref_wall(mask, i, j):
    # cases/Power Law Fluid.py:89-97; a neighbour beyond any edge counts as a wall
    NX, NY = mask.shape
    if mask[i, j] != 1:
        return False
    for k in range(9):
        ix2, iy2 = i - C[k, 0], j - C[k, 1]
        if ix2 < 0 or ix2 > NX - 1 or iy2 < 0 or iy2 > NY - 1 or mask[ix2, iy2] == -1:
            return True
    return False


def _group_map(bc, NX, NY):
    m = np.zeros((NX, NY), bool)
    m[bc.group.nodes[:, 0], bc.group.nodes[:, 1]] = True
    return m


@pytest.mark.parametrize("build", [lambda: cases.contraction(W=80, H=24, power_n=0.7, stream="inside"),
                                   lambda: cases.cavity(N=24, power_n=0.7, u_lid=0.02)])
def test_groups_are_the_reference_predicates(build):
    case = _quiet(build)
    lbf, bce = case["lbf"], case["bce"]
    NX, NY = lbf.NX, lbf.NY
    mask = lbf.mask.arr
    inside = np.array([[_ref_inside(mask, i, j) for j in range(NY)] for i in range(NX)])
    wall = np.array([[_ref_wall(mask, i, j) for j in range(NY)] for i in range(NX)])
    bcs = bce.boundary_conditions
    assert list(bcs)[:3] == ["fluid", "inside", "wall"]
    assert np.array_equal(_group_map(bcs["inside"], NX, NY), inside)
    assert np.array_equal(_group_map(bcs["wall"], NX, NY), wall)
    code, rules = compile_codes(NX, NY, case["macro"].group.nodes, bce.fused_list(), lbf.neighbor_valid())
    assert len(code) == NX * NY


def test_cavity_lid_and_constants():
    # cases/Power Law Fluid.py:40-41, 107-112: K = U^(2-n) (N-1)^n / Re; lid = top row, VelocityBB d = 2
    case = _quiet(cases.cavity, N=32, power_n=0.8, Re=10.0, u_lid=5e-3)
    lid = case["bce"].boundary_conditions["inlet"]
    assert lid.direction == 2 and tuple(lid.velocity_value) == (5e-3, 0.0)
    assert np.array_equal(np.unique(lid.group.nodes[:, 1]), [31]) and len(lid.group.nodes) == 32
    K = 5e-3 ** 1.2 * 31 ** 0.8 / 10.0
    assert case["coll"].fluid_model.m == pytest.approx(K, rel=1e-15)
    assert case["coll"].fluid_model.n == 0.8


def test_unguarded_table_stream_is_refused():
    # the reference's literal contraction wall predicate tests the y range only;
    # without the x-edge columns in the wall group, the inlet and outlet nodes
    # stream from -1 entries (x - c_k beyond the lattice)
    case = _quiet(cases.contraction, W=80, H=24, power_n=0.7, stream="inside")
    lbf, bce = case["lbf"], case["bce"]
    wall = bce.boundary_conditions["wall"]
    keep = wall.group.nodes
    mask = lbf.mask.arr
    narrow = np.array([n for n in keep if 0 < n[0] < lbf.NX - 1 or
                       any(not 0 <= n[1] - C[k, 1] < lbf.NY for k in range(9))])
    wall.group.nodes = narrow
    with pytest.raises(UnsupportedBoundaryList, match="-1 neighbour-table entry"):
        compile_codes(lbf.NX, lbf.NY, case["macro"].group.nodes, bce.fused_list(), lbf.neighbor_valid())
    wall.group.nodes = keep


def test_inside_stream_equals_periodic_stream_in_the_oracle():
    # every -1 entry is bounced, so the table stream gives the periodic list's fields
    a = _quiet(cases.contraction, W=80, H=24, power_n=0.7, stream="inside")
    b = _quiet(cases.contraction, W=80, H=24, power_n=0.7)
    from oracle_bridge import oracle_solver
    sa, sb = oracle_solver(a, np.float32), oracle_solver(b, np.float32)
    sa.run(12)
    sb.run(12)
    assert np.array_equal(sa.lat.f, sb.lat.f) and np.isfinite(sa.lat.f).all()


def _ulp_sensitivity(case, steps):
    """Largest velocity change after `steps` FP32 oracle steps caused by
    flipping the last bit of a random third of the initial populations."""
    from oracle_bridge import oracle_solver
    a = oracle_solver(case, np.float32)
    a.run(steps)
    b = oracle_solver(case, np.float32)
    rng = np.random.default_rng(0)
    b.lat.f[...] = b.lat.f * (1 + np.float32(6e-8) * rng.choice([-1, 0, 1], b.lat.f.shape).astype(np.float32))
    b.run(steps)
    return float(np.abs(a.lat.vel - b.lat.vel).max())


def test_power_law_cavity_at_rest_is_ill_conditioned_in_fp32():
    # a power law with n != 1 maps the near-zero shear rate of the quiescent
    # part of a cavity started at rest (cases/Power Law Fluid.py) through
    # gamma^(n-1): there the FP32 rounding noise of f - feq decides tau (up to
    # the clamp), so one flipped last bit of the initial populations moves the
    # velocity by far more than the FP32 parity bound (1e-6 lattice units);
    # with the reference's n = 1 the same perturbation stays at rounding level.
    # The cavity fixtures therefore use n = 1; the power law's parity rests on
    # driven lattices (porous, contraction), where the shear rate is resolved
    shear_thinning = _quiet(cases.cavity, N=40, power_n=0.7, u_lid=0.02)
    newtonian = _quiet(cases.cavity, N=40, power_n=1.0, u_lid=0.02)
    driven = _quiet(cases.contraction, W=80, H=24, power_n=0.7, stream="inside")
    assert _ulp_sensitivity(shear_thinning, 30) > 1e-5
    assert _ulp_sensitivity(newtonian, 30) < 2e-7
    assert _ulp_sensitivity(driven, 20) < 2e-7
