"""Host-side rules of the multiphase breadth (no GPU): which psi boundary
lists and adhesion maps the fused Shan-Chen step accepts, the halo-grid
classifier, the psi objects' parameters in the C-ABI descriptor."""
import contextlib
import io

import numpy as np
import pytest

from taichi_lbm import _hip as H
from taichi_lbm import cases
from taichi_lbm.BoundaryCondition import BoundaryClassifier, BoundarySpec
from taichi_lbm.SCBoundaryCondition import FixedPsi, OpenNeumannPsi


def _quiet(fn, *a, **k):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


def test_halo_classifier_offset():
    clf = BoundaryClassifier((7, 6), offset=(-1, -1))
    g = clf.get_group(BoundarySpec(lambda I, J: (I == -1) | (J == 4)))
    nodes = {tuple(n) for n in g.nodes}
    assert (-1, -1) in nodes and (-1, 4) in nodes and (5, 4) in nodes
    assert min(n[0] for n in nodes) == -1 and max(n[0] for n in nodes) == 5


def test_fused_plan_accepts_wall_psi_and_refuses_the_rest():
    case = _quiet(cases.shan_chen, NX=40, NY=32, nc=1, seed=1, disk=(20, 16, 6), psi_wall=0.6)
    lbf, bsc = case["lbf"], case["bsc"]
    plan = bsc.fused_plan(lbf, case["fluid"].group.nodes)
    assert len(plan) == 1 and plan[0][1] == np.float32(0.6)
    assert set(map(tuple, plan[0][0])) == set(map(tuple, np.argwhere(lbf.mask.arr == -1)))
    clf = BoundaryClassifier((42, 34), offset=(-1, -1))
    open_bc = OpenNeumannPsi(BoundarySpec(lambda I, J: I == 40), 1)
    open_bc.precompute(clf)
    bsc.add_boundary_condition("open", open_bc)
    with pytest.raises(NotImplementedError, match="granular"):
        bsc.fused_plan(lbf, case["fluid"].group.nodes)
    # FixedPsi on fluid nodes: psi of a fluid node comes from its density in the fused step
    case = _quiet(cases.shan_chen, NX=40, NY=32, nc=1, seed=1)
    bad = FixedPsi(BoundarySpec(lambda I, J: (I == 5) & (J == 5)), 0.3)
    bad.precompute(BoundaryClassifier((42, 34), offset=(-1, -1)))
    case["bsc"].add_boundary_condition("bad", bad)
    with pytest.raises(NotImplementedError):
        case["bsc"].fused_plan(case["lbf"], case["fluid"].group.nodes)


def test_adhesion_codes_need_a_periodic_halo_map():
    case = _quiet(cases.shan_chen, NX=30, NY=20, nc=2, seed=2, disk=(15, 10, 4), gadh=(0.2, 0.1))
    sc = case["lbf"].sc_field
    codes = sc.solid_codes()
    m = case["lbf"].mask.arr.reshape(-1)
    assert np.array_equal(codes != 0, m == -1) and set(np.unique(codes)) <= {0, H.CODE_SC_SOLID}
    sc.fluid_map[0, 5] = ~sc.fluid_map[0, 5]            # the halo no longer wraps
    with pytest.raises(NotImplementedError, match="periodic wrap"):
        sc.solid_codes()


def test_psi_objects_fill_the_descriptor():
    case = _quiet(cases.shan_chen, NX=16, NY=16, nc=1, G=-1.0, seed=3, rho_mean=0.03, amp=0.003,
                  eos=("pr", dict(a=0.35, b=0.2, R=0.3, G=-1.0, rho_cr=1.0, pressure_cr=0.05, temperature_cr=1.1,
                                  omega=0.344)))
    L = H.Lattice()
    case["lbf"].sc_field.fill_lattice(L)
    assert L.psi_kind[0] == H.PSI_PR
    assert list(L.eos[0]) == pytest.approx([0.35, 0.2, 0.3, -1.0, 0.344, 1.1], rel=1e-6)
    case = _quiet(cases.shan_chen, NX=16, NY=16, nc=2, seed=3, disk=(8, 8, 3), gadh=(0.2, -0.1))
    L = H.Lattice()
    case["lbf"].sc_field.fill_lattice(L)
    assert L.sc_adhesion == 1 and list(L.g_adh) == pytest.approx([0.2, -0.1]) and L.psi_kind[1] == H.PSI_SC


def test_huang_collision_descriptor_and_effect():
    """HuangMRTCollision fills k1, k2 = epslion / -8 - k1 and G = g_coh[c][c]
    into the descriptor, refuses a zero self-coupling, and on the GPU test
    lattice (48 x 40, G = -5, k1 = 0.1, epslion = 0.5) changes the oracle's
    populations after 25 steps by about 40 times the parity bound (so the GPU
    parity test of that lattice would see a kernel that dropped the source)."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent))
    from oracle_bridge import oracle_solver
    from taichi_lbm.CollisionEngine import HuangMRTCollision
    with contextlib.redirect_stdout(io.StringIO()):
        case = cases.shan_chen(NX=48, NY=40, nc=1, G=-5.0, seed=27, huang=(0.1, 0.5))
        plain = cases.shan_chen(NX=48, NY=40, nc=1, G=-5.0, seed=27)
    col = case["coll"]
    L = H.Lattice()
    col.fill_lattice(L)
    assert L.huang == 1 and L.collision == H.LBM_MRT
    assert L.huang_k1[0] == pytest.approx(0.1) and L.huang_k2[0] == pytest.approx(0.5 / -8 - 0.1)
    assert L.huang_g[0] == pytest.approx(-5.0)
    with pytest.raises(ValueError):
        HuangMRTCollision({'group': col.group, 'fluid_model': None, 'NX': 48, 'NY': 40, 'num_components': 1,
                           'shearviscosity': [0.1], 'bulkviscosity': [0.1], 'k1': [0.1], 'epslion': [0.5],
                           'g_coh': [[0.0]]})
    a, b = oracle_solver(case, np.float32), oracle_solver(plain, np.float32)
    a.run(25)
    b.run(25)
    d = np.abs(a.lat.f - b.lat.f).max()
    assert d > 10 * 1e-5 * np.abs(b.lat.f).max(), d   # about 40x the bound
