"""Multiphase breadth on the GPU against the oracle: ShanChenForceC2
adhesion at solids (ExternalForce.py:105-116), FixedPsi walls and the open
psi boundaries (SCBoundaryCondition.py:9-123), the cubic-EOS
pseudopotentials (EOS.py:5-95), HuangMRTCollision's source
(CollisionEngine.py:453-515), fused (LDS-tiled Shan-Chen steps) and through
the reference's granular call sequence."""
import contextlib
import io

import numpy as np
import pytest

from oracle_bridge import oracle_solver
from taichi_lbm import cases
from taichi_lbm.stepper import Stepper
from test_gpu_parity import close, compare_state

pytestmark = pytest.mark.gpu

EOS_PARAMS = {   # T = 1 on every node; T_c about 1.1
    "cs": dict(a=1.17, b=4.0, R=0.1, G=-1.0, rho_cr=0.0326, pressure_cr=0.00517, temperature_cr=1.104),
    "vdw": dict(a=0.3, b=0.3, R=0.3, G=-1.0, rho_cr=1.111, pressure_cr=0.0412, temperature_cr=1.185),
    "rk": dict(a=0.45, b=0.3, R=0.3, G=-1.0, rho_cr=1.0, pressure_cr=0.05, temperature_cr=1.1),
    "pr": dict(a=0.35, b=0.2, R=0.3, G=-1.0, rho_cr=1.0, pressure_cr=0.05, temperature_cr=1.1, omega=0.344),
}
RHO_MEAN = {"cs": 0.0326, "vdw": 1.0, "rk": 0.9, "pr": 1.2}


def _quiet(fn, *a, **k):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


CASES = {
    "adhesion_c2": dict(NX=60, NY=44, nc=2, gcross=0.9, seed=21, disk=(30, 22, 7), gadh=(0.35, -0.25)),
    "wall_psi_c1": dict(NX=60, NY=44, nc=1, G=-5.0, seed=22, disk=(30, 22, 7), psi_wall=0.65),
    "wall_psi_adhesion_c2": dict(NX=50, NY=40, nc=2, gcross=0.9, seed=23, disk=(25, 20, 6), gadh=(0.2, 0.2),
                                 psi_wall=0.5),
}
for kind in EOS_PARAMS:
    CASES[f"eos_{kind}_c1"] = dict(NX=45, NY=38, nc=1, G=-1.0, seed=24, rho_mean=0.9 * RHO_MEAN[kind],
                                   amp=0.2 * RHO_MEAN[kind], eos=(kind, EOS_PARAMS[kind]))
CASES["eos_cs_c2"] = dict(NX=40, NY=36, nc=2, gcross=0.9, seed=25, eos=("cs", EOS_PARAMS["cs"]), rho_mean=0.0326,
                         amp=0.003)
CASES["huang_c1"] = dict(NX=48, NY=40, nc=1, G=-5.0, seed=27, huang=(0.1, 0.5))
CASES["huang_c2"] = dict(NX=40, NY=36, nc=2, gcross=0.9, seed=28, huang=(0.05, 0.3))


@pytest.mark.parametrize("name", sorted(CASES))
def test_multiphase_fused_matches_oracle(name):
    print(f"case multiphase fused {name}", flush=True)
    case = _quiet(cases.shan_chen, **CASES[name])
    _quiet(cases.initialise, case)
    st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], boundary_engine_sc=case["bsc"])
    st.run(13)
    st.run(12)
    sol = oracle_solver(case, np.float32)
    sol.run(25)
    compare_state(case["lbf"], sol.lat, ("f", "f2", "rho", "vel", "total_rho", "SCforce", "pressure"))
    close("psi", case["lbf"].sc_field.psi_field.to_numpy(), sol.lat.psi)


@pytest.mark.parametrize("name", ["adhesion_c2", "wall_psi_c1", "eos_cs_c1", "eos_pr_c1", "huang_c1", "huang_c2"])
def test_multiphase_granular_matches_oracle(name):
    print(f"case multiphase granular {name}", flush=True)
    case = _quiet(cases.shan_chen, **CASES[name])
    _quiet(cases.initialise, case)
    lbf, macro, col, bce, bsc = case["lbf"], case["macro"], case["coll"], case["bce"], case["bsc"]
    n = 12
    for _ in range(n):
        macro.density(lbf)
        macro.pressure(lbf)
        lbf.sc_field.update_psi(lbf)
        bsc.apply_boundary_conditions(lbf)
        lbf.sc_field.cau_SCforce(lbf)
        macro.force_density(lbf)
        macro.velocity(lbf)
        col.apply(lbf)
        bce.apply_boundary_conditions(lbf)
    sol = oracle_solver(case, np.float32)
    sol.run(n)
    compare_state(lbf, sol.lat, ("f", "f2", "rho", "vel", "SCforce", "pressure"))
    close("psi", lbf.sc_field.psi_field.to_numpy(), sol.lat.psi)


@pytest.mark.parametrize("kind", ["neumann", "extrapolation", "convective1", "convective2"])
def test_open_psi_boundaries_granular(kind):
    """A periodic Shan-Chen lattice moving at u = (0.01, 0.01) whose psi halo
    columns take an open boundary on both ends (x = -1 with d = 3, x = NX
    with d = 1) after the periodic pair, through the granular sequence."""
    from taichi_lbm.BoundaryCondition import BoundaryClassifier, BoundarySpec
    from taichi_lbm import SCBoundaryCondition as S
    print(f"case psi boundary {kind}", flush=True)
    case = _quiet(cases.shan_chen, NX=40, NY=24, nc=1, G=-5.0, seed=26)
    lbf = case["lbf"]
    NX, NY = lbf.NX, lbf.NY
    cls = {"neumann": S.OpenNeumannPsi, "extrapolation": S.OpenExtrapolationPsi,
           "convective1": S.OpenConvective1orderPsi, "convective2": S.OpenConvective2orderPsi}[kind]
    clf = BoundaryClassifier((NX + 2, NY + 2), offset=(-1, -1))
    for name, x, d in (("left", -1, 3), ("right", NX, 1)):
        bc = cls(BoundarySpec(lambda I, J, x=x: (I == x) & (J >= 0) & (J < NY)), d)
        bc.precompute(clf)
        case["bsc"].add_boundary_condition(name, bc)
    case["vel"][...] = 0.01                 # lambda > 0 for the convective kinds
    _quiet(cases.initialise, case)
    macro, col, bce, bsc = case["macro"], case["coll"], case["bce"], case["bsc"]
    sol = oracle_solver(case, np.float32)
    n = 6
    for _ in range(n):
        macro.density(lbf)
        macro.pressure(lbf)
        lbf.sc_field.update_psi(lbf)
        bsc.apply_boundary_conditions(lbf)
        lbf.sc_field.cau_SCforce(lbf)
        macro.force_density(lbf)
        macro.velocity(lbf)
        col.apply(lbf)
        bce.apply_boundary_conditions(lbf)
    sol.run(n)
    compare_state(lbf, sol.lat, ("f", "rho", "vel", "SCforce"))
    close("psi", lbf.sc_field.psi_field.to_numpy(), sol.lat.psi)
