"""The C++ CPU restatement (CPU baseline) and the NumPy oracle are two
independent restatements of the reference step; they must agree."""
import numpy as np
import pytest

from oracle import cpu_ref
from taichi_lbm import cases
from oracle_bridge import oracle_solver
from tolerances import check


def _compare(case, n, fields=("f", "f2", "rho", "vel", "total_rho")):
    cl = cpu_ref.from_case(case, nthreads=2)
    cl.run(n)
    sol = oracle_solver(case, np.float32)
    sol.run(n)
    for name in fields:
        check(name, getattr(cl, name), getattr(sol.lat, name))


@pytest.mark.parametrize("coll", ["bgk", "trt", "mrt"])
def test_cylinder(coll):
    case = cases.cylinder(NX=48, NY=24, cx=12, cy=11, r=4, tau=0.6, u_in=0.05, collision=coll, tau_bulk=0.8)
    _compare(case, 30)


@pytest.mark.parametrize("coll", ["bgk", "trt", "mrt"])
def test_porous_power_law(coll):
    case = cases.porous(NX=64, NY=48, seed=2, collision=coll, rmin=3.0, rmax=6.0, g=(2e-5, 1e-6))
    _compare(case, 20, ("f", "f2", "rho", "vel", "shear_rate"))


@pytest.mark.parametrize("coll", ["bgk", "mrt"])
def test_reference_cylinder_boundaries(coll):
    """VelocityBB inlet over bounce-back + OpenNeumann outlet (cases/Circular
    cylinder.py's list)."""
    case = cases.cylinder_reference_bcs(NX=60, NY=24, tau=0.7, u_in=0.03, collision=coll, tau_bulk=0.9)
    _compare(case, 30)


def test_cylinder_velocity_bb_inlet():
    case = cases.cylinder(NX=48, NY=24, cx=12, cy=11, r=4, tau=0.6, u_in=0.05, collision="trt", inlet="bb")
    _compare(case, 30)
