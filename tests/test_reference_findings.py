"""Properties of the reference's own semantics that the configurations run
into, shown with the FP64 oracle (no GPU)."""
import contextlib
import io

import numpy as np

from taichi_lbm import cases
from oracle_bridge import oracle_solver


def _quiet(fn, *a, **k):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


def test_velocity_neem_inlet_next_to_walls_diverges():
    """VelocityNEEM (BoundaryCondition.py:149-166) extrapolates the neighbour's
    post-stream f against its rho/u of the step before.  Next to bounce-back
    rows this inlet is unstable: the plain-method channel (here 96 x 32,
    tau = 0.8, u = 0.02, a smaller cousin of the 400 x 100 case) blows up
    within a few hundred steps even in FP64, while the same channel with the
    reference cases' VelocityBB inlet stays bounded."""
    kw = dict(NX=96, NY=32, cx=19, cy=16, r=4, tau=0.8, u_in=0.02, collision="bgk")
    neem = oracle_solver(_quiet(cases.cylinder, inlet="neem", **kw), np.float64)
    bb = oracle_solver(_quiet(cases.cylinder, inlet="bb", **kw), np.float64)
    neem.run(600)
    bb.run(600)
    v = neem.lat.vel
    assert not np.isfinite(v).all() or np.abs(v).max() > 1.0        # blown up
    assert np.isfinite(bb.lat.vel).all() and np.abs(bb.lat.vel).max() < 0.1


def test_reference_cylinder_list_reaches_a_steady_state():
    """cases/Circular cylinder.py's list (VelocityBB + OpenNeumann) at Re ~ 2:
    mass stops changing and the flow settles."""
    case = _quiet(cases.cylinder_reference_bcs, NX=96, NY=32, tau=0.8, u_in=0.02, collision="bgk")
    sol = oracle_solver(case, np.float64)
    sol.run(3000)
    m0, u0 = sol.lat.rho.sum(), sol.lat.vel.copy()
    sol.run(200)
    assert abs(sol.lat.rho.sum() - m0) / m0 < 1e-6
    assert np.abs(sol.lat.vel - u0).max() < 1e-5
