"""Two steps per launch (lbm_run_blocked) against one step per launch
(lbm_run): bit-identical populations, macroscopic fields and relaxation state,
over odd and even step counts, restarts after a readback, tiles cut by the
lattice edges, lattices smaller than one tile, and node rules (two-node
ring)."""
import contextlib
import io

import numpy as np
import pytest

from taichi_lbm import cases
from taichi_lbm.stepper import Stepper

pytestmark = pytest.mark.gpu


def _quiet(fn, *a, **k):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


def _pair(build, power_law, tile_cols=0):
    a, b = _quiet(build), _quiet(build)
    _quiet(cases.initialise, a)
    _quiet(cases.initialise, b)
    sa = Stepper(a["lbf"], a["macro"], a["coll"], a["bce"], power_law=power_law, blocked=True, tile_cols=tile_cols)
    sb = Stepper(b["lbf"], b["macro"], b["coll"], b["bce"], power_law=power_law, blocked=False)
    return a, b, sa, sb


def _same(a, b, names):
    for name in names:
        x, y = getattr(a["lbf"], name).to_numpy(), getattr(b["lbf"], name).to_numpy()
        assert np.array_equal(x, y), (name, np.abs(x - y).max())


@pytest.mark.parametrize("coll,tile_cols", [("bgk", 64), ("trt", 32), ("mrt", 64), ("mrt", 32), ("mrt", 16)])
def test_blocked_equals_one_step_power_law(coll, tile_cols):
    print(f"case porous 200x136 {coll} power law, {tile_cols}-column tiles", flush=True)
    build = lambda: cases.porous(NX=200, NY=136, seed=5, collision=coll, rmin=3.0, rmax=8.0, g=(2e-5, 1e-6))
    a, b, sa, sb = _pair(build, True, tile_cols)
    # the second state buffer starts as NaN: nodes outside the fluid group
    # must keep their state whatever it holds
    import torch
    sa._tau_next = torch.full_like(a["coll"].state_tensor(), float("nan"))
    for n in (7, 10):                     # odd, then even without a readback in between
        sa.run(n)
        sb.run(n)
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho", "shear_rate", "body_force"))
    state = {"bgk": "tau", "trt": "tau_sym", "mrt": "diag"}[coll]
    assert np.array_equal(getattr(a["coll"], state).to_numpy(), getattr(b["coll"], state).to_numpy())
    sa.run(5)                             # restart from materialised populations
    sb.run(5)
    _same(a, b, ("f", "f2", "rho", "vel", "shear_rate"))


def test_blocked_equals_one_step_constant_tau():
    build = lambda: cases.porous(NX=96, NY=200, seed=6, collision="bgk", rmin=3.0, rmax=6.0, g=(1e-5, 0.0))
    a, b, sa, sb = _pair(build, False)
    sa.run(12)
    sb.run(12)
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho"))


@pytest.mark.parametrize("tile_cols", [16, 32, 64])
def test_lattice_smaller_than_a_tile(tile_cols):
    build = lambda: cases.porous(NX=11, NY=30, seed=7, collision="mrt", rmin=1.5, rmax=2.5, g=(1e-5, 0.0))
    a, b, sa, sb = _pair(build, True, tile_cols)
    sa.run(9)
    sb.run(9)
    _same(a, b, ("f", "f2", "rho", "vel"))


@pytest.mark.parametrize("inlet,tile_cols,NX", [("bb", 64, 23), ("neem", 64, 23), ("bb", 32, 21), ("bb", 16, 43)])
def test_node_rules_two_steps(inlet, tile_cols, NX):
    """Node rules (VelocityBB / VelocityNEEM inlet, PressureNEEM outlet) in
    the two-step kernel with a two-node ring; NX - 1 is a multiple of the
    tile rows (11 with 64 columns, 10 with 32, 21 with 16), so the outlet
    row is the first row of a tile and its interior neighbour lies in the
    ring."""
    build = lambda: cases.cylinder(NX=NX, NY=70, cx=9, cy=35, r=4, collision="trt", inlet=inlet)
    a, b, sa, sb = _pair(build, False, tile_cols)
    sa.run(21)
    sb.run(21)
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho"))


def test_reference_cylinder_rules_two_steps():
    build = lambda: cases.cylinder_reference_bcs(NX=58, NY=40, collision="mrt", tau_bulk=0.9)
    a, b, sa, sb = _pair(build, False)
    sa.run(16)
    sb.run(16)
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho"))


def test_large_lattice_edge_tiles():
    """154 x 24 tiles of 13 x 64 nodes, cut by both lattice edges
    (2000 = 153 x 13 + 11, 1500 = 23 x 64 + 28)."""
    print("case porous 2000x1500 mrt power law", flush=True)
    build = lambda: cases.porous(NX=2000, NY=1500, seed=8, collision="mrt", rmin=4.0, rmax=12.0, g=(2e-5, 1e-6))
    a, b, sa, sb = _pair(build, True)
    for n in (7, 10):
        sa.run(n)
        sb.run(n)
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho", "shear_rate"))
    assert np.array_equal(a["coll"].diag.to_numpy(), b["coll"].diag.to_numpy())


@pytest.mark.parametrize("tile_cols", [16, 32, 64])
def test_large_cylinder_node_rules(tile_cols):
    """A two-node ring over thousands of tiles (11 x 64, 10 x 32 or 21 x 16
    nodes): VelocityBB inlet and PressureNEEM outlet rows."""
    print(f"case cylinder 1500x1300 trt, {tile_cols}-column tiles", flush=True)
    build = lambda: cases.cylinder(NX=1500, NY=1300, cx=300, cy=650, r=60, collision="trt", inlet="bb")
    a, b, sa, sb = _pair(build, False, tile_cols)
    sa.run(9)
    sb.run(9)
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho"))


@pytest.mark.parametrize("coll,tile_cols", [("bgk", 64), ("mrt", 32), ("trt", 16)])
def test_pulls_from_outside_the_fluid_group(coll, tile_cols):
    """No BounceBackWall: fluid nodes next to a disk stream the disk's
    unchanged f2 (nodes marked PULLED, staged in LDS by the two-step
    kernel); the disks hold the rest equilibrium w_k."""
    print(f"case porous 150x140 {coll} without walls", flush=True)
    build = lambda: cases.porous(NX=150, NY=140, seed=9, collision=coll, rmin=3.0, rmax=7.0, g=(1e-5, 0.0),
                                 walls=False)
    a, b, sa, sb = _pair(build, False, tile_cols)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    for case in (a, b):
        solid = case["lbf"].mask.arr == -1
        assert solid.any()
        for name in ("f", "f2"):
            x = getattr(case["lbf"], name).to_numpy()
            x[solid] = w
            getattr(case["lbf"], name).from_numpy(x)
    sa.run(11)
    sb.run(11)
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho"))
    assert np.all(np.isfinite(a["lbf"].f2.to_numpy()))


def test_autotuned_tile_width_changes_nothing():
    """tile_cols="auto": the first long run times both widths on ordinary
    steps of the run and keeps the faster; the fields equal the one-step
    path's bit for bit (power law, so the state buffers alternate too)."""
    print("case porous 200x136 mrt power law, autotuned width", flush=True)
    build = lambda: cases.porous(NX=200, NY=136, seed=5, collision="mrt", rmin=3.0, rmax=8.0, g=(2e-5, 1e-6))
    a, b, sa, sb = _pair(build, True, "auto")
    n = Stepper.tune_steps() + 3
    sa.run(n)
    sb.run(n)
    assert set(sa.tune_timings) == set(Stepper.TUNE_CANDIDATES) and sa.tile_cols in Stepper.TUNE_CANDIDATES
    _same(a, b, ("f", "f2", "rho", "vel", "total_rho", "shear_rate", "body_force"))
    assert np.array_equal(a["coll"].diag.to_numpy(), b["coll"].diag.to_numpy())
    sa.run(6)
    sb.run(6)
    _same(a, b, ("f", "f2", "rho", "vel"))
