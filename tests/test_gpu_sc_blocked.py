"""Shan-Chen with two steps per launch (k_sc2: the psi of step n on a
three-node ring, step n on two, psi of step n+1 on one, step n+1 on the
tile, all through LDS) against one step per launch (k_sc): bit-identical
populations, macroscopic fields, forces, psi and pressure observables, over
odd and even step counts, runs split into pieces, tiles cut by the lattice
edges, lattices smaller than one tile, solid disks with bounce-back,
adhesion, fixed wall psi, cubic-EOS pseudopotentials and the Huang MRT
source."""
import contextlib
import io

import numpy as np
import pytest

from taichi_lbm import cases
from taichi_lbm.stepper import Stepper

pytestmark = pytest.mark.gpu

FIELDS = ("f", "f2", "rho", "vel", "total_rho", "SCforce", "body_force", "pressure", "total_pressure")

CASES = {
    "c1_partial_tiles": dict(NX=70, NY=45, nc=1, G=-5.0, seed=31),
    "c2_partial_tiles": dict(NX=70, NY=45, nc=2, gcross=0.9, seed=32),
    "c1_smaller_than_a_tile": dict(NX=7, NY=5, nc=1, G=-5.0, seed=33),
    "c2_smaller_than_a_tile": dict(NX=5, NY=9, nc=2, gcross=0.9, seed=34),
    "c1_disk_wall_psi": dict(NX=60, NY=44, nc=1, G=-5.0, seed=35, disk=(30, 22, 7), psi_wall=0.6),
    "c2_disk_adhesion": dict(NX=60, NY=44, nc=2, gcross=0.9, seed=36, disk=(30, 22, 7), gadh=(0.3, -0.2)),
    "c1_cs_eos": dict(NX=45, NY=38, nc=1, G=-1.0, seed=37, rho_mean=0.0293, amp=0.0065,
                      eos=("cs", dict(a=1.17, b=4.0, R=0.1, G=-1.0, rho_cr=0.0326, pressure_cr=0.00517,
                                      temperature_cr=1.104))),
    "c1_bgk": dict(NX=64, NY=40, nc=1, G=-5.0, seed=38, collision="bgk", nu=0.1),
    "c2_trt": dict(NX=64, NY=40, nc=2, gcross=0.9, seed=39, collision="trt", nu=0.1),
    "c1_huang": dict(NX=45, NY=38, nc=1, G=-5.0, seed=40, huang=(0.1, 0.5)),
}


def _pair(kw):
    out = []
    for blocked in (True, False):
        with contextlib.redirect_stdout(io.StringIO()):
            case = cases.shan_chen(**kw)
            cases.initialise(case)
        out.append((case, Stepper(case["lbf"], case["macro"], case["coll"], case["bce"],
                                  boundary_engine_sc=case["bsc"], blocked=blocked)))
    return out


def _same(a, b):
    for name in FIELDS:
        x, y = getattr(a["lbf"], name).to_numpy(), getattr(b["lbf"], name).to_numpy()
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (name, np.abs(x - y).max())
    x, y = a["lbf"].sc_field.psi_field.to_numpy(), b["lbf"].sc_field.psi_field.to_numpy()
    assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), ("psi", np.abs(x - y).max())


@pytest.mark.parametrize("name", sorted(CASES))
def test_two_steps_per_launch_equal_one(name):
    print(f"case shan-chen two-step {name}", flush=True)
    (a, sa), (b, sb) = _pair(CASES[name])
    for n in (11, 1, 2, 6):
        sa.run(n)
        sb.run(n)
        _same(a, b)
    assert np.isfinite(a["lbf"].f.to_numpy()).all()
