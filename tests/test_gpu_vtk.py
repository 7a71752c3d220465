"""Files written from GPU runs (PostProcessing.py:70-156): the VTK files of a
fused run, parsed back, hold the oracle's fields of the same run within the
FP32 parity bounds of tests/tolerances.py (the ASCII values carry 8
decimals, 5e-9 below those bounds)."""
import contextlib
import io

import numpy as np
import pytest

from oracle_bridge import oracle_solver
from taichi_lbm import cases
from taichi_lbm.PostProcessing import PostProcessingEngine
from taichi_lbm.stepper import Stepper
from test_postprocessing import _vtk_blocks
from tolerances import check

pytestmark = pytest.mark.gpu


def _grid(values, pts, NX, NY):
    """Values in file order -> (NX, NY[, 3]) arrays indexed by node."""
    out = np.zeros((NX, NY) + values.shape[1:])
    out[pts[:, 0].astype(int), pts[:, 1].astype(int)] = values
    return out


def test_vtk_of_a_fused_cylinder_run(tmp_path):
    print("case cylinder 64x32 trt, 40 steps, writeVTK", flush=True)
    with contextlib.redirect_stdout(io.StringIO()):
        case = cases.initialise(cases.cylinder(NX=64, NY=32, cx=16, cy=15, r=5, tau=0.6, u_in=0.05,
                                               collision="trt"))
    n = 40
    Stepper(case["lbf"], case["macro"], case["coll"], case["bce"]).run(n)
    sol = oracle_solver(case, np.float32)
    sol.run(n)
    lbf = case["lbf"]
    with contextlib.redirect_stdout(io.StringIO()):
        PostProcessingEngine(0).writeVTK(str(tmp_path / "cyl"), lbf)
    dims, pts, blocks = _vtk_blocks((tmp_path / "cyl.vtk").read_text())
    assert dims == (64, 32, 1)
    rho = _grid(blocks["density"][1], pts, 64, 32) / lbf.C_rho
    vel = _grid(blocks["velocity"][1], pts, 64, 32) / lbf.Cu
    assert np.all(vel[..., 2] == 0)
    check("total_rho", rho, sol.lat.total_rho)
    check("vel", vel[..., :2], sol.lat.vel)
    # and the file is the run's fields, to the format's 8 decimals
    assert np.abs(rho - lbf.total_rho.to_numpy()).max() <= 5e-9
    assert np.abs(vel[..., :2] - lbf.vel.to_numpy()).max() <= 5e-9


def test_vtk_psi_of_a_fused_shan_chen_run(tmp_path):
    print("case shan-chen 40x36 nc=2, 30 steps, writeVTKpsi", flush=True)
    with contextlib.redirect_stdout(io.StringIO()):
        case = cases.shan_chen(NX=40, NY=36, nc=2, gcross=0.9, seed=7)
        cases.initialise(case)
    n = 30
    Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], boundary_engine_sc=case["bsc"]).run(n)
    sol = oracle_solver(case, np.float32)
    sol.run(n)
    lbf = case["lbf"]
    with contextlib.redirect_stdout(io.StringIO()):
        PostProcessingEngine(0).writeVTKpsi(str(tmp_path / "psi"), lbf)
    dims, pts, blocks = _vtk_blocks((tmp_path / "psi.vtk").read_text())
    assert dims == (42, 38, 1)
    psi0 = _grid(blocks["PSI"][1], pts, 42, 38)
    check("psi", psi0, sol.lat.psi[:, :, 0])
    assert np.abs(psi0 - lbf.sc_field.psi_field.to_numpy()[:, :, 0]).max() <= 5e-9
