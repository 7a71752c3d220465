"""Shifted FP16 / BF16 population storage on a single lattice (``Stepper(...,
storage="f16" | "bf16")``): the fused kernels run on 16-bit populations
g = f - w_k with FP32 arithmetic.

* Against the FP64 vectors of every golden fixture, within the bounds the
  batched 16-bit path is held to (tests/test_gpu_golden.py); the same
  lattice in a 16-bit BatchRunner gives the same velocities bit for bit
  (the tiles run the same per-node code on the same stored inputs).
* Runs without a readback in between equal one run bit for bit; a readback
  in between moves the fields by a rounding of the post-boundary
  populations only.
* Shan-Chen lattices (one-step path with the psi pass) against the FP64
  vectors.
"""
import contextlib
import io

import numpy as np
import pytest

from test_gpu_golden import NAMES, _device_case, _load, _s

pytestmark = pytest.mark.gpu

BOUND = {"f16": 7e-5, "bf16": 5e-4}


def _stepper(case, storage):
    from taichi_lbm.stepper import Stepper
    return Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=case["power_law"],
                   boundary_engine_sc=case.get("bsc"), storage=storage)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_single_lattice_shifted_storage_against_fixtures(storage):
    from golden.make_golden import build
    from taichi_lbm.batch import BatchRunner, spec_from_case
    for name in [n for n in NAMES if not n.startswith("shan_chen")]:
        print(f"golden {storage} Stepper {name}", flush=True)
        fx = _load(name)
        case, steps = _device_case(name)
        _stepper(case, storage).run(steps)
        u = case["lbf"].vel.to_numpy()
        du = np.abs(u[_s(fx, "out64")] - fx["out64_vel"]).max()
        with contextlib.redirect_stdout(io.StringIO()):
            bcase, _ = build(name)
        runner = BatchRunner([spec_from_case(bcase)], storage=storage)
        runner.run(steps)
        ub = runner.fields(0)["vel"]
        runner.close()
        dbatch = np.abs(u - ub).max()
        print(f"  max |u - u_fp64| = {du:.2e}; max |u - u_batch({storage})| = {dbatch:.2e}", flush=True)
        assert np.isfinite(u).all()
        assert du < BOUND[storage]
        assert np.array_equal(u, ub)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_split_runs_and_readback(storage):
    from taichi_lbm import cases
    print(f"split runs {storage}: porous power law 96x80 MRT", flush=True)

    def make():
        with contextlib.redirect_stdout(io.StringIO()):
            c = cases.porous(NX=96, NY=80, seed=3, collision="mrt", rmin=4.0, rmax=9.0, g=(2e-5, 1e-6))
            cases.initialise(c)
        return c

    a, b, c = make(), make(), make()
    sa, sb, sc = _stepper(a, storage), _stepper(b, storage), _stepper(c, storage)
    sa.run(51)
    sb.run(30)
    sb.run(21)
    sc.run(30)
    _ = c["lbf"].vel.to_numpy()            # readback in between
    sc.run(21)
    for name in ("f", "f2", "rho", "vel"):
        x, y = getattr(a["lbf"], name).to_numpy(), getattr(b["lbf"], name).to_numpy()
        assert np.array_equal(x, y), name
    du = np.abs(a["lbf"].vel.to_numpy() - c["lbf"].vel.to_numpy()).max()
    print(f"  readback in between moves u by {du:.2e}", flush=True)
    assert du < BOUND[storage] / 10
    # the FP32 fields after a 16-bit run are decoded populations: granular
    # calls and a new FP32 Stepper continue from them
    from taichi_lbm.stepper import Stepper
    s32 = Stepper(a["lbf"], a["macro"], a["coll"], a["bce"], power_law=True)
    s32.run(2)
    assert np.isfinite(a["lbf"].vel.to_numpy()).all()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("name", [n for n in NAMES if n.startswith("shan_chen")])
def test_shan_chen_shifted_storage_against_fixture(storage, name):
    print(f"golden {storage} Stepper {name}", flush=True)
    fx = _load(name)
    case, steps = _device_case(name)
    _stepper(case, storage).run(steps)
    lbf = case["lbf"]
    u = lbf.vel.to_numpy()
    rho = lbf.rho.to_numpy()
    du = np.abs(u[_s(fx, "out64")] - fx["out64_vel"]).max()
    drho = np.abs(rho[_s(fx, "out64")] - fx["out64_rho"]).max()
    # the shift is w_k (a density of 1): a population's storage error is about
    # eps w_k whatever the density, so the velocity error grows as 1 / rho
    # (the cubic-EOS lattice sits at rho = 0.032)
    r64 = fx["out64_rho"]
    rho_mean = float(r64[r64 > 0].mean())
    print(f"  max |u - u_fp64| = {du:.2e}, max |rho - rho_fp64| = {drho:.2e} (mean rho {rho_mean:.3f})", flush=True)
    assert np.isfinite(u).all()
    assert du < 4 * BOUND[storage] / min(rho_mean, 1.0) and drho < 40 * BOUND[storage]
