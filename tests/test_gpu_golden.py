"""The HIP path against the committed golden fixtures (tests/golden/*.npz).

* Fused steps (`Stepper`, two steps per launch where the lattice allows it)
  against the FP32 oracle's vectors, within the FP32 parity bounds of
  tests/tolerances.py, with the distance to the FP64 vectors stated and
  bounded by 1e-5 lattice units of velocity.
* The reference's own unfused call sequence through the granular kernels,
  against the same vectors.
* The FP64-storage batch path (one-component lattices) against the FP64
  populations and velocities within 1e-12 absolute (the full-size fixtures
  keep ρ and u only): both integrate the same discrete problem in double
  precision, so only rounding-order differences remain (below 2e-15
  measured).
* FP16 / BF16 shifted storage against the FP64 velocities, with the bounds
  the batch tests allow after 200 steps.
"""
import contextlib
import io
import json
from pathlib import Path

import numpy as np
import pytest

from tolerances import check

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def _load(name):
    z = np.load(GOLDEN / f"{name}.npz")
    return {k: z[k] for k in z.files}


def _s(fx, key):
    from golden.make_golden import sub
    k = sub(fx, key)
    return (slice(None, None, k), slice(None, None, k))


def _device_case(name):
    from golden.make_golden import build
    from taichi_lbm import cases
    case, steps = build(name)
    with contextlib.redirect_stdout(io.StringIO()):
        cases.initialise(case)
    return case, steps


@pytest.mark.parametrize("name", NAMES)
def test_fused_against_fixture(name):
    from taichi_lbm import cases
    print(f"golden fused {name}", flush=True)
    fx = _load(name)
    case, steps = _device_case(name)
    cases.stepper(case).run(steps)
    lbf = case["lbf"]
    if "out32_f" in fx:
        check("f", lbf.f.to_numpy(), fx["out32_f"])
    check("rho", lbf.rho.to_numpy()[_s(fx, "out32")], fx["out32_rho"])
    vel = lbf.vel.to_numpy()
    check("vel", vel[_s(fx, "out32")], fx["out32_vel"])
    err64 = np.abs(vel[_s(fx, "out64")].astype(np.float64) - fx["out64_vel"]).max()
    print(f"  max |u - u_fp64| = {err64:.3e}")
    assert err64 < 1e-5


@pytest.mark.parametrize("name", ["cylinder_bgk_48x24", "cylinder_refbcs_bgk_60x24", "porous_powerlaw_mrt_48x40",
                                  "contraction_inside_trt_80x24", "cavity_mrt_40x40", "cylinder_eq_abc_mrt_48x24",
                                  "contraction_abc_trt_80x24"])
def test_granular_sequence_against_fixture(name):
    """lbm_solve's order (cases/Circular cylinder.py:182-193, cases/Power Law
    Fluid.py:160-173) through the one-kernel-per-method entry points; the
    table-stream lists run InsideBoundary (BoundaryCondition.py:435-449)
    and a VelocityBB with direction 3 (inlet) and 2 (lid)."""
    print(f"golden granular {name}", flush=True)
    fx = _load(name)
    meta = json.loads(str(fx["meta"]))
    case, steps = _device_case(name)
    lbf, macro, col, bce = case["lbf"], case["macro"], case["coll"], case["bce"]
    for _ in range(steps):
        macro.density(lbf)
        macro.pressure(lbf)
        macro.force_density(lbf)
        macro.velocity(lbf)
        if meta["power_law"]:
            col.shear_rate(lbf)
            col.relaxation_time(lbf)
        col.apply(lbf)
        bce.apply_boundary_conditions(lbf)
    check("f", lbf.f.to_numpy(), fx["out32_f"])
    check("rho", lbf.rho.to_numpy(), fx["out32_rho"])
    check("vel", lbf.vel.to_numpy(), fx["out32_vel"])


@pytest.mark.parametrize("name", [n for n in NAMES if not n.startswith("shan_chen")])
def test_fp64_batch_against_fixture(name):
    from golden.make_golden import build
    from taichi_lbm.batch import BatchRunner, spec_from_case
    print(f"golden fp64 {name}", flush=True)
    fx = _load(name)
    case, steps = build(name)
    runner = BatchRunner([spec_from_case(case)], storage="f64")
    runner.run(steps)
    out = runner.fields(0)
    e_vel = np.abs(out["vel"][_s(fx, "out64")] - fx["out64_vel"]).max()
    if "out64_f" in fx:
        e_f = np.abs(out["f"] - fx["out64_f"]).max()
        print(f"  max |f - f64| = {e_f:.3e}, max |u - u64| = {e_vel:.3e}")
        assert e_f <= 1e-12 and e_vel <= 1e-12
    else:   # full-size fixtures keep rho and u only (FP64 in the FP64-storage batch)
        e_rho = np.abs(out["rho"][_s(fx, "out64")] - fx["out64_rho"]).max()
        print(f"  max |rho - rho64| = {e_rho:.3e}, max |u - u64| = {e_vel:.3e}")
        assert e_rho <= 1e-12 and e_vel <= 1e-12


def test_fp32_batch_of_all_fixtures():
    """Every one-component fixture lattice in one batch (uneven sizes, BGK,
    TRT and MRT, power law and not, two steps per launch): each lattice
    against its own FP32 vectors."""
    from golden.make_golden import build
    from taichi_lbm.batch import BatchRunner, spec_from_case
    names = [n for n in NAMES if not n.startswith("shan_chen")]
    built = [build(n) for n in names]
    steps = {s for _, s in built}
    for n_steps in sorted(steps):
        group = [(n, c) for n, (c, s) in zip(names, built) if s == n_steps]
        print(f"golden fp32 batch of {len(group)} lattices, {n_steps} steps", flush=True)
        runner = BatchRunner([spec_from_case(c) for _, c in group], storage="f32")
        runner.run(n_steps)
        for q, (name, _) in enumerate(group):
            fx = _load(name)
            out = runner.fields(q)
            print(f" {name}")
            if "out32_f" in fx:
                check("f", out["f"], fx["out32_f"])
            check("vel", out["vel"][_s(fx, "out32")], fx["out32_vel"])


# measured at most 4.7e-5 (FP16) and 3.3e-4 (BF16, above the 1e-4 bar) over the fixtures
@pytest.mark.parametrize("storage,bound", [("f16", 7e-5), ("bf16", 5e-4)])
def test_shifted_storage_against_fixture(storage, bound):
    """FP16 / BF16 shifted population storage (f - w stored, arithmetic in
    FP32): velocity against the FP64 vectors, within about 1.5 times the
    largest error measured over the fixtures."""
    from golden.make_golden import build
    from taichi_lbm.batch import BatchRunner, spec_from_case
    for name in [n for n in NAMES if not n.startswith("shan_chen")]:
        fx = _load(name)
        case, steps = build(name)
        runner = BatchRunner([spec_from_case(case)], storage=storage)
        runner.run(steps)
        du = np.abs(runner.fields(0)["vel"][_s(fx, "out64")] - fx["out64_vel"]).max()
        print(f"  {storage} {name}: max |u - u_fp64| = {du:.2e}", flush=True)
        assert du < bound
