"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the case builders, the NumPy oracle and the C++ restatement against the
committed vectors, on the CPU.

The two full-size 400 x 100 cylinder fixtures keep every second node (FP32)
and every fourth (FP64) in each direction, and no populations.

Tolerances: the builders' inputs bit for bit; the FP32 oracle within the FP32
parity bounds of tests/tolerances.py (bit-identical on the machine that wrote
the fixtures; summation order inside NumPy may differ elsewhere); the FP64
oracle within 1e-12 absolute; the FP32 fixture within 2e-6 lattice units of
the FP64 fixture's velocity (the FP32 rounding of 20-30 steps, about 3e-7
measured), which ties the two halves of each fixture to one discrete problem.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from tolerances import check

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def _load(name):
    z = np.load(GOLDEN / f"{name}.npz")   # allow_pickle stays False
    return {k: z[k] for k in z.files}


def _build(name):
    from golden.make_golden import build
    return build(name)


def _s(fx, key):
    """The nodes of a full field that fixture array `key` keeps."""
    from golden.make_golden import sub
    k = sub(fx, key)
    return (slice(None, None, k), slice(None, None, k))


def test_every_case_has_a_fixture():
    from golden.make_golden import CASES
    assert sorted(CASES) == NAMES


@pytest.mark.parametrize("name", NAMES)
def test_builders_reproduce_fixture_inputs(name):
    fx = _load(name)
    meta = json.loads(str(fx["meta"]))
    case, steps = _build(name)
    assert steps == meta["steps"]
    assert np.array_equal(case["lbf"].mask.arr.astype(np.int8), fx["in_mask"])
    assert np.array_equal(np.asarray(case["rho"], np.float32), fx["in_rho"])
    assert np.array_equal(np.asarray(case["vel"], np.float32), fx["in_vel"])
    if "in_psi" in fx:
        assert np.array_equal(np.asarray(case["psi"], np.float32), fx["in_psi"])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    from golden.make_golden import run_oracle
    fx = _load(name)
    case, steps = _build(name)
    l32 = run_oracle(case, steps, np.float32, params32=False)
    if "out32_f" in fx:
        check("f", l32.f, fx["out32_f"])
    check("rho", l32.rho[_s(fx, "out32")], fx["out32_rho"])
    check("vel", l32.vel[_s(fx, "out32")], fx["out32_vel"])
    l64 = run_oracle(case, steps, np.float64, params32=True)
    assert np.abs(l64.rho[_s(fx, "out64")] - fx["out64_rho"]).max() <= 1e-12
    assert np.abs(l64.vel[_s(fx, "out64")] - fx["out64_vel"]).max() <= 1e-12
    if "out64_f" in fx:
        assert np.abs(l64.f - fx["out64_f"]).max() <= 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_fp32_fixture_is_the_fp64_problem(name):
    from golden.make_golden import sub
    fx = _load(name)
    r = sub(fx, "out64") // sub(fx, "out32")
    err = np.abs(fx["out32_vel"][::r, ::r].astype(np.float64) - fx["out64_vel"]).max()
    print(f"  {name}: max |u32 - u64| = {err:.3e}")
    assert err <= 2e-6
    assert np.all(np.isfinite(fx["out32_vel"])) and np.all(np.isfinite(fx["out32_rho"]))
    assert np.abs(fx["out32_vel"]).max() > 1e-5       # the flow moved


@pytest.mark.parametrize("name", [n for n in NAMES if not n.startswith("shan_chen")])
def test_cpu_restatement_against_fixture(name):
    """The independent C++ restatement (one component, no Shan-Chen)."""
    from oracle import cpu_ref
    fx = _load(name)
    case, steps = _build(name)
    cl = cpu_ref.from_case(case, nthreads=2)
    cl.run(steps)
    if "out32_f" in fx:
        check("f", cl.f, fx["out32_f"])
    check("rho", cl.rho[_s(fx, "out32")], fx["out32_rho"])
    check("vel", cl.vel[_s(fx, "out32")], fx["out32_vel"])
