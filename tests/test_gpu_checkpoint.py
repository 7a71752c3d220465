"""Checkpoint / resume: a run continued from a saved state equals the
uninterrupted run bit for bit (power-law relaxation state, Shan-Chen psi,
step counter)."""
import contextlib
import io

import numpy as np
import pytest

from taichi_lbm import cases, checkpoint
from taichi_lbm.stepper import Stepper

pytestmark = pytest.mark.gpu


def _quiet(fn, *a, **k):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


def _stepper(case, **kw):
    return Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], boundary_engine_sc=case.get("bsc"),
                   advance_time=True, **kw)


@pytest.mark.parametrize("kind", ["porous_power_law", "shan_chen_2"])
def test_resume_equals_uninterrupted(kind, tmp_path):
    print(f"case checkpoint {kind}", flush=True)
    if kind == "porous_power_law":
        build = lambda: cases.porous(NX=96, NY=80, seed=3, collision="mrt", rmin=4.0, rmax=9.0, g=(2e-5, 1e-6))
        extra = dict(power_law=True)
        names = ("f", "f2", "rho", "vel", "total_rho", "shear_rate", "body_force")
    else:
        build = lambda: cases.shan_chen(NX=70, NY=45, nc=2, gcross=0.9, seed=5)
        extra = {}
        names = ("f", "f2", "rho", "vel", "total_rho", "SCforce", "pressure", "total_pressure")
    a = _quiet(cases.initialise, _quiet(build))
    _stepper(a, **extra).run(23)
    b = _quiet(cases.initialise, _quiet(build))
    _stepper(b, **extra).run(10)
    path = tmp_path / "state.npz"
    checkpoint.save(path, b["lbf"], b["coll"])
    c = _quiet(build)                      # fresh lattice: geometry and models from the case, state from the file
    checkpoint.load(path, c["lbf"], c["coll"])
    assert c["lbf"].time[None] == 10
    _stepper(c, **extra).run(13)
    assert c["lbf"].time[None] == 23 == a["lbf"].time[None]
    for name in names:
        assert np.array_equal(getattr(a["lbf"], name).to_numpy(), getattr(c["lbf"], name).to_numpy()), name
    if kind == "porous_power_law":
        assert np.array_equal(a["coll"].diag.to_numpy(), c["coll"].diag.to_numpy())
    else:
        assert np.array_equal(a["lbf"].sc_field.psi_field.to_numpy(), c["lbf"].sc_field.psi_field.to_numpy())
