"""lbm_run_blocked_noalloc (the entry point that allocates nothing) against
lbm_run_blocked: bit-identical fields on two calls in a row, the second on
a smaller lattice, both power-law runs whose second relaxation-state buffer
lives in the library's fixed device workspace."""
import contextlib
import io

import numpy as np
import pytest

from taichi_lbm import cases
from taichi_lbm.stepper import Stepper

pytestmark = pytest.mark.gpu

FIELDS = ("f", "f2", "rho", "vel", "shear_rate")


def _pair(build):
    out = []
    for no_alloc in (False, True):
        with contextlib.redirect_stdout(io.StringIO()):
            case = build()
            cases.initialise(case)
        out.append((case, Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=True,
                                  blocked=True, no_alloc=no_alloc)))
    return out


def _same(a, b, coll_state):
    for name in FIELDS:
        x, y = getattr(a["lbf"], name).to_numpy(), getattr(b["lbf"], name).to_numpy()
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (name, np.abs(x - y).max())
    x, y = coll_state(a["coll"]), coll_state(b["coll"])
    assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), np.abs(x - y).max()


def test_no_alloc_equals_blocked_on_two_calls_smaller_second():
    import torch
    # first call: a porous MRT power-law lattice (131 072 state values)
    print("case porous 512x256 mrt power law, 17 steps (odd: a copy-back of the workspace)", flush=True)
    (a, sa), (b, sb) = _pair(lambda: cases.porous(NX=512, NY=256, seed=7, collision="mrt"))
    sa.run(17)
    sb.run(17)
    torch.cuda.synchronize()
    _same(a, b, lambda c: c.diag.to_numpy()[..., 7])
    # second call, smaller: a power-law contraction with node rules (two-node ring)
    print("case contraction 160x48 trt power law n=0.6, 22 steps", flush=True)
    (c, sc), (d, sd) = _pair(lambda: cases.contraction(W=160, H=48, power_n=0.6))
    sc.run(22)
    sd.run(22)
    torch.cuda.synchronize()
    _same(c, d, lambda k: k.tau_sym.to_numpy())
    # the workspace keeps the last call's contents: a further run of the first
    # lattice still matches
    print("case porous 512x256 again, 6 more steps", flush=True)
    sa.run(6)
    sb.run(6)
    torch.cuda.synchronize()
    _same(a, b, lambda k: k.diag.to_numpy()[..., 7])
