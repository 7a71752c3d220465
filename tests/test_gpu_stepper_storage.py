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
* Stochastic rounding of the stores (rounding="stochastic"): a steady
  channel in 16-bit storage against the same channel in FP32 after 20k
  steps; two runs of one lattice are equal bit for bit, a different step
  count changes the rounding, and a BatchRunner (eager runs and graph
  replays, whose step counts live in device memory) rounds the lattice as
  its Stepper does.
"""
import contextlib
import io

import numpy as np
import pytest

from test_gpu_golden import NAMES, _device_case, _load, _s

pytestmark = pytest.mark.gpu

BOUND = {"f16": 7e-5, "bf16": 5e-4}


def _stepper(case, storage, rounding="nearest"):
    from taichi_lbm.stepper import Stepper
    return Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=case["power_law"],
                   boundary_engine_sc=case.get("bsc"), storage=storage, rounding=rounding)


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


def _channel(u_in=0.02):
    from taichi_lbm import cases
    with contextlib.redirect_stdout(io.StringIO()):
        c = cases.cylinder(NX=120, NY=40, cx=30, cy=20, r=5, tau=0.8, u_in=u_in, inlet="bb", outlet="abc",
                           name="sr-channel")
        cases.initialise(c)
    return c


def _channel_error(storage, rounding, u_in, steps=20000):
    from taichi_lbm.stepper import Stepper
    ref, c16 = _channel(u_in), _channel(u_in)
    Stepper(ref["lbf"], ref["macro"], ref["coll"], ref["bce"]).run(steps)
    Stepper(c16["lbf"], c16["macro"], c16["coll"], c16["bce"], storage=storage, rounding=rounding).run(steps)
    u32, u16 = ref["lbf"].vel.to_numpy(), c16["lbf"].vel.to_numpy()
    assert np.isfinite(u16).all()
    return float(np.abs(u16 - u32).max())


# A steady channel (Re < 2, 20k steps; the FP32 velocity moves by 2e-7 over
# the next 20k) in 16-bit storage against FP32 storage.  Rounded to nearest,
# a steady flow rounds every population the same way every step and the
# bias stays in the solution; stochastic stores leave about one unit in the
# last place of noise instead.  Measured (max |u - u_f32|):
#   u_in = 0.02 (max |u| 0.037): FP16 6.1e-5 -> 1.7e-5, BF16 3.1e-4 -> 1.5e-4
#   u_in = 0.005 (max |u| 0.009): FP16 1.6e-5 -> 6.2e-6, BF16 1.05e-4 -> 3.3e-5
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_stochastic_rounding_steady_channel(storage):
    for u_in, bar in ((0.02, storage == "f16"), (0.005, True)):
        print(f"steady channel 120x40 u_in {u_in}, 20000 steps, {storage} against f32", flush=True)
        e_near = _channel_error(storage, "nearest", u_in)
        e_sr = _channel_error(storage, "stochastic", u_in)
        print(f"  max |u_{storage} - u_f32|: nearest {e_near:.2e}, stochastic {e_sr:.2e}", flush=True)
        assert e_sr < 0.7 * e_near
        if bar:
            assert e_sr < 1e-4      # the velocity bar


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_stochastic_rounding_is_deterministic(storage):
    from taichi_lbm.stepper import Stepper
    print(f"rounding seeds {storage}: repeat and shifted step count", flush=True)
    a, b, c = _channel(), _channel(), _channel()
    for case in (a, b, c):
        st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], storage=storage, rounding="stochastic")
        if case is c:
            case["lbf"]._round_step = 1000     # the same state, other seeds
        st.run(200)
    ua, ub, uc = (x["lbf"].vel.to_numpy() for x in (a, b, c))
    assert np.array_equal(ua, ub)
    d = np.abs(ua - uc).max()
    print(f"  other seeds move u by {d:.2e}", flush=True)
    assert 0 < d < (1e-4 if storage == "f16" else 1e-3)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_stochastic_rounding_batch_matches_stepper(storage):
    from golden.make_golden import build
    from taichi_lbm.batch import BatchRunner, spec_from_case
    from taichi_lbm.stepper import Stepper
    name = "cylinder_trt_vbb_48x24"
    print(f"stochastic {storage}: Stepper against BatchRunner (eager + graph replays), {name}", flush=True)
    case, _ = _device_case(name)
    st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=case["power_law"],
                 storage=storage, rounding="stochastic")
    for n in (10, 20, 20):
        st.run(n)
    u = case["lbf"].vel.to_numpy()
    with contextlib.redirect_stdout(io.StringIO()):
        bcase, _ = build(name)
    import torch
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        runner = BatchRunner([spec_from_case(bcase)], storage=storage, rounding="stochastic")
        runner.run(10)
        g = runner.capture(20)
        g.replay()
        g.replay()
        ub = runner.fields(0)["vel"]
        g.close()
        runner.close()
    torch.cuda.synchronize()
    print(f"  max |u_stepper - u_batch| = {np.abs(u - ub).max():.2e}", flush=True)
    assert np.array_equal(u, ub)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_stochastic_rounding_shan_chen(storage):
    """Shan-Chen lattices in 16-bit storage (psi pass + one-step kernel)
    round stochastically too: two runs equal bit for bit, and the fields stay
    within the FP64 fixture bounds of rounding to nearest."""
    name = "shan_chen_c2_32x24"
    print(f"stochastic {storage} Stepper {name}", flush=True)
    fx = _load(name)
    outs = []
    for _ in range(2):
        case, steps = _device_case(name)
        _stepper(case, storage, rounding="stochastic").run(steps)
        outs.append((case["lbf"].vel.to_numpy(), case["lbf"].rho.to_numpy()))
    (u, rho), (u2, rho2) = outs
    assert np.array_equal(u, u2) and np.array_equal(rho, rho2)
    du = np.abs(u[_s(fx, "out64")] - fx["out64_vel"]).max()
    drho = np.abs(rho[_s(fx, "out64")] - fx["out64_rho"]).max()
    print(f"  max |u - u_fp64| = {du:.2e}, max |rho - rho_fp64| = {drho:.2e}", flush=True)
    assert np.isfinite(u).all()
    assert du < 4 * BOUND[storage] and drho < 40 * BOUND[storage]
