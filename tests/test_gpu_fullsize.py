"""Parity at full configuration size.

The 8192 x 8192 porous lattice the headline figure is measured on (MRT, and
TRT with magic 1/4; power law n = 0.7, tau clamped to [0.51, 2.0], gravity
1e-6, seeded disks of radius 8-32 to solid fraction 0.3; `cases/Power Law
Fluid.py`'s step order):

* 10 fused steps on the GPU (two per launch) against 10 steps of the C++
  restatement of the reference's unfused kernels (oracle/cpu_ref.cpp) on the
  same lattice, within the FP32 parity bounds of tests/tolerances.py;
* the fields do not depend on the tile width: 64- and 32-column tiles give
  the same populations bit for bit;
* FP64 mass (the sum of rho over the fluid group) is conserved to FP32
  rounding (bounce-back walls and periodic edges move mass, never create it):
  relative drift below 1e-6 over the 10 steps.

The 2048 x 2048 two-component Shan-Chen separation (MRT, G_cross = 0.9,
densities 1 +- 1 % seeded per component; `cases/Bubble or two-phase
separation.py`'s step order): 40 steps of the LDS-tiled Shan-Chen kernels
(two steps per launch) against 40 steps of the C++ restatement (its
Shan-Chen step agrees with the NumPy oracle, tests/test_cpu_restatement.py),
every field within the FP32 parity bounds.

The 512 batched cylinder channels: a sample of lattices against the same
lattices run alone (bit for bit) and against the C++ restatement.

Served contraction solves at the serving workload's request sizes: each
result equal to the same solve run alone, bit for bit.
"""
import contextlib
import io
import os

import numpy as np
import pytest

from tolerances import check

pytestmark = pytest.mark.gpu

N, STEPS = 8192, 10


def _case(collision):
    from taichi_lbm import cases
    with contextlib.redirect_stdout(io.StringIO()):
        return cases.porous(NX=N, NY=N, seed=1, collision=collision)


def _gpu_run(case, tile_cols):
    from taichi_lbm import cases
    from taichi_lbm.stepper import Stepper
    with contextlib.redirect_stdout(io.StringIO()):
        cases.initialise(case)
    st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=True, tile_cols=tile_cols)
    st.run(STEPS)
    return st


@pytest.mark.parametrize("collision", ["mrt", "trt"])
def test_headline_lattice_matches_cpu_restatement(collision):
    from oracle import cpu_ref
    print(f"case porous {N}x{N} {collision} power law, {STEPS} steps", flush=True)
    case = _case(collision)
    threads = min(16, os.cpu_count() or 1)
    cl = cpu_ref.from_case(case, nthreads=threads)
    cl.run(STEPS)
    st = _gpu_run(case, 64)
    lbf = case["lbf"]
    f_w64 = lbf.f.to_numpy()
    check("f", f_w64, cl.f)
    check("rho", lbf.rho.to_numpy(), cl.rho)
    check("vel", lbf.vel.to_numpy(), cl.vel)
    check("shear_rate", lbf.shear_rate.to_numpy(), cl.shear_rate)
    red = st.reduce()
    assert red["nonfinite"] == 0
    mass = float(red["mass"][0])
    m_init = float(np.asarray(case["rho"], np.float64)[lbf.mask.arr == 1].sum())
    print(f"  FP64 mass {mass:.6f}, initial {m_init:.6f}, relative drift {abs(mass - m_init) / m_init:.2e}")
    assert abs(mass - m_init) / m_init < 1e-6
    del st, cl
    # the same lattice on 32-column tiles
    case32 = _case(collision)
    _gpu_run(case32, 32)
    assert np.array_equal(case32["lbf"].f.to_numpy(), f_w64)


def test_shan_chen_2048_two_components_matches_oracle():
    import os
    from oracle import cpu_ref
    from taichi_lbm import cases
    n = 40
    threads = min(16, os.cpu_count() or 1)
    print(f"case shan-chen 2048x2048x2 MRT, {n} steps (C++ restatement on {threads} threads)", flush=True)
    with contextlib.redirect_stdout(io.StringIO()):
        case = cases.shan_chen(NX=2048, NY=2048, nc=2, gcross=0.9, seed=3)
    cl = cpu_ref.from_case(case, nthreads=threads)
    cl.run(n)
    with contextlib.redirect_stdout(io.StringIO()):
        cases.initialise(case)
    st = cases.stepper(case)
    st.run(n)
    lbf = case["lbf"]
    for name in ("f", "rho", "vel", "total_rho", "SCforce", "body_force"):
        check(name, getattr(lbf, name).to_numpy(), getattr(cl, name))
    check("psi", lbf.sc_field.psi_field.to_numpy(), cl.psi)
    assert st.reduce()["nonfinite"] == 0


def test_512_batched_channels_sample():
    """The 512-lattice batch of independent cylinder channels (NX ~ U{128..1024},
    NY ~ U{64..256}, r ~ U[4, NY/8], u ~ U[0.01, 0.08], tau ~ U[0.52, 0.8], BGK
    or TRT by a seeded coin; VelocityBB inlet, PressureNEEM outlet), 20 steps
    at two per launch: a seeded sample of 8 lattices (the largest among them)
    equals each lattice run alone through the Stepper bit for bit, and 3 of
    them agree with the C++ restatement within the FP32 parity bounds."""
    from oracle import cpu_ref
    from taichi_lbm import cases
    from taichi_lbm.batch import BatchRunner, spec_from_case
    from taichi_lbm.stepper import Stepper
    n, count = 20, 512
    print(f"case {count} batched channels, {n} steps", flush=True)
    rng = np.random.default_rng(1)
    kws = []
    for q in range(count):
        NX, NY = int(rng.integers(128, 1025)), int(rng.integers(64, 257))
        r, u = float(rng.uniform(4, NY / 8)), float(rng.uniform(0.01, 0.08))
        tau, coll = float(rng.uniform(0.52, 0.8)), ("bgk" if rng.random() < 0.5 else "trt")
        kws.append(dict(NX=NX, NY=NY, cx=max(NX / 5, r + 4), cy=NY / 2, r=r, tau=tau, u_in=u,
                        collision=coll, name=f"b{q}", inlet="bb"))

    def build(q):
        with contextlib.redirect_stdout(io.StringIO()):
            return cases.cylinder(**kws[q])

    runner = BatchRunner([spec_from_case(build(q)) for q in range(count)], storage="f32")
    runner.run(n)
    largest = max(range(count), key=lambda q: kws[q]["NX"] * kws[q]["NY"])
    sample = sorted({largest, *np.random.default_rng(2).choice(count, 7, replace=False).tolist()})
    for i, q in enumerate(sample):
        print(f"  lattice {q}: {kws[q]['NX']} x {kws[q]['NY']} {kws[q]['collision']}", flush=True)
        got = runner.fields(q)
        c = build(q)
        if i < 3:
            cl = cpu_ref.from_case(c, nthreads=min(16, os.cpu_count() or 1))
            cl.run(n)
            for name in ("f", "rho", "vel"):
                check(name, got[name], getattr(cl, name))
        with contextlib.redirect_stdout(io.StringIO()):
            cases.initialise(c)
        Stepper(c["lbf"], c["macro"], c["coll"], c["bce"]).run(n)
        for name in ("f", "f2", "rho", "vel"):
            assert np.array_equal(got[name].astype(np.float32), getattr(c["lbf"], name).to_numpy()), (q, name)


def test_served_contraction_solves_at_full_request_sizes():
    """Contraction-flow power-law requests at the request sizes of the serving
    workload (widths 256-2048, heights 64-512, 4:1 contraction, n in
    [0.5, 1.5]), arriving while others run: each result equals the same solve
    run alone bit for bit."""
    from taichi_lbm.batch import spec_from_case
    from taichi_lbm.serving import Request, SolveServer, solve_alone
    from taichi_lbm import cases
    sizes = [(2048, 512, 0.7), (256, 64, 1.5), (1024, 256, 0.5), (640, 160, 1.0), (1536, 384, 1.2)]
    print(f"case {len(sizes)} served contraction solves, 2048 x 512 the largest", flush=True)
    specs = []
    with contextlib.redirect_stdout(io.StringIO()):
        for q, (W, Hh, n) in enumerate(sizes):
            specs.append(spec_from_case(cases.contraction(W=W, H=Hh, power_n=n, name=f"r{q}")))
    arrivals = [0.0, 0.0, 0.01, 0.02, 0.03]
    max_steps = 600
    reqs = [Request(q, arrivals[q], s, max_steps=max_steps, tol=1e-12) for q, s in enumerate(specs)]
    srv = SolveServer(storage="f32", hbm_budget_bytes=1 << 32, max_slots=8, tick_steps=100)
    stats = srv.serve(reqs)
    srv.close()
    assert len(stats.completed) == len(sizes)
    for r in stats.completed:
        u, steps, res = solve_alone(specs[r.rid], max_steps, 1e-12, 100, "f32")
        print(f"  request {r.rid} {sizes[r.rid][:2]}: steps {r.steps} vs {steps}", flush=True)
        assert steps == r.steps
        assert np.all(np.isfinite(u))
        assert np.array_equal(u, r.result)
