"""Continuous batching: each result is the same alone and in the mix, and
admission respects the HBM budget."""
import contextlib
import io

import numpy as np
import pytest

from taichi_lbm import cases
from taichi_lbm.batch import footprint_bytes, spec_from_case
from taichi_lbm.serving import Request, SolveServer, solve_alone

pytestmark = pytest.mark.gpu


def _specs():
    out = []
    with contextlib.redirect_stdout(io.StringIO()):
        for q, (W, Hh, n) in enumerate([(96, 32, 0.7), (128, 40, 1.3), (80, 24, 1.0), (112, 48, 0.5)]):
            out.append(spec_from_case(cases.contraction(W=W, H=Hh, power_n=n, u_in=0.03, name=f"s{q}")))
    return out


def test_alone_equals_in_the_mix():
    specs = _specs()
    arrivals = [0.0, 0.0, 0.02, 0.05]
    reqs = [Request(q, arrivals[q], s, max_steps=1200, tol=1e-9) for q, s in enumerate(specs)]
    srv = SolveServer(storage="f32", hbm_budget_bytes=1 << 28, max_slots=8, tick_steps=100)
    stats = srv.serve(reqs)
    srv.close()
    assert len(stats.completed) == 4
    # four admissions and four retirements, but the captured tick takes most
    # of them in place
    print(f"  ticks {stats.ticks}, graphs captured {stats.captures}", flush=True)
    assert 1 <= stats.captures < 8
    for r in stats.completed:
        u, steps, res = solve_alone(specs[r.rid], 1200, 1e-9, 100, "f32")
        print(f"  request {r.rid}: steps {r.steps} vs {steps}, residual {r.residual:.2e}", flush=True)
        assert steps == r.steps
        assert np.array_equal(u, r.result)
        assert np.all(np.isfinite(u))


def test_admission_respects_budget():
    specs = _specs()
    budget = footprint_bytes(specs[3]) + footprint_bytes(specs[1]) + 1
    reqs = [Request(q, 0.0, s, max_steps=400, tol=0.0) for q, s in enumerate(specs)]
    srv = SolveServer(storage="f32", hbm_budget_bytes=budget, max_slots=8, tick_steps=100)
    stats = srv.serve(reqs)
    srv.close()
    assert len(stats.completed) == 4
    assert stats.peak_bytes <= budget
    assert stats.peak_resident <= 3
    assert all(r.steps == 400 for r in stats.completed)
    # later requests waited for memory: their admission is after the first finish
    first_finish = min(r.finished for r in stats.completed)
    assert max(r.admitted for r in stats.completed) >= first_finish - 1e-9


# 600 steps from rest (transients), against the FP64 path.  Round to nearest
# measured 7.7e-5 / 3.9e-5 (FP16) and 1.6e-3 / 9.6e-4 (BF16); stochastic
# rounding, whose gain is on steady flows, adds a little noise here.
SERVER_BOUND = {("f16", "nearest"): 1.2e-4, ("bf16", "nearest"): 2.4e-3,
                ("f16", "stochastic"): 1.2e-4, ("bf16", "stochastic"): 2.4e-3}


@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_shifted_storage_runs_in_the_server(rounding):
    specs = _specs()[:2]
    arrivals = [0.0, 0.01]
    for storage in ("f16", "bf16"):
        print(f"server {storage} {rounding}", flush=True)
        reqs = [Request(q, arrivals[q], s, max_steps=600, tol=0.0) for q, s in enumerate(specs)]
        srv = SolveServer(storage=storage, hbm_budget_bytes=1 << 28, max_slots=4, tick_steps=100, rounding=rounding)
        stats = srv.serve(reqs)
        srv.close()
        for r in stats.completed:
            u64, _, _ = solve_alone(specs[r.rid], 600, 0.0, 100, "f64")
            err = np.abs(r.result - u64).max()
            print(f"  {storage} {rounding} request {r.rid}: max |u - u_fp64| = {err:.2e}", flush=True)
            assert err < SERVER_BOUND[(storage, rounding)]
            # alone, with the same launches (two steps each): the same fields
            # bit for bit -- the stochastic seeds follow each lattice's own
            # step count wherever and whenever it was admitted
            u, _, _ = solve_alone(specs[r.rid], 600, 0.0, 100, storage, blocked=True, rounding=rounding)
            assert np.array_equal(u, r.result)


def test_diverged_solve_is_retired_without_disturbing_the_mix():
    """The 400 x 100 cylinder with the reference's VelocityNEEM inlet as stated
    blows up after about 170 steps (every restatement does); in a mix with a
    healthy solve it is retired at the first tick whose velocity is not
    finite, as failed, and the healthy solve still equals its run alone."""
    with contextlib.redirect_stdout(io.StringIO()):
        bad = spec_from_case(cases.cylinder(inlet="neem", name="diverging"))
    good = _specs()[0]
    reqs = [Request(0, 0.0, bad, max_steps=2000, tol=0.0), Request(1, 0.0, good, max_steps=2000, tol=0.0)]
    srv = SolveServer(storage="f32", hbm_budget_bytes=1 << 28, max_slots=4, tick_steps=40)
    stats = srv.serve(reqs)
    srv.close()
    done = {r.rid: r for r in stats.completed}
    print(f"  diverging solve retired after {done[0].steps} steps (failed={done[0].failed})", flush=True)
    assert done[0].failed and not done[0].converged
    assert done[0].steps < 1000
    assert not np.all(np.isfinite(done[0].result))
    assert not done[1].failed and done[1].steps == 2000
    assert stats.summary()["failed"] == 1
    u, steps, res = solve_alone(good, 2000, 0.0, 40, "f32")
    assert steps == 2000 and np.array_equal(u, done[1].result)
    u, steps, res = solve_alone(bad, 2000, 0.0, 40, "f32")
    assert steps == done[0].steps and np.isnan(res)


def test_serving_request_converges_and_retires_early():
    """A small request of the serving workload (cases.serving_contraction:
    PressureABC outlet, u_in = 0.02) reaches the 1e-6 residual well within
    the 50k-step budget and is retired at once; its velocity equals the same
    solve alone bit for bit and the FP64 path's at that step count within
    1e-5 lattice units."""
    with contextlib.redirect_stdout(io.StringIO()):
        spec = spec_from_case(cases.serving_contraction(256, 64, 0.5, name="small"))
        other = spec_from_case(cases.serving_contraction(320, 80, 1.0, name="other"))
    reqs = [Request(0, 0.0, spec, max_steps=50_000, tol=1e-6), Request(1, 0.0, other, max_steps=4000, tol=0.0)]
    srv = SolveServer(storage="f32", hbm_budget_bytes=1 << 28, max_slots=4, tick_steps=200)
    stats = srv.serve(reqs)
    srv.close()
    r = {q.rid: q for q in stats.completed}[0]
    print(f"  converged {r.converged} after {r.steps} steps, residual {r.residual:.2e}", flush=True)
    assert r.converged and r.steps < 50_000
    u, steps, res = solve_alone(spec, 50_000, 1e-6, 200, "f32")
    assert steps == r.steps and np.array_equal(u, r.result)
    u64, _, _ = solve_alone(spec, r.steps, 0.0, 200, "f64")
    err = np.abs(u - u64).max()
    print(f"  max |u - u_fp64| at convergence = {err:.2e} (max |u| {np.abs(u64).max():.3f})", flush=True)
    assert err < 1e-5
