"""Batched lattices (one persistent launch per step, graph replay) against the
single-lattice fused path and the oracle."""
import numpy as np
import pytest

from taichi_lbm import cases
from taichi_lbm.batch import BatchRunner, autotune_grid, autotune_tiles, spec_from_case
from taichi_lbm.stepper import Stepper
from oracle_bridge import oracle_solver
from tolerances import check

pytestmark = pytest.mark.gpu


def _mix():
    out = []
    params = [(48, 24, "bgk", 0.6, 0.05), (40, 32, "trt", 0.7, 0.03), (64, 20, "mrt", 0.55, 0.08),
              (33, 17, "bgk", 0.8, 0.01), (50, 26, "trt", 0.52, 0.06)]
    for q, (NX, NY, coll, tau, u) in enumerate(params):
        r = max(2, NY // 8)
        # the first three use the VelocityBB inlet (stable over long runs), the
        # last two the VelocityNEEM inlet (node rules of both kinds in a batch)
        out.append(cases.cylinder(NX=NX, NY=NY, cx=NX // 5, cy=NY // 2, r=r, tau=tau, u_in=u, collision=coll,
                                  tau_bulk=0.8, name=f"c{q}", inlet="bb" if q < 3 else "neem"))
    return out


def test_batch_equals_single_lattice_runs():
    import torch
    cs = _mix()
    runner = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    n = 40
    runner.run(15)
    runner.run(n - 15)
    for q, c in enumerate(cs):
        print(f"lattice {q}", flush=True)
        cases.initialise(c)
        st = Stepper(c["lbf"], c["macro"], c["coll"], c["bce"])
        st.run(n)
        got = runner.fields(q)
        for name in ("f", "f2", "rho", "vel"):
            ref = getattr(c["lbf"], name).to_numpy()
            assert np.array_equal(got[name].astype(np.float32), ref), (q, name)
    torch.cuda.synchronize()


def test_batch_matches_oracle_fp32():
    cs = _mix()[:2]
    runner = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    n = 30
    runner.run(n)
    for q, c in enumerate(cs):
        sol = oracle_solver(c, np.float32)
        sol.run(n)
        got = runner.fields(q)
        for name in ("f", "f2", "rho", "vel"):
            check(name, got[name], getattr(sol.lat, name))


def test_graph_replay_equals_run():
    import torch
    cs = _mix()
    a = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    b = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        a.run(2)
        b.run(2)
        g = b.capture(8)
        for rep in range(4):
            g.replay()
            torch.cuda.synchronize()          # eager work between replays
            _ = b.residuals()
        a.run(32)
        torch.cuda.synchronize()
        for q in range(len(cs)):
            fa, fb = a.fields(q), b.fields(q)
            for name in ("f", "f2", "rho", "vel"):
                assert np.array_equal(fa[name], fb[name]), (q, name)
        g.close()


def test_fp64_batch_matches_fp64_oracle():
    cs = _mix()[2:4]
    runner = BatchRunner([spec_from_case(c) for c in cs], storage="f64")
    n = 25
    runner.run(n)
    for q, c in enumerate(cs):
        sol = oracle_solver(c, np.float64, params32=True)
        sol.run(n)
        got = runner.fields(q)
        for name in ("f", "f2"):
            err = np.abs(got[name] - getattr(sol.lat, name)).max()
            print(f"  lattice {q} {name}: |fp64 hip - fp64 oracle| = {err:.2e}")
            assert err < 1e-13
        assert np.abs(got["vel"] - sol.lat.vel).max() < 1e-7   # vel is written back as fp32


@pytest.mark.parametrize("storage,bound", [("f16", 2e-4), ("bf16", 2e-3)])
def test_shifted_storage_error_vs_fp64(storage, bound):
    cs = _mix()[:3]
    lo = BatchRunner([spec_from_case(c) for c in cs], storage=storage)
    hi = BatchRunner([spec_from_case(c) for c in cs], storage="f64")
    lo.run(200)
    hi.run(200)
    for q in range(len(cs)):
        du = np.abs(lo.fields(q)["vel"] - hi.fields(q)["vel"]).max()
        print(f"  {storage} lattice {q}: max |u - u_fp64| = {du:.2e}")
        assert du < bound


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_shifted_two_step_alone_equals_mix(storage):
    """Shifted 16-bit storage with two steps per launch: step n stays in LDS
    at FP32, so the populations are rounded to 16 bits once per two steps
    (the one-step path rounds every step and gives other fields).  Every
    tile recomputes its ring from the same stored inputs, so a lattice's
    fields do not depend on the tiling or on the other lattices: alone and in
    a batch they are equal bit for bit.  Both paths' errors against FP64 are
    printed."""
    cs = _mix()[:3]
    specs = [spec_from_case(c) for c in cs]
    mix = BatchRunner(specs, storage=storage, blocked=True)
    alone = BatchRunner([specs[1]], storage=storage, blocked=True)
    one = BatchRunner([specs[1]], storage=storage, blocked=False)
    hi = BatchRunner([specs[1]], storage="f64")
    assert mix.mode > 0 and alone.mode > 0
    for r in (mix, alone, one, hi):
        r.run(120)
    u_mix, u_alone = mix.fields(1)["vel"], alone.fields(0)["vel"]
    assert np.array_equal(u_mix.view(np.uint32), u_alone.view(np.uint32))
    e2 = np.abs(u_alone - hi.fields(0)["vel"]).max()
    e1 = np.abs(one.fields(0)["vel"] - hi.fields(0)["vel"]).max()
    print(f"  {storage}: max |u - u_fp64| two steps per launch {e2:.2e}, one step per launch {e1:.2e}")
    assert e2 < (2e-4 if storage == "f16" else 2e-3) and e1 < (2e-4 if storage == "f16" else 2e-3)
    for r in (mix, alone, one, hi):
        r.close()


def test_deactivated_lattice_is_frozen_and_others_unchanged():
    cs = _mix()
    a = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    b = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    a.run(10)
    b.run(10)
    before = b.fields(1)
    b.set_active(1, 0)
    a.run(20)
    b.run(20)
    after = b.fields(1)
    assert np.array_equal(before["f2"], after["f2"])
    for q in (0, 2, 3, 4):
        assert np.array_equal(a.fields(q)["f2"], b.fields(q)["f2"]), q


def test_residuals_match_host():
    cs = _mix()[:2]
    r = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    r.run(5)
    r.residuals()
    u_prev = [r.fields(q)["vel"].astype(np.float64) for q in range(2)]
    r.run(7)
    res = r.residuals()
    for q in range(2):
        u = r.fields(q)["vel"].astype(np.float64)
        act = (r.slots[q]["spec"].code.reshape(u.shape[:2]) & 0x100) != 0
        d = ((u - u_prev[q]) ** 2).sum(-1)[act].sum()
        assert np.isclose(res[q, 0], d, rtol=1e-6)


def test_autotune_picks_a_candidate():
    cs = _mix()
    r = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    r.run(1)
    best, timings = autotune_grid(r, candidates=(2, 8), steps=4)
    assert best in timings and best % 8 == 0


@pytest.mark.parametrize("tile_cols,ring", [(16, 0), (32, 0), (64, 0), (32, 2), (64, 2)])
def test_blocked_batch_equals_one_step_batch(tile_cols, ring):
    """Two-step batched launches of channels with node rules give the fields
    of one-step batched launches, bit for bit: with a one-node ring on
    shifted tile grids (mode 3; with 32 columns the 13-row tiles put the
    40-row channel's outlet row on a tile's first row unless its grid is
    shifted) and with a two-node ring (mode 2)."""
    cs = _mix()
    a = BatchRunner([spec_from_case(c) for c in cs], storage="f32", blocked=True, tile_cols=tile_cols)
    if ring:
        a.set_ring(ring)
    b = BatchRunner([spec_from_case(c) for c in cs], storage="f32", blocked=False)
    assert a.mode == (3 if ring == 0 else 2) and b.mode == 0
    for n in (7, 10):
        a.run(n)
        b.run(n)
    for q in range(len(cs)):
        fa, fb = a.fields(q), b.fields(q)
        for name in ("f", "f2", "rho", "vel", "total_rho"):
            assert np.array_equal(fa[name], fb[name]), (q, name)


def test_blocked_batch_without_rules_and_graph():
    """Lattices without node rules (ring width 1), run eagerly and replayed
    from a graph, against one-step launches."""
    import torch
    import contextlib, io
    with contextlib.redirect_stdout(io.StringIO()):
        cs = [cases.porous(NX=90 + 13 * q, NY=70 + 20 * q, seed=q, collision=("bgk", "trt", "mrt")[q % 3],
                           rmin=3.0, rmax=6.0, g=(1e-5, 0.0)) for q in range(4)]
    specs = [spec_from_case(c, power_law=False) for c in cs]
    a = BatchRunner(specs, storage="f32", blocked=True)
    b = BatchRunner(specs, storage="f32", blocked=False)
    assert a.mode == 1
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        best, timings = autotune_tiles(a, candidates=(16, 32, 64), steps=2)   # switches widths before capture
        assert best in (16, 32, 64) and set(timings) == {16, 32, 64}
        b.run(3 * (2 + 2))
        a.run(2)
        b.run(2)
        g = a.capture(12)
        for _ in range(3):
            g.replay()
            torch.cuda.synchronize()
        b.run(36)
        torch.cuda.synchronize()
        for q in range(len(cs)):
            fa, fb = a.fields(q), b.fields(q)
            for name in ("f", "f2", "rho", "vel"):
                assert np.array_equal(fa[name], fb[name]), (q, name)
        g.close()


@pytest.mark.parametrize("tile_cols", [32, 64])
def test_open_neumann_outlets_on_shifted_grids(tile_cols):
    """The reference cylinder's boundary list (VelocityBB inlet, OpenNeumann
    outlet reading x - c_1) on lattices whose outlet row NX - 1 is a multiple
    of the tile rows (13 with 32 columns, 5 with 64), so mode 3 shifts their
    tile grids; against one-step launches."""
    import contextlib, io
    with contextlib.redirect_stdout(io.StringIO()):
        cs = [cases.cylinder_reference_bcs(NX=NX, NY=NY, name=f"r{q}")
              for q, (NX, NY) in enumerate([(66, 40), (131, 48), (79, 36)])]
    specs = [spec_from_case(c) for c in cs]
    a = BatchRunner(specs, storage="f32", blocked=True, tile_cols=tile_cols)
    b = BatchRunner(specs, storage="f32", blocked=False)
    assert a.mode == 3
    for n in (9, 12):
        a.run(n)
        b.run(n)
    for q in range(len(cs)):
        fa, fb = a.fields(q), b.fields(q)
        for name in ("f", "f2", "rho", "vel", "total_rho"):
            assert np.array_equal(fa[name], fb[name]), (q, name)


def test_power_law_batch_two_steps():
    """Power-law lattices take two steps per launch in a batch (their
    relaxation state alternates between tau and tau2 and is copied back after
    an odd number of launches): bit-identical to one-step batched launches,
    relaxation state included, run eagerly and replayed from a graph."""
    import torch
    import contextlib, io
    with contextlib.redirect_stdout(io.StringIO()):
        cs = [cases.contraction(W=W, H=Hh, power_n=n, u_in=0.03, name=f"p{q}")
              for q, (W, Hh, n) in enumerate([(96, 32, 0.7), (128, 40, 1.3), (80, 24, 0.5)])]
        cs.append(cases.porous(NX=70, NY=90, seed=3, collision="mrt", rmin=3.0, rmax=6.0, g=(1e-5, 0.0)))
    specs = [spec_from_case(c) for c in cs]
    a = BatchRunner(specs, storage="f32", blocked=True)
    b = BatchRunner(specs, storage="f32", blocked=False)
    assert a.mode in (2, 3) and b.mode == 0

    def same():
        for q in range(len(cs)):
            fa, fb = a.fields(q), b.fields(q)
            for name in ("f", "f2", "rho", "vel", "total_rho"):
                assert np.array_equal(fa[name], fb[name]), (q, name)
        assert torch.equal(a.tau, b.tau)

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for n in (6, 8, 7):       # 3 launches (state copied back), 4, 3 + one single step
            a.run(n)
            b.run(n)
        a.run(1)                  # back to an even parity for the capture
        b.run(1)
        torch.cuda.synchronize()
        same()
        g = a.capture(8)
        for _ in range(2):
            g.replay()
            torch.cuda.synchronize()
        b.run(16)
        torch.cuda.synchronize()
        same()
        g.close()


def test_force_variants_equal_single_runs():
    """A batch without body forces runs kernels built without the force
    terms (BatchRunner.unforced); one lattice with gravity switches the batch
    to the kernels that read the force switch at run time.  Either way every
    lattice equals its own Stepper run (which takes the forced or unforced
    two-step build) bit for bit."""
    import contextlib
    import io
    cs = _mix()[:3]
    with contextlib.redirect_stdout(io.StringIO()):
        porous = cases.porous(NX=64, NY=48, seed=2, collision="mrt", rmin=3.0, rmax=6.0, g=(2e-5, 1e-6))
    plain = BatchRunner([spec_from_case(c) for c in cs], storage="f32")
    mixed = BatchRunner([spec_from_case(c) for c in cs + [porous]], storage="f32")
    assert plain.unforced and not mixed.unforced
    n = 24
    plain.run(n)
    mixed.run(n)
    for q, c in enumerate(cs + [porous]):
        print(f"lattice {q}", flush=True)
        with contextlib.redirect_stdout(io.StringIO()):
            cases.initialise(c)
        Stepper(c["lbf"], c["macro"], c["coll"], c["bce"], power_law=c["power_law"]).run(n)
        ref = c["lbf"].vel.to_numpy()
        for r, slot in ((mixed, q),) + (((plain, q),) if q < len(cs) else ()):
            assert np.array_equal(r.fields(slot)["vel"], ref), q
    plain.close()
    mixed.close()
