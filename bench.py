"""Benchmark: fused D2Q9 stream+collide on one MI355X.

Workload (the single-lattice roofline configuration): an 8192 x 8192 FP32
lattice, periodic, seeded random non-overlapping solid disks (radius 8-32,
solid area fraction 0.3), body force gx = 1e-6, power-law rheology n = 0.7
(m = 0.01) with tau clamped to [0.51, 2.0]; MRT (headline) and TRT with
magic 1/4.  One step = one fused launch over the lattice.

Prints one JSON line: MLUPS over fluid nodes (value), the roofline of the
fused kernel (algorithmic bytes: 72 B of populations + 8 B of relaxation
state per fluid node per step, over the kernel's average duration measured
with HIP events on its stream), and a CPU baseline: the C++/OpenMP
restatement of the reference's unfused algorithm (oracle/cpu_ref.cpp) timed
on the host for a bounded number of steps of the same lattice.
"""
import argparse
import contextlib
import copy
import io
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "taichi-lbm_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS & % HBM roofline; batched solves/s + p50/p99 latency vs plain path"
HBM_PEAK_GBS = 8000.0
BYTES_POWER_LAW = 80   # 9 reads + 9 writes x 4 B, + tau read + write (4 + 4 B)
BYTES_POWER_LAW_SHIFTED = 44   # 9 reads + 9 writes x 2 B (FP16 / BF16 shifted), + tau read + write


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(N, collision, seed):
    from taichi_lbm import cases
    with contextlib.redirect_stdout(sys.stderr):
        case = cases.porous(NX=N, NY=N, seed=seed, collision=collision)
    return case


def time_gpu(case, steps, warmup, storage="f32", rounding="nearest"):
    import torch
    from taichi_lbm import cases
    from taichi_lbm.stepper import Stepper
    with contextlib.redirect_stdout(sys.stderr):
        cases.initialise(case)
    # the tile width is tuned in a pass of its own before the warm-up (its
    # steps are ordinary steps of the run; the fields do not depend on the
    # width), so the timed kernel is the tuned one whatever --warmup is
    st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=True, tile_cols="auto",
                 storage=storage, rounding=rounding)
    st.run(Stepper.tune_steps())
    st.run(warmup)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()      # lbm_run launches on this stream
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    st.run(steps)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    red = st.reduce()
    red["tile_cols"] = st.tile_cols or 32
    red["tile_timings_ms"] = {w: round(v, 4) for w, v in st.tune_timings.items()}
    red["tune_steps"] = Stepper.tune_steps()
    return ms, red


def kernel_symbol(tile_cols):
    """The MRT single-component FP32 two-step kernel a run with these tile
    columns launches (as the profiler names it): mode 9 is the one-node ring
    without node rules (1) built for lattices with a body force (8)."""
    return f"k_fused2<2, 1, 0, 9, {tile_cols}, true>"


def kernel_label(tile_cols):
    return f"k_fused2 (two steps per launch, LDS-staged, {tile_cols}-column tiles)"


def copy_ceiling_gbs(nbytes=4 << 30, reps=10):
    import ctypes
    import torch
    from taichi_lbm import _hip as H
    n = nbytes // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    s = torch.cuda.current_stream()
    for _ in range(2):
        H.check(H.lib().lbm_copy_f32(a.data_ptr(), b.data_ptr(), n, s.cuda_stream))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        H.check(H.lib().lbm_copy_f32(a.data_ptr(), b.data_ptr(), n, s.cuda_stream))
    e1.record(s)
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / sec / 1e9


def cpu_baseline(case, steps, one_thread_steps=3):
    """The reference-algorithm CPU restatement on the same lattice: `steps`
    steps on up to 16 threads (the value), then a short single-thread sample;
    the host's CPU model and available CPUs are recorded."""
    from oracle import cpu_ref
    nthreads = min(16, os.cpu_count() or 1)
    cl = cpu_ref.from_case(case, nthreads=nthreads)
    cl.run(1)                                  # warm-up (first touch)
    t = time.perf_counter()
    cl.run(steps)
    dt = time.perf_counter() - t
    n_fluid = len(case["fluid"].group.nodes)
    # the same lattice continued on one thread, a short sample
    cl.p.nthreads = 1
    t = time.perf_counter()
    cl.run(one_thread_steps)
    dt1 = time.perf_counter() - t
    model = ""
    try:
        model = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(n_fluid * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": nthreads,
            "kind": "port",
            "sample": f"{steps} steps (after 1 warm-up step) of the same {case['lbf'].NX}x{case['lbf'].NY} "
                      f"MRT power-law lattice; C++/OpenMP restatement of the reference's unfused kernel "
                      f"sequence (AoS layout, group-index loops), not Taichi",
            "seconds": round(dt, 2),
            "one_thread": {"value": round(n_fluid * one_thread_steps / dt1 / 1e6, 2), "unit": "MLUPS",
                           "steps": one_thread_steps, "seconds": round(dt1, 2)},
            "host": {"cpu_model": model, "cpus_available": len(os.sched_getaffinity(0))}}


def batched_leg(n_lattices, seed, steps, plain_steps, solve_steps, graph_steps=100, n_fp64=16,
                storages=("f32", "f16", "bf16"), bar=1e-4):
    """512 independent cylinder channels (NX ~ U{128..1024}, NY ~ U{64..256},
    r ~ U[4, NY/8] at (max(NX/5, r + 4), NY/2), u_in ~ U[0.01, 0.08], tau ~ U[0.52, 0.8],
    BGK or TRT (magic 1/4) by a seeded coin; bounce-back walls, VelocityBB
    inlet, PressureNEEM outlet).

    Batched: two steps per launch over the tiles of every lattice, replayed
    from a graph, with FP32 populations and with shifted FP16 / BF16 storage
    (f - w in 16 bits, FP32 arithmetic).  Plain FP32 path, one lattice per
    launch, four ways: one step per launch (lbm_run) and two steps per launch
    (lbm_run_blocked), each launched from the host and replayed from a graph
    that holds every lattice's launches.  Accuracy: every lattice is run to
    solve_steps in each batched format, the non-finite ones are counted, and
    a sample of n_fp64 finite lattices is rerun in FP64; the largest velocity
    difference is reported per format (the plain and batched FP32 paths give
    bit-identical fields, so the FP32 number describes both).  The speedup at
    matched accuracy compares the fastest batched format whose error meets
    the bar with the fastest plain path."""
    import ctypes
    import numpy as np
    import torch
    from taichi_lbm import _hip as H
    from taichi_lbm import cases
    from taichi_lbm.batch import BatchRunner, autotune_grid, autotune_tiles, spec_from_case, stored_nodes
    rng = np.random.default_rng(seed)
    specs = []
    t0 = time.time()
    with contextlib.redirect_stdout(io.StringIO()):
        for q in range(n_lattices):
            NX, NY = int(rng.integers(128, 1025)), int(rng.integers(64, 257))
            r, u = float(rng.uniform(4, NY / 8)), float(rng.uniform(0.01, 0.08))
            tau, coll = float(rng.uniform(0.52, 0.8)), ("bgk" if rng.random() < 0.5 else "trt")
            c = cases.cylinder(NX=NX, NY=NY, cx=max(NX / 5, r + 4), cy=NY / 2, r=r, tau=tau, u_in=u,
                               collision=coll, name=f"b{q}", inlet="bb")
            specs.append(spec_from_case(c))
    nodes = sum(s.N for s in specs)
    fluid = sum(int(((s.code & H.CODE_ACTIVE) != 0).sum()) for s in specs)
    log(f"batched setup {time.time() - t0:.1f}s: {n_lattices} lattices, {nodes} nodes")
    stream = torch.cuda.Stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps):
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    out = {"lattices": n_lattices, "nodes": nodes, "fluid_nodes": fluid,
           "stored_nodes": sum(stored_nodes(sp) for sp in specs), "steps_per_solve": solve_steps,
           "graph_steps": graph_steps}
    ms_b, tile_cols, grid, mode, unforced = {}, 0, 0, None, None
    with torch.cuda.stream(stream):
        for storage in storages:
            runner = BatchRunner(specs, storage=storage, grid_blocks=grid)
            runner.run(2)
            mode = runner.mode if mode is None else mode
            if storage == storages[0]:
                if runner.mode == 0:    # the persistent one-step grid is tuned; two-step launches have a block per tile
                    grid, timings = autotune_grid(runner, candidates=(2, 4, 8, 16), steps=10)
                    log(f"autotuned grid {grid}: {timings}")
                else:                   # two-step launches: the tile width is tuned
                    tile_cols, timings = autotune_tiles(runner, candidates=(16, 32, 64), steps=10)
                    log(f"autotuned tile columns {tile_cols}: {timings}")
            elif tile_cols:
                runner.set_tile_cols(tile_cols)
                runner.fit_tables()
            runner.run(10)
            if unforced is None:
                unforced = runner.unforced
            g = runner.capture(graph_steps)
            g.replay()
            reps = max(1, steps // graph_steps)
            ms_b[storage] = timed(g.replay, reps) / (reps * graph_steps)
            log(f"batched {storage}: {ms_b[storage]:.4f} ms/step")
            g.close()
            runner.close()
        out["timed_steps_batched"] = max(1, steps // graph_steps) * graph_steps
        # the plain FP32 path: each lattice on its own buffers, as given (rows not padded)
        plain = BatchRunner(specs, storage="f32", blocked=False, align_rows=False)
        plain.run(2)
        lib = H.lib()
        s = stream.cuda_stream
        par = plain.parity
        slots = sorted(plain.slots)
        outp = ctypes.c_int32()

        def rules(l):
            return ctypes.cast(ctypes.addressof(plain.neem) + l * H.MAX_NEEM * ctypes.sizeof(H.NeemRule),
                               ctypes.POINTER(H.NeemRule))

        def one_step(n):        # lbm_run per lattice: n launches each
            for l in slots:
                H.check(lib.lbm_run(ctypes.byref(plain.lats[l]), ctypes.byref(plain.flds[l]), rules(l), n, par, 0, s))

        def two_step(n):        # lbm_run_blocked per lattice: n / 2 launches each
            for l in slots:
                H.check(lib.lbm_run_blocked(ctypes.byref(plain.lats[l]), ctypes.byref(plain.flds[l]), rules(l),
                                            n, par, 0, 0, None, ctypes.byref(outp), s))
        ms_p = {}
        for name, fn in (("one_step", one_step), ("two_step", two_step)):
            fn(4)                                               # warm-up
            ms_p[name] = timed(lambda: fn(plain_steps), 1) / plain_steps
            # the same launches captured once per 4 steps (the parity comes back
            # to where it started) and replayed
            cg = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.graph(cg, stream=stream):
                fn(4)
            cg.replay()
            reps = max(1, plain_steps // 4)
            ms_p[name + "_graph"] = timed(cg.replay, reps) / (reps * 4)
            del cg
            log(f"plain {name}: {ms_p[name]:.4f} ms/step host-launched, {ms_p[name + '_graph']:.4f} from a graph")
        plain.close()
        # accuracy at solve_steps, per batched format, against FP64 on a sample
        vel = {}
        for storage in storages:
            runner = BatchRunner(specs, storage=storage, grid_blocks=grid)
            runner.run(solve_steps)
            vel[storage] = [runner.velocity(l) for l in range(n_lattices)]
            runner.close()
        finite = {st: [bool(np.isfinite(u).all()) for u in v] for st, v in vel.items()}
        sample = [l for l in range(n_lattices) if all(finite[st][l] for st in storages)][:n_fp64]
        r64 = BatchRunner([specs[l] for l in sample], storage="f64", grid_blocks=grid)
        r64.run(solve_steps)
        u64 = [r64.velocity(q) for q in range(len(sample))]
        r64.close()
    err = {st: (max(float(np.abs(vel[st][l] - u64[q]).max()) for q, l in enumerate(sample)) if sample else None)
           for st in storages}
    umax = max(float(np.abs(u).max()) for u in u64) if u64 else None
    sps = lambda ms: n_lattices / (ms * 1e-3 * solve_steps)
    best_plain = min(ms_p, key=ms_p.get)
    ok = [st for st in storages if err[st] is not None and err[st] <= bar]
    best_b = min(ok, key=lambda st: ms_b[st]) if ok else None
    out.update({
        "ms_per_step_batched": {st: round(v, 4) for st, v in ms_b.items()},
        "ms_per_step_plain": {k: round(v, 4) for k, v in ms_p.items()},
        "timed_steps_plain": plain_steps,
        "mlups_batched": {st: round(fluid / (v * 1e-3) / 1e6, 1) for st, v in ms_b.items()},
        "hbm_frac_batched": {st: round(fluid * 72 * (0.5 if st in ("f16", "bf16") else 1.0) / (v * 1e-3) / 1e9
                                       / HBM_PEAK_GBS, 4) for st, v in ms_b.items()},
        "solves_per_s_batched": {st: round(sps(v), 3) for st, v in ms_b.items()},
        "solves_per_s_plain": {k: round(sps(v), 3) for k, v in ms_p.items()},
        "speedup": round(ms_p["one_step"] / ms_b["f32"], 2),
        "speedup_vs_best_plain_f32": round(ms_p[best_plain] / ms_b["f32"], 2),
        "best_plain": best_plain,
        "max_abs_u_err_vs_fp64": err, "max_abs_u_fp64": umax, "accuracy_bar": bar,
        "matched_accuracy": {"batched": best_b, "plain": best_plain,
                             "speedup": round(ms_p[best_plain] / ms_b[best_b], 2) if best_b else None},
        "launch_mode": mode, "unforced_kernels": unforced, "grid_blocks": grid, "tile_cols": tile_cols,
        "finite_after_solve": {st: int(sum(v)) for st, v in finite.items()}, "fp64_sample": len(sample),
    })
    return out


def serving_leg(n_req, rate, seed, max_steps, tol, tick, storages=("f32", "f16", "bf16"), n_fp64=4,
                short_steps=2000):
    """Poisson stream of 4:1 contraction-flow power-law solves (width ~ U{256..2048},
    height ~ U{64..512}, n ~ U[0.5, 1.5]; TRT magic 1/4, u_in = 0.02,
    nu = 0.4 at the inlet shear rate, tau in [0.55, 2.0], VelocityBB inlet,
    PressureABC outlet: cases.SERVING_CONTRACTION), each run until
    ||u - u_(-tick)||/||u|| < tol or max_steps.
    Continuous batching (SolveServer) against the plain path (each solve alone,
    one after another in arrival order); shifted FP16/BF16 storage against
    FP32; errors against the FP64 path on the first n_fp64 requests, at the
    step count they ran and at short_steps, beside the FP64 path's own
    spread under a 1e-12 perturbation of the initial densities."""
    import numpy as np
    import torch
    from taichi_lbm import cases
    from taichi_lbm.batch import spec_from_case
    from taichi_lbm.serving import Request, SolveServer, solve_alone
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    specs = []
    with contextlib.redirect_stdout(io.StringIO()):
        for q in range(n_req):
            W, Hh, n = int(rng.integers(256, 2049)), int(rng.integers(64, 513)), float(rng.uniform(0.5, 1.5))
            specs.append(spec_from_case(cases.serving_contraction(W, Hh, n, name=f"r{q}")))
    nodes = [s.N for s in specs]
    out = {"requests": n_req, "rate_per_s": rate, "tol": tol, "max_steps": max_steps, "tick_steps": tick,
           "mean_nodes": int(np.mean(nodes))}
    results = {}
    for storage in storages:
        reqs = [Request(q, float(arrivals[q]), specs[q], max_steps=max_steps, tol=tol) for q in range(n_req)]
        srv = SolveServer(storage=storage, hbm_budget_bytes=24 << 30, max_slots=n_req, tick_steps=tick)
        st = srv.serve(reqs)
        srv.close()
        summ = st.summary()
        summ["mean_steps"] = float(np.mean([r.steps for r in st.completed]))
        out[f"batched_{storage}"] = summ
        results[storage] = {r.rid: (r.result, r.steps) for r in st.completed}
        if storage == "f32":
            fp32_completed = list(st.completed)
        log(f"serving {storage}: {summ}")
    # plain path: each request alone, one after another in arrival order (FP32)
    svc, steps_alone, same, nonfinite = [], [], True, 0
    for q in range(n_req):
        t0 = time.perf_counter()
        u, nst, _ = solve_alone(specs[q], max_steps, tol, tick, "f32")
        torch.cuda.synchronize()
        svc.append(time.perf_counter() - t0)
        steps_alone.append(nst)
        ub, sb = results["f32"][q]
        # bit for bit (a diverged solve's NaNs compare equal to themselves)
        same &= (nst == sb) and u.shape == ub.shape and np.array_equal(
            np.ascontiguousarray(u).view(np.uint32), np.ascontiguousarray(ub).view(np.uint32))
        nonfinite += int(not np.all(np.isfinite(u)))
    fin, lat = 0.0, []
    for q in range(n_req):
        start = max(float(arrivals[q]), fin)
        fin = start + svc[q]
        lat.append(fin - float(arrivals[q]))
    out["plain_f32"] = {"solves_per_s": round(n_req / fin, 3), "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                        "p99_latency_s": round(float(np.percentile(lat, 99)), 4), "wall_s": round(fin, 3),
                        "mean_service_s": round(float(np.mean(svc)), 4)}
    out["alone_equals_mix_f32"] = bool(same)
    out["nonfinite_results_f32"] = nonfinite
    b = out["batched_f32"]
    out["speedup_solves_per_s"] = round(b["solves_per_s"] / out["plain_f32"]["solves_per_s"], 2)
    # precision: max |u - u_fp64| at the step count the FP32 solve ran, on the
    # requests that converged (steady flows) and, apart, on a few that ran to
    # the step limit (still evolving or unsteady)
    conv = {r.rid for r in fp32_completed if r.converged}
    out["converged_requests"] = sorted(conv)
    sample = sorted(conv)[:n_fp64] + [q for q in range(n_req) if q not in conv][:2]
    err = {s: [] for s in storages}
    err_conv = {s: [] for s in storages}
    err_short = {s: [] for s in storages}
    sens = []
    per_request = []
    for q in sample:
        nst = results["f32"][q][1]
        u64, _, _ = solve_alone(specs[q], nst, 0.0, tick, "f64")
        row = {"rid": q, "size": [specs[q].NX, specs[q].NY], "steps": nst, "converged": q in conv,
               "max_abs_u_fp64": float(np.abs(u64).max())}
        for s in storages:
            u, ns = results[s][q]
            if ns == nst:
                e = float(np.abs(u - u64).max())
            else:   # a different stopping step: compare at the FP32 step count
                u2, _, _ = solve_alone(specs[q], nst, 0.0, tick, s, blocked=True)
                e = float(np.abs(u2 - u64).max())
            err[s].append(e)
            if q in conv:
                err_conv[s].append(e)
            row[s] = e
            row[s + "_steps"] = ns
        per_request.append(row)
        # the flow's own sensitivity: FP64 again from densities perturbed by
        # 1e-12 (relative); an unsteady flow amplifies any difference, the
        # storage format's rounding included
        sp = copy.copy(specs[q])
        sp.rho0 = specs[q].rho0 * (1.0 + 1e-12 * rng.standard_normal(specs[q].rho0.shape))
        u64p, _, _ = solve_alone(sp, nst, 0.0, tick, "f64")
        sens.append(float(np.abs(u64p - u64).max()))
        # a short horizon, before any instability has grown
        u64s, _, _ = solve_alone(specs[q], short_steps, 0.0, tick, "f64")
        for s in storages:   # two steps per launch, as the server runs them
            us, _, _ = solve_alone(specs[q], short_steps, 0.0, tick, s, blocked=True)
            err_short[s].append(float(np.abs(us - u64s).max()))
    # the same converged requests with round-to-nearest 16-bit stores (the
    # server rounds stochastically)
    err_near = {s: [] for s in storages if s != "f32"}
    for q in sorted(conv)[:n_fp64]:
        nst = results["f32"][q][1]
        u64, _, _ = solve_alone(specs[q], nst, 0.0, tick, "f64")
        for s in err_near:
            un, _, _ = solve_alone(specs[q], nst, 0.0, tick, s, blocked=True, rounding="nearest")
            err_near[s].append(float(np.abs(un - u64).max()))
    out["rounding_16bit"] = "stochastic"
    out["max_abs_u_err_vs_fp64_converged_nearest"] = {s: max(v) if v else None for s, v in err_near.items()}
    out["max_abs_u_err_vs_fp64"] = {s: max(v) if v else None for s, v in err.items()}
    out["max_abs_u_err_vs_fp64_converged"] = {s: max(v) if v else None for s, v in err_conv.items()}
    out["precision_requests"] = per_request
    out["fp64_sensitivity_1e-12"] = max(sens) if sens else None
    out["max_abs_u_err_vs_fp64_short"] = {"steps": short_steps,
                                          **{s: max(v) if v else None for s, v in err_short.items()}}
    return out


def configs_leg(cyl_steps=10_000, sc_N=2048, sc_steps=20_000, sc_every=5000):
    """The two other single-lattice configurations, timed with HIP events on
    the launch stream:
    - the 400 x 100 cylinder channel (BGK tau = 0.6, r = 10 at (80, 50),
      bounce-back walls, VelocityBB inlet u = 0.05 over a bounce-back inlet
      column, PressureNEEM outlet rho = 1), cyl_steps steps: us/step and MLUPS
      (its 1.4 MB of populations stay in the caches, so no HBM fraction),
      through the Stepper (one launch per two steps from the host) and as a
      one-lattice batch replayed from a HIP graph (about the same: the
      launches are bound by their own latency on the GPU, not by the host);
    - two-component Shan-Chen separation, sc_N^2 periodic, MRT nu = 1/6,
      G = [[0, 0.9], [0.9, 0]], rho_c = 1 + U(-0.01, 0.01) seeded per
      component, sc_steps steps: ms/step, MLUPS, the algorithmic HBM fraction
      (72 B x 2 components per node and step; psi traffic not credited) and
      the FP64 mass of each component every sc_every steps against its
      initial sum."""
    import numpy as np
    import torch
    from taichi_lbm import cases
    from taichi_lbm.stepper import Stepper
    out = {}
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    with contextlib.redirect_stdout(io.StringIO()):
        case = cases.cylinder(inlet="bb")
        cases.initialise(case)
    st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"])
    st.run(20)
    torch.cuda.synchronize()
    e0.record(s)
    st.run(cyl_steps)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / cyl_steps
    nf = len(case["fluid"].group.nodes)
    red = st.reduce()
    out["cylinder_400x100"] = {"steps": cyl_steps + 20, "us_per_step": round(ms * 1e3, 2),
                               "mlups": round(nf / (ms * 1e-3) / 1e6, 1), "fluid_nodes": nf,
                               "nonfinite": red["nonfinite"]}
    del st, case
    # the same lattice as a one-lattice batch replayed from a HIP graph (no
    # per-launch host cost), on a stream of its own (graphs are not captured
    # on the legacy default stream)
    from taichi_lbm.batch import BatchRunner, spec_from_case
    with contextlib.redirect_stdout(io.StringIO()):
        spec = spec_from_case(cases.cylinder(inlet="bb"))
    gs = torch.cuda.Stream()
    with torch.cuda.stream(gs):
        r = BatchRunner([spec], storage="f32")
        r.run(20)
        g = r.capture(200)
        g.replay()
        torch.cuda.synchronize()
        reps = max(1, cyl_steps // 200)
        e0.record(gs)
        for _ in range(reps):
            g.replay()
        e1.record(gs)
        torch.cuda.synchronize()
        msg = e0.elapsed_time(e1) / (reps * 200)
        g.close()
        u = r.velocity(0)
        r.close()
    out["cylinder_400x100"].update({"graph_us_per_step": round(msg * 1e3, 2),
                                    "graph_mlups": round(nf / (msg * 1e-3) / 1e6, 1),
                                    "graph_steps": reps * 200 + 220, "graph_finite": bool(np.isfinite(u).all())})
    torch.cuda.empty_cache()

    with contextlib.redirect_stdout(io.StringIO()):
        case = cases.shan_chen(NX=sc_N, NY=sc_N, nc=2, gcross=0.9, seed=3)
        cases.initialise(case)
    m0 = case["rho"].astype(np.float64).sum(axis=(0, 1))
    st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], boundary_engine_sc=case["bsc"])
    st.run(20)                 # warm-up: the first run also compiles the node codes on the host
    done, t_ms, drift = 20, 0.0, []
    while done < sc_steps + 20:
        n = min(sc_every, sc_steps + 20 - done)
        torch.cuda.synchronize()
        e0.record(s)
        st.run(n)
        e1.record(s)
        torch.cuda.synchronize()
        t_ms += e0.elapsed_time(e1)
        done += n
        red = st.reduce()
        drift.append([done] + [float(x) for x in np.abs(np.array(red["mass"]) - m0) / m0])
    ms = t_ms / sc_steps
    nodes = sc_N * sc_N
    out[f"shan_chen_{sc_N}x{sc_N}x2"] = {
        "steps": sc_steps + 20, "timed_steps": sc_steps, "ms_per_step": round(ms, 4), "mlups": round(nodes / (ms * 1e-3) / 1e6, 1),
        "hbm_frac": round(nodes * 72 * 2 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "mass_drift_rel": drift, "nonfinite": red["nonfinite"]}
    del st, case
    torch.cuda.empty_cache()
    out.update(ib_leg())
    out.update(sc_batch_leg())
    return out


def sc_batch_leg(n_lattices=64, steps=200, seed=11):
    """Many two-component Shan-Chen lattices (MRT nu = 1/6, G_cross = 0.9,
    periodic, sides U{128..384} seeded) advanced by one ShanChenBatch launch
    per step, against each lattice's own Stepper (two steps per launch where
    its kernel allows, one launch per lattice): ms per step of the whole set,
    timed with HIP events; the fields of a seeded sample equal their runs
    alone bit for bit (tested in tests/test_gpu_sc_batch.py)."""
    import numpy as np
    import torch
    from taichi_lbm import cases
    from taichi_lbm.batch import ShanChenBatch
    rng = np.random.default_rng(seed)
    sides = rng.integers(128, 385, (n_lattices, 2))
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def make():
        out = []
        with contextlib.redirect_stdout(io.StringIO()):
            for q, (nx, ny) in enumerate(sides):
                c = cases.shan_chen(NX=int(nx), NY=int(ny), nc=2, gcross=0.9, seed=100 + q)
                cases.initialise(c)
                out.append(c)
        return out

    def timed(fn):
        fn(10)
        torch.cuda.synchronize()
        e0.record(s)
        fn(steps)
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    cs = make()
    nodes = int(sum(int(a) * int(b) for a, b in sides))
    sb = ShanChenBatch([cases.stepper(c) for c in cs])
    ms_b = timed(sb.run)
    finite = all(bool(np.isfinite(c["lbf"].vel.to_numpy()).all()) for c in cs[:8])
    sb.close()
    del cs, sb
    torch.cuda.empty_cache()
    cs = make()
    solo = [cases.stepper(c) for c in cs]
    ms_p = timed(lambda k: [st.run(k) for st in solo])
    del cs, solo
    torch.cuda.empty_cache()
    return {f"shan_chen_batch_{n_lattices}": {
        "lattices": n_lattices, "nodes": nodes, "components": 2, "timed_steps": steps,
        "ms_per_step_batched": round(ms_b, 4), "ms_per_step_plain": round(ms_p, 4),
        "speedup": round(ms_p / ms_b, 2), "mlups_batched": round(nodes * 2 / (ms_b * 1e-3) / 1e6, 1),
        "hbm_frac_batched": round(nodes * 72 * 2 / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "plain": "each lattice's Stepper.run (k_sc2, two steps per launch), one lattice after another",
        "finite": finite}}


def ib_leg(NX=2048, NY=1024, n_particles=256, steps=200, mdf=3):
    """Immersed-boundary particles: an all-fluid periodic NX x NY channel
    (BGK tau = 0.7, g = 1e-6) with n_particles circular particles (radius
    U[6, 10], speed 0.01, turning 0.002 rad/step, one crossing the edge),
    coupled by multi-direct forcing with `mdf` iterations through the
    granular sequence density -> force_density -> velocity ->
    SphereInfoUpdata -> MultiDirectForcingLoop -> ComputeParticleForce ->
    collide -> boundaries, `steps` steps timed with HIP events on the launch
    stream; the IB calls alone are timed on the same state."""
    import numpy as np
    import torch
    from taichi_lbm import cases
    from taichi_lbm.ibm import IBDEMCouplerEngine, IBEngine
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with contextlib.redirect_stdout(io.StringIO()):
        case = cases.ib_particles(NX=NX, NY=NY, n_particles=n_particles, radius=(6.0, 10.0), u_particle=0.01,
                                  seed=2)
        cases.initialise(case)
        lbf, macro, coll, bce, sp = case["lbf"], case["macro"], case["coll"], case["bce"], case["sp"]
        sph = case["spheres"]
        n = n_particles
        sp.init_Balls(n, np.arange(n), sph["pos"], sph["vel"], np.zeros((n, 2)), np.zeros(n), sph["radius"],
                      sph["angle"])
        sp.init_IB_nodes()
    coupler, ibe = IBDEMCouplerEngine(lbf), IBEngine(mdf)

    def ib_part(t):
        pos, vel, ang = cases.ib_particle_state(case, t)
        coupler.SphereInfoUpdata(sp, pos, vel, ang)
        ibe.MultiDirectForcingLoop(sp, lbf)
        ibe.ComputeParticleForce(sp)

    def step(t):
        macro.density(lbf)
        macro.force_density(lbf)
        macro.velocity(lbf)
        ib_part(t)
        coll.apply(lbf)
        bce.apply_boundary_conditions(lbf)

    from taichi_lbm.stepper import Stepper
    fused = Stepper(lbf, macro, coll, bce, ib_engine=ibe, sphere_field=sp)

    def fused_step(t):
        pos, vel, ang = cases.ib_particle_state(case, t)
        coupler.SphereInfoUpdata(sp, pos, vel, ang)
        fused.run(1)
        ibe.ComputeParticleForce(sp)

    def fused_run(t0, n):
        # particles held in place over n steps: one lbm_ib_run launch sequence
        pos, vel, ang = cases.ib_particle_state(case, t0)
        coupler.SphereInfoUpdata(sp, pos, vel, ang)
        fused.run(n)

    for t in range(10):
        step(t)
    torch.cuda.synchronize()
    e0.record(s)
    for t in range(10, 10 + steps):
        step(t)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    e0.record(s)
    for t in range(steps):
        ib_part(10 + steps + t)
    e1.record(s)
    torch.cuda.synchronize()
    ib_us = e0.elapsed_time(e1) / steps * 1e3
    for t in range(10):
        fused_step(t)
    torch.cuda.synchronize()
    e0.record(s)
    for t in range(steps):
        fused_step(t)
    e1.record(s)
    torch.cuda.synchronize()
    ms_fused = e0.elapsed_time(e1) / steps
    e0.record(s)
    fused_run(0, steps)
    e1.record(s)
    torch.cuda.synchronize()
    ms_fixed = e0.elapsed_time(e1) / steps
    markers = int(sp.IBInsiderCount[None])
    u = lbf.vel.to_numpy()
    force = coupler.GetSphereForce(sp)
    out = {f"ib_particles_{NX}x{NY}_{n_particles}": {
        "steps": steps + 10, "ms_per_step": round(ms, 4), "mlups": round(NX * NY / (ms * 1e-3) / 1e6, 1),
        "ib_us_per_step": round(ib_us, 1), "markers": markers, "mdf_iterations": mdf,
        "path": "granular sequence (one HIP kernel per reference kernel) with the IB kernels of csrc/lbm_ibm.hip",
        "fused_ms_per_step": round(ms_fused, 4), "fused_mlups": round(NX * NY / (ms_fused * 1e-3) / 1e6, 1),
        "fused_path": "Stepper(ib_engine, sphere_field): lbm_ib_run, particles moved every step",
        "fused_fixed_ms_per_step": round(ms_fixed, 4),
        "fused_fixed_path": "the same with the particles held over one run of all steps",
        "finite": bool(np.isfinite(u).all() and np.isfinite(force).all()),
        "max_abs_particle_force": float(np.abs(force).max())}}
    del case, sp
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-trt", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-lattices", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=1000)          # timed graph-replayed steps
    ap.add_argument("--batch-plain-steps", type=int, default=100)     # timed steps of the plain paths
    ap.add_argument("--batch-solve-steps", type=int, default=10_000)  # steps of the accuracy run
    ap.add_argument("--only-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--only-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--only-configs", action="store_true")
    ap.add_argument("--only-ib", action="store_true")
    ap.add_argument("--only-sc-batch", action="store_true")
    ap.add_argument("--configs-cyl-steps", type=int, default=10_000)
    ap.add_argument("--configs-sc-steps", type=int, default=20_000)
    ap.add_argument("--headline-only", action="store_true",
                    help="only the 8192^2 MRT power-law timing (for profiling runs)")
    args = ap.parse_args()
    if args.headline_only:
        args.no_cpu = args.no_trt = args.no_batched = args.no_serving = args.no_configs = True
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: there is no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)

    if args.only_serving:
        print(json.dumps(serving_leg(args.serve_requests, 4.0, 5, 50_000, 1e-6, 200)), flush=True)
        return
    if args.only_ib:
        print(json.dumps(ib_leg()), flush=True)
        return
    if args.only_sc_batch:
        print(json.dumps(sc_batch_leg()), flush=True)
        return
    if args.only_configs:
        print(json.dumps(configs_leg(args.configs_cyl_steps, sc_steps=args.configs_sc_steps)), flush=True)
        return
    if args.only_batched:
        print(json.dumps(batched_leg(args.batch_lattices, 4, args.batch_steps, args.batch_plain_steps, args.batch_solve_steps)), flush=True)
        return
    t0 = time.time()
    case = build_case(args.size, "mrt", args.seed)
    N = args.size * args.size
    n_fluid = len(case["fluid"].group.nodes)
    log(f"setup {time.time() - t0:.1f}s: {N} nodes, {n_fluid} fluid ({n_fluid / N:.3f})")
    ms, red = time_gpu(case, args.steps, args.warmup)
    log(f"MRT: {ms:.4f} ms/step, mass {red['mass'][0]:.6e}, nonfinite {red['nonfinite']}")
    mlups = n_fluid / (ms * 1e-3) / 1e6
    achieved = n_fluid * BYTES_POWER_LAW / (ms * 1e-3) / 1e9
    extra = {"mlups_all_nodes": round(N / (ms * 1e-3) / 1e6, 1), "fluid_fraction": round(n_fluid / N, 4),
             "tile_cols": red["tile_cols"], "tile_timings_ms": red["tile_timings_ms"],
             "tuning_pass_steps": red["tune_steps"],
             "n_disks": case["spec"]["n_disks"], "mass_fp64": red["mass"][0], "nonfinite": red["nonfinite"]}

    if not args.no_trt:
        from taichi_lbm import cases as C
        lbf = case["lbf"]
        tcase = dict(case)
        tcase["coll"] = C.make_collision("trt", case["fluid"].group, args.size, args.size, 1, [1.0 / 6.0],
                                         case["coll"].fluid_model, magic=0.25)
        tcase["coll"].tau_min, tcase["coll"].tau_max = 0.51, 2.0
        lbf._st = None
        lbf._finisher = None
        case["coll"]._state = None
        torch.cuda.empty_cache()
        ms_t, red_t = time_gpu(tcase, args.steps, args.warmup)
        log(f"TRT: {ms_t:.4f} ms/step")
        extra["trt_ms_per_step"] = round(ms_t, 5)
        extra["trt_tile_cols"] = red_t["tile_cols"]
        extra["trt_mlups"] = round(n_fluid / (ms_t * 1e-3) / 1e6, 1)
        extra["trt_hbm_frac"] = round(n_fluid * BYTES_POWER_LAW / (ms_t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        tcase["coll"]._state = None
        lbf._st = None
        torch.cuda.empty_cache()

    if not args.no_trt:
        # the same MRT lattice with shifted 16-bit populations (FP32 arithmetic):
        # 36 B of populations + 8 B of power-law state per fluid node-step
        lbf = case["lbf"]
        for storage, rounding in (("f16", "nearest"), ("bf16", "nearest"), ("f16", "stochastic"),
                                  ("bf16", "stochastic")):
            lbf._st = None
            lbf._finisher = None
            lbf._pop_owner = None
            case["coll"]._state = None
            torch.cuda.empty_cache()
            ms_s, red_s = time_gpu(case, args.steps, args.warmup, storage=storage, rounding=rounding)
            log(f"MRT {storage} ({rounding} rounding): {ms_s:.4f} ms/step")
            if rounding == "stochastic":
                extra[f"{storage}_stochastic_ms_per_step"] = round(ms_s, 5)
                extra[f"{storage}_stochastic_nonfinite"] = red_s["nonfinite"]
                continue
            extra[f"{storage}_ms_per_step"] = round(ms_s, 5)
            extra[f"{storage}_tile_cols"] = red_s["tile_cols"]
            extra[f"{storage}_mlups"] = round(n_fluid / (ms_s * 1e-3) / 1e6, 1)
            extra[f"{storage}_hbm_frac"] = round(n_fluid * BYTES_POWER_LAW_SHIFTED / (ms_s * 1e-3) / 1e9 /
                                                 HBM_PEAK_GBS, 4)
            extra[f"{storage}_nonfinite"] = red_s["nonfinite"]
        lbf._st = None
        case["coll"]._state = None
        torch.cuda.empty_cache()

    batched = None
    if not args.no_batched:
        case["lbf"]._st = None
        case["coll"]._state = None
        torch.cuda.empty_cache()
        batched = batched_leg(args.batch_lattices, 4, args.batch_steps, args.batch_plain_steps, args.batch_solve_steps)
        log(f"batched: {batched}")

    serving = None
    if not args.no_serving:
        torch.cuda.empty_cache()
        serving = serving_leg(args.serve_requests, 4.0, 5, 50_000, 1e-6, 200)
        log(f"serving: {serving}")

    configs = None
    if not args.no_configs:
        torch.cuda.empty_cache()
        configs = configs_leg(args.configs_cyl_steps, sc_steps=args.configs_sc_steps)
        log(f"configs: {configs}")

    try:
        extra["copy_ceiling_gbs"] = round(copy_ceiling_gbs(), 1)
    except Exception as e:   # the bandwidth probe is informative only
        log("copy ceiling probe failed:", e)

    traffic, steps_per_launch = None, 2
    pmc = ROOT / "profiles" / "pmc_traffic.json"
    if pmc.exists():
        try:
            rec = json.loads(pmc.read_text())
            # the record's counters belong to one kernel (tile width): use them
            # only if that is the kernel this run timed
            if kernel_symbol(red["tile_cols"]) in rec.get("kernel", ""):
                traffic = rec.get("mrt_step_bytes_per_launch")
                steps_per_launch = rec.get("steps_per_launch", steps_per_launch)
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        log("CPU baseline ...")
        cpu = cpu_baseline(case, args.cpu_steps)
        log(f"CPU: {cpu}")

    out = {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"porous-{args.size}x{args.size}-mrt-powerlaw", "lattice": [args.size, args.size],
                   "collision": "MRT", "power_law_n": 0.7, "power_law_m": 0.01, "tau_clamp": [0.51, 2.0],
                   "gravity": [1e-6, 0.0], "solid_fraction": 0.3, "disk_radius": [8, 32], "seed": args.seed,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_fluid_node": BYTES_POWER_LAW, "steps_per_launch": steps_per_launch,
                     "kernel": kernel_label(red["tile_cols"])},
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "configs": configs,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
