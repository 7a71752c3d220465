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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(N, collision, seed):
    from taichi_lbm import cases
    with contextlib.redirect_stdout(sys.stderr):
        case = cases.porous(NX=N, NY=N, seed=seed, collision=collision)
    return case


def time_gpu(case, steps, warmup):
    import torch
    from taichi_lbm import cases
    from taichi_lbm.stepper import Stepper
    with contextlib.redirect_stdout(sys.stderr):
        cases.initialise(case)
    st = Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=True)
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
    return ms, red


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


def cpu_baseline(case, steps):
    """The reference-algorithm CPU restatement on the same lattice."""
    from oracle import cpu_ref
    nthreads = min(16, os.cpu_count() or 1)
    cl = cpu_ref.from_case(case, nthreads=nthreads)
    cl.run(1)                                  # warm-up (first touch)
    t = time.perf_counter()
    cl.run(steps)
    dt = time.perf_counter() - t
    n_fluid = len(case["fluid"].group.nodes)
    return {"value": round(n_fluid * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": nthreads,
            "kind": "port",
            "sample": f"{steps} steps (after 1 warm-up step) of the same {case['lbf'].NX}x{case['lbf'].NY} "
                      f"MRT power-law lattice; C++/OpenMP restatement of the reference's unfused kernel "
                      f"sequence (AoS layout, group-index loops), not Taichi",
            "seconds": round(dt, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-trt", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)

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
        extra["trt_mlups"] = round(n_fluid / (ms_t * 1e-3) / 1e6, 1)
        extra["trt_hbm_frac"] = round(n_fluid * BYTES_POWER_LAW / (ms_t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        tcase["coll"]._state = None
        lbf._st = None
        torch.cuda.empty_cache()

    try:
        extra["copy_ceiling_gbs"] = round(copy_ceiling_gbs(), 1)
    except Exception as e:   # the bandwidth probe is informative only
        log("copy ceiling probe failed:", e)

    traffic = None
    pmc = ROOT / "profiles" / "pmc_traffic.json"
    if pmc.exists():
        try:
            traffic = json.loads(pmc.read_text()).get("mrt_step_bytes_per_launch")
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
                     "bytes_per_fluid_node": BYTES_POWER_LAW},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
