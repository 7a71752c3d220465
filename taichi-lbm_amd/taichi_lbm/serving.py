"""Serving: continuous batching of independent solves on one GPU.

Requests (one lattice each, e.g. a contraction-flow power-law solve) arrive
over time.  The server keeps admitted solves resident in a BatchRunner and
advances all of them together, ``tick_steps`` steps per tick (a replayed HIP
graph).  Between ticks it

* checks each solve's convergence -- the relative velocity change over the
  tick, ||u - u_prev||_2 / ||u||_2 over fluid nodes in FP64, against
  ``tol`` -- and its step budget, retires finished solves (reading their
  velocity field) and frees their HBM; a solve whose velocity is no longer
  finite (it diverged) is retired at once as failed, so it frees its slot
  without disturbing the solves batched with it;
* admits waiting requests in arrival order while their footprint fits the
  HBM budget and a slot is free, so work that arrives while others run
  joins the next tick.

Each solve's result is independent of what else is resident (the batched
kernel touches only the lattice's own arena segment); the tests check that
a solve returns the same field alone and in a mix.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _hip as H
from .batch import BatchRunner, LatticeSpec, footprint_bytes, plane_stride, stored_nodes
from ._device import torch


@dataclass
class Request:
    rid: int
    arrival: float                  # seconds after the server starts
    spec: LatticeSpec
    max_steps: int = 50_000
    tol: float = 1e-6
    # filled by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    steps: int = 0
    residual: float = float("inf")
    converged: bool = False
    failed: bool = False            # diverged: non-finite velocity at a tick
    rejected: Optional[str] = None  # why it can never be admitted (then failed, never run)
    result: Optional[np.ndarray] = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


@dataclass
class ServeStats:
    completed: List[Request] = field(default_factory=list)
    wall: float = 0.0
    ticks: int = 0
    captures: int = 0        # graphs captured (a change that fits the captured tick needs none)
    peak_resident: int = 0
    peak_bytes: int = 0

    def summary(self):
        lat = np.array([r.latency for r in self.completed])
        return {
            "solves": len(self.completed),
            "wall_s": round(self.wall, 3),
            "solves_per_s": round(len(self.completed) / self.wall, 3) if self.wall > 0 else None,
            "p50_latency_s": round(float(np.percentile(lat, 50)), 4) if len(lat) else None,
            "p99_latency_s": round(float(np.percentile(lat, 99)), 4) if len(lat) else None,
            "converged": int(sum(r.converged for r in self.completed)),
            "failed": int(sum(r.failed for r in self.completed)),
            "rejected": int(sum(r.rejected is not None for r in self.completed)),
            "ticks": self.ticks,
            "captures": self.captures,
            "peak_resident": self.peak_resident,
            "peak_resident_bytes": self.peak_bytes,
        }


class SolveServer:
    def __init__(self, storage: str = "f32", hbm_budget_bytes: int = 8 << 30, max_slots: int = 256,
                 tick_steps: int = 200, grid_blocks: int = 0, keep_results: bool = True, tile_cols: int = 0,
                 rounding: str = "stochastic"):
        if tick_steps % 4:
            raise ValueError("tick_steps must be a multiple of 4 (graph replays of two-step launches keep one "
                             "parity schedule)")
        self.storage = storage
        self.budget = int(hbm_budget_bytes)
        # arena sized from the budget: bytes per node of a resident lattice
        per_node = footprint_bytes(64, storage) // 64
        self.capacity_nodes = self.budget // per_node // 64 * 64
        # two steps per launch where the storage allows it; the tick's graph
        # is captured again whenever the resident set changes (the two-step
        # tiling, its table and mode follow the resident lattices)
        # 16-bit storage rounds stochastically by default: requests run to a
        # steady state, where rounding to nearest would push every population
        # the same way every step
        self.runner = BatchRunner([], storage=storage, grid_blocks=grid_blocks,
                                  capacity_nodes=self.capacity_nodes, max_slots=max_slots, blocked=True,
                                  tile_cols=tile_cols, rounding=rounding)
        self.changed = False
        self.tick_steps = int(tick_steps)
        self.keep_results = keep_results
        self.graph = None
        self.resident: Dict[int, Request] = {}      # slot -> request
        self.used_bytes = 0

    def _admit(self, req: Request, now: float):
        slot = self.runner.add(req.spec)
        self.changed = True
        req.admitted = now
        self.resident[slot] = req
        self.used_bytes += footprint_bytes(req.spec, self.storage)

    def _retire(self, slot: int, now: float):
        req = self.resident.pop(slot)
        if self.keep_results:
            req.result = self.runner.velocity(slot)
        self.runner.remove(slot)
        self.changed = True
        self.used_bytes -= footprint_bytes(req.spec, self.storage)
        req.finished = now

    def serve(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
              on_tick: Optional[Callable] = None) -> ServeStats:
        """Runs until every request has finished; arrivals are honoured in real time."""
        t = torch()
        stats = ServeStats()
        pending = []
        for req in sorted(requests, key=lambda r: r.arrival):
            # a request that cannot fit even an empty server would hold the
            # arrival-order queue forever: it is finished at once as failed
            why = never_fits(req.spec, self.storage, self.budget, self.capacity_nodes, self.runner.max_slots,
                             self.runner.align_rows)
            if why is None:
                pending.append(req)
            else:
                req.rejected, req.failed, req.finished = why, True, req.arrival
                stats.completed.append(req)
        stream = t.cuda.Stream()
        t0 = clock()
        with t.cuda.stream(stream):
            while pending or self.resident:
                now = clock() - t0
                # admission in arrival order while the footprint fits
                while pending and pending[0].arrival <= now and self.runner.fits(pending[0].spec) and \
                        self.used_bytes + footprint_bytes(pending[0].spec, self.storage) <= self.budget:
                    self._admit(pending.pop(0), now)
                if not self.resident:
                    if pending:        # idle until the next arrival
                        time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
                    continue
                stats.peak_resident = max(stats.peak_resident, len(self.resident))
                stats.peak_bytes = max(stats.peak_bytes, self.used_bytes)
                if self.changed and self.graph is not None and not self._graph_holds():
                    self.graph.close()      # releases the batch's tables for the new tiling
                    self.graph = None
                self.changed = False
                if self.graph is None:
                    # the tick runs eagerly, then is captured with room for
                    # about a quarter more tiles: later changes to the resident
                    # set are committed into the captured tables (a fresh
                    # lattice reads in place at its first launch) until they
                    # outgrow them or change the kernel variant
                    self.runner.run(self.tick_steps)
                    tiles = H.lib().lbm_batch_tiles2(self.runner.h)
                    self.runner.fit_tables(margin=max(64, tiles // 4))
                    self.graph = self.runner.capture(self.tick_steps)
                    stats.captures += 1
                else:
                    self.graph.replay()
                stats.ticks += 1
                res = self.runner.residuals()            # synchronises the stream
                now = clock() - t0
                for slot, req in list(self.resident.items()):
                    req.steps += self.tick_steps
                    req.residual, req.failed = _residual(*res[slot])
                    if req.residual < req.tol:
                        req.converged = True
                    if req.converged or req.failed or req.steps >= req.max_steps:
                        self._retire(slot, now)
                        stats.completed.append(req)
                if on_tick:
                    on_tick(self, stats)
        stats.wall = clock() - t0
        return stats

    def _graph_holds(self) -> bool:
        """Commits the changed resident set; True if the captured tick still
        runs it (same kernel variant, tiles within the captured tables)."""
        try:
            return self.runner._kernel_mode() == self.graph.mode
        except H.LBMError:      # the two-step table is full after the capture
            return False

    def close(self):
        if self.graph is not None:
            self.graph.close()
            self.graph = None
        self.runner.close()


def never_fits(spec: LatticeSpec, storage: str, budget_bytes: int, capacity_nodes: int, max_slots: int = 1,
               align_rows: bool = True) -> Optional[str]:
    """Why a request can never be admitted by a server with this HBM budget
    and arena (None if an otherwise empty server admits it)."""
    need = footprint_bytes(spec, storage)
    if need > budget_bytes:
        return f"footprint {need} B exceeds the HBM budget of {budget_bytes} B"
    nodes = plane_stride(stored_nodes(spec) if align_rows else spec.N)
    if nodes > capacity_nodes:
        return f"{nodes} stored nodes exceed the arena of {capacity_nodes} nodes"
    if max_slots < 1:
        return "the server has no slots"
    return None


def _residual(d2: float, u2: float):
    """(||u - u_prev|| / ||u||, diverged) from the FP64 sums of a tick."""
    if not (np.isfinite(d2) and np.isfinite(u2)):
        return float("nan"), True
    return (float(np.sqrt(d2 / u2)) if u2 > 0 else float("inf")), False


def solve_alone(spec: LatticeSpec, max_steps: int, tol: float, tick_steps: int, storage: str = "f32",
                grid_blocks: int = 0, blocked: bool = False, rounding: str = "stochastic"):
    """The same request on its own: one lattice, the same tick / residual
    schedule (and the same stop on divergence); by default the plain path
    (one step per launch), which gives the same fields bit for bit as two
    steps per launch (blocked=True)."""
    r = BatchRunner([spec], storage=storage, grid_blocks=grid_blocks, blocked=blocked, rounding=rounding)
    steps, res = 0, float("inf")
    while True:
        r.run(tick_steps)
        steps += tick_steps
        res, failed = _residual(*r.residuals()[0])
        if res < tol or failed or steps >= max_steps:
            break
    u = r.velocity(0)
    r.close()
    return u, steps, res
