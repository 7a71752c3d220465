"""Stepper: the fused form of a case script's ``lbm_solve()``.

The reference advances a lattice with 9-14 Taichi kernels per step
(cases/Circular cylinder.py:182-193, cases/Power Law Fluid.py:160-173,
cases/Bubble or two-phase separation.py:197-212).  ``Stepper.run(n)`` performs
n such steps with one HIP launch per step (two with Shan-Chen), through
``lbm_run`` of the C-ABI, and leaves the LBField exactly as n calls of the
granular sequence would (the final boundary pass runs lazily, on the first
read of ``f`` / ``rho`` / ``vel``).

Equivalent granular sequence (what each constructor flag selects):
    density -> pressure -> [update_psi -> psi halo -> cau_SCforce]
    -> force_density -> velocity -> [shear_rate -> relaxation_time]
    -> collide -> boundary list -> [time_updata]
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _hip as H
from ._codes import compile_codes
from ._device import device, stream, torch
from .lbm import gfields, group_device, lattice_struct

_STORAGE = {"f32": H.STORE_F32, "f16": H.STORE_F16_SHIFT, "bf16": H.STORE_BF16_SHIFT}
ROUNDING = {"nearest": H.ROUND_NEAREST, "stochastic": H.ROUND_STOCHASTIC}


def _same_nodes(g1, g2, NY) -> bool:
    """The two groups hold the same set of nodes (groups are sets,
    BoundaryCondition.py:38-39)."""
    if g1 is g2:
        return True
    a, b = np.asarray(g1.nodes, np.int64), np.asarray(g2.nodes, np.int64)
    if len(a) != len(b):
        return False
    return np.array_equal(np.sort(a[:, 0] * NY + a[:, 1]), np.sort(b[:, 0] * NY + b[:, 1]))


class Stepper:
    def __init__(self, lb_field, macroscopic_engine, collision_engine, boundary_engine, power_law=False,
                 boundary_engine_sc=None, advance_time=False, storage="f32", blocked=True, tile_cols=0,
                 no_alloc=False, ib_engine=None, sphere_field=None, rounding="nearest"):
        """blocked: run two steps per launch where the lattice allows it
        (lbm_run_blocked; bit-identical to one step per launch), on tiles of
        tile_cols = 16, 32 or 64 columns (0: the library's default; "auto":
        the first run of at least tune_steps() steps times TUNE_STEPS steps
        with each of TUNE_CANDIDATES, twice, and keeps the fastest -- the
        fields do not depend on the width, so those steps are ordinary steps
        of the run).  no_alloc: two-step runs go through
        lbm_run_blocked_noalloc, which keeps the power-law state's second
        buffer in the library's fixed device workspace (nothing is allocated
        per run; lattices up to 2^22 state values).  storage: "f32", or "f16" /
        "bf16" for shifted 16-bit populations g = f - w_k (FP32 arithmetic,
        half the population bytes): the FP32 populations are encoded at the
        first run (and again after anything else writes them), the kernels
        run on the 16-bit buffers, and the boundary pass of a readback
        decodes them into the LBField's FP32 fields (active nodes).  Runs
        without a readback in between equal one run; a readback rounds the
        post-boundary populations, which the next run starts from, to the
        storage format.  rounding: how a step's populations are rounded to
        16-bit storage -- "nearest" (even), or "stochastic" (gfx950's
        stochastic conversions seeded by node and step count: unbiased, so a
        steady flow does not accumulate the same rounding every step; the
        steady-state velocity error drops about tenfold, while short
        transients see a little more noise).  ib_engine, sphere_field: immersed-boundary particles
        (ibm.py IBEngine, SphereIB) coupled into every step, as a case's
        lbm_solve() with ib_engine.MultiDirectForcingLoop(sphere_field,
        lb_field) between MacroscopicEngine.velocity and the collision
        (lbm_ib_run: one fused step per launch plus the forcing at the
        markers' stencil nodes); the markers are the last SphereIBUpdate's
        (IBDEMCouplerEngine.SphereInfoUpdata before each run moves them).
        FP32 storage, no node rules, Shan-Chen or power law."""
        self.lbf = lb_field
        self.blocked = bool(blocked)
        self.no_alloc = bool(no_alloc)
        self._tune = tile_cols == "auto"
        self.tile_cols = 0 if self._tune else int(tile_cols)
        self.tune_timings = {}
        self._tau_next = None
        self.macro = macroscopic_engine
        self.coll = collision_engine
        self.bce = boundary_engine
        self.power_law = bool(power_law)
        self.advance_time = bool(advance_time)
        if storage not in _STORAGE:
            raise ValueError(f"storage must be one of {sorted(_STORAGE)}")
        self.storage_name = storage
        self.storage = _STORAGE[storage]
        if rounding not in ROUNDING:
            raise ValueError(f"rounding must be one of {sorted(ROUNDING)}")
        self.rounding = ROUNDING[rounding]
        self._pop16 = None       # (2, nc*9, ps) shifted 16-bit populations (storage f16 / bf16)
        self.sc = lb_field.sc_field
        self.ib_engine, self.sphere_field = ib_engine, sphere_field
        self._ib_scratch = None
        if (ib_engine is None) != (sphere_field is None):
            raise ValueError("IB coupling needs both ib_engine and sphere_field")
        if ib_engine is not None and (self.sc is not None or self.power_law or storage != "f32"):
            raise NotImplementedError("IB-coupled fused steps need FP32 storage without Shan-Chen or a power law")
        self._psi_fixed = []
        if self.sc is not None:
            if boundary_engine_sc is None:
                raise NotImplementedError("the fused Shan-Chen path needs the psi boundaries "
                                          "PeriodicPsi_LR then PeriodicPsi_BT")
            # FixedPsi over walls: constant psi on non-fluid nodes (component 0)
            self._psi_fixed = boundary_engine_sc.fused_plan(lb_field, macroscopic_engine.group.nodes)
        if not _same_nodes(macroscopic_engine.group, collision_engine.group, lb_field.NY):
            raise NotImplementedError("the macroscopic and collision groups differ")
        self._compiled_version = None
        self._compile()

    def _compile(self):
        lbf = self.lbf
        code, rules = compile_codes(lbf.NX, lbf.NY, self.macro.group.nodes, self.bce.fused_list(),
                                    lbf.neighbor_valid())
        if self.sc is not None and getattr(self.sc, "adhesion", False):
            code = code | self.sc.solid_codes()      # LBM_CODE_SC_SOLID where fluid_strategy is false
        self.code_host = code
        self.neem_host = rules
        t = torch()
        buf = np.zeros(lbf.ps, np.uint16)
        buf[:lbf.N] = code
        self.code = t.as_tensor(buf.view(np.int16), device=device())
        arr = (H.NeemRule * max(1, len(rules)))()
        for e, (kind, d, ux, uy, rho) in enumerate(rules):
            arr[e].kind, arr[e].direction, arr[e].ux, arr[e].uy, arr[e].rho = kind, d, ux, uy, rho
        self.neem = arr
        if getattr(self, "ib_engine", None) is not None and rules:
            raise NotImplementedError("IB-coupled fused steps need a boundary list without node rules "
                                      "(NEEM, VelocityBB, OpenNeumann, ...): use the granular methods")
        self._compiled_version = lbf.geometry_version

    def _tau_mode(self):
        if self.power_law:
            return H.TAU_POWER_LAW
        return H.TAU_CONST if self.coll._uniform else H.TAU_FIELD

    # ---- shifted 16-bit storage ---------------------------------------------
    def _shift_weights(self):
        t = torch()
        w = np.array([4, 1, 1, 1, 1, 1 / 4, 1 / 4, 1 / 4, 1 / 4], np.float32) / np.float32(9.0)
        return t.as_tensor(np.tile(w, self.lbf.nc)[:, None], device=device())

    def _encode16(self):
        """FP32 populations (both buffers) -> g = f - w_k in FP16 / BF16, rounded
        to nearest even as the kernels store them."""
        t = torch()
        lbf = self.lbf
        lbf.materialize()
        pop = lbf.state()["pop"]
        dt = t.float16 if self.storage_name == "f16" else t.bfloat16
        if self._pop16 is None or self._pop16.shape != pop.shape:
            self._pop16 = t.empty(pop.shape, dtype=dt, device=pop.device)
        w = self._shift_weights()
        for b in range(2):
            self._pop16[b].copy_((pop[b] - w).to(dt))
        lbf._pop_owner = self

    def _decode16(self):
        """16-bit populations -> FP32 fields of the LBField (active nodes; the
        kernels leave the others' populations untouched)."""
        t = torch()
        lbf = self.lbf
        pop = lbf.state()["pop"]
        w = self._shift_weights()
        active = ((self.code.to(t.int32) & H.CODE_ACTIVE) != 0)[None, :]
        for b in range(2):
            pop[b].copy_(t.where(active, self._pop16[b].float() + w, pop[b]))

    def _structs(self, tau_mode):
        lbf = self.lbf
        st = lbf.state()
        L = lattice_struct(lbf, collision=self.coll, sc=self.sc, storage=self.storage, tau_mode=tau_mode,
                           neem=self.neem_host)
        L.rounding = self.rounding
        F = H.Fields()
        pops = st["pop"] if self.storage == H.STORE_F32 else self._pop16
        F.f[0], F.f[1] = H.ptr(pops[0]), H.ptr(pops[1])
        F.macro[0], F.macro[1] = H.ptr(st["mac"][0]), H.ptr(st["mac"][1])
        F.tau = H.ptr(self.coll.state_tensor())
        F.body_force = H.ptr(lbf.aux("body_force"))
        F.shear_rate = H.ptr(lbf.aux("shear_rate")) if tau_mode == H.TAU_POWER_LAW else None
        if self.sc is not None:
            planes = self.sc.planes_for_run()
            for nodes, value in self._psi_fixed:
                if len(nodes):
                    flat = torch().as_tensor(nodes[:, 0] * lbf.NY + nodes[:, 1], device=device())
                    planes[:, 0, flat] = value
            F.psi[0], F.psi[1] = H.ptr(planes[0]), H.ptr(planes[1])
            F.sc_force = H.ptr(lbf.aux("SCforce"))
            if any(L.psi_kind[c] != H.PSI_SC for c in range(lbf.nc)):
                F.temperature = H.ptr(lbf.aux("T"))
        F.code = H.ptr(self.code)
        return L, F

    TUNE_CANDIDATES = (32, 64)    # the first is the library's default width
    TUNE_STEPS = 12           # timed steps per sample (even: two-step launches only)
    TUNE_MARGIN = 0.02        # another width must beat the default by this much (short samples
                              # read a few percent slow; closer calls keep the default)

    @classmethod
    def tune_steps(cls) -> int:
        """Steps a tuning run takes (tile_cols="auto")."""
        return 2 * len(cls.TUNE_CANDIDATES) * (2 + cls.TUNE_STEPS)

    def run(self, nsteps: int):
        """Advance ``nsteps`` steps (asynchronous on the current stream; a
        tuning run with tile_cols="auto" synchronises)."""
        nsteps = int(nsteps)
        if nsteps <= 0:
            return
        if self._tune and self.blocked and nsteps >= self.tune_steps():
            # each width twice, in the order A B B A (the first runs of a
            # lattice are slower: first touches of its buffers); each sample
            # after one launch of that width (which also loads its code)
            self._tune = False
            t = torch()
            s = t.cuda.current_stream()
            samples = {w: [] for w in self.TUNE_CANDIDATES}
            for w in self.TUNE_CANDIDATES + tuple(reversed(self.TUNE_CANDIDATES)):
                self.tile_cols = w
                e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
                self._run(2)
                e0.record(s)
                self._run(self.TUNE_STEPS)
                e1.record(s)
                t.cuda.synchronize()
                samples[w].append(e0.elapsed_time(e1) / self.TUNE_STEPS)
            self.tune_timings = {w: min(v) for w, v in samples.items()}
            best = min(self.tune_timings, key=self.tune_timings.get)
            default = self.TUNE_CANDIDATES[0]
            if self.tune_timings[best] > (1.0 - self.TUNE_MARGIN) * self.tune_timings[default]:
                best = default
            self.tile_cols = best
            nsteps -= self.tune_steps()
            if nsteps <= 0:
                return
        self._run(nsteps)

    def _begin(self, nsteps: int):
        """Prepares a run of nsteps: compiles the codes if the geometry changed,
        encodes 16-bit storage if needed; returns the structs, the tau mode,
        the parity of the buffer the run reads and whether its first step
        reads f in place."""
        lbf = self.lbf
        if self._compiled_version != lbf.geometry_version:
            self._compile()
        if self.storage != H.STORE_F32 and lbf._pop_owner is not self:
            self._encode16()
        st = lbf.state()
        tau_mode = self._tau_mode()
        L, F = self._structs(tau_mode)
        if st["f_valid"]:
            parity, first_local = 1 - st["f2_idx"], 1
        else:
            parity, first_local = st["f2_idx"], 0
        if self.sc is not None and nsteps == 1:
            # a launch writes its psi into plane set parity ^ 1; a run of two
            # or more leaves the psi of its last two steps in both sets
            self.sc.keep_previous(parity ^ 1)
        return L, F, tau_mode, parity, first_local

    def _run(self, nsteps: int):
        lbf = self.lbf
        L, F, tau_mode, parity, first_local = self._begin(nsteps)
        st = lbf.state()
        if self.ib_engine is not None:
            lbf = self.lbf
            if self._ib_scratch is None:
                t = torch()
                self._ib_scratch = t.zeros((3 + 3 * lbf.nc, lbf.ps), dtype=t.float32, device=device())
            sc = self._ib_scratch
            G = H.GFields()
            G.vel = H.ptr(sc[0])
            G.total_rho = H.ptr(sc[2])
            G.rho = H.ptr(sc[3])
            G.body_force = H.ptr(sc[3 + lbf.nc])
            out = ctypes.c_int32(-1)
            H.check(H.lib().lbm_ib_run(ctypes.byref(L), ctypes.byref(F), ctypes.byref(self.sphere_field.abi()),
                                       ctypes.byref(G), self.ib_engine.MDFIteration, nsteps, parity, first_local,
                                       ctypes.byref(out), stream()), "lbm_ib_run")
            st["f2_idx"] = int(out.value)
        elif self.blocked and self.no_alloc:
            out = ctypes.c_int32(-1)
            H.check(H.lib().lbm_run_blocked_noalloc(ctypes.byref(L), ctypes.byref(F), self.neem, nsteps, parity,
                                                    first_local, self.tile_cols, ctypes.byref(out), stream()),
                    "lbm_run_blocked_noalloc")
            st["f2_idx"] = int(out.value)
        elif self.blocked:
            tau_next = None
            if tau_mode == H.TAU_POWER_LAW:
                state = self.coll.state_tensor()
                if self._tau_next is None or self._tau_next.shape != state.shape:
                    self._tau_next = torch().empty_like(state)
                tau_next = H.ptr(self._tau_next)
            out = ctypes.c_int32(-1)
            H.check(H.lib().lbm_run_blocked(ctypes.byref(L), ctypes.byref(F), self.neem, nsteps, parity, first_local,
                                            self.tile_cols, tau_next, ctypes.byref(out), stream()),
                    "lbm_run_blocked")
            st["f2_idx"] = int(out.value)
        else:
            H.check(H.lib().lbm_run(ctypes.byref(L), ctypes.byref(F), self.neem, nsteps, parity, first_local,
                                    stream()), "lbm_run")
            st["f2_idx"] = (parity + nsteps) % 2
        self._end(nsteps, L, F, tau_mode)

    def _end(self, nsteps: int, L, F, tau_mode):
        """After a run that left the newest f2 in buffer state["f2_idx"]: the
        LBField's lazy state (boundary pass and pressure observable of a
        readback)."""
        lbf = self.lbf
        st = lbf.state()
        shifted = self.storage != H.STORE_F32
        st["f_valid"] = False
        lbf._pop_owner = self if shifted else None
        if tau_mode == H.TAU_POWER_LAW:
            self.coll._uniform = False
        if self.sc is not None:
            self.sc.after_run(st["f2_idx"])
        if self.advance_time:
            lbf.time[None] += nsteps
        lbf._round_step += nsteps
        keep = (L, F, self.neem)
        last = st["f2_idx"]

        def finish():
            H.check(H.lib().lbm_finish(ctypes.byref(keep[0]), ctypes.byref(keep[1]), keep[2], last, stream()),
                    "lbm_finish")
            if shifted:
                self._decode16()
        lbf._finisher = finish

        def refresh_pressure():
            # MacroscopicEngine.pressure of the last step: it ran after density and
            # before update_psi and the boundary pass, so it sees rho of that step
            # without the NEEM overrides and psi of the step before (plane set 1 - last)
            if self.sc is not None:
                self.sc._halo_from_planes(1 - last)
            g = group_device(self.macro.group)
            H.check(H.lib().lbm_g_pressure(ctypes.byref(lattice_struct(lbf, sc=self.sc)),
                                           ctypes.byref(gfields(lbf, sc=self.sc)), g[0], g[1], stream()),
                    "pressure")
            if self.sc is not None:
                self.sc._halo_valid = False
        lbf._pressure_refresh = refresh_pressure

    # ---- FP64 diagnostics ---------------------------------------------------
    def reduce(self):
        """FP64 sums over the fluid group: {'mass': [per component], 'momentum': (x, y), 'nonfinite': n}."""
        lbf = self.lbf
        lbf.materialize()
        L = lattice_struct(lbf)
        out = (ctypes.c_double * (lbf.nc + 3))()
        H.check(H.lib().lbm_reduce_macro(ctypes.byref(L), H.ptr(lbf._mac()), H.ptr(self.code), out, stream()),
                "lbm_reduce_macro")
        v = list(out)
        return {"mass": v[:lbf.nc], "momentum": (v[lbf.nc], v[lbf.nc + 1]), "nonfinite": int(v[lbf.nc + 2])}
