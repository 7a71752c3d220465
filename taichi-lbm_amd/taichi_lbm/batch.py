"""Batched lattices: many independent solves in one persistent launch per step.

The reference runs one lattice per sequence of Taichi kernels.  ``BatchRunner``
keeps any number of independent lattices (different sizes, collision models,
relaxation times and boundary values) resident in device arenas and advances
all active ones with one launch per step (``lbm_batch_run``) or a replayed
HIP graph (``lbm_batch_graph``).  Each lattice follows exactly the semantics
of ``Stepper.run`` (the same per-node body runs in both kernels).

Lattices can be added and removed between launches (slots + a first-fit
arena allocator), which is what the serving loop (serving.py) builds on.

Population storage is selectable: ``"f32"``, ``"f16"`` / ``"bf16"`` (shifted:
f - w_k is stored, arithmetic in fp32) or ``"f64"`` (fp64 throughout, the
high-precision path).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _hip as H
from ._codes import compile_codes
from ._device import device, plane_stride, stream, torch

STORAGE = {"f32": H.STORE_F32, "f16": H.STORE_F16_SHIFT, "bf16": H.STORE_BF16_SHIFT, "f64": H.STORE_F64}
POP_BYTES = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}
_W64 = np.array([4, 1, 1, 1, 1, 1 / 4, 1 / 4, 1 / 4, 1 / 4]) / 9.0
_W32 = np.array([4, 1, 1, 1, 1, .25, .25, .25, .25], np.float32) / np.float32(9)
_CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
_CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def _torch_dtype(storage):
    t = torch()
    return {"f32": t.float32, "f16": t.float16, "bf16": t.bfloat16, "f64": t.float64}[storage]


@dataclass
class LatticeSpec:
    """Host description of one lattice of a batch (compiled from the drop-in API)."""
    NX: int
    NY: int
    code: np.ndarray                  # uint16 (NX*NY,)
    neem: list                        # [(kind, direction, ux, uy, rho)]
    lattice: H.Lattice                # parameters (collision, tau, fluid model, gravity)
    rho0: np.ndarray                  # (NX, NY) initial density (fluid nodes)
    u0: np.ndarray                    # (NX, NY, 2) initial velocity
    tag: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.NX * self.NY


def spec_from_case(case, power_law=None) -> LatticeSpec:
    """Compile a case built through the drop-in API (cases.py) into a LatticeSpec."""
    from .lbm import lattice_struct
    lbf, coll = case["lbf"], case["coll"]
    if lbf.nc != 1 or lbf.sc_field is not None:
        raise NotImplementedError("batched lattices are single-component without Shan-Chen")
    code, rules = compile_codes(lbf.NX, lbf.NY, case["macro"].group.nodes, case["bce"].fused_list(),
                                lbf.neighbor_valid())
    pl = case["power_law"] if power_law is None else power_law
    L = lattice_struct(lbf, collision=coll, tau_mode=H.TAU_POWER_LAW if pl else H.TAU_CONST, neem=rules)
    return LatticeSpec(lbf.NX, lbf.NY, code, rules, L, case["rho"][..., 0].copy(), case["vel"].copy(),
                       tag=dict(case.get("spec", {})))


ROW_ALIGN = 32     # nodes: a 128-byte line of FP32 populations


def _no_wrap_across_rows(spec: LatticeSpec) -> bool:
    """True when no active node streams across the lattice's j boundary: every
    population that would wrap from column NY-1 to 0 or back is bounced."""
    code = spec.code.reshape(spec.NX, spec.NY)
    for j, ks in ((0, (2, 5, 6)), (spec.NY - 1, (4, 7, 8))):   # pulls with cy = +1 at j = 0, cy = -1 at NY-1
        col = code[:, j].astype(np.uint32)
        act = (col & H.CODE_ACTIVE) != 0
        for k in ks:
            if np.any(act & (((col >> (k - 1)) & 1) == 0)):
                return False
    return True


def pad_rows(spec: LatticeSpec, align: int = ROW_ALIGN) -> LatticeSpec:
    """The lattice stored with rows of a multiple of `align` nodes: columns
    NY..NY'-1 are added outside the fluid group.  Only where no active node
    streams across the j boundary (the new columns then are never read and the
    fields are those of the unpadded lattice, bit for bit); otherwise, or when
    NY is already a multiple, the spec itself.  Aligned rows let every tile
    row of the two-step kernel start on a cache line."""
    NYp = -(-spec.NY // align) * align
    if NYp == spec.NY or spec.NY < 2 or not _no_wrap_across_rows(spec):
        return spec
    code = np.zeros((spec.NX, NYp), np.uint16)
    code[:, :spec.NY] = spec.code.reshape(spec.NX, spec.NY)
    rho0 = np.ones((spec.NX, NYp), spec.rho0.dtype)
    rho0[:, :spec.NY] = spec.rho0
    u0 = np.zeros((spec.NX, NYp, 2), spec.u0.dtype)
    u0[:, :spec.NY] = spec.u0
    L = H.Lattice()
    ctypes.memmove(ctypes.byref(L), ctypes.byref(spec.lattice), ctypes.sizeof(H.Lattice))
    L.ny = NYp
    return LatticeSpec(spec.NX, NYp, code.reshape(-1), spec.neem, L, rho0, u0, tag=dict(spec.tag, view_ny=spec.NY))


def stored_nodes(spec: LatticeSpec) -> int:
    """Nodes a lattice occupies in a batch (rows padded where that is exact)."""
    return pad_rows(spec).N


def footprint_bytes(spec_or_n, storage="f32") -> int:
    """Device bytes one lattice occupies in a batch (its arena segment)."""
    n = spec_or_n if isinstance(spec_or_n, int) else stored_nodes(spec_or_n)
    ps = plane_stride(n)
    r = 8 if storage == "f64" else 4          # relaxation state and macro planes: fp64 with fp64 storage
    return ps * (2 * 9 * POP_BYTES[storage] + 2 * 4 * r + 2 * 4 + 2 + 2 * r)


def two_step_tiles(specs, tile_cols: int = 0) -> int:
    """Tiles of the two-step launch for these lattices: tile_cols = 16, 32 or
    64 columns (0: the library's default) by lbm_tile_rows rows (ring width 2
    when any lattice has node rules)."""
    ring = 2 if any(len(s.neem) for s in specs) else 1
    tj = tile_cols or 32
    ti = H.lib().lbm_tile_rows(ring, tile_cols)
    H.check(min(ti, 0), "lbm_tile_rows")
    return sum(-(-s.NX // ti) * -(-s.NY // tj) for s in specs)


class _Arena:
    """First-fit allocator over node offsets (units of 64 nodes)."""

    def __init__(self, capacity: int):
        self.free = [(0, capacity)]

    def alloc(self, n: int) -> Optional[int]:
        for q, (s, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(q)
                else:
                    self.free[q] = (s + n, l - n)
                return s
        return None

    def release(self, s: int, n: int):
        self.free.append((s, n))
        self.free.sort()
        merged = []
        for a, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((a, l))
        self.free = merged


class BatchRunner:
    """Resident lattices advanced together.

    ``BatchRunner(specs)`` sizes the arena for the given lattices and adds them;
    ``BatchRunner([], capacity_nodes=..., max_slots=...)`` starts empty for
    dynamic use (``add`` / ``remove``).  ``rounding`` ("nearest" or
    "stochastic") is the 16-bit stores' rounding, as ``Stepper``'s: each
    lattice's stochastic seeds follow its own step count (kept in device
    memory and advanced by every run and graph replay), so a lattice rounds
    the same alone and in the batch."""

    def __init__(self, specs: Sequence[LatticeSpec] = (), storage: str = "f32", grid_blocks: int = 0,
                 capacity_nodes: int = 0, max_slots: int = 0, blocked: bool = True, tile_cols: int = 0,
                 align_rows: bool = True, rounding: str = "nearest"):
        from .stepper import ROUNDING
        if storage not in STORAGE:
            raise ValueError(f"storage must be one of {sorted(STORAGE)}")
        if rounding not in ROUNDING:
            raise ValueError(f"rounding must be one of {sorted(ROUNDING)}")
        self.rounding = ROUNDING[rounding]
        t = torch()
        dev = device()
        specs = list(specs)
        self.storage = storage
        self.grid = int(grid_blocks)
        self.blocked = bool(blocked)    # two steps per launch where the batch allows it
        self.align_rows = bool(align_rows)   # rows padded to whole cache lines where that is exact
        cap = int(capacity_nodes) or sum(plane_stride(stored_nodes(s) if align_rows else s.N) for s in specs)
        cap = max(64, (cap + 63) // 64 * 64)
        self.capacity = cap
        self.max_slots = int(max_slots) or max(1, len(specs))
        pdt = _torch_dtype(storage)
        rdt = t.float64 if storage == "f64" else t.float32
        # arenas indexed by node offset: lattice l at offset o uses populations
        # [9*o, 9*(o+ps)), macro [4*o, 4*(o+ps)), code / tau [o, o+ps), u_old [2*o, ...);
        # macro planes and relaxation state are fp64 with fp64 storage
        self.popseg = t.zeros((2, 9 * cap), dtype=pdt, device=dev)
        self.macseg = t.zeros((2, 4 * cap), dtype=rdt, device=dev)
        self.velseg = t.zeros(2 * cap, dtype=t.float32, device=dev)
        self.code = t.zeros(cap, dtype=t.int16, device=dev)
        self.tau = t.zeros(cap, dtype=rdt, device=dev)
        # second relaxation-state buffer: power-law lattices take two steps per launch
        self.tau2 = t.zeros(cap, dtype=rdt, device=dev) if self.blocked else None
        self.arena = _Arena(cap)
        n = self.max_slots
        self.lats = (H.Lattice * n)()
        self.flds = (H.Fields * n)()
        self.neem = (H.NeemRule * (n * H.MAX_NEEM))()
        for l in range(n):            # empty slots: 1x1 placeholders, inactive
            L = self.lats[l]
            L.nx = L.ny = L.nc = 1
            L.plane_stride = 64
            L.storage = STORAGE[storage]
            L.tau_min, L.tau_max = 0.5000001, 4.0
            self._point(l, 0)
        h = ctypes.c_void_p()
        H.check(H.lib().lbm_batch_create(self.lats, self.flds, self.neem, n, ctypes.byref(h)), "lbm_batch_create")
        self.h = h
        for l in range(n):
            H.check(H.lib().lbm_batch_set(h, l, 0, -1, None), "lbm_batch_set")
        H.check(H.lib().lbm_batch_reserve(h, cap // 256 + n + 8), "lbm_batch_reserve")
        self.tile_cols = int(tile_cols)
        H.check(H.lib().lbm_batch_tile_cols(h, self.tile_cols), "lbm_batch_tile_cols")
        if self.blocked and specs:
            # room for any tile width, so that autotune_tiles can switch (a
            # commit before capture grows the table further if it must)
            need = max(two_step_tiles(specs, w) for w in (16, 32, 64))
            H.check(H.lib().lbm_batch_reserve2(h, need + 64), "lbm_batch_reserve2")
        self.slots: Dict[int, dict] = {}
        self._open_graphs = 0
        self.parity = 0
        self.finished = False
        self._dirty = True
        self._pending_local = set()    # slots whose first launch reads in place
        self.ps = {}
        for s in specs:
            self.add(s)
        self.commit()

    # ---- layout --------------------------------------------------------------
    def _point(self, l, o, ps=64):
        F = self.flds[l]
        pe = self.popseg.element_size()
        F.f[0] = self.popseg[0].data_ptr() + 9 * o * pe
        F.f[1] = self.popseg[1].data_ptr() + 9 * o * pe
        me = self.macseg.element_size()
        F.macro[0] = self.macseg[0].data_ptr() + 4 * o * me
        F.macro[1] = self.macseg[1].data_ptr() + 4 * o * me
        F.tau = self.tau.data_ptr() + o * self.tau.element_size()
        F.tau2 = self.tau2.data_ptr() + o * self.tau2.element_size() if self.tau2 is not None else None
        F.code = self.code.data_ptr() + o * 2

    def _vel_ptr(self, o):
        return self.velseg.data_ptr() + 2 * o * 4

    # ---- lattices --------------------------------------------------------------
    def fits(self, spec: LatticeSpec) -> bool:
        n = stored_nodes(spec) if self.align_rows else spec.N
        return len(self.slots) < self.max_slots and any(l >= plane_stride(n) for _, l in self.arena.free)

    def add(self, spec: LatticeSpec) -> int:
        """Makes the lattice resident (f = feq(u0, rho0), init_LBM) and active
        from the next launch on; returns its slot."""
        given = spec
        if self.align_rows:
            spec = pad_rows(spec)
        ps = plane_stride(spec.N)
        slot = next((l for l in range(self.max_slots) if l not in self.slots), None)
        if slot is None:
            raise RuntimeError("no free batch slot")
        o = self.arena.alloc(ps)
        if o is None:
            raise RuntimeError("batch arena full")
        # "spec" as given, "stored" as laid out in the arena (rows possibly padded)
        self.slots[slot] = {"spec": given, "stored": spec, "off": o, "ps": ps}
        self.ps[slot] = ps
        L = self.lats[slot]
        ctypes.memmove(ctypes.byref(L), ctypes.byref(spec.lattice), ctypes.sizeof(H.Lattice))
        L.plane_stride = ps
        L.storage = STORAGE[self.storage]
        L.rounding = self.rounding
        L.n_neem = len(spec.neem)
        self._point(slot, o, ps)
        for e, (kind, d, ux, uy, rho) in enumerate(spec.neem):
            r = self.neem[slot * H.MAX_NEEM + e]
            r.kind, r.direction, r.ux, r.uy, r.rho = kind, d, ux, uy, rho
        self._init_slot(slot)
        rules = ctypes.cast(ctypes.addressof(self.neem) + slot * H.MAX_NEEM * ctypes.sizeof(H.NeemRule),
                            ctypes.POINTER(H.NeemRule))
        H.check(H.lib().lbm_batch_update(self.h, slot, ctypes.byref(L), ctypes.byref(self.flds[slot]), rules),
                "lbm_batch_update")
        H.check(H.lib().lbm_batch_set(self.h, slot, 1, 0, self._vel_ptr(o)), "lbm_batch_set")
        # a new lattice starts at step 0 (the seeds of 16-bit stochastic rounding)
        H.check(H.lib().lbm_batch_set_round_step(self.h, slot, 0, stream()), "lbm_batch_set_round_step")
        self._pending_local.add(slot)
        self._dirty = True
        return slot

    def remove(self, slot: int):
        info = self.slots.pop(slot)
        H.check(H.lib().lbm_batch_set(self.h, slot, 0, -1, None), "lbm_batch_set")
        self.arena.release(info["off"], info["ps"])
        self._pending_local.discard(slot)
        self._dirty = True

    def _init_slot(self, slot):
        """init_LBM: f = feq(u0, rho0) on active nodes into the populations the
        next launch reads; relaxation state = initTau's value; u_old = 0."""
        t = torch()
        info = self.slots[slot]
        s, o, ps = info["stored"], info["off"], info["ps"]
        f64 = self.storage == "f64"
        dt = np.float64 if f64 else np.float32
        w = _W64 if f64 else _W32
        act = (s.code & H.CODE_ACTIVE) != 0
        rho = s.rho0.reshape(-1).astype(dt)
        ux = s.u0[..., 0].reshape(-1).astype(dt)
        uy = s.u0[..., 1].reshape(-1).astype(dt)
        uv = ux * ux + uy * uy
        host = np.zeros((9, ps), dt)
        for k in range(9):
            eu = _CX[k] * ux + _CY[k] * uy
            fe = w[k] * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uv)
            if self.storage in ("f16", "bf16"):
                fe = fe - w[k]                                  # shifted storage
            host[k, :s.N] = np.where(act, fe, 0)
        self.popseg[self.parity, 9 * o: 9 * (o + ps)].copy_(t.as_tensor(host.reshape(-1)).to(self.popseg.dtype))
        self.popseg[self.parity ^ 1, 9 * o: 9 * (o + ps)].zero_()
        self.macseg[:, 4 * o: 4 * (o + ps)].zero_()
        self.velseg[2 * o: 2 * (o + ps)].zero_()
        code = np.zeros(ps, np.uint16)
        code[:s.N] = s.code
        self.code[o:o + ps].copy_(t.as_tensor(code.view(np.int16)))
        L = s.lattice
        state = {0: L.tau[0], 1: L.tau[0], 2: L.s_shear[0]}[L.collision]
        tau = np.zeros(ps, dt)
        tau[:s.N] = np.where(act, state, 1.0 if L.collision == 2 else 0.0)
        self.tau[o:o + ps].copy_(t.as_tensor(tau))

    # ---- control -------------------------------------------------------------
    def commit(self):
        if self._dirty:
            H.check(H.lib().lbm_batch_commit(self.h, stream()), "lbm_batch_commit")
            self._dirty = False

    @property
    def tiles(self):
        return H.lib().lbm_batch_tiles(self.h)

    def set_active(self, slot, active):
        info = self.slots[slot]
        H.check(H.lib().lbm_batch_set(self.h, slot, int(active), -1, self._vel_ptr(info["off"])), "lbm_batch_set")
        self._dirty = True

    def _after_launches(self):
        if self._pending_local:
            for slot in self._pending_local:
                H.check(H.lib().lbm_batch_set(self.h, slot, 1, -1, self._vel_ptr(self.slots[slot]["off"])),
                        "lbm_batch_set")
            self._pending_local.clear()
            self._dirty = True

    def fit_tables(self, margin: int = 64):
        """Size the two-step tile table to the current tiling plus `margin`
        entries (a graph captured afterwards launches one block per entry;
        before any capture)."""
        self.commit()
        H.check(H.lib().lbm_batch_reserve2(self.h, H.lib().lbm_batch_tiles2(self.h) + int(margin)),
                "lbm_batch_reserve2")

    def set_ring(self, ring: int):
        """0: node rules with a one-node ring where the lattices' tile grids
        can be shifted to allow it (mode 3), else a two-node ring; 2: always
        a two-node ring."""
        H.check(H.lib().lbm_batch_ring(self.h, int(ring)), "lbm_batch_ring")
        self._dirty = True

    def set_tile_cols(self, tile_cols: int):
        """Two-step tile width (16, 32 or 64; 0: default) from the next commit on."""
        H.check(H.lib().lbm_batch_tile_cols(self.h, int(tile_cols)), "lbm_batch_tile_cols")
        self.tile_cols = int(tile_cols)
        self._dirty = True

    @property
    def mode(self) -> int:
        """0: one step per launch; 1: two-step launches, one-node ring; 2:
        two-node ring (node rules); 3: node rules on a one-node ring with
        shifted tile grids."""
        return self._kernel_mode() & 7

    @property
    def unforced(self) -> bool:
        """The two-step kernels are built without the force terms (no
        active lattice has a body force)."""
        return bool(self._kernel_mode() & 16)

    def _kernel_mode(self) -> int:
        self.commit()
        return H.lib().lbm_batch_mode(self.h) if self.blocked else 0

    def run(self, nsteps: int):
        if nsteps <= 0:
            return
        self.commit()
        if self.blocked:
            out = ctypes.c_int32(-1)
            H.check(H.lib().lbm_batch_run_blocked(self.h, int(nsteps), self.parity, self.grid, ctypes.byref(out),
                                                  stream()), "lbm_batch_run_blocked")
            self.parity = int(out.value)
        else:
            H.check(H.lib().lbm_batch_run(self.h, int(nsteps), self.parity, self.grid, stream()), "lbm_batch_run")
            self.parity ^= (int(nsteps) & 1)
        self.finished = False
        self._after_launches()

    def capture(self, nsteps: int) -> "BatchGraph":
        """A graph of `nsteps` batched steps from the current parity.  Every
        replay must start on the same parity: nsteps even with one step per
        launch, a multiple of 4 with two-step launches.  Descriptor and
        tile-list changes committed between replays are picked up."""
        s = stream()
        if s == 0:
            raise RuntimeError("capture needs a non-default stream (torch.cuda.stream(...))")
        mode = self.mode
        if nsteps % (4 if mode else 2):
            raise ValueError(f"capture a multiple of {4 if mode else 2} steps")
        g = ctypes.c_void_p()
        if self.blocked:
            out = ctypes.c_int32(-1)
            H.check(H.lib().lbm_batch_graph_blocked(self.h, int(nsteps), self.parity, self.grid, s, ctypes.byref(g),
                                                    ctypes.byref(out)), "lbm_batch_graph_blocked")
            assert int(out.value) == self.parity
        else:
            H.check(H.lib().lbm_batch_graph(self.h, int(nsteps), self.parity, self.grid, s, ctypes.byref(g)),
                    "lbm_batch_graph")
        self._open_graphs += 1
        return BatchGraph(self, g, int(nsteps), self._kernel_mode())

    def _graph_closed(self):
        self._open_graphs -= 1
        if self._open_graphs == 0 and getattr(self, "h", None):
            H.check(H.lib().lbm_batch_release(self.h), "lbm_batch_release")

    def finish(self):
        """Boundary pass for every active lattice (f, rho, vel as the reference reads them)."""
        if not self.finished:
            self.commit()
            H.check(H.lib().lbm_batch_finish(self.h, self.parity, self.grid, stream()), "lbm_batch_finish")
            self.finished = True

    def residuals(self) -> np.ndarray:
        """Per slot: (sum |u - u_old|^2, sum |u|^2) in FP64 over its active nodes; u_old <- u.

        u is the velocity field as the reference reads it after a step (the
        boundary pass's overrides included), as in the error() of
        cases/Power Law Fluid.py:183-192."""
        self.finish()
        out = (ctypes.c_double * (2 * self.max_slots))()
        H.check(H.lib().lbm_batch_residual(self.h, self.parity, out, stream()), "lbm_batch_residual")
        return np.array(out).reshape(-1, 2)

    # ---- readback --------------------------------------------------------------
    def fields(self, slot: int):
        """Reference-shaped numpy arrays of a lattice after the boundary pass:
        f, f2 (NX, NY, 1, 9), rho (NX, NY, 1), total_rho (NX, NY), vel (NX, NY, 2)."""
        if slot in self._pending_local:
            raise RuntimeError("run the batch at least one step before reading a newly added lattice")
        self.finish()
        t = torch()
        info = self.slots[slot]
        s, o, ps = info["stored"], info["off"], info["ps"]
        NX, NY, p = s.NX, s.NY, self.parity
        VY = info["spec"].NY   # rows padded in the arena: the caller sees NX x VY

        def pops(q):
            a = self.popseg[q, 9 * o: 9 * (o + ps)].reshape(9, ps)[:, :s.N].to(t.float64).cpu().numpy()
            if self.storage in ("f16", "bf16"):
                act = (s.code & H.CODE_ACTIVE) != 0
                a = a + np.where(act[None, :], _W32.astype(np.float64)[:, None], 0.0)
            return np.ascontiguousarray(a.reshape(9, NX, NY).transpose(1, 2, 0)[:, :VY, None, :])

        m = self.macseg[p, 4 * o: 4 * (o + ps)].reshape(4, ps)[:, :s.N].cpu().numpy().reshape(4, NX, NY)[:, :, :VY]
        return {"f2": pops(p), "f": pops(p ^ 1), "rho": np.ascontiguousarray(m[0][..., None]),
                "total_rho": np.ascontiguousarray(m[1]), "vel": np.stack([m[2], m[3]], -1)}

    def velocity(self, slot: int) -> np.ndarray:
        return self.fields(slot)["vel"]

    def close(self):
        if getattr(self, "h", None):
            H.lib().lbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchGraph:
    """An instantiated graph of `nsteps` batched steps."""

    def __init__(self, runner: BatchRunner, g, nsteps: int, mode: int = 0):
        self.runner, self.g, self.nsteps, self.mode = runner, g, nsteps, mode
        self.parity = runner.parity

    def replay(self):
        r = self.runner
        if r.parity != self.parity:
            raise RuntimeError("the batch parity moved since capture (an odd eager run in between)")
        if r._kernel_mode() != self.mode:
            raise RuntimeError("the batch's launch mode changed since capture (power law, node rules or a "
                               "body force admitted); capture a new graph")
        r.commit()
        H.check(H.lib().lbm_graph_launch(self.g, stream()), "lbm_graph_launch")
        r.finished = False
        r._after_launches()

    def close(self):
        if self.g:
            H.lib().lbm_graph_destroy(self.g)
            self.g = None
            self.runner._graph_closed()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_tiles(runner: BatchRunner, candidates=(16, 32, 64), steps: int = 10):
    """Times the two-step tile width on the runner's own batch (steps a
    multiple of 2) and keeps the fastest; before any graph capture.  Advances
    the batch by len(candidates) * (steps + 2) steps."""
    t = torch()
    timings = {}
    for w in candidates:
        runner.set_tile_cols(w)
        runner.run(2)
        s = t.cuda.current_stream()
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record(s)
        runner.run(steps)
        e1.record(s)
        t.cuda.synchronize()
        timings[w] = e0.elapsed_time(e1) / steps
    best = min(timings, key=timings.get)
    runner.set_tile_cols(best)
    runner.fit_tables()
    return best, timings


def autotune_grid(runner: BatchRunner, candidates=(2, 4, 8, 16, 32), steps: int = 10,
                  cus: Optional[int] = None):
    """Times the persistent grid size (workgroups per CU x CUs, a multiple of 8)
    on the runner's own batch and keeps the fastest.  Advances the batch by
    len(candidates) * (steps + 2) steps."""
    t = torch()
    cus = cus or t.cuda.get_device_properties(t.cuda.current_device()).multi_processor_count
    best = None
    timings = {}
    for per_cu in candidates:
        grid = max(8, (per_cu * cus) // 8 * 8)
        runner.grid = grid
        runner.run(2)
        s = t.cuda.current_stream()
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record(s)
        runner.run(steps)
        e1.record(s)
        t.cuda.synchronize()
        timings[grid] = e0.elapsed_time(e1) / steps
        if best is None or timings[grid] < timings[best]:
            best = grid
    runner.grid = best
    return best, timings


class ShanChenBatch:
    """Many Shan-Chen lattices advanced together: one launch per step over the
    tiles of all of them (lbm_sc_batch_run), instead of one launch per
    lattice and step.

    ``ShanChenBatch([stepper, ...])`` takes fused Steppers of Shan-Chen
    lattices (``Stepper(..., boundary_engine_sc=...)``, FP32 storage, the
    same num_components; collision models, sizes, rules and psi objects may
    differ).  ``run(n)`` advances every lattice by n steps and leaves each
    LBField exactly as its own ``Stepper.run(n)`` would: the per-tile step is
    the one-lattice kernel's, so the fields are equal bit for bit.  Each
    lattice keeps its own buffers; reading one (``lb_field.vel.to_numpy()``)
    between runs is allowed."""

    def __init__(self, steppers):
        self.steppers = list(steppers)
        if not self.steppers:
            raise ValueError("a Shan-Chen batch needs at least one lattice")
        nc = self.steppers[0].lbf.nc
        for st in self.steppers:
            if st.sc is None:
                raise ValueError("every lattice of a ShanChenBatch needs a Shan-Chen force")
            if st.storage != H.STORE_F32:
                raise NotImplementedError("Shan-Chen batches store FP32 populations")
            if st.lbf.nc != nc:
                raise ValueError("every lattice of a ShanChenBatch needs the same num_components")
        self.h = None
        self._key = None

    def _handle(self, structs, offsets):
        """The batch handle for these buffers and relative parities (rebuilt
        only when a lattice's buffers or ping-pong phase changed)."""
        key = tuple((tuple(F.f), tuple(F.macro), tuple(F.psi), F.tau, F.code, F.temperature, o)
                    for (L, F), o in zip(structs, offsets))
        n = len(structs)
        lats = (H.Lattice * n)()
        flds = (H.Fields * n)()
        neem = (H.NeemRule * (n * H.MAX_NEEM))()
        for l, ((L, F), st) in enumerate(zip(structs, self.steppers)):
            ctypes.memmove(ctypes.byref(lats[l]), ctypes.byref(L), ctypes.sizeof(H.Lattice))
            ctypes.memmove(ctypes.byref(flds[l]), ctypes.byref(F), ctypes.sizeof(H.Fields))
            for e in range(L.n_neem):
                ctypes.memmove(ctypes.byref(neem[l * H.MAX_NEEM + e]), ctypes.byref(st.neem[e]),
                               ctypes.sizeof(H.NeemRule))
        # parameters (relaxation, forces, rules) are part of the descriptors too
        key = key + (bytes(lats), bytes(neem))
        if self.h is not None and key == self._key:
            return self.h
        self.close()
        off = (ctypes.c_int32 * n)(*offsets)
        h = ctypes.c_void_p()
        H.check(H.lib().lbm_sc_batch_create(lats, flds, neem, off, n, ctypes.byref(h)), "lbm_sc_batch_create")
        self.h, self._key = h, key
        return h

    def tiles(self) -> int:
        return H.lib().lbm_sc_batch_tiles(self.h) if self.h is not None else 0

    def run(self, nsteps: int):
        """Advances every lattice by nsteps steps (asynchronous on the current stream)."""
        nsteps = int(nsteps)
        if nsteps <= 0:
            return
        states = [st.lbf.state() for st in self.steppers]
        if len({s["f_valid"] for s in states}) > 1:
            # some lattices were read back: run the others' pending boundary
            # pass too, so that every first step reads f in place (a readback
            # does not change a lattice's later steps)
            for st in self.steppers:
                st.lbf.materialize()
        begun = [st._begin(nsteps) for st in self.steppers]
        first_local = begun[0][4]
        p0 = begun[0][3]
        offsets = [b[3] ^ p0 for b in begun]
        h = self._handle([(b[0], b[1]) for b in begun], offsets)
        out = ctypes.c_int32(-1)
        H.check(H.lib().lbm_sc_batch_run(h, nsteps, p0, first_local, ctypes.byref(out), stream()),
                "lbm_sc_batch_run")
        for st, s, off, (L, F, tau_mode, _, _) in zip(self.steppers, states, offsets, begun):
            s["f2_idx"] = int(out.value) ^ off      # a launch of two steps flips the buffers once
            st._end(nsteps, L, F, tau_mode)

    def close(self):
        if self.h is not None:
            H.lib().lbm_sc_batch_destroy(self.h)
            self.h = None
            self._key = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
