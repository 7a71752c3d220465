"""ctypes front end of the C++ CPU restatement (cpu_ref.cpp) -- test
infrastructure and CPU baseline only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libcpu_ref.so"
ASAN_LIB = HERE / "_build" / "libcpu_ref_asan.so"   # `make asan`; CPU_REF_LIB selects it


class Params(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nc", C.c_int32), ("coll", C.c_int32),
                ("power_law", C.c_int32), ("nthreads", C.c_int32), ("magic", C.c_float), ("s_bulk", C.c_float),
                ("visc_m", C.c_float), ("visc_nm1", C.c_float), ("tau_min", C.c_float), ("tau_max", C.c_float),
                ("gx", C.c_float), ("gy", C.c_float), ("sc", C.c_int32), ("g_coh", C.c_float * 4),
                ("psi_rho0", C.c_float * 2)]


class BC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("direction", C.c_int32), ("ux", C.c_float), ("uy", C.c_float),
                ("rho", C.c_float), ("nodes", C.c_void_p), ("count", C.c_int32)]


KIND = {"periodic": 1, "inside": 2, "bounceback": 3, "velocity_neem": 4, "pressure_neem": 5, "velocity_bb": 6,
        "open_neumann": 7, "pressure_abc": 8, "velocity_eq": 9}


def build(target: str = "all"):
    subprocess.run(["make", "-s", "-C", str(HERE), target], check=True)
    return ASAN_LIB if target == "asan" else LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = Path(os.environ.get("CPU_REF_LIB", str(LIB)))
        if path == LIB and not LIB.exists():
            build()
        _lib = C.CDLL(str(path))
        _lib.cref_run.restype = C.c_int
        _lib.cref_max_threads.restype = C.c_int
    return _lib


class CpuLattice:
    """State of one lattice in the reference's AoS layout."""

    def __init__(self, NX, NY, nc, coll, group, bcs, neighbor, *, power_law=False, magic=0.25, s_bulk=1.0,
                 visc=(0.0, 0.0), tau_clamp=(0.5000001, 4.0), g=(0.0, 0.0), nthreads=0, sc=0, g_coh=(0.0,) * 4,
                 psi_rho0=(1.0, 1.0)):
        """sc: 0, 1 (ShanChenForceC1) or 2 (ShanChenForceC2) with SC_psi(psi_rho0), g_coh
        the C1 constant ([0]) or the C2 matrix (row-major), on a periodic
        lattice that is all fluid for the force."""
        f32 = np.float32
        self.p = Params(NX, NY, nc, coll, int(power_law), nthreads, f32(magic), f32(s_bulk), f32(visc[0]),
                        f32(visc[1]), f32(tau_clamp[0]), f32(tau_clamp[1]), f32(g[0]), f32(g[1]), int(sc),
                        (C.c_float * 4)(*[f32(v) for v in g_coh]), (C.c_float * 2)(*[f32(v) for v in psi_rho0]))
        self.SCforce = np.zeros((NX, NY, nc, 2), f32)
        self.psi = np.zeros((NX + 2, NY + 2, nc), f32)
        self.f = np.zeros((NX, NY, nc, 9), f32)
        self.f2 = np.zeros_like(self.f)
        self.rho = np.zeros((NX, NY, nc), f32)
        self.total_rho = np.zeros((NX, NY), f32)
        self.vel = np.zeros((NX, NY, 2), f32)
        self.body_force = np.zeros((NX, NY, nc, 2), f32)
        self.shear_rate = np.zeros((NX, NY, nc), f32)
        self.tau = np.zeros((NX, NY, nc), f32)
        self.pressure = np.zeros((NX, NY, nc), f32)
        self.total_pressure = np.zeros((NX, NY), f32)
        self.neighbor = np.ascontiguousarray(neighbor, np.int32)
        self.group = np.ascontiguousarray(group, np.int32)
        self._bc_nodes = [np.ascontiguousarray(b["nodes"], np.int32) for b in bcs]
        arr = (BC * max(1, len(bcs)))()
        for q, b in enumerate(bcs):
            arr[q].kind = KIND[b["kind"]]
            arr[q].direction = int(b.get("direction") or 0)
            arr[q].ux, arr[q].uy = (f32(v) for v in b.get("u", (0.0, 0.0)))
            arr[q].rho = f32(b.get("rho", 0.0))
            arr[q].nodes = self._bc_nodes[q].ctypes.data
            arr[q].count = len(self._bc_nodes[q])
        self._bcs = arr
        self._nbc = len(bcs)

    def run(self, nsteps):
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        rc = lib().cref_run(C.byref(self.p), P(self.f), P(self.f2), P(self.rho), P(self.total_rho), P(self.vel),
                            P(self.body_force), P(self.shear_rate), P(self.tau), P(self.pressure),
                            P(self.total_pressure), P(self.neighbor), P(self.group), C.c_int32(len(self.group)),
                            self._bcs, C.c_int32(self._nbc), C.c_int32(nsteps), P(self.SCforce), P(self.psi))
        if rc != 0:
            raise RuntimeError(f"cref_run failed ({rc})")


def max_threads():
    return int(lib().cref_max_threads())


def from_case(case, nthreads=0):
    """CpuLattice for a case dict built through the drop-in API (duck-typed:
    reads groups, boundary list, collision and fluid-model parameters), with
    f = f2 = feq(u, rho) on the fluid group as init_LBM leaves them."""
    from . import lbm_oracle as O
    lbf, coll = case["lbf"], case["coll"]
    bcs = []
    for bc in case["bce"].boundary_conditions.values():
        if bc.kind == "fluid":
            continue
        d = {"kind": bc.kind, "nodes": bc.group.nodes, "direction": getattr(bc, "direction", 0)}
        if bc.kind in ("velocity_neem", "velocity_bb", "velocity_eq"):
            d["u"] = bc.velocity_value
        if bc.kind in ("pressure_neem", "pressure_abc"):
            d["rho"] = bc.rho_value
        bcs.append(d)
    kind = coll.KIND
    fm = coll.fluid_model
    visc = (fm.m, fm.n - 1.0) if hasattr(fm, "n") else (fm.nu, 0.0)
    s_bulk = 1.0
    if kind == 2:
        tb = np.float32(coll.nu_bulk[0]) * np.float32(3.0) + np.float32(0.5)
        tb = min(max(tb, np.float32(coll.tau_min)), np.float32(coll.tau_max))
        s_bulk = np.float32(1.0) / np.float32(tb)
    group = case["fluid"].group.nodes
    sc = {}
    spec = case.get("sc_spec")
    if spec is not None:
        obj = lbf.sc_field
        psis = [obj.psi] if spec["kind"] == "c1" else [obj.psi1, obj.psi2]
        if (any(type(q).__name__ != "SC_psi" for q in psis) or getattr(obj, "fluid_map", None) is not None
                or any(type(b).__name__ not in ("PeriodicPsi_LR", "PeriodicPsi_BT")
                       for b in case["bsc"].boundary_conditions.values())):
            raise NotImplementedError("the C++ restatement takes SC_psi on periodic all-fluid lattices only")
        g_coh = ([float(spec["g_coh"])] + [0.0] * 3 if spec["kind"] == "c1"
                 else [float(v) for v in np.asarray(spec["g_coh"], np.float64).reshape(-1)])
        sc = dict(sc=1 if spec["kind"] == "c1" else 2, g_coh=g_coh,
                  psi_rho0=[float(v) for v in spec["rho0"]] + [1.0] * (2 - len(spec["rho0"])))
    cl = CpuLattice(lbf.NX, lbf.NY, lbf.nc, kind, group, bcs, lbf.neighbor.arr, power_law=case["power_law"],
                    magic=float(getattr(coll, "magic", [0.25])[0]), s_bulk=float(s_bulk), visc=visc,
                    tau_clamp=(coll.tau_min, coll.tau_max), g=lbf.gravity_force, nthreads=nthreads, **sc)
    # initTau: the collision's per-node state on the group
    st = coll.init_state_values()
    I, J = group[:, 0], group[:, 1]
    for c in range(lbf.nc):
        cl.tau[I, J, c] = st[c]
    cl.rho[...] = case["rho"]
    cl.vel[...] = case["vel"]
    for c in range(lbf.nc):
        fe = O.f_eq(cl.vel[I, J].astype(np.float32), cl.rho[I, J, c].astype(np.float32), np.float32)
        cl.f[I, J, c] = fe
        cl.f2[I, J, c] = fe
    return cl
