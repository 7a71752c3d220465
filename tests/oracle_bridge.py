"""Rebuilds a case made through the drop-in API as an oracle Solver with the
same groups, parameters and initial state (test-only)."""
from __future__ import annotations

import numpy as np

from oracle import lbm_oracle as O


def _r32(x):
    return float(np.float32(x))


def oracle_solver(case, dtype=np.float32, rho=None, vel=None, f=None, psi=None, params32=False):
    """params32: give an FP64 oracle the FP32-rounded parameters the HIP library
    receives (relaxation state as initTau computes it in FP32, boundary
    velocities, gravity, power-law constants), so that it integrates the same
    discrete problem as the HIP FP64 path."""
    lbf = case["lbf"]
    coll = case["coll"]
    g = tuple(_r32(v) for v in lbf.gravity_force) if params32 else lbf.gravity_force
    lat = O.Lattice(lbf.NX, lbf.NY, lbf.nc, dtype=dtype, gravity=g)
    lat.mask[...] = lbf.mask.arr
    lat.neighbor[...] = lbf.neighbor.arr
    group = case["fluid"].group.nodes.astype(np.int64)
    fm = coll.fluid_model
    if type(fm).__name__ == "PowerLawFluid":
        if params32:
            model = O.FluidModel("power", m=_r32(fm.m), n=1.0 + _r32(fm.n - 1.0))
        else:
            model = O.FluidModel("power", m=fm.m, n=fm.n)
    else:
        model = O.FluidModel("newtonian", nu=fm.nu)
    kind = {0: "bgk", 1: "trt", 2: "mrt"}[coll.KIND]
    magic = list(getattr(coll, "magic", [])) or None
    tmin, tmax = coll.tau_min, coll.tau_max
    if params32:
        magic = [_r32(v) for v in magic] if magic else None
        tmin, tmax = _r32(tmin), _r32(tmax)
    args = dict(magic=magic, nu_bulk=list(getattr(coll, "nu_bulk", [])) or None, tau_min=tmin, tau_max=tmax)
    if type(coll).__name__ == "HuangMRTCollision":
        args["huang"] = dict(k1=[float(v) for v in coll.k1], epslion=list(coll.epslion),
                             g=[float(v) for v in coll.g_coh])
    oc = O.Collision(kind, group, model, list(coll.nu), **args)
    oc.init_tau(lat)
    if params32 and np.dtype(dtype) != np.float32:
        lat32 = O.Lattice(lbf.NX, lbf.NY, lbf.nc, dtype=np.float32)
        oc32 = O.Collision(kind, group, model, list(coll.nu), **args)
        oc32.init_tau(lat32)
        for attr in ("tau", "tau_antisym", "diag"):
            if getattr(oc32, attr, None) is not None:
                setattr(oc, attr, getattr(oc32, attr).astype(dtype))
    bcs = []
    for name, bc in case["bce"].boundary_conditions.items():
        k = bc.kind
        val = None
        if k in ("velocity_neem", "velocity_bb", "velocity_eq"):
            val = tuple(_r32(v) for v in bc.velocity_value) if params32 else bc.velocity_value
        elif k in ("pressure_neem", "pressure_abc"):
            val = bc.rho_value
        bcs.append(O.BC(k, bc.group.nodes.astype(np.int64), getattr(bc, "direction", None), val))
    dt = np.dtype(dtype)
    lat.rho[...] = (case["rho"] if rho is None else rho).astype(dt)
    lat.vel[...] = (case["vel"] if vel is None else vel).astype(dt)
    sc = None
    if case.get("sc_spec") is not None:
        s = case["sc_spec"]
        lat.psi = np.zeros((lbf.NX + 2, lbf.NY + 2, lbf.nc), dt)
        psi = case.get("psi") if psi is None else psi
        if psi is not None:
            lat.psi[...] = psi.astype(dt)
        lat.T[...] = 1.0
        obj = lbf.sc_field
        psis = [obj.psi] * lbf.nc if s["kind"] == "c1" else [obj.psi1, obj.psi2]
        eos = [None if type(q).__name__ == "SC_psi" else
               O.EOS({"vdW_psi": "vdw", "RK_psi": "rk", "PR_psi": "pr", "CS_psi": "cs"}[type(q).__name__],
                     q.a, q.b, q.R, q.G, q.omega, q.temperature_cr or 1.0) for q in psis]
        sc = O.ShanChen(s["kind"], group, s["g_coh"], s["rho0"], g_adh=getattr(obj, "g_adh", None),
                        fluid=getattr(obj, "fluid_map", None), eos=eos if any(e is not None for e in eos) else None)
    psi_bcs = None
    if case.get("bsc") is not None:
        psi_bcs = []
        kinds = {"OpenNeumannPsi": "neumann", "OpenConvective1orderPsi": "convective1",
                 "OpenConvective2orderPsi": "convective2", "OpenExtrapolationPsi": "extrapolation", "FixedPsi": "fixed"}
        for b in case["bsc"].boundary_conditions.values():
            name = type(b).__name__
            if name in ("PeriodicPsi_LR", "PeriodicPsi_BT"):
                if name == "PeriodicPsi_LR":
                    psi_bcs.append("periodic")
                continue
            psi_bcs.append(O.PsiBC(kinds[name], np.asarray(b.group.nodes, np.int64), getattr(b, "direction", None),
                                   getattr(b, "psi_value", 0.0)))
    if f is None:
        O.init_lbm(lat, group)
    else:
        lat.f[...] = f.astype(dt)
        lat.f2[...] = f.astype(dt)
    solver = O.Solver(lat, group, oc, bcs, power_law=case.get("power_law", False), sc=sc, psi_bcs=psi_bcs)
    return solver
