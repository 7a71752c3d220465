"""Workload builders: the benchmark and test configurations set up through the
drop-in API, the way the reference's case scripts do (cases/*.py: mask ->
groups -> boundary list -> collision -> init_hydro -> init_LBM).

Every builder returns a dict with the LBField, engines, collision and the
Stepper arguments, plus a plain description (``spec``) of what was built so a
checker can rebuild the same case independently.
"""
from __future__ import annotations

import math

import numpy as np

from .BoundaryCondition import (BounceBackWall, BoundaryClassifier, BoundaryEngine, BoundarySpec, FluidBoundary,
                                InsideBoundary, PeriodicAllBoundary, PressureNEEM, VelocityNEEM)
from .CollisionEngine import BGKCollision, MRTCollision, TRTCollision
from .FluidModels import NewtonianFluid, PowerLawFluid
from .lbm import LBField, MacroscopicEngine

C = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])


def solid_neighbour_map(mask: np.ndarray, y_walls: bool = False) -> np.ndarray:
    """Fluid nodes with a solid neighbour x - c_k (and, with y_walls, a
    neighbour beyond the bottom/top rows): the bounce-back wall group."""
    from ._codes import shifted
    NX, NY = mask.shape
    solid = mask == -1
    out = np.zeros((NX, NY), bool)
    for k in range(1, 9):
        out |= shifted(solid, int(C[k, 0]), int(C[k, 1]), False)
        if y_walls and C[k, 1] != 0:
            edge = np.zeros((NX, NY), bool)
            edge[:, 0 if C[k, 1] > 0 else NY - 1] = True
            out |= edge
    return out & (mask == 1)


def make_collision(kind, group, NX, NY, nc, nu, fluid_model=None, magic=0.25, nu_bulk=None):
    params = {'group': group, 'fluid_model': fluid_model or NewtonianFluid(nu[0]), 'NX': NX, 'NY': NY,
              'num_components': nc, 'shearviscosity': list(nu), 'magic': [magic] * nc,
              'bulkviscosity': list(nu_bulk if nu_bulk is not None else nu)}
    return {"bgk": BGKCollision, "trt": TRTCollision, "mrt": MRTCollision}[kind](params)


def _finish(lbf, bce, fluid_bc, coll, rho0=1.0, u0=(0.0, 0.0)):
    """init_hydro (rho, u on the fluid group) + init_LBM."""
    macro = MacroscopicEngine(fluid_bc.group)
    nodes = fluid_bc.group.nodes
    rho = np.zeros((lbf.NX, lbf.NY, lbf.nc), np.float32)
    vel = np.zeros((lbf.NX, lbf.NY, 2), np.float32)
    rho[nodes[:, 0], nodes[:, 1], :] = rho0 / lbf.nc
    vel[nodes[:, 0], nodes[:, 1], :] = u0
    return macro, rho, vel


def cylinder(NX=400, NY=100, cx=80, cy=50, r=10, tau=0.6, u_in=0.05, rho_out=1.0, collision="bgk", magic=0.25,
             tau_bulk=None, name="cylinder", inlet="neem", outlet="neem"):
    """Cylinder in a channel: bounce-back top/bottom rows and cylinder,
    PressureNEEM outlet (d=1) on 0 < j < NY-1, f = w at start.  The inlet
    (i = 0, 0 < j < NY-1, d = 3) is

      "neem"  VelocityNEEM (BoundaryCondition.py:149-166), the plain-method
              configuration as stated.  With the reference's NEEM semantics
              (post-stream f of the neighbour against its rho/u of the step
              before) this inlet next to the bounce-back rows diverges after a
              few hundred steps (in FP64 as in FP32);
      "bb"    VelocityBB over a bounce-back inlet column, the inlet every
              reference case uses (cases/Circular cylinder.py:91-118,
              cases/contraction flow.py:110-138); stable;
      "eq"    VelocityEquilibrium (BoundaryCondition.py:132-145): the node's
              populations replaced by feq(u_in, rho_w).
    outlet="abc" replaces the PressureNEEM outlet with PressureABC over a
    bounced outlet column (BoundaryCondition.py:177-205)."""
    if inlet not in ("neem", "bb", "eq"):
        raise ValueError("inlet must be 'neem', 'bb' or 'eq'")
    lbf = LBField(name, NX, NY, 1)
    bce = BoundaryEngine()
    clf = BoundaryClassifier((NX, NY))
    bce.Mask_cricle_identify(lbf, cx, cy, r)
    m = lbf.mask.arr
    fluid = FluidBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    fluid.precompute(clf)
    bce.add_boundary_condition("fluid", fluid)
    periodic = PeriodicAllBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    periodic.precompute(clf)
    bce.add_boundary_condition("periodic", periodic)
    wmap = solid_neighbour_map(m, y_walls=True)
    if inlet == "bb":
        wmap[0, :] |= m[0, :] == 1
    if outlet == "abc":
        wmap[NX - 1, :] |= m[NX - 1, :] == 1
    wall = BounceBackWall(BoundarySpec(lambda I, J: wmap[I, J]))
    wall.precompute(clf)
    bce.add_boundary_condition("wall", wall)
    inlet_spec = BoundarySpec(lambda I, J: (m[I, J] == 1) & (I == 0) & (J > 0) & (J < NY - 1))
    if inlet == "bb":
        from .BoundaryCondition import VelocityBB
        inlet_bc = VelocityBB(inlet_spec, (u_in, 0.0), 3)
    elif inlet == "eq":
        from .BoundaryCondition import VelocityEquilibrium
        inlet_bc = VelocityEquilibrium(inlet_spec, (u_in, 0.0), 3)
    else:
        inlet_bc = VelocityNEEM(inlet_spec, (u_in, 0.0), 3)
    inlet_bc.precompute(clf)
    bce.add_boundary_condition("inlet", inlet_bc)
    outlet_spec = BoundarySpec(lambda I, J: (m[I, J] == 1) & (I == NX - 1) & (J > 0) & (J < NY - 1))
    if outlet == "abc":
        from .BoundaryCondition import PressureABC
        outlet_bc = PressureABC(outlet_spec, rho_out, 1)
    else:
        outlet_bc = PressureNEEM(outlet_spec, rho_out, 1)
    outlet = outlet_bc
    outlet.precompute(clf)
    bce.add_boundary_condition("outlet", outlet)
    lbf.neighbor_classify()
    nu = (tau - 0.5) / 3.0
    nu_b = None if tau_bulk is None else [(tau_bulk - 0.5) / 3.0]
    coll = make_collision(collision, fluid.group, NX, NY, 1, [nu], magic=magic, nu_bulk=nu_b)
    macro, rho, vel = _finish(lbf, bce, fluid, coll)
    return dict(lbf=lbf, bce=bce, macro=macro, coll=coll, fluid=fluid, rho=rho, vel=vel, power_law=False,
                spec=dict(kind="cylinder", NX=NX, NY=NY, tau=tau, u_in=u_in, collision=collision, magic=magic,
                          inlet=inlet, outlet=outlet_bc.kind))


def place_disks(NX, NY, seed, rmin=8.0, rmax=32.0, solid_fraction=0.3, margin=2.0, max_tries=10_000_000):
    """Random sequential placement of non-overlapping disks (seeded) until the
    disk area reaches solid_fraction of the lattice; disks keep `margin` nodes
    from the edges.  Returns an (m, 3) array of (x, y, r)."""
    rng = np.random.default_rng(seed)
    target = solid_fraction * NX * NY
    cell = 2 * rmax + 2
    gx, gy = int(math.ceil(NX / cell)), int(math.ceil(NY / cell))
    grid = [[] for _ in range(gx * gy)]
    disks = []
    area = 0.0
    tries = 0
    while area < target and tries < max_tries:
        tries += 1
        r = rng.uniform(rmin, rmax)
        x = rng.uniform(r + margin, NX - 1 - r - margin)
        y = rng.uniform(r + margin, NY - 1 - r - margin)
        cx_, cy_ = int(x // cell), int(y // cell)
        ok = True
        for ax in range(max(0, cx_ - 1), min(gx, cx_ + 2)):
            for ay in range(max(0, cy_ - 1), min(gy, cy_ + 2)):
                for (x2, y2, r2) in grid[ax * gy + ay]:
                    if (x - x2) ** 2 + (y - y2) ** 2 < (r + r2 + 1.0) ** 2:
                        ok = False
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid[cx_ * gy + cy_].append((x, y, r))
            disks.append((x, y, r))
            area += math.pi * r * r
    return np.array(disks, np.float64).reshape(-1, 3)


def rasterize_disks(NX, NY, disks) -> np.ndarray:
    """mask (1 fluid, -1 solid) with Mask_cricle_identify's f32 test per disk."""
    mask = np.ones((NX, NY), np.int32)
    f32 = np.float32
    for x, y, r in disks:
        i0, i1 = max(0, int(x - r) - 1), min(NX, int(x + r) + 2)
        j0, j1 = max(0, int(y - r) - 1), min(NY, int(y + r) + 2)
        I, J = np.meshgrid(np.arange(i0, i1), np.arange(j0, j1), indexing="ij")
        dx = I.astype(f32) - f32(x)
        dy = J.astype(f32) - f32(y)
        hit = dx * dx + dy * dy < f32(r) * f32(r)
        mask[i0:i1, j0:j1][hit] = -1
    return mask


def porous(NX=8192, NY=8192, seed=1, collision="mrt", power_n=0.7, power_m=0.01, tau_min=0.51, tau_max=2.0,
           g=(1e-6, 0.0), nu0=1.0 / 6.0, nu_bulk=1.0 / 6.0, magic=0.25, solid_fraction=0.3, rmin=8.0, rmax=32.0,
           name="porous", walls=True):
    """Periodic lattice with seeded random non-overlapping solid disks, body
    force g, power-law rheology with tau clamped to [tau_min, tau_max].
    walls=False leaves out the BounceBackWall entry: fluid nodes next to a
    disk then stream the disk's unchanged f2 (what the reference does with
    such a list)."""
    lbf = LBField(name, NX, NY, 1)
    disks = place_disks(NX, NY, seed, rmin, rmax, solid_fraction)
    lbf.mask.arr[...] = rasterize_disks(NX, NY, disks)
    lbf._geometry_changed()
    bce = BoundaryEngine()
    clf = BoundaryClassifier((NX, NY))
    m = lbf.mask.arr
    fluid = FluidBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    fluid.precompute(clf)
    bce.add_boundary_condition("fluid", fluid)
    periodic = PeriodicAllBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    periodic.precompute(clf)
    bce.add_boundary_condition("periodic", periodic)
    if walls:
        wmap = solid_neighbour_map(m)
        wall = BounceBackWall(BoundarySpec(lambda I, J: wmap[I, J]))
        wall.precompute(clf)
        bce.add_boundary_condition("wall", wall)
    lbf.neighbor_classify()
    lbf.set_gravity(g)
    coll = make_collision(collision, fluid.group, NX, NY, 1, [nu0], PowerLawFluid(power_m, power_n), magic=magic,
                          nu_bulk=[nu_bulk])
    coll.tau_min, coll.tau_max = tau_min, tau_max
    macro, rho, vel = _finish(lbf, bce, fluid, coll)
    return dict(lbf=lbf, bce=bce, macro=macro, coll=coll, fluid=fluid, rho=rho, vel=vel, power_law=True,
                disks=disks,
                spec=dict(kind="porous", NX=NX, NY=NY, seed=seed, collision=collision, power_n=power_n,
                          power_m=power_m, tau_min=tau_min, tau_max=tau_max, g=g, nu0=nu0, nu_bulk=nu_bulk,
                          magic=magic, n_disks=len(disks)))


def initialise(case):
    """init_hydro + init_LBM on the device (needs a GPU)."""
    lbf = case["lbf"]
    lbf.rho.from_numpy(case["rho"])
    lbf.vel.from_numpy(case["vel"])
    lbf.init_LBM(case["coll"], case["fluid"].group)
    if case.get("psi") is not None:
        lbf.sc_field.psi_field.from_numpy(case["psi"])
        lbf.T.fill(1.0)
    return case


def stepper(case):
    """The fused Stepper for a built case."""
    from .stepper import Stepper
    return Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=case["power_law"],
                   boundary_engine_sc=case.get("bsc"))


def shan_chen(NX=128, NY=128, nc=1, G=-5.0, gcross=0.9, seed=0, collision="mrt", nu=1.0 / 6.0, rho_mean=None,
              amp=None, name="shan_chen", disk=None, gadh=None, psi_wall=None, eos=None, huang=None):
    """Periodic Shan-Chen lattice.  nc=1: ShanChenForceC1 with cohesion G, rho =
    rho_mean + amp*U(0,1) (cases/Bubble or two-phase separation.py:174);
    nc=2: ShanChenForceC2 with G = [[0, gcross], [gcross, 0]] and each rho_c =
    rho_mean (default 1) + amp*U(-1, 1), seeded per component.  psi = SC_psi(rho0=1).
    disk=(cx, cy, r): a solid disk (Mask_cricle_identify) with a
    BounceBackWall group around it; psi stays 0 on its nodes (update_psi
    writes the group only) unless psi_wall gives a FixedPsi over them
    (SCBoundaryCondition.py:110-123, after the periodic pair).  gadh=(a0, a1)
    (nc=2): ShanChenForceC2 adhesion at the disk, fluid_strategy = mask == 1
    (periodic on the halo; ExternalForce.py:105-116).  eos=(kind, params)
    with kind "vdw" | "rk" | "pr" | "cs": the cubic-EOS pseudopotential of
    EOS.py in place of SC_psi, at T = 1.  huang=(k1, epslion): the MRT
    collision with Huang's source (HuangMRTCollision, the same k1 and
    epslion for every component, G = the component's own coupling)."""
    from .EOS import CS_psi, PR_psi, RK_psi, SC_psi, vdW_psi
    from .ExternalForce import ShanChenForceC1, ShanChenForceC2
    from .SCBoundaryCondition import BoundaryEngineSC, FixedPsi, PeriodicPsi_BT, PeriodicPsi_LR
    import contextlib
    import io
    if eos is not None:
        with contextlib.redirect_stdout(io.StringIO()):
            psi_obj = {"vdw": vdW_psi, "rk": RK_psi, "pr": PR_psi, "cs": CS_psi}[eos[0]](dict(eos[1]))
    else:
        psi_obj = SC_psi(1.0)
    lbf = LBField(name, NX, NY, nc)
    bce = BoundaryEngine()
    clf = BoundaryClassifier((NX, NY))
    if disk is not None:
        bce.Mask_cricle_identify(lbf, *disk)
    m = lbf.mask.arr
    fluid = FluidBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    fluid.precompute(clf)
    bce.add_boundary_condition("fluid", fluid)
    stream_bc = PeriodicAllBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    stream_bc.precompute(clf)
    bce.add_boundary_condition("stream", stream_bc)
    if disk is not None:
        wmap = solid_neighbour_map(m)
        wall = BounceBackWall(BoundarySpec(lambda I, J: wmap[I, J]))
        wall.precompute(clf)
        bce.add_boundary_condition("wall", wall)
    lbf.neighbor_classify()
    coll = make_collision(collision, fluid.group, NX, NY, nc, [nu] * nc, nu_bulk=[nu] * nc)
    if huang is not None:
        from .CollisionEngine import HuangMRTCollision
        if collision != "mrt":
            raise ValueError("HuangMRTCollision is an MRT collision")
        Gself = np.diag([G] * nc) if nc == 1 else np.array([[G, gcross], [gcross, G]])
        coll = HuangMRTCollision({'group': fluid.group, 'fluid_model': coll.fluid_model, 'NX': NX, 'NY': NY,
                                  'num_components': nc, 'shearviscosity': [nu] * nc, 'bulkviscosity': [nu] * nc,
                                  'k1': [huang[0]] * nc, 'epslion': [huang[1]] * nc, 'g_coh': Gself})
    rng = np.random.default_rng(seed)
    rho = np.zeros((NX, NY, nc), np.float32)
    if nc == 1:
        rho_mean = math.log(2.0) if rho_mean is None else rho_mean
        amp = 0.1 if amp is None else amp
        rho[..., 0] = rho_mean + amp * rng.random((NX, NY))
        sc = ShanChenForceC1({'lb_field': lbf, 'g_coh': G, 'group': fluid.group, 'psi': psi_obj})
        g_spec = G
    else:
        amp = 0.01 if amp is None else amp
        rho_mean = 1.0 if rho_mean is None else rho_mean
        for c in range(nc):
            rho[..., c] = rho_mean + amp * rng.uniform(-1.0, 1.0, (NX, NY))
        Gm = np.array([[0.0, gcross], [gcross, 0.0]])
        strategy = None if gadh is None else (lambda I, J: m[I % NX, J % NY] == 1)
        sc = ShanChenForceC2({'lb_field': lbf, 'g_coh': Gm, 'gadh': list(gadh) if gadh else [0.0, 0.0],
                              'group': fluid.group, 'psi_list': [psi_obj, psi_obj], 'fluid_strategy': strategy})
        g_spec = Gm
    lbf.sc_field = sc
    bsc = BoundaryEngineSC()
    bsc.add_boundary_condition("periodicLR", PeriodicPsi_LR())
    bsc.add_boundary_condition("periodicBT", PeriodicPsi_BT())
    if psi_wall is not None:
        clf_psi = BoundaryClassifier((NX + 2, NY + 2), offset=(-1, -1))
        wall_psi = FixedPsi(BoundarySpec(lambda I, J: (I >= 0) & (I < NX) & (J >= 0) & (J < NY) &
                                         (m[np.clip(I, 0, NX - 1), np.clip(J, 0, NY - 1)] == -1)), psi_wall)
        wall_psi.precompute(clf_psi)
        bsc.add_boundary_condition("wall_psi", wall_psi)
    psi = np.zeros((NX + 2, NY + 2, nc), np.float32)
    psi[1:-1, 1:-1, :] = psi_obj.get_psi(rho, np.float32(1.0))
    psi[1:-1, 1:-1, :][m == -1] = 0.0
    psi = np.pad(psi[1:-1, 1:-1, :], ((1, 1), (1, 1), (0, 0)), mode="wrap")
    macro = MacroscopicEngine(fluid.group)
    vel = np.zeros((NX, NY, 2), np.float32)
    return dict(lbf=lbf, bce=bce, macro=macro, coll=coll, fluid=fluid, rho=rho, vel=vel, power_law=False,
                psi=psi, bsc=bsc, sc_spec=dict(kind="c1" if nc == 1 else "c2", g_coh=g_spec, rho0=(1.0,) * nc),
                spec=dict(kind="shan_chen", NX=NX, NY=NY, nc=nc, G=G, gcross=gcross, seed=seed,
                          collision=collision, nu=nu, disk=disk, gadh=gadh, psi_wall=psi_wall, eos=eos,
                          huang=huang))



def reference_inside_map(mask: np.ndarray) -> np.ndarray:
    """The InsideBoundary group of cases/contraction flow.py:96-108 and
    cases/Power Law Fluid.py:75-87: fluid nodes with at least one neighbour
    x - c_k strictly inside the lattice (0 < i < NX-1, 0 < j < NY-1) that
    is fluid."""
    from ._codes import shifted
    NX, NY = mask.shape
    interior_fluid = np.zeros((NX, NY), bool)
    interior_fluid[1:NX - 1, 1:NY - 1] = mask[1:NX - 1, 1:NY - 1] == 1
    out = np.zeros((NX, NY), bool)
    for k in range(9):
        out |= shifted(interior_fluid, int(C[k, 0]), int(C[k, 1]), False)
    return out & (mask == 1)


def edge_or_solid_map(mask: np.ndarray) -> np.ndarray:
    """The BounceBackWall group of cases/Power Law Fluid.py:89-99 (and of
    cases/contraction flow.py:111-123, whose predicate reads mask[-1, j] and
    mask[NX, j] out of range at the inlet and outlet columns; here a
    neighbour beyond any lattice edge counts as a wall): fluid nodes with a
    neighbour x - c_k outside the lattice or solid."""
    NX, NY = mask.shape
    out = solid_neighbour_map(mask, y_walls=True)
    edge = np.zeros((NX, NY), bool)
    edge[[0, NX - 1], :] = True
    return out | (edge & (mask == 1))


def contraction(W=512, H=128, power_n=1.0, u_in=0.02, nu_ref=0.1, collision="trt", magic=0.25, tau_min=0.51,
                tau_max=2.0, name="contraction", inlet="bb", stream="periodic", outlet="neem"):
    """4:1 planar contraction scaled from cases/contraction flow.py:16-33,80-81
    (a 40 x 8 channel with solid blocks over x in (25, 35) leaving the gap
    y in (3, 5)): VelocityBB inlet over a bounce-back inlet column (d=3,
    u_in; as the reference case, :110-138; inlet="neem" gives a VelocityNEEM
    inlet instead, which diverges like the cylinder's), PressureNEEM outlet
    (d=1, rho=1), bounce-back walls, power-law fluid with m chosen so that the
    apparent viscosity is nu_ref at the inlet shear rate u_in / (H/2), tau
    clamped to [tau_min, tau_max].  stream="inside" keeps the reference's
    boundary list (cases/contraction flow.py:96-149): InsideBoundary over its
    interior-neighbour group (table stream), then BounceBackWall over the
    fluid nodes next to a solid or beyond a lattice edge, VelocityBB,
    PressureNEEM; every -1 table entry the table stream would read is then
    bounced.  stream="periodic" streams with PeriodicAllBoundary instead and
    bounces the inlet column: the same fields.  outlet="abc" replaces the
    PressureNEEM outlet with PressureABC (BoundaryCondition.py:177-205,
    anti-bounce-back over a bounced outlet column), which imposes rho = 1
    there; PressureNEEM extrapolates the outlet density from the node itself,
    so the channel's mass drifts and the flow never becomes steady."""
    NX, NY = int(W), int(H)
    lbf = LBField(name, NX, NY, 1)
    bce = BoundaryEngine()
    clf = BoundaryClassifier((NX, NY))
    sx, sy = NX / 40.0, NY / 8.0          # lattice nodes per reference length unit
    bce.Mask_rectangle_identify(lbf, 25 * sx - 0.5, 35 * sx - 0.5, 5 * sy - 0.5, 8 * sy - 0.5)
    bce.Mask_rectangle_identify(lbf, 25 * sx - 0.5, 35 * sx - 0.5, 0 * sy - 0.5, 3 * sy - 0.5)
    m = lbf.mask.arr
    fluid = FluidBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    fluid.precompute(clf)
    bce.add_boundary_condition("fluid", fluid)
    if stream == "inside":
        imap = reference_inside_map(m)
        inside = InsideBoundary(BoundarySpec(lambda I, J: imap[I, J]))
        inside.precompute(clf)
        bce.add_boundary_condition("inside", inside)
        wmap = edge_or_solid_map(m)
    else:
        periodic = PeriodicAllBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
        periodic.precompute(clf)
        bce.add_boundary_condition("periodic", periodic)
        wmap = solid_neighbour_map(m, y_walls=True)
        if inlet == "bb":
            wmap[0, :] |= m[0, :] == 1
        if outlet == "abc":
            wmap[NX - 1, :] |= m[NX - 1, :] == 1
    wall = BounceBackWall(BoundarySpec(lambda I, J: wmap[I, J]))
    wall.precompute(clf)
    bce.add_boundary_condition("wall", wall)
    inlet_spec = BoundarySpec(lambda I, J: (m[I, J] == 1) & (I == 0) & (J > 0) & (J < NY - 1))
    if inlet == "bb":
        from .BoundaryCondition import VelocityBB
        inlet_bc = VelocityBB(inlet_spec, (u_in, 0.0), 3)
    else:
        inlet_bc = VelocityNEEM(inlet_spec, (u_in, 0.0), 3)
    inlet_bc.precompute(clf)
    bce.add_boundary_condition("inlet", inlet_bc)
    outlet_spec = BoundarySpec(lambda I, J: (m[I, J] == 1) & (I == NX - 1) & (J > 0) & (J < NY - 1))
    if outlet == "abc":
        from .BoundaryCondition import PressureABC
        outlet = PressureABC(outlet_spec, 1.0, 1)
    else:
        outlet = PressureNEEM(outlet_spec, 1.0, 1)
    outlet.precompute(clf)
    bce.add_boundary_condition("outlet", outlet)
    lbf.neighbor_classify()
    gamma_ref = u_in / (NY / 2.0)
    power_m = nu_ref * gamma_ref ** (1.0 - power_n)
    coll = make_collision(collision, fluid.group, NX, NY, 1, [nu_ref], PowerLawFluid(power_m, power_n), magic=magic,
                          nu_bulk=[nu_ref])
    coll.tau_min, coll.tau_max = tau_min, tau_max
    macro, rho, vel = _finish(lbf, bce, fluid, coll)
    return dict(lbf=lbf, bce=bce, macro=macro, coll=coll, fluid=fluid, rho=rho, vel=vel, power_law=True,
                spec=dict(kind="contraction", NX=NX, NY=NY, power_n=power_n, power_m=power_m, u_in=u_in,
                          nu_ref=nu_ref, collision=collision, tau_min=tau_min, tau_max=tau_max, inlet=inlet,
                          stream=stream, outlet=outlet if isinstance(outlet, str) else outlet.kind))


# The serving workload's request: a 4:1 power-law contraction whose free
# parameters are chosen so that small requests reach a steady state within
# the 50k-step budget: an anti-bounce-back pressure outlet (PressureABC) holds
# the outlet density (with PressureNEEM the channel's mass drifts for good),
# u_in = 0.02 keeps the FP32 rounding jitter of the velocity (about 2e-8)
# below the 1e-6 relative residual, and nu_ref = 0.4 with tau >= 0.55 keeps
# the shear-thinning corners of the largest requests stable (with nu_ref =
# 0.3 the 1840 x 455 and 2048 x 512 requests at n = 0.5 diverge).
SERVING_CONTRACTION = dict(u_in=0.02, nu_ref=0.4, collision="trt", magic=0.25, tau_min=0.55, tau_max=2.0,
                           outlet="abc")


def serving_contraction(W, H, power_n, name="request"):
    """One request of the serving workload (SERVING_CONTRACTION)."""
    return contraction(W=W, H=H, power_n=power_n, name=name, **SERVING_CONTRACTION)


def cavity(N=256, power_n=1.0, Re=10.0, u_lid=5e-3, collision="mrt", magic=0.25, name="cavity"):
    """Power-law lid-driven cavity of cases/Power Law Fluid.py:14-141: an
    N x N lattice without solids, InsideBoundary over the interior-neighbour
    group, BounceBackWall over the nodes next to a lattice edge, a VelocityBB
    lid (direction 2, u = (u_lid, 0)) over the top row j = N-1; power-law
    consistency K = u_lid^(2-n) (N-1)^n / Re (:40-41, with the reference's
    lattice velocity Ma cs = u_lid), shearviscosity = bulkviscosity = [K],
    magic 1/4, rho = 1, u = 0, default tau clamps."""
    from .BoundaryCondition import VelocityBB
    NX = NY = int(N)
    lbf = LBField(name, NX, NY, 1)
    bce = BoundaryEngine()
    clf = BoundaryClassifier((NX, NY))
    m = lbf.mask.arr
    fluid = FluidBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    fluid.precompute(clf)
    bce.add_boundary_condition("fluid", fluid)
    imap = reference_inside_map(m)
    inside = InsideBoundary(BoundarySpec(lambda I, J: imap[I, J]))
    inside.precompute(clf)
    bce.add_boundary_condition("inside", inside)
    wmap = edge_or_solid_map(m)
    wall = BounceBackWall(BoundarySpec(lambda I, J: wmap[I, J]))
    wall.precompute(clf)
    bce.add_boundary_condition("wall", wall)
    lid = VelocityBB(BoundarySpec(lambda I, J: (J == NY - 1) & (I >= 0)), (u_lid, 0.0), 2)
    lid.precompute(clf)
    bce.add_boundary_condition("inlet", lid)
    lbf.neighbor_classify()
    power_K = u_lid ** (2.0 - power_n) * (NX - 1.0) ** power_n / Re
    coll = make_collision(collision, fluid.group, NX, NY, 1, [power_K], PowerLawFluid(power_K, power_n), magic=magic,
                          nu_bulk=[power_K])
    macro, rho, vel = _finish(lbf, bce, fluid, coll)
    return dict(lbf=lbf, bce=bce, macro=macro, coll=coll, fluid=fluid, rho=rho, vel=vel, power_law=True,
                spec=dict(kind="cavity", NX=NX, NY=NY, power_n=power_n, power_K=power_K, Re=Re, u_lid=u_lid,
                          collision=collision))


def cylinder_reference_bcs(NX=600, NY=160, tau=0.62, u_in=0.0173, collision="mrt", tau_bulk=None,
                           name="cylinder_ref", outlet="neumann"):
    """The boundary list of cases/Circular cylinder.py:72-133 with bounds-checked
    predicates: cylinder at (3/8 NX - 1/2, NY/2 - 1/2), radius NY/10 - 1/2;
    periodic streaming; bounce-back on the first/last columns and around the
    cylinder; VelocityBB inlet (d=3) on column 0; OpenNeumann outlet (d=1) on
    the last column (outlet="convective1" | "extrapolation": the
    OpenConvective1order / OpenExtrapolation outlet of :337-413 there
    instead).  (The reference's wall predicate reads the mask out of
    range at the top/bottom rows; here those reads count as fluid, i.e. the
    channel is periodic in y as the streaming is.)"""
    from .BoundaryCondition import OpenNeumann, VelocityBB
    lbf = LBField(name, NX, NY, 1)
    bce = BoundaryEngine()
    clf = BoundaryClassifier((NX, NY))
    bce.Mask_cricle_identify(lbf, 3 * NX / 8 - 0.5, NY / 2 - 0.5, NY / 10 - 0.5)
    m = lbf.mask.arr
    fluid = FluidBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    fluid.precompute(clf)
    bce.add_boundary_condition("fluid", fluid)
    periodic = PeriodicAllBoundary(BoundarySpec(lambda I, J: m[I, J] == 1))
    periodic.precompute(clf)
    bce.add_boundary_condition("periodic", periodic)
    wmap = solid_neighbour_map(m)
    wmap[0, :] |= m[0, :] == 1
    wmap[-1, :] |= m[-1, :] == 1
    wall = BounceBackWall(BoundarySpec(lambda I, J: wmap[I, J]))
    wall.precompute(clf)
    bce.add_boundary_condition("wall", wall)
    inlet = VelocityBB(BoundarySpec(lambda I, J: (m[I, J] == 1) & (I == 0)), (u_in, 0.0), 3)
    inlet.precompute(clf)
    bce.add_boundary_condition("inlet", inlet)
    from .BoundaryCondition import OpenConvective1order, OpenExtrapolation
    out_cls = {"neumann": OpenNeumann, "convective1": OpenConvective1order,
               "extrapolation": OpenExtrapolation}[outlet]
    outlet = out_cls(BoundarySpec(lambda I, J: (m[I, J] == 1) & (I == NX - 1)), 1)
    outlet.precompute(clf)
    bce.add_boundary_condition("outlet", outlet)
    lbf.neighbor_classify()
    nu = (tau - 0.5) / 3.0
    nu_b = None if tau_bulk is None else [(tau_bulk - 0.5) / 3.0]
    coll = make_collision(collision, fluid.group, NX, NY, 1, [nu], nu_bulk=nu_b)
    macro, rho, vel = _finish(lbf, bce, fluid, coll, u0=(u_in, 0.0))
    return dict(lbf=lbf, bce=bce, macro=macro, coll=coll, fluid=fluid, rho=rho, vel=vel, power_law=False,
                spec=dict(kind="cylinder_ref", NX=NX, NY=NY, tau=tau, u_in=u_in, collision=collision))


def ib_particles(NX=256, NY=128, n_particles=6, radius=(6.0, 10.0), u_particle=0.01, tau=0.7, collision="bgk",
                 g=(1e-6, 0.0), seed=0, max_num_ib=None, name="ib_particles"):
    """Immersed-boundary particles in a periodic channel: an all-fluid lattice
    with periodic streaming and body force g, and ``n_particles`` circular
    particles (ibm.py Sphere) of seeded radii in ``radius`` and velocities of
    size ``u_particle``, the first one crossing the left edge of the lattice.
    The particles couple to the fluid by multi-direct forcing (IBEngine),
    called between MacroscopicEngine.velocity and the collision."""
    from .ibm import SphereIB
    lbf = LBField(name, NX, NY, 1)
    bce = BoundaryEngine()
    clf = BoundaryClassifier((NX, NY))
    fluid = FluidBoundary(BoundarySpec(lambda I, J: np.ones(I.shape, bool)))
    fluid.precompute(clf)
    bce.add_boundary_condition("fluid", fluid)
    periodic = PeriodicAllBoundary(BoundarySpec(lambda I, J: np.ones(I.shape, bool)))
    periodic.precompute(clf)
    bce.add_boundary_condition("periodic", periodic)
    lbf.neighbor_classify()
    lbf.set_gravity(g)
    nu = (tau - 0.5) / 3.0
    coll = make_collision(collision, fluid.group, NX, NY, 1, [nu])
    macro, rho, vel = _finish(lbf, bce, fluid, coll)
    rng = np.random.default_rng(seed)
    r = rng.uniform(radius[0], radius[1], n_particles).astype(np.float32)
    pos = np.stack([rng.uniform(2.5 * r, NX - 1 - 2.5 * r), rng.uniform(2.5 * r, NY - 1 - 2.5 * r)], 1)
    pos[0, 0] = 0.5 * r[0]                       # crosses the edge i = 0
    ang = rng.uniform(0, 2 * math.pi, n_particles)
    vdir = rng.uniform(0, 2 * math.pi, n_particles)
    pvel = u_particle * np.stack([np.cos(vdir), np.sin(vdir)], 1)
    K = max_num_ib or int(math.ceil(2 * math.pi * radius[1])) + 1
    sp = SphereIB(n_particles, K)
    spheres = dict(pos=pos.astype(np.float32), vel=pvel.astype(np.float32), radius=r,
                   angle=ang.astype(np.float32), omega=np.float32(0.002))
    return dict(lbf=lbf, bce=bce, macro=macro, coll=coll, fluid=fluid, rho=rho, vel=vel, power_law=False,
                sp=sp, spheres=spheres,
                spec=dict(kind="ib_particles", NX=NX, NY=NY, n_particles=n_particles, radius=radius,
                          u_particle=u_particle, tau=tau, collision=collision, g=g, seed=seed, max_num_ib=K))


def ib_particle_state(case, step):
    """Prescribed particle motion of an ib_particles case at ``step``: each
    particle translates with its velocity and turns at ``omega`` rad/step."""
    s = case["spheres"]
    pos = (s["pos"] + np.float32(step) * s["vel"]).astype(np.float32)
    ang = (s["angle"] + np.float32(step) * s["omega"]).astype(np.float32)
    return pos, s["vel"], ang
