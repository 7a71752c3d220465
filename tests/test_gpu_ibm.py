"""Immersed-boundary particles on the GPU (csrc/lbm_ibm.hip through the
taichi_lbm.ibm drop-in) against the IBM oracle (oracle/ibm_oracle.py) on the
same inputs: marker placement, the marker list, the multi-direct forcing
loop, particle forces, and an IB-coupled granular run of 30 steps against the
oracle's step sequence.

Tolerances: marker positions within 4e-5 (FP32 of coordinates up to ~250
plus the cosf/sinf of the device); velocities and forces 1e-6 absolute
plus 1e-5 relative (tests/tolerances.py), as for the lattice fields.  The
spread sums onto a node in an unspecified order (float atomics, as the
reference's parallel +=), so nothing here is compared bit for bit.
"""
import numpy as np
import pytest

from taichi_lbm import cases
from taichi_lbm.ibm import IBDEMCouplerEngine, IBEngine, SphereIB
from taichi_lbm._hip import LBMError
from oracle import ibm_oracle as I
from oracle_bridge import oracle_solver
from tolerances import check

pytestmark = pytest.mark.gpu


def _close(name, got, ref, atol=1e-6, rtol=1e-5):
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (name, got.shape, ref.shape)
    err = np.abs(got - ref)
    bound = atol + rtol * np.abs(ref).max(initial=0.0)
    print(f"  {name}: max abs err {err.max(initial=0.0):.3e}", flush=True)
    assert np.all(err <= bound), f"{name}: {err.max():.3e} > {bound:.3e}"


def _oracle_spheres(sp_dev, balls):
    o = I.Spheres(sp_dev.Max_Number, sp_dev.Max_NumIB, np.float32)
    o.init_balls(*balls)
    return o


def _balls(NX, NY):
    pos = np.array([[30.0, 20.0], [2.0, 22.0], [-30.0, 20.0], [60.5, 33.25], [NX - 3.0, NY - 2.5]], np.float32)
    vel = np.array([[0.01, -0.02], [0.03, 0.0], [0.5, 0.5], [-0.004, 0.007], [0.0, -0.01]], np.float32)
    radius = np.array([6.0, 5.0, 4.0, 7.3, 4.6], np.float32)
    angle = np.array([0.0, 0.4, 0.0, -1.3, 2.9], np.float32)
    n = len(pos)
    return (n, np.arange(n), pos, vel, np.zeros((n, 2), np.float32), np.zeros(n, np.float32), radius, angle)


def test_markers_and_marker_list_match_oracle():
    print("case IB markers 5 spheres on 80x48", flush=True)
    NX, NY = 80, 48
    lbf = cases.ib_particles(NX=NX, NY=NY, n_particles=1)["lbf"]
    sp = SphereIB(8, 64)
    balls = _balls(NX, NY)
    sp.init_Balls(*balls)
    sp.init_IB_nodes()
    o = _oracle_spheres(sp, balls)
    I.init_ib_nodes(o)
    assert np.array_equal(sp.Sphere.NumIB.to_numpy()[:5], o.num_ib[:5])
    _close("init pos", sp.IBNodes.pos.to_numpy(), o.node_pos, atol=4e-5)
    coupler = IBDEMCouplerEngine(lbf)
    for step in range(3):
        pos = balls[2] + step * np.float32(1.5) * balls[3]
        ang = balls[7] + np.float32(0.1 * step)
        coupler.SphereInfoUpdata(sp, pos, balls[3], ang)
        o.pos[:5], o.vel[:5], o.angle[:5] = pos, balls[3], ang
        I.sphere_ib_update(o, NX, NY)
        cnt = sp.IBInsiderCount[None]
        assert cnt == o.count
        g = sp.IBNodeInsiderGroup.to_numpy()
        assert np.array_equal(g, o.group), "marker list (order and cleared tail)"
        _close("pos", sp.IBNodes.pos.to_numpy(), o.node_pos, atol=4e-5)
        assert np.array_equal(sp.IBNodes.vel_desired.to_numpy(), o.node_vel_desired)
    # the computer_IB_nodes_pos host form agrees with the device markers
    _close("host pos", sp.computer_IB_nodes_pos(3, 7), sp.IBNodes.pos.to_numpy()[3, 7], atol=4e-5)


@pytest.mark.parametrize("nc", [1, 2])
def test_multi_direct_forcing_matches_oracle(nc):
    print(f"case IB multi-direct forcing nc={nc}", flush=True)
    NX, NY = 80, 48
    rng = np.random.default_rng(11 + nc)
    from taichi_lbm.lbm import LBField
    lbf = LBField("ib", NX, NY, nc)
    rho = rng.uniform(0.4, 0.6, (NX, NY, nc)).astype(np.float32)
    vel = rng.normal(0, 0.01, (NX, NY, 2)).astype(np.float32)
    bf = rng.normal(0, 1e-5, (NX, NY, nc, 2)).astype(np.float32)
    lbf.rho.from_numpy(rho)
    lbf.total_rho.from_numpy(rho.sum(axis=2))
    lbf.vel.from_numpy(vel)
    lbf.body_force.from_numpy(bf)
    sp = SphereIB(8, 64)
    balls = _balls(NX, NY)
    sp.init_Balls(*balls)
    sp.init_IB_nodes()
    coupler = IBDEMCouplerEngine(lbf)
    coupler.SphereIBUpdate(sp)
    ibe = IBEngine(3)
    ibe.MultiDirectForcingLoop(sp, lbf)
    ibe.ComputeParticleForce(sp)
    o = _oracle_spheres(sp, balls)
    I.init_ib_nodes(o)
    I.sphere_ib_update(o, NX, NY)
    ovel, obf, orho, otot = vel.copy(), bf.copy(), rho.copy(), rho.sum(axis=2)
    I.multi_direct_forcing(o, 3, ovel, obf, orho, otot)
    I.particle_force(o)
    _close("lattice vel", lbf.vel.to_numpy(), ovel)
    _close("body_force", lbf.body_force.to_numpy(), obf)
    _close("node vel", sp.IBNodes.vel.to_numpy(), o.node_vel)
    _close("node force", sp.IBNodes.force.to_numpy(), o.node_force)
    _close("node force_temp", sp.IBNodes.force_temp.to_numpy(), o.node_force_temp)
    _close("sphere force", coupler.GetSphereForce(sp), o.force[:5], rtol=1e-4)
    _close("sphere torque", coupler.GetSphereTorque(sp), o.torque[:5], atol=1e-5, rtol=1e-4)
    assert coupler.GetSphereForce(sp).dtype == np.float64 and coupler.GetSphereForce(sp).shape == (5, 2)
    assert np.array_equal(coupler.GetSphereForce(sp)[2], [0.0, 0.0])       # outside: untouched
    # the split form (reset, then interpolate / spread calls) equals the loop
    lbf.vel.from_numpy(vel)
    lbf.body_force.from_numpy(bf)
    sp.reset_force()
    for _ in range(3):
        ibe.InterpolateBodyVelocities(sphere_field=sp, lb_filed=lbf)
        ibe.SpreadBodyForces(sphere_field=sp, lb_filed=lbf)
    _close("split vel", lbf.vel.to_numpy(), ovel)
    _close("split node force", sp.IBNodes.force.to_numpy(), o.node_force)


def test_capacity_is_checked_before_launch():
    sp = SphereIB(2, 10)
    sp.init_Balls(1, [0], [(20.0, 20.0)], [(0.0, 0.0)], [(0.0, 0.0)], [0.0], [5.0], [0.0])
    with pytest.raises(LBMError):
        sp.init_IB_nodes()                 # 31 markers > Max_NumIB = 10
    with pytest.raises(ValueError):
        sp.init_Balls(3, [0, 1, 2], np.zeros((3, 2)), np.zeros((3, 2)), np.zeros((3, 2)), np.zeros(3),
                      np.ones(3), np.zeros(3))


@pytest.mark.parametrize("coll", ["bgk", "mrt"])
def test_ib_coupled_granular_run_matches_oracle(coll):
    print(f"case IB-coupled periodic channel 96x64 {coll}, 30 steps", flush=True)
    case = cases.ib_particles(NX=96, NY=64, n_particles=3, radius=(5.0, 8.0), u_particle=0.02, collision=coll,
                              g=(2e-6, 0.0), seed=4)
    cases.initialise(case)
    lbf, macro, col, bce, sp = case["lbf"], case["macro"], case["coll"], case["bce"], case["sp"]
    s = case["spheres"]
    n = len(s["radius"])
    balls = (n, np.arange(n), s["pos"], s["vel"], np.zeros((n, 2), np.float32), np.zeros(n, np.float32),
             s["radius"], s["angle"])
    sp.init_Balls(*balls)
    sp.init_IB_nodes()
    coupler = IBDEMCouplerEngine(lbf)
    ibe = IBEngine(2)
    sol = oracle_solver(case, np.float32)
    lat = sol.lat
    o = _oracle_spheres(sp, balls)
    I.init_ib_nodes(o)
    from oracle import lbm_oracle as O
    steps = 30
    for step in range(steps):
        pos, pvel, ang = cases.ib_particle_state(case, step)
        macro.density(lbf)
        macro.force_density(lbf)
        macro.velocity(lbf)
        coupler.SphereInfoUpdata(sp, pos, pvel, ang)
        ibe.MultiDirectForcingLoop(sp, lbf)
        ibe.ComputeParticleForce(sp)
        col.apply(lbf)
        bce.apply_boundary_conditions(lbf)
        # the oracle's step with the same IB coupling
        O.density(lat, sol.group)
        O.force_density(lat, sol.group)
        O.velocity(lat, sol.group)
        o.pos[:n], o.vel[:n], o.angle[:n] = pos, pvel, ang
        I.sphere_ib_update(o, lat.NX, lat.NY)
        I.multi_direct_forcing(o, 2, lat.vel, lat.body_force, lat.rho, lat.total_rho)
        I.particle_force(o)
        sol.collision.apply(lat)
        O.apply_bcs(lat, sol.bcs)
    for name in ("f", "f2", "rho", "vel", "body_force"):
        check(name, getattr(lbf, name).to_numpy(), getattr(lat, name), growth=3.0)
    _close("sphere force", coupler.GetSphereForce(sp), o.force[:n], atol=1e-5, rtol=1e-3)
    _close("sphere torque", coupler.GetSphereTorque(sp), o.torque[:n], atol=1e-4, rtol=1e-3)
    # the particles push the fluid: the coupled run differs from one without particles
    assert np.abs(lbf.vel.to_numpy()).max() > 1e-3


@pytest.mark.parametrize("coll", ["bgk", "mrt"])
def test_ib_coupled_fused_stepper_matches_granular(coll):
    """The fused IB-coupled step (Stepper(..., ib_engine, sphere_field):
    lbm_ib_run) against the granular sequence with MultiDirectForcingLoop
    between velocity and collide, on the same moving particles: the same
    discrete problem, so the fields agree within the FP32 parity bounds
    (the spread force enters the fused step as a force term, so its sums
    round differently)."""
    from taichi_lbm.stepper import Stepper
    print(f"case IB-coupled fused vs granular 96x64 {coll}, 30 steps", flush=True)

    def make():
        case = cases.ib_particles(NX=96, NY=64, n_particles=3, radius=(5.0, 8.0), u_particle=0.02,
                                  collision=coll, g=(2e-6, 0.0), seed=4)
        cases.initialise(case)
        s = case["spheres"]
        n = len(s["radius"])
        case["sp"].init_Balls(n, np.arange(n), s["pos"], s["vel"], np.zeros((n, 2), np.float32),
                              np.zeros(n, np.float32), s["radius"], s["angle"])
        case["sp"].init_IB_nodes()
        return case

    g, fz = make(), make()
    ibe_g, ibe_f = IBEngine(2), IBEngine(2)
    cg, cf = IBDEMCouplerEngine(g["lbf"]), IBDEMCouplerEngine(fz["lbf"])
    st = Stepper(fz["lbf"], fz["macro"], fz["coll"], fz["bce"], ib_engine=ibe_f, sphere_field=fz["sp"])
    for step in range(30):
        pos, pvel, ang = cases.ib_particle_state(g, step)
        g["macro"].density(g["lbf"])
        g["macro"].force_density(g["lbf"])
        g["macro"].velocity(g["lbf"])
        cg.SphereInfoUpdata(g["sp"], pos, pvel, ang)
        ibe_g.MultiDirectForcingLoop(g["sp"], g["lbf"])
        ibe_g.ComputeParticleForce(g["sp"])
        g["coll"].apply(g["lbf"])
        g["bce"].apply_boundary_conditions(g["lbf"])
        cf.SphereInfoUpdata(fz["sp"], pos, pvel, ang)
        st.run(1)
        ibe_f.ComputeParticleForce(fz["sp"])
        if step == 14:       # a readback in between
            _ = fz["lbf"].vel.to_numpy()
    for name in ("f", "f2", "rho", "vel"):
        check(name, getattr(fz["lbf"], name).to_numpy(), getattr(g["lbf"], name).to_numpy(), growth=3.0)
    _close("sphere force", cf.GetSphereForce(fz["sp"]), cg.GetSphereForce(g["sp"]), atol=1e-5, rtol=1e-3)
    _close("sphere torque", cf.GetSphereTorque(fz["sp"]), cg.GetSphereTorque(g["sp"]), atol=1e-4, rtol=1e-3)
    # without the particles the flow would stay uniform: the coupling acted
    assert np.abs(fz["lbf"].vel.to_numpy()).max() > 1e-3


def test_ib_fused_stepper_refuses_what_it_cannot_run():
    from taichi_lbm.stepper import Stepper
    case = cases.ib_particles(NX=40, NY=32, n_particles=1, radius=(3.0, 4.0))
    with pytest.raises(NotImplementedError):
        Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], power_law=True, ib_engine=IBEngine(1),
                sphere_field=case["sp"])
    with pytest.raises(ValueError):
        Stepper(case["lbf"], case["macro"], case["coll"], case["bce"], ib_engine=IBEngine(1))
