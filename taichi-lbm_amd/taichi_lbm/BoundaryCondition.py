"""Node grouping and boundary conditions -- drop-in for
taichi_lbm/BoundaryCondition.py of the reference.

Geometry predicates are evaluated on the host.  A predicate is called once
with broadcastable integer index arrays ``I`` (NX, 1) and ``J`` (1, NY) and
must return a boolean array that broadcasts to (NX, NY); predicates written
for scalars (Python ``if``) are evaluated node by node instead.  Field reads inside predicates are
bounds-checked (the reference's read out of range silently).
"""
from __future__ import annotations

import ctypes
import numpy as np

from . import _hip as H
from ._codes import BCSpec, UNKNOWN
from ._device import _Scalar, device, stream, torch
from .lbm import gfields, group_device, lattice_struct


class _GroupNodes:
    """The ``group.group`` vector field of MaskAndGroup (BoundaryCondition.py:11)."""

    def __init__(self, owner):
        self._o = owner

    @property
    def shape(self):
        return (self._o.capacity,)

    def to_numpy(self):
        out = np.zeros((self._o.capacity, 2), np.int32)
        out[:len(self._o.nodes)] = self._o.nodes
        return out

    def __getitem__(self, m):
        if not 0 <= m < self._o.capacity:
            raise IndexError(m)
        return tuple(self.to_numpy()[m])


class MaskAndGroup:
    """BoundaryCondition.py:8-12: a node list of capacity NX*NY plus a count.

    The reference compacts with atomic_add, so its order is arbitrary; here the
    nodes are in (i, j) lexicographic order.
    """

    def __init__(self, NX, NY, nodes=None):
        self.NX, self.NY = int(NX), int(NY)
        self.capacity = self.NX * self.NY
        self.nodes = np.zeros((0, 2), np.int32) if nodes is None else np.asarray(nodes, np.int32).reshape(-1, 2)
        self.count = _Scalar(len(self.nodes))
        self.group = _GroupNodes(self)
        self._dev = None

    def device_nodes(self):
        if len(self.nodes) == 0:
            return None
        if self._dev is None:
            self._dev = torch().as_tensor(np.ascontiguousarray(self.nodes), device=device())
        return self._dev

    def flat(self):
        return self.nodes[:, 0].astype(np.int64) * self.NY + self.nodes[:, 1]


class BoundarySpec:
    """BoundaryCondition.py:14-20."""

    def __init__(self, geometry_fn, direction=None, value_fn=None):
        self.geometry_fn = geometry_fn
        self.direction = direction
        self.value_fn = value_fn


def evaluate_predicate(fn, NX: int, NY: int) -> np.ndarray:
    """Boolean (NX, NY) map of a geometry predicate (vectorised, else per node)."""
    I = np.arange(NX)[:, None]
    J = np.arange(NY)[None, :]
    try:
        res = np.broadcast_to(np.asarray(fn(I, J)), (NX, NY))
        return res.astype(bool)
    except (ValueError, TypeError):
        out = np.zeros((NX, NY), bool)
        for i in range(NX):
            for j in range(NY):
                out[i, j] = bool(fn(i, j))
        return out


class BoundaryClassifier:
    """BoundaryCondition.py:22-39.  ``shape`` is any object with ``.shape`` (the
    reference passes a Taichi field) or an (NX, NY) tuple; ``offset`` (or the
    object's own ``.offset``) shifts the index range, as the psi halo field of
    shape (NX+2, NY+2) with offset (-1, -1) does (the classifier of the psi
    boundaries, cases/Bubble or two-phase separation.py:149-150): the group
    then holds (i, j) from the offset on."""

    def __init__(self, shape, offset=None):
        self.shape = shape
        sh = shape.shape if hasattr(shape, "shape") else tuple(shape)
        self.NX, self.NY = int(sh[0]), int(sh[1])
        off = offset if offset is not None else getattr(shape, "offset", (0, 0))
        self.offset = (int(off[0]), int(off[1]))

    def get_group(self, spec: BoundarySpec) -> MaskAndGroup:
        ox, oy = self.offset
        fn = spec.geometry_fn if (ox, oy) == (0, 0) else (lambda I, J: spec.geometry_fn(I + ox, J + oy))
        hit = evaluate_predicate(fn, self.NX, self.NY)
        nodes = np.argwhere(hit).astype(np.int32) + np.array([ox, oy], np.int32)
        return MaskAndGroup(self.NX, self.NY, nodes)


_C = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])


def check_granular_bounds(kind, nodes, NX, NY, direction=None, table=None, name="boundary"):
    """Raises LBMError unless every address a granular boundary kernel forms
    from this group lies inside the NX x NY lattice: the group's nodes; for
    the neighbour-table kinds (BounceBackWall, InsideBoundary) every table
    entry of those nodes is (-1, -1) or a node of the lattice; for the node
    rules that read x - c_d (VelocityNEEM, PressureNEEM, OpenNeumann,
    PressureABC) that interior neighbour.  The reference reads out of range silently
    (BoundaryCondition.py:154-166, :215-226, :321-334, :446-449); the kernels
    trust these addresses, so they are checked here before any launch."""
    nodes = np.asarray(nodes, np.int64).reshape(-1, 2)
    if len(nodes) == 0:
        return
    I, J = nodes[:, 0], nodes[:, 1]
    if I.min() < 0 or I.max() >= NX or J.min() < 0 or J.max() >= NY:
        raise H.LBMError(f"{name}: group node outside the {NX} x {NY} lattice")
    if kind in (H.BC_INSIDE, H.BC_BOUNCEBACK):
        nb = np.asarray(table)[I, J].reshape(-1, 9, 2).astype(np.int64)
        none = (nb[..., 0] == -1) & (nb[..., 1] == -1)
        inside = (nb[..., 0] >= 0) & (nb[..., 0] < NX) & (nb[..., 1] >= 0) & (nb[..., 1] < NY)
        bad = ~(none | inside)
        if bad.any():
            m, k = np.argwhere(bad)[0]
            raise H.LBMError(f"{name}: neighbour table entry {tuple(nb[m, k])} of node {tuple(nodes[m])}, "
                             f"population {k}, is neither -1 nor a node of the lattice")
    elif kind in (H.BC_VELOCITY_NEEM, H.BC_PRESSURE_NEEM, H.BC_OPEN_NEUMANN, H.BC_PRESSURE_ABC, H.BC_VELOCITY_EQ,
                  H.BC_OPEN_CONVECTIVE1, H.BC_OPEN_EXTRAPOLATION):
        if direction not in (1, 2, 3, 4):
            raise H.LBMError(f"{name}: direction {direction} (1..4)")
        if kind == H.BC_VELOCITY_EQ:     # reads its own node only
            return
        for s in ((1, 2) if kind == H.BC_OPEN_EXTRAPOLATION else (1,)):   # x - c_d (and x - 2 c_d)
            I2, J2 = I - s * _C[direction, 0], J - s * _C[direction, 1]
            out = (I2 < 0) | (I2 >= NX) | (J2 < 0) | (J2 >= NY)
            if out.any():
                m = int(np.argmax(out))
                raise H.LBMError(f"{name}: node {tuple(nodes[m])} reads its interior neighbour "
                                 f"{(int(I2[m]), int(J2[m]))} outside the lattice (direction {direction})")


class BoundaryCondition:
    """BoundaryCondition.py:41-54."""

    kind = "fluid"

    def __init__(self, spec: BoundarySpec):
        self.spec = spec
        self.group = None

    def precompute(self, classifier: BoundaryClassifier):
        self.group = classifier.get_group(self.spec)

    def apply(self, lb_field):
        pass

    def fused_spec(self, name="") -> BCSpec:
        return BCSpec(self.kind, self.group.nodes, name=name)

    def _granular(self, lb_field, kind, rule=None):
        need_nb = kind in (H.BC_INSIDE, H.BC_BOUNCEBACK)
        key = (kind, id(self.group), len(self.group.nodes), lb_field.geometry_version,
               rule.direction if rule is not None else 0)
        if getattr(self, "_checked", None) != key:
            check_granular_bounds(kind, self.group.nodes, lb_field.NX, lb_field.NY,
                                  rule.direction if rule is not None else None,
                                  lb_field.neighbor.arr if need_nb else None, type(self).__name__)
            self._checked = key
        lb_field.materialize()
        g = group_device(self.group)
        H.check(H.lib().lbm_g_boundary(ctypes.byref(lattice_struct(lb_field)),
                                       ctypes.byref(gfields(lb_field, neighbor=need_nb)),
                                       kind, ctypes.byref(rule) if rule is not None else None, g[0], g[1],
                                       stream()), type(self).__name__)
        lb_field._pops_written()


class PlaneBoundary(BoundaryCondition):
    """BoundaryCondition.py:57-72."""

    def __init__(self, spec: BoundarySpec, direction):
        super().__init__(spec)
        self.direction = direction
        self.unknow = list(UNKNOWN[direction]) if direction in UNKNOWN else None


class VelocityBoundary(PlaneBoundary):
    def __init__(self, spec, velocity_value, direction):
        super().__init__(spec, direction)
        self.velocity_value = tuple(float(v) for v in velocity_value)


class PressureBoundary(PlaneBoundary):
    def __init__(self, spec, rho_value, direction):
        super().__init__(spec, direction)
        self.rho_value = float(rho_value)


class VelocityNEEM(VelocityBoundary):
    """BoundaryCondition.py:149-166."""

    kind = "velocity_neem"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = H.NEEM_VELOCITY, int(self.direction)
        r.ux, r.uy = (float(np.float32(v)) for v in self.velocity_value)
        return r

    def fused_spec(self, name=""):
        ux, uy = (float(np.float32(v)) for v in self.velocity_value)
        return BCSpec(self.kind, self.group.nodes, int(self.direction), ux, uy, 0.0, name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_VELOCITY_NEEM, self.rule())


class PressureNEEM(PressureBoundary):
    """BoundaryCondition.py:209-226."""

    kind = "pressure_neem"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = H.NEEM_PRESSURE, int(self.direction)
        r.rho = float(np.float32(self.rho_value))
        return r

    def fused_spec(self, name=""):
        return BCSpec(self.kind, self.group.nodes, int(self.direction), 0.0, 0.0,
                      float(np.float32(self.rho_value)), name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_PRESSURE_NEEM, self.rule())


class VelocityEquilibrium(VelocityBoundary):
    """BoundaryCondition.py:132-145: vel = u_w, rho_w from the node's populations
    (Zou-He form), then all nine populations := feq(u_w, rho_w)."""

    kind = "velocity_eq"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = H.RULE_VELOCITY_EQ, int(self.direction)
        r.ux, r.uy = (float(np.float32(v)) for v in self.velocity_value)
        return r

    def fused_spec(self, name=""):
        ux, uy = (float(np.float32(v)) for v in self.velocity_value)
        return BCSpec(self.kind, self.group.nodes, int(self.direction), ux, uy, 0.0, name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_VELOCITY_EQ, self.rule())


class PressureABC(PressureBoundary):
    """BoundaryCondition.py:177-205: anti-bounce-back pressure outlet.  The
    velocity is extrapolated, u_w = 1.5 u(x) - 0.5 u(x - c_d), total_rho :=
    rho_value, and each unknown population (written by an earlier
    bounce-back) becomes -f_k + 2 w_k rho_value (1 + 4.5 (c_k . u_w)^2 -
    1.5 u_w^2), which imposes rho_value."""

    kind = "pressure_abc"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = H.RULE_PRESSURE_ABC, int(self.direction)
        r.rho = float(np.float32(self.rho_value))
        return r

    def fused_spec(self, name=""):
        return BCSpec(self.kind, self.group.nodes, int(self.direction), 0.0, 0.0,
                      float(np.float32(self.rho_value)), name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_PRESSURE_ABC, self.rule())


class VelocityBB(VelocityBoundary):
    """BoundaryCondition.py:80-102: adds 6 w_d rho c_d . u_w to the unknown
    populations (after an earlier bounce-back has written them)."""

    kind = "velocity_bb"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = H.RULE_VELOCITY_BB, int(self.direction)
        r.ux, r.uy = (float(np.float32(v)) for v in self.velocity_value)
        return r

    def fused_spec(self, name=""):
        ux, uy = (float(np.float32(v)) for v in self.velocity_value)
        return BCSpec(self.kind, self.group.nodes, int(self.direction), ux, uy, 0.0, name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_VELOCITY_BB, self.rule())


class OpenNeumann(PlaneBoundary):
    """BoundaryCondition.py:315-334: zero-gradient outlet -- copies the unknown
    populations from x - c_d, then recomputes density and velocity at x."""

    kind = "open_neumann"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = H.RULE_OPEN_NEUMANN, int(self.direction)
        return r

    def fused_spec(self, name=""):
        return BCSpec(self.kind, self.group.nodes, int(self.direction), name=name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_OPEN_NEUMANN, self.rule())


class BounceBackWall(BoundaryCondition):
    """BoundaryCondition.py:231-259 (halfway bounce-back through the neighbour table)."""

    kind = "bounceback"

    def __init__(self, spec: BoundarySpec):
        super().__init__(spec)
        self.bounce_map = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_BOUNCEBACK)


class PeriodicAllBoundary(BoundaryCondition):
    """BoundaryCondition.py:419-432 (periodic pull streaming)."""

    kind = "periodic"

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_PERIODIC)


class InsideBoundary(BoundaryCondition):
    """BoundaryCondition.py:435-449 (pull streaming through the neighbour table)."""

    kind = "inside"

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_INSIDE)


class FluidBoundary(BoundaryCondition):
    """BoundaryCondition.py:452-456 (no-op entry holding the fluid group)."""

    kind = "fluid"


class OpenConvective1order(PlaneBoundary):
    """BoundaryCondition.py:337-359: convective outlet -- the unknown
    populations relax towards x - c_d's, f := (f + l f(x - c_d)) / (1 + l)
    with l = |u(x - c_d)| of the step's velocity pass, then density and
    velocity at x (fused and granular)."""

    kind = "open_convective1"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = H.RULE_OPEN_CONVECTIVE1, int(self.direction)
        return r

    def fused_spec(self, name=""):
        return BCSpec(self.kind, self.group.nodes, int(self.direction), name=name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_OPEN_CONVECTIVE1, self.rule())


class OpenExtrapolation(PlaneBoundary):
    """BoundaryCondition.py:394-413: the unknown populations extrapolated
    linearly from x - c_d and x - 2 c_d (density and velocity left as they
    are, as in the reference).  Granular only: the fused step's node rules
    read one node inward, this one reads two."""

    kind = "open_extrapolation"

    def rule(self):
        r = H.NeemRule()
        r.kind, r.direction = 0, int(self.direction)
        return r

    def fused_spec(self, name=""):
        return BCSpec(self.kind, self.group.nodes, int(self.direction), name=name)

    def apply(self, lb_field):
        self._granular(lb_field, H.BC_OPEN_EXTRAPOLATION, self.rule())


def _out_of_scope(name, where, why):
    def init(self, *a, **k):
        raise NotImplementedError(f"{name} ({where}) is not built: {why}")
    return type(name, (BoundaryCondition,), {"__init__": init})


OpenConvective2order = _out_of_scope(
    "OpenConvective2order", "BoundaryCondition.py:362-390",
    "it needs ShanChenForceC1's psi and rewrites the interior neighbour's force and velocity, and its three "
    "passes over the group run as parallel loop iterations in the reference (a race)")


class BoundaryEngine:
    """BoundaryCondition.py:459-483: an ordered registry applied in insertion order."""

    def __init__(self):
        self.boundary_conditions = {}

    def Mask_rectangle_identify(self, lb_field, xmin, xmax, ymin, ymax):
        """BoundaryCondition.py:464-469 (strict inequalities, f32 bounds)."""
        f32 = np.float32
        I, J = np.meshgrid(np.arange(lb_field.NX), np.arange(lb_field.NY), indexing="ij")
        hit = (I < f32(xmax)) & (I > f32(xmin)) & (J < f32(ymax)) & (J > f32(ymin))
        m = lb_field.mask.arr
        m[hit] = -1
        lb_field._geometry_changed()

    def Mask_cricle_identify(self, lb_field, x, y, r):
        """BoundaryCondition.py:471-475: (i-x)^2 + (j-y)^2 < r^2 in f32."""
        f32 = np.float32
        I, J = np.meshgrid(np.arange(lb_field.NX), np.arange(lb_field.NY), indexing="ij")
        dx = I.astype(f32) - f32(x)
        dy = J.astype(f32) - f32(y)
        hit = dx * dx + dy * dy < f32(r) * f32(r)
        lb_field.mask.arr[hit] = -1
        lb_field._geometry_changed()

    Mask_circle_identify = Mask_cricle_identify   # README.md:105 spelling

    def add_boundary_condition(self, name, bc):
        """BoundaryCondition.py:477-479 (a repeated name replaces the entry in place)."""
        self.boundary_conditions[name] = bc
        print(f"  add boundary {name}, size: {bc.group.count[None]}")

    def apply_boundary_conditions(self, lb_field):
        """BoundaryCondition.py:481-483 (granular: one kernel per entry)."""
        for bc in self.boundary_conditions.values():
            bc.apply(lb_field)

    def fused_list(self):
        return [bc.fused_spec(name) for name, bc in self.boundary_conditions.items()]

    def writing_boundary(self, lb_field, path="boundary.png"):
        """BoundaryCondition.py:486-510: paints each group in insertion order into a PNG."""
        from PIL import Image
        colors = [(0.0, 0.0, 0.0), (52 / 255, 152 / 255, 219 / 255), (1.0, 0.0, 0.0),
                  (143 / 255, 24 / 255, 172 / 255), (255 / 255, 190 / 255, 53 / 255), (1.0, 0.0, 0.0)]
        img = np.zeros((lb_field.NX, lb_field.NY, 3), np.float32)
        for num, bc in enumerate(self.boundary_conditions.values()):
            nodes = bc.group.nodes
            img[nodes[:, 0], nodes[:, 1]] = colors[num]   # the reference indexes colors[num] too
        Image.fromarray((img * 255).astype(np.uint8)).save(path)
        print("writing boundary finish")
