"""Compiles the reference's ordered boundary list into one uint16 code per node.

The reference streams and applies boundaries as a list of kernels run in
insertion order (BoundaryEngine.apply_boundary_conditions,
BoundaryCondition.py:481-483); each later kernel overwrites what earlier ones
wrote.  The fused kernel needs, per node and population, only the *last*
writer.  This module replays the list symbolically on the host:

  PeriodicAllBoundary (:424-432)  every population: pull from x - c_k (wrapped)
  InsideBoundary      (:440-449)  pull from the neighbour table; a -1 entry
                                  reads outside the field (undefined) unless a
                                  later rule overwrites it
  BounceBackWall      (:244-259)  pull where the table has a neighbour, else
                                  halfway bounce-back f2(x, opp k)
  VelocityNEEM        (:154-166)  all nine populations
  PressureNEEM        (:215-226)  the three unknown populations
  VelocityBB          (:85-102)   adds 6 w_k rho(x) c_k.u_w to the three unknowns
  OpenNeumann         (:321-334)  the three unknowns := f(x - c_d)
  FluidBoundary / base (:51-54, :452-456)  nothing

A table pull with a valid neighbour reads the same node as a wrapped pull, so
each population resolves to PULL or BOUNCE, plus an optional node rule (the
last four kinds) applied after it.  NEEM and OpenNeumann rules re-gather their
interior neighbour x - c_d, whose populations must then be final (no later
boundary rewrites them) and carry no rule of their own.  Lists whose reference
semantics the fused kernel cannot reproduce are rejected with an explanation
(never silently approximated).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from . import _hip as H

C = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])
UNKNOWN = {1: (3, 6, 7), 3: (1, 5, 8), 4: (2, 5, 6), 2: (4, 7, 8)}

NONE, PULL, BOUNCE, INVALID = 0, 1, 2, 3


class UnsupportedBoundaryList(ValueError):
    """The boundary list has reference semantics the fused path does not reproduce."""


@dataclass
class BCSpec:
    """Host description of one boundary-list entry."""
    kind: str                 # fluid | periodic | inside | bounceback | velocity_neem | pressure_neem
                              # | velocity_bb | open_neumann | pressure_abc | velocity_eq
                              # | open_convective1 | open_extrapolation (granular only)
    nodes: np.ndarray         # (m, 2) int (i, j)
    direction: int = 0
    ux: float = 0.0
    uy: float = 0.0
    rho: float = 0.0
    name: str = ""


def shifted(a: np.ndarray, dx: int, dy: int, fill) -> np.ndarray:
    """b[i, j] = a[i - dx, j - dy] where that index is inside the array, else fill."""
    NX, NY = a.shape
    b = np.full_like(a, fill)
    xs, xd = (slice(0, NX - dx), slice(dx, NX)) if dx >= 0 else (slice(-dx, NX), slice(0, NX + dx))
    ys, yd = (slice(0, NY - dy), slice(dy, NY)) if dy >= 0 else (slice(-dy, NY), slice(0, NY + dy))
    b[xd, yd] = a[xs, ys]
    return b


def neighbor_valid(mask: np.ndarray) -> np.ndarray:
    """(NX, NY, 9) bool: x - c_k is inside the lattice and not solid -- the
    entries of LBField.neighbor_classify (lbm.py:127-142) that are not -1."""
    fluid = mask != -1
    return np.stack([shifted(fluid, int(C[k, 0]), int(C[k, 1]), False) for k in range(9)], axis=-1)


def neighbor_table(mask: np.ndarray) -> np.ndarray:
    """LBField.neighbor_classify, lbm.py:127-142 (non-periodic; solid nodes left 0)."""
    NX, NY = mask.shape
    tmp = np.zeros((9, 2, NX, NY), np.int32)
    fluid = mask != -1
    I = np.arange(NX, dtype=np.int32)[:, None]
    J = np.arange(NY, dtype=np.int32)[None, :]
    for k in range(9):
        ok = shifted(fluid, int(C[k, 0]), int(C[k, 1]), False)   # fluid at x - c_k, inside the lattice
        tmp[k, 0] = np.where(fluid, np.where(ok, I - C[k, 0], -1), 0)
        tmp[k, 1] = np.where(fluid, np.where(ok, J - C[k, 1], -1), 0)
    return np.ascontiguousarray(tmp.transpose(2, 3, 0, 1))


def compile_codes(NX: int, NY: int, active_nodes: np.ndarray, bcs: Sequence[BCSpec],
                  neighbor: np.ndarray) -> Tuple[np.ndarray, List[Tuple]]:
    """Returns (code[NX*NY] uint16, neem_rules[(kind, direction, ux, uy, rho)]).

    ``neighbor`` is either the (NX, NY, 9, 2) table of neighbor_classify or
    the (NX, NY, 9) boolean map of its valid (not -1) entries."""
    N = NX * NY
    active = np.zeros(N, bool)
    if len(active_nodes):
        a = np.asarray(active_nodes).reshape(-1, 2)
        active[a[:, 0] * NY + a[:, 1]] = True
    rule = np.full((N, 9), NONE, np.int8)
    neem_of = np.zeros(N, np.int32)          # rule index + 1 at the node
    pre_rule = {}                            # rule idx -> rules of its nodes before it ran
    nb_rule_at = {}                          # rule idx -> rules of its interior neighbours then
    neem_rules: List[Tuple] = []
    neem_nodes: List[Tuple[np.ndarray, np.ndarray]] = []   # (nodes, interior neighbours or None)
    rule_kinds = {"velocity_neem": H.NEEM_VELOCITY, "pressure_neem": H.NEEM_PRESSURE,
                  "velocity_bb": H.RULE_VELOCITY_BB, "open_neumann": H.RULE_OPEN_NEUMANN,
                  "pressure_abc": H.RULE_PRESSURE_ABC, "velocity_eq": H.RULE_VELOCITY_EQ,
                  "open_convective1": H.RULE_OPEN_CONVECTIVE1}
    nb_ok = (neighbor.reshape(N, 9) if neighbor.ndim == 3 else neighbor.reshape(N, 9, 2)[:, :, 0] != -1)

    for bc in bcs:
        nodes = np.asarray(bc.nodes, np.int64).reshape(-1, 2)
        if bc.kind == "fluid" or len(nodes) == 0:
            continue
        flat = nodes[:, 0] * NY + nodes[:, 1]
        if not active[flat].all():
            bad = nodes[~active[flat]][0]
            raise UnsupportedBoundaryList(
                f"boundary '{bc.name}' writes node ({bad[0]}, {bad[1]}) outside the fluid group; the fused "
                "path only streams and collides fluid-group nodes")
        if np.any(neem_of[flat] != 0):
            raise UnsupportedBoundaryList(
                f"boundary '{bc.name}' overwrites a node already set by a NEEM / VelocityBB / OpenNeumann rule; "
                "the fused path applies node rules last")
        if bc.kind == "periodic":
            rule[flat, :] = PULL
        elif bc.kind in ("inside", "bounceback"):
            valid = nb_ok[flat]
            rule[flat, :] = np.where(valid, PULL, INVALID if bc.kind == "inside" else BOUNCE)
        elif bc.kind in rule_kinds:
            d = int(bc.direction)
            if d not in UNKNOWN:
                raise UnsupportedBoundaryList(f"boundary '{bc.name}': direction must be 1..4, got {d}")
            if len(neem_rules) >= H.MAX_NEEM:
                raise UnsupportedBoundaryList(f"more than {H.MAX_NEEM} node rules")
            e = len(neem_rules)
            pre_rule[e] = rule[flat].copy()
            nbf = None
            if bc.kind not in ("velocity_bb", "velocity_eq"):     # rules that read x - c_d
                I2, J2 = nodes[:, 0] - C[d, 0], nodes[:, 1] - C[d, 1]
                if np.any((I2 < 0) | (I2 >= NX) | (J2 < 0) | (J2 >= NY)):
                    raise UnsupportedBoundaryList(
                        f"boundary '{bc.name}': the interior neighbour x - c_{d} of a node lies outside the lattice")
                nbf = I2 * NY + J2
                if not active[nbf].all():
                    raise UnsupportedBoundaryList(f"boundary '{bc.name}': interior neighbour outside the fluid group")
                nb_rule_at[e] = rule[nbf].copy()
            neem_of[flat] = e + 1
            neem_rules.append((rule_kinds[bc.kind], d, float(bc.ux), float(bc.uy), float(bc.rho)))
            neem_nodes.append((flat, nbf))
        else:
            raise UnsupportedBoundaryList(f"boundary kind '{bc.kind}' ('{bc.name}') has no fused form")

    # every population of every active node must have a writer that the fused
    # kernel can reproduce (NEEM nodes: before their rule)
    final = rule.copy()
    for e, (flat, nbf) in enumerate(neem_nodes):
        final[flat] = pre_rule[e]
        if nbf is None:
            continue
        if np.any(neem_of[nbf] != 0):
            raise UnsupportedBoundaryList("a NEEM / OpenNeumann node's interior neighbour carries a node rule too")
        if not np.array_equal(rule[nbf], nb_rule_at[e]):
            raise UnsupportedBoundaryList(
                "a boundary after a NEEM / OpenNeumann rule rewrites that rule's interior neighbour")
    act = np.nonzero(active)[0]
    missing = final[act] == NONE
    if missing.any():
        n, k = act[np.nonzero(missing.any(1))[0][0]], int(np.nonzero(missing[missing.any(1)][0])[0][0])
        raise UnsupportedBoundaryList(
            f"population {k} of node ({n // NY}, {n % NY}) is written by no boundary condition (the reference "
            "would keep its previous value)")
    invalid = final[act] == INVALID
    if invalid.any():
        n = act[np.nonzero(invalid.any(1))[0][0]]
        raise UnsupportedBoundaryList(
            f"node ({n // NY}, {n % NY}) streams from a -1 neighbour-table entry (InsideBoundary without a "
            "later BounceBackWall reads outside the field in the reference)")

    code = np.zeros(N, np.uint16)
    bits = np.zeros(N, np.uint32)
    for k in range(1, 9):
        bits |= (final[:, k] == BOUNCE).astype(np.uint32) << (k - 1)
    code[active] = (bits[active] | H.CODE_ACTIVE).astype(np.uint16)
    # nodes outside the group that an active node pulls a population from
    # (its unchanged f2): the two-step kernel stages them in LDS
    pulled = np.zeros(N, bool)
    I, J = act // NY, act % NY
    for k in range(1, 9):
        plain = final[act, k] == PULL
        src = ((I[plain] - C[k, 0]) % NX) * NY + (J[plain] - C[k, 1]) % NY
        pulled[src] = True
    pulled &= ~active
    code[pulled] |= np.uint16(H.CODE_PULLED)
    for e, (flat, nbf) in enumerate(neem_nodes):
        kind = neem_rules[e][0]
        code[flat] |= np.uint16((e + 1) << H.CODE_NEEM_SHIFT)
        if kind in (H.NEEM_VELOCITY, H.NEEM_PRESSURE, H.RULE_PRESSURE_ABC, H.RULE_OPEN_CONVECTIVE1):
            code[nbf] |= np.uint16(H.CODE_KEEP_MACRO)        # neighbour's rho/u of the previous step
        if kind in (H.NEEM_PRESSURE, H.RULE_VELOCITY_BB, H.RULE_PRESSURE_ABC):
            code[flat] |= np.uint16(H.CODE_KEEP_MACRO)       # own rho (ABC: u) of the previous step
    return code, neem_rules
