"""Host checks that keep the granular boundary kernels' addresses inside the
lattice (BoundaryCondition.check_granular_bounds; no GPU).  The reference
reads out of range silently; here a bad group or table raises before launch."""
import numpy as np
import pytest

from taichi_lbm import _hip as H
from taichi_lbm.BoundaryCondition import check_granular_bounds
from taichi_lbm.lbm import neighbor_table


def _table(NX, NY):
    mask = np.ones((NX, NY), np.int32)
    mask[3, 2] = -1
    return np.asarray(neighbor_table(mask))


def test_valid_groups_pass():
    NX, NY = 6, 5
    tab = _table(NX, NY)
    every = np.argwhere(np.ones((NX, NY), bool))
    check_granular_bounds(H.BC_BOUNCEBACK, every, NX, NY, table=tab)
    check_granular_bounds(H.BC_INSIDE, every, NX, NY, table=tab)
    check_granular_bounds(H.BC_PERIODIC, every, NX, NY)
    check_granular_bounds(H.BC_VELOCITY_BB, [[0, j] for j in range(NY)], NX, NY, direction=3)
    check_granular_bounds(H.BC_VELOCITY_NEEM, [[0, j] for j in range(NY)], NX, NY, direction=3)
    check_granular_bounds(H.BC_PRESSURE_NEEM, [[NX - 1, j] for j in range(NY)], NX, NY, direction=1)
    check_granular_bounds(H.BC_OPEN_NEUMANN, [[i, NY - 1] for i in range(NX)], NX, NY, direction=2)


@pytest.mark.parametrize("kind,d,line", [
    (H.BC_VELOCITY_NEEM, 1, "left"),      # x - c_1 = (-1, j)
    (H.BC_PRESSURE_NEEM, 3, "right"),     # x - c_3 = (NX, j)
    (H.BC_OPEN_NEUMANN, 4, "top"),        # x - c_4 = (i, NY)
    (H.BC_VELOCITY_NEEM, 5, "left"),      # not a plane direction
])
def test_rule_reading_outside_is_rejected(kind, d, line):
    NX, NY = 6, 5
    nodes = {"left": [[0, j] for j in range(NY)], "right": [[NX - 1, j] for j in range(NY)],
             "top": [[i, NY - 1] for i in range(NX)]}[line]
    with pytest.raises(H.LBMError):
        check_granular_bounds(kind, nodes, NX, NY, direction=d)


def test_edited_table_entry_is_rejected():
    NX, NY = 6, 5
    tab = _table(NX, NY).copy()
    tab[2, 2, 5] = (NX, 1)                  # neither -1 nor a lattice node
    with pytest.raises(H.LBMError, match="neighbour table"):
        check_granular_bounds(H.BC_INSIDE, [[2, 2]], NX, NY, table=tab)
    tab[2, 2, 5] = (-1, 3)                  # half of a -1 pair
    with pytest.raises(H.LBMError):
        check_granular_bounds(H.BC_BOUNCEBACK, [[2, 2]], NX, NY, table=tab)


def test_group_node_outside_is_rejected():
    with pytest.raises(H.LBMError, match="outside"):
        check_granular_bounds(H.BC_PERIODIC, [[6, 0]], 6, 5)
