"""Host-side geometry and boundary-list compilation (no GPU).

Classification and the neighbour table are compared bit-exactly (groups as
sets) with the oracle's scalar restatement of the reference kernels; the
compiled node codes are checked by replaying the boundary list in the oracle
on tagged populations."""
import numpy as np
import pytest

from oracle import lbm_oracle as O
from taichi_lbm import cases
from taichi_lbm import _hip as H
from taichi_lbm._codes import BCSpec, UnsupportedBoundaryList, compile_codes, neighbor_table


def _set(a):
    return {tuple(x) for x in np.asarray(a).reshape(-1, 2).tolist()}


def test_mask_and_groups_match_scalar_restatement():
    NX, NY = 60, 30
    case = cases.cylinder(NX=NX, NY=NY, cx=17.3, cy=14.6, r=5.2)
    lat = O.Lattice(NX, NY, 1)
    O.mask_circle_identify(lat, 17.3, 14.6, 5.2)
    assert np.array_equal(lat.mask, case["lbf"].mask.arr)
    m = lat.mask

    def wall(i, j):   # scalar predicate in the reference's style, bounds-checked
        if m[i, j] != 1:
            return False
        for k in range(9):
            i2, j2 = i - O.C[k, 0], j - O.C[k, 1]
            if j2 < 0 or j2 >= NY:
                return True
            if 0 <= i2 < NX and m[i2, j2] == -1:
                return True
        return False

    preds = {"fluid": lambda i, j: m[i, j] == 1, "periodic": lambda i, j: m[i, j] == 1, "wall": wall,
             "inlet": lambda i, j: m[i, j] == 1 and i == 0 and 0 < j < NY - 1,
             "outlet": lambda i, j: m[i, j] == 1 and i == NX - 1 and 0 < j < NY - 1}
    for name, bc in case["bce"].boundary_conditions.items():
        assert _set(bc.group.nodes) == _set(O.classify_scalar(NX, NY, preds[name])), name
        assert bc.group.count[None] == len(bc.group.nodes)


def test_rectangle_mask():
    from taichi_lbm import BoundaryEngine, LBField
    lbf = LBField("r", 40, 20, 1)
    BoundaryEngine().Mask_rectangle_identify(lbf, 9.5, 20.5, 3.5, 11.5)
    lat = O.Lattice(40, 20, 1)
    O.mask_rectangle_identify(lat, 9.5, 20.5, 3.5, 11.5)
    assert np.array_equal(lbf.mask.arr, lat.mask)
    assert (lat.mask == -1).sum() == 11 * 8


def test_neighbor_table_matches_scalar_restatement():
    rng = np.random.default_rng(0)
    lat = O.Lattice(17, 13, 1)
    lat.mask[rng.random((17, 13)) < 0.2] = -1
    O.neighbor_classify(lat)
    assert np.array_equal(neighbor_table(lat.mask), lat.neighbor)


def test_scalar_predicates_are_accepted():
    from taichi_lbm import BoundaryClassifier, BoundarySpec
    clf = BoundaryClassifier((6, 5))

    def pred(i, j):          # Python control flow: evaluated node by node
        if i == 2:
            return True
        return j == 4
    g = clf.get_group(BoundarySpec(pred))
    assert _set(g.nodes) == {(2, j) for j in range(5)} | {(i, 4) for i in range(6)}


def test_out_of_range_field_read_raises():
    case = cases.cylinder(NX=20, NY=10, cx=5, cy=5, r=2)
    mask = case["lbf"].mask
    from taichi_lbm import BoundaryClassifier, BoundarySpec
    with pytest.raises(IndexError):
        BoundaryClassifier((20, 10)).get_group(BoundarySpec(lambda I, J: mask[I - 1, J] == 1))


def _replay(NX, NY, mask, bcs_oracle, active):
    """Oracle replay of the boundary list with f2(x, k) = unique tag."""
    lat = O.Lattice(NX, NY, 1, dtype=np.float64)
    lat.mask[...] = mask
    O.neighbor_classify(lat)
    tags = np.arange(NX * NY * 9, dtype=np.float64).reshape(NX, NY, 1, 9) + 1.0
    lat.f2[...] = tags
    lat.f[...] = np.nan
    O.apply_bcs(lat, bcs_oracle)
    return lat.f[..., 0, :], lat


def _expected_from_codes(code, NX, NY, mask):
    exp = np.full((NX, NY, 9), np.nan)
    for i in range(NX):
        for j in range(NY):
            c = int(code[i * NY + j])
            if not c & H.CODE_ACTIVE:
                continue
            for k in range(9):
                if k > 0 and (c >> (k - 1)) & 1:
                    src = (i, j, O.OPP[k])
                else:
                    src = ((i - O.C[k, 0]) % NX, (j - O.C[k, 1]) % NY, k)
                exp[i, j, k] = (src[0] * NY + src[1]) * 9 + src[2] + 1.0
    return exp


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_codes_reproduce_stream_and_bounce_rules(seed):
    rng = np.random.default_rng(seed)
    NX, NY = 14, 11
    mask = np.ones((NX, NY), np.int32)
    mask[rng.random((NX, NY)) < 0.15] = -1
    nb = neighbor_table(mask)
    fluid = np.argwhere(mask == 1)
    wall = np.argwhere((mask == 1) & np.any(nb[..., 0] == -1, axis=2))
    inside = np.argwhere((mask == 1) & np.any(nb[..., 0] != -1, axis=2))
    order = [("periodic", fluid), ("bounceback", wall)] if seed != 2 else \
        [("inside", inside), ("bounceback", wall)]
    if seed == 1:
        order = [("bounceback", wall), ("periodic", fluid[fluid[:, 0] % 2 == 0]), ("bounceback", wall[::2])]
        # nodes of odd columns not on the wall get no writer -> rejected
        with pytest.raises(UnsupportedBoundaryList):
            compile_codes(NX, NY, fluid, [BCSpec(k, g) for k, g in order], nb)
        order = [("periodic", fluid), ("bounceback", wall[::2]), ("periodic", fluid[::3]), ("bounceback", wall[1::2])]
    code, rules = compile_codes(NX, NY, fluid, [BCSpec(k, g) for k, g in order], nb)
    assert rules == []
    got, _ = _replay(NX, NY, mask, [O.BC(k, g) for k, g in order], fluid)
    exp = _expected_from_codes(code, NX, NY, mask)
    act = mask == 1
    assert np.array_equal(got[act], exp[act])
    # PULLED marks exactly the solid nodes whose f2 the replay streams into an
    # active node (tag t = node * 9 + k + 1)
    src_nodes = (got[act].astype(np.int64).reshape(-1) - 1) // 9
    pulled = np.zeros(NX * NY, bool)
    pulled[src_nodes] = True
    pulled &= (mask == -1).reshape(-1)
    assert np.array_equal((code & H.CODE_PULLED) != 0, pulled)
    assert np.all(code[(mask == -1).reshape(-1)] & ~np.uint16(H.CODE_PULLED) == 0)
    if seed == 1:
        assert pulled.any()     # periodic rules over wall nodes pull from solid nodes


def test_cylinder_codes_flags():
    NX, NY = 40, 20
    case = cases.cylinder(NX=NX, NY=NY, cx=10, cy=10, r=3)
    lbf, bce = case["lbf"], case["bce"]
    code, rules = compile_codes(NX, NY, case["fluid"].group.nodes, bce.fused_list(), lbf.neighbor.arr)
    assert [r[:2] for r in rules] == [(H.NEEM_VELOCITY, 3), (H.NEEM_PRESSURE, 1)]
    c2 = code.reshape(NX, NY)
    e = (c2 & H.CODE_NEEM_MASK) >> H.CODE_NEEM_SHIFT
    assert set(map(tuple, np.argwhere(e == 1))) == {(0, j) for j in range(1, NY - 1)}
    assert set(map(tuple, np.argwhere(e == 2))) == {(NX - 1, j) for j in range(1, NY - 1)}
    keep = np.argwhere(c2 & H.CODE_KEEP_MACRO)
    assert {(1, j) for j in range(1, NY - 1)} <= set(map(tuple, keep))
    assert {(NX - 1, j) for j in range(1, NY - 1)} <= set(map(tuple, keep))   # own rho (PressureNEEM)
    # corners of the channel bounce the x-populations off the lattice edge (non-periodic table)
    assert c2[0, 0] & (1 << 0)      # k = 1 at (0, 0): source (-1, 0) outside -> bounce
    assert c2[5, 5] == H.CODE_ACTIVE


def test_rejections():
    NX, NY = 8, 6
    mask = np.ones((NX, NY), np.int32)
    mask[3, 3] = -1
    nb = neighbor_table(mask)
    fluid = np.argwhere(mask == 1)
    allnodes = np.argwhere(np.ones((NX, NY), bool))
    with pytest.raises(UnsupportedBoundaryList, match="outside the fluid group"):
        compile_codes(NX, NY, fluid, [BCSpec("periodic", allnodes)], nb)
    with pytest.raises(UnsupportedBoundaryList, match="-1"):
        compile_codes(NX, NY, fluid, [BCSpec("inside", fluid)], nb)
    line = np.array([[0, j] for j in range(1, NY - 1)])
    line2 = np.array([[1, j] for j in range(1, NY - 1)])
    with pytest.raises(UnsupportedBoundaryList, match="NEEM"):
        compile_codes(NX, NY, fluid, [BCSpec("periodic", fluid), BCSpec("velocity_neem", line, 3, 0.05),
                                      BCSpec("velocity_neem", line2, 3, 0.05)], nb)
    with pytest.raises(UnsupportedBoundaryList, match="outside the lattice"):
        compile_codes(NX, NY, fluid, [BCSpec("periodic", fluid), BCSpec("velocity_neem", line, 1, 0.05)], nb)


def test_reference_cylinder_codes():
    """cases/Circular cylinder.py's list: VelocityBB inlet adds to the bounced
    unknowns of column 0 (own rho kept), OpenNeumann outlet re-gathers column
    NX-2 (which is plain streaming and final)."""
    NX, NY = 60, 20
    case = cases.cylinder_reference_bcs(NX=NX, NY=NY)
    lbf, bce = case["lbf"], case["bce"]
    code, rules = compile_codes(NX, NY, case["fluid"].group.nodes, bce.fused_list(), lbf.neighbor.arr)
    assert [r[:2] for r in rules] == [(H.RULE_VELOCITY_BB, 3), (H.RULE_OPEN_NEUMANN, 1)]
    c2 = code.reshape(NX, NY)
    e = (c2 & H.CODE_NEEM_MASK) >> H.CODE_NEEM_SHIFT
    assert set(map(tuple, np.argwhere(e == 1))) == {(0, j) for j in range(NY)}
    assert set(map(tuple, np.argwhere(e == 2))) == {(NX - 1, j) for j in range(NY)}
    assert np.all(c2[0] & H.CODE_KEEP_MACRO)                      # VelocityBB reads its own rho
    assert not np.any(c2[NX - 2] & H.CODE_KEEP_MACRO)
    for k in (1, 5, 8):                                            # unknowns of the inlet bounce first
        assert np.all(c2[0] & (1 << (k - 1)))
    for k in (3, 6, 7):
        assert np.all(c2[NX - 1] & (1 << (k - 1)))


def test_rule_rejections():
    NX, NY = 8, 6
    mask = np.ones((NX, NY), np.int32)
    nb = neighbor_table(mask)
    fluid = np.argwhere(mask == 1)
    col = lambda i: np.array([[i, j] for j in range(NY)])
    # OpenNeumann's neighbour rewritten by a later boundary
    with pytest.raises(UnsupportedBoundaryList, match="rewrites"):
        compile_codes(NX, NY, fluid, [BCSpec("periodic", fluid), BCSpec("open_neumann", col(NX - 1), 1),
                                      BCSpec("bounceback", col(NX - 2))], nb)
    # OpenNeumann whose neighbour carries a VelocityBB rule
    with pytest.raises(UnsupportedBoundaryList, match="node rule"):
        compile_codes(NX, NY, fluid, [BCSpec("periodic", fluid), BCSpec("velocity_bb", col(NX - 2), 1, 0.01),
                                      BCSpec("open_neumann", col(NX - 1), 1)], nb)
    # a later boundary overwriting a VelocityBB node
    with pytest.raises(UnsupportedBoundaryList, match="node rule"):
        compile_codes(NX, NY, fluid, [BCSpec("periodic", fluid), BCSpec("velocity_bb", col(0), 3, 0.01),
                                      BCSpec("bounceback", col(0))], nb)


def test_row_padding_only_without_wrapping_pulls():
    """A batch lattice is stored with rows padded to 32 nodes only when no
    active node streams across the j boundary (checked by brute force over
    the boundary columns here)."""
    from taichi_lbm.batch import LatticeSpec, pad_rows
    NX, NY = 9, 21
    mask = np.ones((NX, NY), np.int32)
    mask[:, 0] = mask[:, -1] = -1            # solid walls: the rows next to them bounce
    nb = neighbor_table(mask)
    fluid = np.argwhere(mask == 1)
    wall = np.argwhere((mask == 1) & np.any(nb[..., 0] == -1, axis=2))
    code, _ = compile_codes(NX, NY, fluid, [BCSpec("periodic", fluid), BCSpec("bounceback", wall)], nb)
    L = H.Lattice()
    L.nx, L.ny, L.nc = NX, NY, 1
    spec = LatticeSpec(NX, NY, code, [], L, np.ones((NX, NY)), np.zeros((NX, NY, 2)))
    padded = pad_rows(spec)
    assert padded.NY == 32 and padded.lattice.ny == 32 and padded.tag["view_ny"] == NY
    c2 = padded.code.reshape(NX, 32)
    assert np.array_equal(c2[:, :NY].reshape(-1), code) and not c2[:, NY:].any()
    # the same lattice periodic across j (no walls): fluid nodes at j = 0 pull from j = NY-1
    mask[:] = 1
    nb = neighbor_table(mask)
    fluid = np.argwhere(mask == 1)
    code, _ = compile_codes(NX, NY, fluid, [BCSpec("periodic", fluid)], nb)
    spec = LatticeSpec(NX, NY, code, [], L, np.ones((NX, NY)), np.zeros((NX, NY, 2)))
    assert pad_rows(spec) is spec
