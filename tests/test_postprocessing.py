"""VTK output layout (PostProcessing.py:70-120) on a duck-typed field holder."""
import numpy as np

from taichi_lbm.PostProcessing import PostProcessingEngine


class _F:
    def __init__(self, a):
        self.a = a

    def to_numpy(self):
        return self.a


class _Lat:
    def __init__(self, NX, NY):
        rng = np.random.default_rng(0)
        self.NX, self.NY, self.nc = NX, NY, 1
        self.total_rho = _F(rng.random((NX, NY)))
        self.vel = _F(rng.random((NX, NY, 2)))
        self.T = _F(np.ones((NX, NY)))
        self.C_rho, self.Cu = 1000.0, 0.5


def test_write_vtk_layout(tmp_path):
    lat = _Lat(3, 2)
    PostProcessingEngine(0).writeVTK(str(tmp_path / "out"), lat)
    lines = (tmp_path / "out.vtk").read_text().splitlines()
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[4] == "DATASET STRUCTURED_GRID"
    assert lines[5] == "DIMENSIONS 3 2 1" and lines[6] == "POINTS 6 double"
    pts = [tuple(map(float, l.split())) for l in lines[7:13]]
    assert pts == [(0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 1, 0), (1, 1, 0), (2, 1, 0)]   # x fastest
    assert lines[14] == "POINT_DATA 6" and lines[15] == "SCALARS density double"
    dens = np.array(list(map(float, lines[17:23])))
    assert np.allclose(dens, lat.total_rho.a.T.ravel() * 1000.0, atol=1e-7)
    assert lines[23] == "VECTORS velocity double"
    vel = np.array([list(map(float, l.split())) for l in lines[24:30]])
    assert np.allclose(vel[:, 0], lat.vel.a[:, :, 0].T.ravel() * 0.5, atol=1e-7)


def test_normalised_velocity_image():
    lat = _Lat(4, 3)
    img = PostProcessingEngine(0).post_vel(lat)
    assert img.shape == (4, 3) and img.min() == 0.0 and img.max() == 1.0


def _vtk_blocks(text):
    """Parse the ASCII STRUCTURED_GRID layout: dimensions, points and the
    named data blocks (scalars as 1-D arrays, vectors as (n, 3))."""
    lines = text.splitlines()
    dims = tuple(int(v) for v in lines[5].split()[1:])
    npts = int(lines[6].split()[1])
    pts = np.array([list(map(float, l.split())) for l in lines[7:7 + npts]])
    i = lines.index(f"POINT_DATA {npts}") + 1
    blocks = {}
    while i < len(lines):
        head = lines[i].split()
        if head[0] == "SCALARS":
            assert lines[i + 1].startswith("LOOKUP_TABLE ")
            blocks[head[1]] = (lines[i + 1].split()[1], np.array(list(map(float, lines[i + 2:i + 2 + npts]))))
            i += 2 + npts
        elif head[0] == "VECTORS":
            blocks[head[1]] = (None, np.array([list(map(float, l.split())) for l in lines[i + 1:i + 1 + npts]]))
            i += 1 + npts
        else:
            raise AssertionError(f"unexpected line {lines[i]!r}")
    return dims, pts, blocks


def test_write_vtk_temperature_block(tmp_path):
    """Tflag adds SCALARS T (lookup table T_table) between density and
    velocity (PostProcessing.py:104-107)."""
    lat = _Lat(4, 3)
    lat.T = _F(np.arange(12, dtype=np.float64).reshape(4, 3) / 7.0)
    PostProcessingEngine(0).writeVTK(str(tmp_path / "t"), lat, Tflag=True)
    dims, pts, blocks = _vtk_blocks((tmp_path / "t.vtk").read_text())
    assert dims == (4, 3, 1) and list(blocks) == ["density", "T", "velocity"]
    assert blocks["density"][0] == "density_table" and blocks["T"][0] == "T_table"
    assert np.allclose(blocks["T"][1], lat.T.a.T.ravel(), atol=5e-9)
    # point p is node (x, y) = pts[p][:2], and every block is in that order
    for p in (0, 5, 11):
        x, y = int(pts[p, 0]), int(pts[p, 1])
        assert abs(blocks["T"][1][p] - lat.T.a[x, y]) < 5e-9
        assert abs(blocks["velocity"][1][p, 1] - 0.5 * lat.vel.a[x, y, 1]) < 5e-9


class _Psi:
    def __init__(self, a):
        self.psi_field = _F(a)


def test_write_vtk_psi_layout(tmp_path):
    """writeVTKpsi (PostProcessing.py:123-156): component 0 of psi on the
    (NX+2) x (NY+2) halo grid, halo index 0 at point coordinate 0."""
    NX, NY = 5, 3
    lat = _Lat(NX, NY)
    psi = np.random.default_rng(1).random((NX + 2, NY + 2, 2))
    lat.sc_field = _Psi(psi)
    PostProcessingEngine(0).writeVTKpsi(str(tmp_path / "psi"), lat)
    text = (tmp_path / "psi.vtk").read_text()
    dims, pts, blocks = _vtk_blocks(text)
    assert dims == (NX + 2, NY + 2, 1) and len(pts) == (NX + 2) * (NY + 2)
    assert list(blocks) == ["PSI"] and blocks["PSI"][0] == "T_table"
    vals = blocks["PSI"][1]
    for p in range(len(pts)):
        x, y = int(pts[p, 0]), int(pts[p, 1])
        assert abs(vals[p] - psi[x, y, 0]) < 5e-9
    assert not np.allclose(vals, psi[:, :, 1].T.ravel(), atol=1e-3)   # component 1 is not written


def test_writing_boundary_png(tmp_path):
    """writing_boundary (BoundaryCondition.py:486-510): an NX x NY RGB image,
    each boundary group painted in the colour of its insertion index, later
    groups over earlier ones, nodes in no group black."""
    from PIL import Image
    from taichi_lbm import cases
    case = cases.cylinder(NX=40, NY=20, cx=12, cy=10, r=4)
    bce = case["bce"]
    path = tmp_path / "boundary.png"
    bce.writing_boundary(case["lbf"], path=str(path))
    img = np.asarray(Image.open(path))
    assert img.shape == (40, 20, 3) and img.dtype == np.uint8
    palette = [(0, 0, 0), (51, 152, 219), (255, 0, 0), (143, 24, 172), (255, 190, 53), (255, 0, 0)]
    owner = -np.ones((40, 20), int)
    for num, bc in enumerate(bce.boundary_conditions.values()):
        nodes = bc.group.nodes
        owner[nodes[:, 0], nodes[:, 1]] = num
    assert (owner >= 0).any() and (owner < 0).any()
    for num in range(-1, len(bce.boundary_conditions)):
        sel = owner == num
        if sel.any():
            want = (0, 0, 0) if num < 0 else palette[num]
            # 52/255 of the reference's colour list comes back as 51 after
            # the float32 x 255 truncation; the others are exact
            assert (np.abs(img[sel].astype(int) - np.array(want)) <= 1).all(), num
