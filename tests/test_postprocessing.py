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
