"""Host-side post-processing -- drop-in for taichi_lbm/PostProcessing.py.

Images are min-max normalised fields read back with ``to_numpy``; the VTK
writers produce the reference's ASCII STRUCTURED_GRID layout
(PostProcessing.py:70-156) so existing ParaView pipelines read the files
unchanged.
"""
from __future__ import annotations

import numpy as np


class PostProcessingEngine:
    """PostProcessing.py:5-156."""

    def __init__(self, num_curves):
        self.num_curves = num_curves
        self.fig = None
        if num_curves:
            import matplotlib.pyplot as plt
            self.fig, axes = plt.subplots(num_curves, 1, figsize=(8, 6 * num_curves))
            self.axes = [axes] if num_curves == 1 else list(axes)
            plt.ion()

    def update_plot(self, data_list):
        for ax, (x, y) in zip(self.axes, data_list):
            ax.clear()
            ax.plot(x, y, lw=2)
            ax.set_xlabel("X")
            ax.set_ylabel("Pressure")
            ax.grid(True)
        self.fig.canvas.draw_idle()

    def show(self):
        import matplotlib.pyplot as plt
        plt.show(block=False)

    def close(self):
        if self.fig is not None:
            import matplotlib.pyplot as plt
            plt.close(self.fig)

    @staticmethod
    def nomalized_field(data):
        lo, hi = np.min(data), np.max(data)
        return (data - lo) / (hi - lo)

    def post_pressure(self, lb_field):
        return self.nomalized_field(lb_field.total_pressure.to_numpy())

    def post_MC_pressure(self, lb_field):
        p = lb_field.pressure.to_numpy()
        return np.concatenate([self.nomalized_field(p[:, :, c]) for c in range(lb_field.nc)], axis=1)

    def post_denstiy(self, lb_field):
        return self.nomalized_field(lb_field.rho.to_numpy()[:, :, 0])

    def post_vel(self, lb_field):
        v = lb_field.vel.to_numpy()
        return self.nomalized_field(np.hypot(v[:, :, 0], v[:, :, 1]))

    @staticmethod
    def _grid(fout, NX, NY):
        fout.write("# vtk DataFile Version 3.0\n")
        fout.write("Hydrodynamics representation\n")
        fout.write("ASCII\n\n")
        fout.write("DATASET STRUCTURED_GRID\n")
        fout.write(f"DIMENSIONS {NX} {NY} 1\n")
        fout.write(f"POINTS {NX * NY} double\n")
        X, Y = np.meshgrid(np.arange(NX), np.arange(NY))     # y slowest, as the reference writes
        np.savetxt(fout, np.column_stack((X.ravel(), Y.ravel(), np.zeros(NX * NY))), fmt="%.0f")
        fout.write("\n")
        fout.write(f"POINT_DATA {NX * NY}\n")

    def writeVTK(self, fname, lb_field, Tflag=None):
        """Density (x C_rho), optional T, velocity (x Cu) on the NX x NY grid."""
        rho = lb_field.total_rho.to_numpy().T.ravel()
        vel = lb_field.vel.to_numpy()
        filename = fname + ".vtk"
        with open(filename, "w") as fout:
            self._grid(fout, lb_field.NX, lb_field.NY)
            fout.write("SCALARS density double\n")
            fout.write("LOOKUP_TABLE density_table\n")
            np.savetxt(fout, rho * lb_field.C_rho, fmt="%.8f")
            if Tflag is not None:
                fout.write("SCALARS T double\n")
                fout.write("LOOKUP_TABLE T_table\n")
                np.savetxt(fout, lb_field.T.to_numpy().T.ravel(), fmt="%.8f")
            fout.write("VECTORS velocity double\n")
            vx, vy = vel[:, :, 0].T.ravel(), vel[:, :, 1].T.ravel()
            np.savetxt(fout, np.column_stack((vx * lb_field.Cu, vy * lb_field.Cu, np.zeros_like(vx))), fmt="%.8f")
        print(filename)

    def writeVTKpsi(self, fname, lb_field):
        """psi of component 0 on the (NX+2) x (NY+2) halo grid."""
        psi = lb_field.sc_field.psi_field.to_numpy()[:, :, 0].T.ravel()
        filename = fname + ".vtk"
        with open(filename, "w") as fout:
            self._grid(fout, lb_field.NX + 2, lb_field.NY + 2)
            fout.write("SCALARS PSI double\n")
            fout.write("LOOKUP_TABLE T_table\n")
            np.savetxt(fout, psi, fmt="%.8f")
        print(filename)
