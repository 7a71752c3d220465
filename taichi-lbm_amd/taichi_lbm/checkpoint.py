"""Checkpoint / resume of one lattice (NumPy ``.npz``).

The reference has no restart: its VTK writers are output-only
(PostProcessing.py:70-120) and a case's state lives in Taichi fields until
the process ends.  ``save`` writes what a run continues from -- the
reference-visible fields after lbm_solve() (f, f2, rho, total_rho, vel, the
force and observable fields), the collision's relaxation state (tau /
tau_sym / 1/tau of MRT, per node), Shan-Chen psi with its halo and the step
counter ``lb_field.time`` -- and ``load`` puts them back into a lattice of
the same shape.  A run continued from a checkpoint gives the fields of the
uninterrupted run bit for bit (16-bit storage too: the file keeps the step
count that seeds its stochastic rounding).  The file holds plain arrays only (loaded
with ``allow_pickle=False``).
"""
from __future__ import annotations

import numpy as np

from ._device import torch

FORMAT = 1
_FIELDS = ("f", "f2", "rho", "total_rho", "vel", "body_force", "SCforce", "shear_rate", "pressure",
           "total_pressure", "T")


def save(path, lb_field, collision=None) -> None:
    """Writes the lattice's state (and the collision's relaxation state) to ``path``."""
    lbf = lb_field
    lbf.materialize()
    arrays = {"format": np.int32(FORMAT), "shape": np.array([lbf.NX, lbf.NY, lbf.nc], np.int64),
              "time": np.int64(lbf.time[None]), "round_step": np.int64(lbf._round_step)}
    for name in _FIELDS:
        arrays[name] = getattr(lbf, name).to_numpy()
    if collision is not None:
        arrays["relaxation_state"] = collision.state_tensor()[:, :lbf.NX * lbf.NY].cpu().numpy()
        arrays["relaxation_uniform"] = np.bool_(collision._uniform)
    if lbf.sc_field is not None:
        arrays["psi"] = lbf.sc_field.psi_field.to_numpy()
    np.savez(path, **arrays)


def load(path, lb_field, collision=None) -> None:
    """Restores a state written by ``save`` into a lattice of the same shape
    (and the same geometry, collision model and boundary list: those are
    code, not state).  Raises ValueError on a shape or content mismatch."""
    lbf = lb_field
    with np.load(path, allow_pickle=False) as z:
        if "format" not in z.files or int(z["format"]) != FORMAT:
            raise ValueError(f"{path}: not a checkpoint of format {FORMAT}")
        shape = tuple(int(v) for v in z["shape"])
        if shape != (lbf.NX, lbf.NY, lbf.nc):
            raise ValueError(f"{path}: lattice {shape} does not match NX, NY, nc = {(lbf.NX, lbf.NY, lbf.nc)}")
        if collision is not None and "relaxation_state" not in z.files:
            raise ValueError(f"{path}: no relaxation state was saved")
        if lbf.sc_field is not None and "psi" not in z.files:
            raise ValueError(f"{path}: no Shan-Chen psi was saved")
        data = {k: z[k] for k in z.files}
    lbf.materialize()
    for name in _FIELDS:
        getattr(lbf, name).from_numpy(data[name])
    lbf.time[None] = int(data["time"])
    lbf._round_step = int(data["round_step"]) if "round_step" in data else 0
    if collision is not None:
        st = collision.state_tensor()
        st[:, :lbf.NX * lbf.NY].copy_(torch().as_tensor(data["relaxation_state"]))
        collision._uniform = bool(data["relaxation_uniform"])
    if lbf.sc_field is not None:
        lbf.sc_field.psi_field.from_numpy(data["psi"])
