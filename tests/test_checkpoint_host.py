"""Checkpoint files: what load() refuses before touching the device."""
import contextlib
import io

import numpy as np
import pytest

from taichi_lbm import cases, checkpoint


def _case():
    with contextlib.redirect_stdout(io.StringIO()):
        return cases.cylinder(NX=40, NY=20, cx=10, cy=10, r=3)


def test_load_rejects_foreign_files(tmp_path):
    case = _case()
    p = tmp_path / "x.npz"
    np.savez(p, f=np.zeros(3))
    with pytest.raises(ValueError, match="not a checkpoint"):
        checkpoint.load(p, case["lbf"], case["coll"])


def test_load_rejects_other_shapes(tmp_path):
    case = _case()
    p = tmp_path / "y.npz"
    np.savez(p, format=np.int32(checkpoint.FORMAT), shape=np.array([41, 20, 1]))
    with pytest.raises(ValueError, match="does not match"):
        checkpoint.load(p, case["lbf"], case["coll"])
    np.savez(p, format=np.int32(checkpoint.FORMAT), shape=np.array([40, 20, 1]))
    with pytest.raises(ValueError, match="no relaxation state"):
        checkpoint.load(p, case["lbf"], case["coll"])
