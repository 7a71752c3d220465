"""Host-side rules of the solve server (no GPU)."""
import math

from taichi_lbm.serving import _residual


def test_residual_and_divergence_rule():
    assert _residual(4.0, 16.0) == (0.5, False)
    assert _residual(0.0, 0.0) == (math.inf, False)          # a field at rest never converges by this test
    for d2, u2 in ((math.nan, 1.0), (1.0, math.inf), (math.inf, math.inf), (math.nan, math.nan)):
        res, failed = _residual(d2, u2)
        assert failed and math.isnan(res)
        assert not res < 1e-6                                   # a failed solve is never "converged"


def _spec(NX, NY):
    import numpy as np
    from taichi_lbm import _hip as H
    from taichi_lbm.batch import LatticeSpec
    code = np.full(NX * NY, H.CODE_ACTIVE, np.uint16)
    return LatticeSpec(NX, NY, code, [], H.Lattice(), np.ones((NX, NY), np.float32), np.zeros((NX, NY, 2), np.float32))


def test_request_that_can_never_fit_is_named():
    # SolveServer.serve finishes such a request at once as failed instead of
    # letting it hold the arrival-order queue (and the loop) forever
    from taichi_lbm.batch import footprint_bytes
    from taichi_lbm.serving import never_fits
    small, big = _spec(64, 32), _spec(2048, 512)
    budget = 4 * footprint_bytes(small)
    assert never_fits(small, "f32", budget, 1 << 20) is None
    assert "HBM budget" in never_fits(big, "f32", budget, 1 << 30)
    assert "arena" in never_fits(big, "f32", 1 << 40, 4096)
    assert never_fits(small, "f32", budget, 1 << 20, max_slots=0) is not None
