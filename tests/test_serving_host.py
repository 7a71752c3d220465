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
