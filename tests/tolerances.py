"""Parity tolerances between two FP32 evaluations of the same step.

* Populations and zeroth moments (f, f2, rho, total_rho, pressure, psi):
  max |a - b| <= 1e-5 * max |ref|  (relative to the field's scale).
* First moments and forces (vel, body_force, SCforce): absolute 1e-6 lattice
  units (plus 1e-5 of the field's scale).  u = sum_k f_k c_k / rho is a
  cancellation of O(1) populations, so its FP32 rounding floor is ~1e-7 in
  absolute terms however slow the flow; a relative bound would test nothing
  but that floor.
* Shear rate: absolute 1e-6 (plus 1e-4 relative): it is a second moment of
  f - feq, a cancellation of O(0.1) numbers.
* Relaxation state tau = clamp(3 m g^(n-1) + 0.5): the shear-rate bound
  propagated through d tau / d g = 3 m (n-1) g^(n-2), plus 1e-5 relative.
* Long runs (10^4 steps) compare two FP32 restatements whose per-step
  rounding differences accumulate: `growth` widens every bound (10 there,
  against sqrt(10^4) = 100 for an unbiased random walk); the tests also state
  each run's distance to the FP64 path.
"""
import numpy as np

POP_RTOL = 1e-5
MOMENT_ATOL = 1e-6
SHEAR_ATOL = 1e-6

SCALE_FIELDS = {"f", "f2", "rho", "total_rho", "pressure", "total_pressure", "psi"}
MOMENT_FIELDS = {"vel", "body_force", "SCforce"}


def check(name, got, ref, log=True, growth=1.0):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (name, got.shape, ref.shape)
    assert np.all(np.isfinite(got)), f"{name}: non-finite values"
    scale = max(float(np.abs(ref).max()), 1e-30)
    err = np.abs(got - ref)
    if name in SCALE_FIELDS:
        bound = POP_RTOL * scale
    elif name in MOMENT_FIELDS:
        bound = MOMENT_ATOL + POP_RTOL * scale
    elif name == "shear_rate":
        bound = SHEAR_ATOL + 1e-4 * np.abs(ref)
    else:
        bound = POP_RTOL * scale
    bound = bound * growth
    ok = np.all(err <= bound)
    if log:
        print(f"  {name}: max abs err {err.max():.3e} (scale {scale:.3e})", flush=True)
    assert ok, f"{name}: max err {err.max():.3e} exceeds the bound (scale {scale:.3e})"


def check_tau(name, got, ref, shear_ref, m, n, tau_min, tau_max, is_rate=False):
    """tau (or 1/tau for MRT when is_rate) against the reference, with the
    shear-rate floor propagated through the power law."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    g = np.maximum(np.asarray(shear_ref, np.float64), 1e-30)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        dtau = np.abs(3 * m * (n - 1) * g ** (n - 2)) * SHEAR_ATOL * 2
    dtau = np.nan_to_num(dtau, nan=np.inf, posinf=np.inf)
    tau_ref = 1.0 / ref if is_rate else ref
    tol = dtau + 1e-5 * np.abs(tau_ref)
    if is_rate:
        tol = tol / np.maximum(tau_ref, tau_min) ** 2
    err = np.abs(got - ref)
    print(f"  {name}: max abs err {err.max():.3e}; nodes over the flat 1e-5 bound "
          f"{int((err > 1e-5 * np.abs(ref)).sum())}", flush=True)
    assert np.all(err <= tol), f"{name}: err {err.max():.3e}"
