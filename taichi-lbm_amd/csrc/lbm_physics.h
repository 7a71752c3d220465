// Per-node D2Q9 physics shared by the fused and the granular kernels.
// Each function restates one reference formula (file:line cited inline).
#pragma once

#include "lbm_common.h"

#include <string>

namespace lbm {

// NEEMMethod.rhow, BoundaryCondition.py:118-129.
template <typename R>
__device__ __forceinline__ R neem_rhow(int d, R ux, R uy, const R (&f)[9]) {
  const R sx = f[0] + f[2] + f[4];   // populations with c_x = 0
  const R sy = f[0] + f[1] + f[3];   // populations with c_y = 0
  const R out1 = f[1] + f[5] + f[8], out3 = f[3] + f[6] + f[7];
  const R out2 = f[2] + f[5] + f[6], out4 = f[4] + f[7] + f[8];
  switch (d) {
    case 2: return R(1) / (R(1) + uy) * (sy + R(2) * out2);
    case 4: return R(1) / (R(1) - uy) * (sy + R(2) * out4);
    case 1: return R(1) / (R(1) + ux) * (sx + R(2) * out1);
    case 3: return R(1) / (R(1) - ux) * (sx + R(2) * out3);
  }
  return R(0);
}

__device__ __forceinline__ bool neem_unknown(int d, int k) {
  // PlaneBoundary.unknow, BoundaryCondition.py:63-72
  switch (d) {
    case 1: return k == 3 || k == 6 || k == 7;
    case 3: return k == 1 || k == 5 || k == 8;
    case 4: return k == 2 || k == 5 || k == 6;
    case 2: return k == 4 || k == 7 || k == 8;
  }
  return false;
}

// M f for the MRT moment basis (CollisionEngine.py:273-282, :379-385).
template <typename R>
__device__ __forceinline__ void mrt_m(const R* f, R* m) {
  const R s14 = f[1] + f[2] + f[3] + f[4];
  const R s58 = f[5] + f[6] + f[7] + f[8];
  m[0] = f[0] + s14 + s58;
  m[1] = R(-4) * f[0] - s14 + R(2) * s58;
  m[2] = R(4) * f[0] - R(2) * s14 + s58;
  m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
  m[4] = R(-2) * f[1] + R(2) * f[3] + f[5] - f[6] - f[7] + f[8];
  m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
  m[6] = R(-2) * f[2] + R(2) * f[4] + f[5] + f[6] - f[7] - f[8];
  m[7] = f[1] - f[2] + f[3] - f[4];
  m[8] = f[5] - f[6] + f[7] - f[8];
}

// f = M^-1 m with the exact inverse M^T diag(1/9,1/36,1/36,1/6,1/12,1/6,1/12,1/4,1/4).
template <typename R>
__device__ __forceinline__ void mrt_minv(const R* m, R* f) {
  const R a0 = m[0] * (R(1) / R(9)), a1 = m[1] * (R(1) / R(36)), a2 = m[2] * (R(1) / R(36));
  const R a3 = m[3] * (R(1) / R(6)), a4 = m[4] * (R(1) / R(12));
  const R a5 = m[5] * (R(1) / R(6)), a6 = m[6] * (R(1) / R(12));
  const R a7 = m[7] * R(0.25), a8 = m[8] * R(0.25);
  const R b = a0 - a1 - R(2) * a2;   // common to k = 1..4
  const R e = a0 + R(2) * a1 + a2;   // common to k = 5..8
  const R x = a3 + a4, y = a5 + a6;
  f[0] = a0 - R(4) * a1 + R(4) * a2;
  f[1] = b + a3 - R(2) * a4 + a7;
  f[2] = b + a5 - R(2) * a6 - a7;
  f[3] = b - a3 + R(2) * a4 + a7;
  f[4] = b - a5 + R(2) * a6 - a7;
  f[5] = e + x + y + a8;
  f[6] = e - x + y - a8;
  f[7] = e - x - y + a8;
  f[8] = e + x - y - a8;
}

// Strain-rate magnitude from a (filtered) non-equilibrium part, shared by
// BGK/TRT (CollisionEngine.py:86-92) and MRT (:359-377).
template <typename R>
__device__ __forceinline__ R shear_from(const R* q, R cof) {
  const R Sxx = (q[1] + q[3] + q[5] + q[6] + q[7] + q[8]) * cof;
  const R Syy = (q[2] + q[4] + q[5] + q[6] + q[7] + q[8]) * cof;
  const R Sxy = (q[5] - q[6] + q[7] - q[8]) * cof;
  return sqrt(R(2) * (Sxx * Sxx + Syy * Syy + R(2) * Sxy * Sxy));
}

template <int COLL, typename R>
__device__ __forceinline__ void collide(const R (&f)[9], R (&out)[9], R rho, R ux, R uy, R Fx, R Fy, bool force,
                                        R tau, R tau_anti, R s_shear, R s_bulk) {
  const R uv = ux * ux + uy * uy;
  if constexpr (COLL == LBM_BGK) {
    const R omega = rcp(tau);
    const R uF = ux * Fx + uy * Fy;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const R fe = feq_k(k, ux, uy, rho, uv);
      R v = f[k] + (-omega * (f[k] - fe));
      if (force) v = v + (R(1.0) - R(0.5) * omega) * fforce_k(k, Fx, Fy, ux, uy, uF);
      out[k] = v;
    }
  } else if constexpr (COLL == LBM_TRT) {
    const R ws = rcp(tau), wa = rcp(tau_anti);
    const R uF = ux * Fx + uy * Fy;
    R fe[9], ff[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      fe[k] = feq_k(k, ux, uy, rho, uv);
      ff[k] = force ? fforce_k(k, Fx, Fy, ux, uy, uF) : R(0.0);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int o = opp(k);
      const R fs = k == 0 ? f[0] : (f[k] + f[o]) / R(2.0);
      const R es = k == 0 ? fe[0] : (fe[k] + fe[o]) / R(2.0);
      const R Fs = k == 0 ? ff[0] : (ff[k] + ff[o]) / R(2.0);
      const R fa = f[k] - fs, ea = fe[k] - es, Fa = ff[k] - Fs;
      const R coll = -ws * (fs - es) - wa * (fa - ea);
      R v = f[k] + coll;
      if (force) v = v + ((R(1.0) - R(0.5) * ws) * Fs + (R(1.0) - R(0.5) * wa) * Fa);
      out[k] = v;
    }
  } else {
    R m[9], m2[9];
    mrt_m(f, m);
    const R meq1 = rho * (-R(2.0) + R(3.0) * uv), meq2 = rho * (R(1.0) - R(3.0) * uv);
    const R meq3 = rho * ux, meq5 = rho * uy;
    const R meq7 = rho * (ux * ux - uy * uy), meq8 = rho * (ux * uy);
    // S = diag(0, s_bulk, 1, 0, 1, 0, 1, s, s) (CollisionEngine.py:301-318)
    m2[0] = m[0];
    m2[1] = m[1] - s_bulk * (m[1] - meq1);
    m2[2] = m[2] - (m[2] - meq2);
    m2[3] = m[3];
    m2[4] = m[4] - (m[4] + meq3);
    m2[5] = m[5];
    m2[6] = m[6] - (m[6] + meq5);
    m2[7] = m[7] - s_shear * (m[7] - meq7);
    m2[8] = m[8] - s_shear * (m[8] - meq8);
    if (force) {
      const R Fu = Fx * ux + Fy * uy;
      m2[1] = m2[1] + (R(1.0) - s_bulk / R(2.0)) * (R(6.0) * Fu);
      m2[2] = m2[2] + R(0.5) * (-R(6.0) * Fu);
      m2[3] = m2[3] + Fx;
      m2[4] = m2[4] + R(0.5) * (-Fx);
      m2[5] = m2[5] + Fy;
      m2[6] = m2[6] + R(0.5) * (-Fy);
      m2[7] = m2[7] + (R(1.0) - s_shear / R(2.0)) * (R(2.0) * (Fx * ux - Fy * uy));
      m2[8] = m2[8] + (R(1.0) - s_shear / R(2.0)) * (Fy * ux + Fx * uy);
    }
    mrt_minv(m2, out);
  }
}

// Power-law shear rate of one component (CollisionEngine.py:80-92, 166-178,
// 332-377).  `tau_state` is the stored relaxation state of the previous step:
// BGK tau, TRT tau_sym, MRT diag[7] (= 1/tau).
//
// The reference sums the non-equilibrium part over the populations:
//   Sxx = cof sum_k cx^2 q_k,  Syy = cof sum_k cy^2 q_k,  Sxy = cof sum_k cx cy q_k
// with q = f - feq (BGK/TRT, cof = -1.5/(rho tau)) or q = M^-1 S M (f - feq)
// (MRT, cof = -1.5/rho).  These are evaluated here in moment form:
//   sum_k cx^2 feq = rho (1/3 + ux^2),  sum_k cy^2 feq = rho (1/3 + uy^2),
//   sum_k cx cy feq = rho ux uy;
// and, for MRT, sum_k {cx^2, cy^2, cx cy} (M^-1 v)_k = {v0 2/3 + v1/6 + v7/2,
// v0 2/3 + v1/6 - v7/2, v8} (the column sums of M^-1 = M^T diag(1/9, 1/36,
// 1/36, 1/6, 1/12, 1/6, 1/12, 1/4, 1/4)), where v = S M (f - feq) has
// v0 = 0, v1 = s_bulk (e - e_eq), v7 = s (p_xx - p_xx,eq), v8 = s (p_xy - p_xy,eq).
// The same quantities with about a tenth of the arithmetic of the matrix form.
template <int COLL, typename R>
__device__ __forceinline__ R shear_rate_of(const R (&f)[9], R rho, R ux, R uy, R tau_state, R s_bulk) {
  const R s14 = f[1] + f[2] + f[3] + f[4];
  const R s58 = f[5] + f[6] + f[7] + f[8];
  const R pxy = f[5] - f[6] + f[7] - f[8];
  R Sxx, Syy, Sxy;
  if constexpr (COLL == LBM_MRT) {
    const R uv = ux * ux + uy * uy;
    const R e = R(-4) * f[0] - s14 + R(2) * s58;
    const R pxx = f[1] - f[2] + f[3] - f[4];
    const R m1 = e - rho * (-R(2) + R(3) * uv);
    const R m7 = pxx - rho * (ux * ux - uy * uy);
    const R m8 = pxy - rho * (ux * uy);
    const R cof = -R(1.5) * rcp(rho);
    const R a = s_bulk * m1 * (R(1) / R(6));
    const R b = tau_state * m7 * R(0.5);
    Sxx = (a + b) * cof;
    Syy = (a - b) * cof;
    Sxy = tau_state * m8 * cof;
  } else {
    const R cof = -R(1.5) * rcp(rho * tau_state);
    const R pxx = f[1] + f[3] + s58;     // sum_k cx^2 f_k
    const R pyy = f[2] + f[4] + s58;     // sum_k cy^2 f_k
    Sxx = (pxx - rho * (R(1) / R(3) + ux * ux)) * cof;
    Syy = (pyy - rho * (R(1) / R(3) + uy * uy)) * cof;
    Sxy = (pxy - rho * (ux * uy)) * cof;
  }
  return sqrt(R(2) * (Sxx * Sxx + Syy * Syy + R(2) * Sxy * Sxy));
}

// SC_psi.get_psi, EOS.py:40-47.
template <typename R>
__device__ __forceinline__ R sc_psi(R rho, R rho0) { return rho0 * (R(1) - exp(-rho / rho0)); }

// error reporting shared by every translation unit of the library
extern thread_local std::string g_err;
int fail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int hip_check(hipError_t e, const char* what);

}  // namespace lbm
