// Per-node D2Q9 physics shared by the fused and the granular kernels.
// Each function restates one reference formula (file:line cited inline).
#pragma once

#include "lbm_common.h"

#include <string>

namespace lbm {

// NEEMMethod.rhow, BoundaryCondition.py:118-129.
__device__ __forceinline__ float neem_rhow(int d, float ux, float uy, const float* f) {
  switch (d) {
    case 2: return 1.0f / (1.0f + uy) * (f[0] + f[1] + f[3] + 2.0f * (f[2] + f[5] + f[6]));
    case 4: return 1.0f / (1.0f - uy) * (f[0] + f[1] + f[3] + 2.0f * (f[4] + f[7] + f[8]));
    case 1: return 1.0f / (1.0f + ux) * (f[0] + f[2] + f[4] + 2.0f * (f[1] + f[5] + f[8]));
    case 3: return 1.0f / (1.0f - ux) * (f[0] + f[2] + f[4] + 2.0f * (f[3] + f[6] + f[7]));
  }
  return 0.0f;
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
__device__ __forceinline__ void mrt_m(const float* f, float* m) {
  const float s14 = f[1] + f[2] + f[3] + f[4];
  const float s58 = f[5] + f[6] + f[7] + f[8];
  m[0] = f[0] + s14 + s58;
  m[1] = -4.0f * f[0] - s14 + 2.0f * s58;
  m[2] = 4.0f * f[0] - 2.0f * s14 + s58;
  m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
  m[4] = -2.0f * f[1] + 2.0f * f[3] + f[5] - f[6] - f[7] + f[8];
  m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
  m[6] = -2.0f * f[2] + 2.0f * f[4] + f[5] + f[6] - f[7] - f[8];
  m[7] = f[1] - f[2] + f[3] - f[4];
  m[8] = f[5] - f[6] + f[7] - f[8];
}

// f = M^-1 m with the exact inverse M^T diag(1/9,1/36,1/36,1/6,1/12,1/6,1/12,1/4,1/4).
__device__ __forceinline__ void mrt_minv(const float* m, float* f) {
  const float a0 = m[0] * (1.0f / 9.0f), a1 = m[1] * (1.0f / 36.0f), a2 = m[2] * (1.0f / 36.0f);
  const float a3 = m[3] * (1.0f / 6.0f), a4 = m[4] * (1.0f / 12.0f);
  const float a5 = m[5] * (1.0f / 6.0f), a6 = m[6] * (1.0f / 12.0f);
  const float a7 = m[7] * 0.25f, a8 = m[8] * 0.25f;
  const float b = a0 - a1 - 2.0f * a2;   // common to k = 1..4
  const float e = a0 + 2.0f * a1 + a2;   // common to k = 5..8
  const float x = a3 + a4, y = a5 + a6;
  f[0] = a0 - 4.0f * a1 + 4.0f * a2;
  f[1] = b + a3 - 2.0f * a4 + a7;
  f[2] = b + a5 - 2.0f * a6 - a7;
  f[3] = b - a3 + 2.0f * a4 + a7;
  f[4] = b - a5 + 2.0f * a6 - a7;
  f[5] = e + x + y + a8;
  f[6] = e - x + y - a8;
  f[7] = e - x - y + a8;
  f[8] = e + x - y - a8;
}

// Strain-rate magnitude from a (filtered) non-equilibrium part, shared by
// BGK/TRT (CollisionEngine.py:86-92) and MRT (:359-377).
__device__ __forceinline__ float shear_from(const float* q, float cof) {
  const float Sxx = (q[1] + q[3] + q[5] + q[6] + q[7] + q[8]) * cof;
  const float Syy = (q[2] + q[4] + q[5] + q[6] + q[7] + q[8]) * cof;
  const float Sxy = (q[5] - q[6] + q[7] - q[8]) * cof;
  return sqrtf(2.0f * (Sxx * Sxx + Syy * Syy + 2.0f * Sxy * Sxy));
}

template <int COLL>
__device__ __forceinline__ void collide(const float (&f)[9], float (&out)[9], float rho, float ux, float uy,
                                        float Fx, float Fy, bool force, float tau, float tau_anti,
                                        float s_shear, float s_bulk) {
  const float uv = ux * ux + uy * uy;
  if constexpr (COLL == LBM_BGK) {
    const float omega = 1.0f / tau;
    const float uF = ux * Fx + uy * Fy;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const float fe = feq_k(k, ux, uy, rho, uv);
      float v = f[k] + (-omega * (f[k] - fe));
      if (force) v = v + (1.0f - 0.5f * omega) * fforce_k(k, Fx, Fy, ux, uy, uF);
      out[k] = v;
    }
  } else if constexpr (COLL == LBM_TRT) {
    const float ws = 1.0f / tau, wa = 1.0f / tau_anti;
    const float uF = ux * Fx + uy * Fy;
    float fe[9], ff[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      fe[k] = feq_k(k, ux, uy, rho, uv);
      ff[k] = force ? fforce_k(k, Fx, Fy, ux, uy, uF) : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int o = opp(k);
      const float fs = k == 0 ? f[0] : (f[k] + f[o]) / 2.0f;
      const float es = k == 0 ? fe[0] : (fe[k] + fe[o]) / 2.0f;
      const float Fs = k == 0 ? ff[0] : (ff[k] + ff[o]) / 2.0f;
      const float fa = f[k] - fs, ea = fe[k] - es, Fa = ff[k] - Fs;
      const float coll = -ws * (fs - es) - wa * (fa - ea);
      float v = f[k] + coll;
      if (force) v = v + ((1.0f - 0.5f * ws) * Fs + (1.0f - 0.5f * wa) * Fa);
      out[k] = v;
    }
  } else {
    float m[9], m2[9];
    mrt_m(f, m);
    const float meq1 = rho * (-2.0f + 3.0f * uv), meq2 = rho * (1.0f - 3.0f * uv);
    const float meq3 = rho * ux, meq5 = rho * uy;
    const float meq7 = rho * (ux * ux - uy * uy), meq8 = rho * (ux * uy);
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
      const float Fu = Fx * ux + Fy * uy;
      m2[1] = m2[1] + (1.0f - s_bulk / 2.0f) * (6.0f * Fu);
      m2[2] = m2[2] + 0.5f * (-6.0f * Fu);
      m2[3] = m2[3] + Fx;
      m2[4] = m2[4] + 0.5f * (-Fx);
      m2[5] = m2[5] + Fy;
      m2[6] = m2[6] + 0.5f * (-Fy);
      m2[7] = m2[7] + (1.0f - s_shear / 2.0f) * (2.0f * (Fx * ux - Fy * uy));
      m2[8] = m2[8] + (1.0f - s_shear / 2.0f) * (Fy * ux + Fx * uy);
    }
    mrt_minv(m2, out);
  }
}

// Power-law shear rate of one component (CollisionEngine.py:80-92, 166-178,
// 332-377).  `tau_state` is the stored relaxation state of the previous step:
// BGK tau, TRT tau_sym, MRT diag[7] (= 1/tau).
template <int COLL>
__device__ __forceinline__ float shear_rate_of(const float (&f)[9], float rho, float ux, float uy,
                                               float tau_state, float s_bulk) {
  const float uv = ux * ux + uy * uy;
  float q[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) q[k] = f[k] - feq_k(k, ux, uy, rho, uv);
  if constexpr (COLL == LBM_MRT) {
    float m[9], sm[9];
    mrt_m(q, m);
    sm[0] = 0.0f * m[0];
    sm[1] = s_bulk * m[1];
    sm[2] = m[2];
    sm[3] = 0.0f * m[3];
    sm[4] = m[4];
    sm[5] = 0.0f * m[5];
    sm[6] = m[6];
    sm[7] = tau_state * m[7];
    sm[8] = tau_state * m[8];
    float qq[9];
    mrt_minv(sm, qq);
    return shear_from(qq, -1.5f / rho);
  } else {
    return shear_from(q, -1.5f / rho / tau_state);
  }
}

// SC_psi.get_psi, EOS.py:40-47.
__device__ __forceinline__ float sc_psi(float rho, float rho0) { return rho0 * (1.0f - expf(-rho / rho0)); }

// error reporting shared by every translation unit of the library
extern thread_local std::string g_err;
int fail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int hip_check(hipError_t e, const char* what);

}  // namespace lbm
