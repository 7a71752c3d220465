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

// M f for the MRT moment basis (CollisionEngine.py:273-282, :379-385),
// from shared pair sums and differences (m0 = rho, m3 = jx, m5 = jy).
template <typename R>
__device__ __forceinline__ void mrt_m(const R* f, R* m) {
  const R s13 = f[1] + f[3], s24 = f[2] + f[4], d13 = f[1] - f[3], d24 = f[2] - f[4];
  const R s57 = f[5] + f[7], s68 = f[6] + f[8], d57 = f[5] - f[7], d68 = f[6] - f[8];
  const R s14 = s13 + s24, s58 = s57 + s68;
  const R dx = d57 - d68, dy = d57 + d68;   // f5 - f6 - f7 + f8, f5 + f6 - f7 - f8
  m[0] = f[0] + s14 + s58;
  m[1] = R(-4) * f[0] - s14 + R(2) * s58;
  m[2] = R(4) * f[0] - R(2) * s14 + s58;
  m[3] = d13 + dx;
  m[4] = R(-2) * d13 + dx;
  m[5] = d24 + dy;
  m[6] = R(-2) * d24 + dy;
  m[7] = s13 - s24;
  m[8] = s57 - s68;
}

// f = M^-1 m with the exact inverse M^T diag(1/9,1/36,1/36,1/6,1/12,1/6,1/12,1/4,1/4),
// the diagonal folded into the shared sums.
template <typename R>
__device__ __forceinline__ void mrt_minv(const R* m, R* f) {
  const R b = m[0] * (R(1) / R(9)) - m[1] * (R(1) / R(36)) - m[2] * (R(1) / R(18));   // k = 1..4
  const R e = m[0] * (R(1) / R(9)) + m[1] * (R(1) / R(18)) + m[2] * (R(1) / R(36));   // k = 5..8
  const R u = (m[3] - m[4]) * (R(1) / R(6)), v = (m[5] - m[6]) * (R(1) / R(6));
  const R x = m[3] * (R(1) / R(6)) + m[4] * (R(1) / R(12));
  const R y = m[5] * (R(1) / R(6)) + m[6] * (R(1) / R(12));
  const R bp = b + m[7] * R(0.25), bm = b - m[7] * R(0.25);
  const R ep = e + m[8] * R(0.25), em = e - m[8] * R(0.25);
  const R xpy = x + y, xmy = x - y;
  f[0] = (m[0] - m[1] + m[2]) * (R(1) / R(9));
  f[1] = bp + u;
  f[2] = bm + v;
  f[3] = bp - u;
  f[4] = bm - v;
  f[5] = ep + xpy;
  f[6] = em - xmy;
  f[7] = ep - xpy;
  f[8] = em + xmy;
}

// MRT equilibrium moments that the relaxation reads (CollisionEngine.py:388-405),
// and the non-equilibrium parts of e, p_xx, p_xy shared by the shear rate and
// the relaxation.
template <typename R>
struct MrtNeq {
  R meq2, meq3, meq5;   // relaxed with rate 1: the post-collision moment is the equilibrium
  R n1, n7, n8;         // m1 - meq1, m7 - meq7, m8 - meq8
};
template <typename R>
__device__ __forceinline__ MrtNeq<R> mrt_neq(const R* m, R rho, R ux, R uy) {
  const R uv = ux * ux + uy * uy;
  MrtNeq<R> q;
  q.meq2 = rho * (R(1.0) - R(3.0) * uv);
  q.meq3 = rho * ux;
  q.meq5 = rho * uy;
  q.n1 = m[1] - rho * (-R(2.0) + R(3.0) * uv);
  q.n7 = m[7] - rho * (ux * ux - uy * uy);
  q.n8 = m[8] - rho * (ux * uy);
  return q;
}

// MRT collision from the moments: S = diag(0, s_bulk, 1, 0, 1, 0, 1, s, s)
// (CollisionEngine.py:301-318), Guo forcing (1 - S/2) M F in moment space.
// HuangMRTCollision.m_Q (CollisionEngine.py:466-478): the source moments
// Q1 = 3 (k1 + 2 k2) |F|^2 / (G psi^2), Q2 = -Q1, Q7 = k1 (Fx^2 - Fy^2) /
// (G psi^2), Q8 = k1 Fx Fy / (G psi^2), F the component's body force, psi
// its own psi and G its coupling with itself; the collision adds S Q
// (:507), S the node's MRT rates (the reference indexes the rates per
// component only, :507, which the per-node MRT state does not have).
template <typename R>
struct HuangQ {
  R q1, q7, q8;
};
template <typename R>
__device__ __forceinline__ HuangQ<R> huang_q(R Fx, R Fy, R psi, R k1, R k2, R g) {
  const R F2 = Fx * Fx + Fy * Fy;
  const R b = g * (psi * psi);
  return HuangQ<R>{R(3) * (k1 + R(2) * k2) * F2 / b, k1 * (Fx * Fx - Fy * Fy) / b, k1 * Fx * Fy / b};
}

template <typename R>
__device__ __forceinline__ void mrt_collide_m(const R* m, const MrtNeq<R>& q, R (&out)[9], R ux, R uy, R Fx, R Fy,
                                              bool force, R s_shear, R s_bulk, const HuangQ<R>* hq = nullptr) {
  R m2[9];
  m2[0] = m[0];
  m2[1] = m[1] - s_bulk * q.n1;
  m2[2] = q.meq2;
  m2[3] = m[3];
  m2[4] = -q.meq3;
  m2[5] = m[5];
  m2[6] = -q.meq5;
  m2[7] = m[7] - s_shear * q.n7;
  m2[8] = m[8] - s_shear * q.n8;
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
  if (hq) {
    // HuangMRTCollision: + S Q (S_2 = 1)
    m2[1] = m2[1] + s_bulk * hq->q1;
    m2[2] = m2[2] - hq->q1;
    m2[7] = m2[7] + s_shear * hq->q7;
    m2[8] = m2[8] + s_shear * hq->q8;
  }
  mrt_minv(m2, out);
}

// Squared strain-rate magnitude 2 (Sxx^2 + Syy^2 + 2 Sxy^2) from the
// non-equilibrium moments n1, n7, n8 (BGK CollisionEngine.py:80-92, TRT
// :166-178, MRT :359-377).  The reference sums cof x {cx^2, cy^2, cx cy} over
// a non-equilibrium part v of the populations; as moments, Sxx + Syy =
// cof v1 / 3, Sxx - Syy = cof v7, Sxy = cof v8, so the square is
// cof^2 (v1^2 / 9 + v7^2 + 4 v8^2).  MRT: v = S M (f - feq), i.e. v1 =
// s_bulk n1, v7 = s n7, v8 = s n8 with s the previous step's rate, and cof =
// -1.5 / rho; BGK / TRT: v = f - feq and cof = -1.5 / (rho tau) -- the same
// with s = s_bulk = 1 / tau of the previous step.
template <typename R>
__device__ __forceinline__ R mrt_shear2(const MrtNeq<R>& q, R inv_rho, R s, R s_bulk) {
  const R cof = R(1.5) * inv_rho;
  const R v1 = s_bulk * q.n1 * (R(1) / R(3)), v7 = s * q.n7, v8 = s * q.n8;
  return (cof * cof) * (v1 * v1 + v7 * v7 + R(4) * (v8 * v8));
}

// Strain-rate magnitude from a (filtered) non-equilibrium part in population
// form, as the reference sums it (BGK/TRT CollisionEngine.py:86-92, MRT
// :359-377); the granular kernels' shear_rate.
template <typename R>
__device__ __forceinline__ R shear_from(const R* q, R cof) {
  const R Sxx = (q[1] + q[3] + q[5] + q[6] + q[7] + q[8]) * cof;
  const R Syy = (q[2] + q[4] + q[5] + q[6] + q[7] + q[8]) * cof;
  const R Sxy = (q[5] - q[6] + q[7] - q[8]) * cof;
  return sqrt(R(2) * (Sxx * Sxx + Syy * Syy + R(2) * Sxy * Sxy));
}

// BGK / TRT collision over the direction pairs (k, opp k) (BGK
// CollisionEngine.py:95-103, TRT :181-200, Guo forcing :20-25).  With
// fs, fa = (f_k +- f_opp)/2 and the matching parts of feq and of the force
// term, es = w rho (1 - 1.5 u.u + 4.5 (c.u)^2), ea = 3 w rho c.u,
// Fs = w (9 (c.u)(c.F) - 3 u.F), Fa = 3 w c.F:
//   f_k'   = f_k   + S + A,   f_opp' = f_opp + S - A,
//   S = -ws (fs - es) + (1 - ws/2) Fs,   A = -wa (fa - ea) + (1 - wa/2) Fa;
// BGK is ws = wa = 1/tau.  Rest population: S alone with c = 0.
template <typename R>
__device__ __forceinline__ void pair_collide(const R (&f)[9], R (&out)[9], R rho, R ux, R uy, R Fx, R Fy,
                                             bool force, R ws, R wa) {
  const R A0 = R(1) - R(1.5) * (ux * ux + uy * uy);
  const R uF3 = R(3) * (ux * Fx + uy * Fy);
  const R hs = R(1) - R(0.5) * ws, ha = R(1) - R(0.5) * wa;
  {
    const R w = R(4.0 / 9.0);
    R v = f[0] - ws * (f[0] - (w * rho) * A0);
    if (force) v = v + hs * (w * -uF3);
    out[0] = v;
  }
#pragma unroll
  for (int k : {1, 2, 5, 6}) {
    const int o = opp(k);
    const R w = wgt<R>(k), wr = w * rho;
    const R cu = (R)cx(k) * ux + (R)cy(k) * uy;
    const R es = wr * (A0 + R(4.5) * cu * cu), ea = R(3) * wr * cu;
    const R fs = (f[k] + f[o]) * R(0.5), fa = (f[k] - f[o]) * R(0.5);
    R S = -ws * (fs - es), An = -wa * (fa - ea);
    if (force) {
      const R cF = (R)cx(k) * Fx + (R)cy(k) * Fy;
      S = S + hs * (w * (R(9) * cu * cF - uF3));
      An = An + ha * (R(3) * w * cF);
    }
    out[k] = f[k] + S + An;
    out[o] = f[o] + S - An;
  }
}

template <int COLL, typename R>
__device__ __forceinline__ void collide(const R (&f)[9], R (&out)[9], R rho, R ux, R uy, R Fx, R Fy, bool force,
                                        R tau, R tau_anti, R s_shear, R s_bulk) {
  if constexpr (COLL == LBM_BGK) {
    const R omega = rcp(tau);
    pair_collide(f, out, rho, ux, uy, Fx, Fy, force, omega, omega);
  } else if constexpr (COLL == LBM_TRT) {
    pair_collide(f, out, rho, ux, uy, Fx, Fy, force, rcp(tau), rcp(tau_anti));
  } else {
    R m[9];
    mrt_m(f, m);
    mrt_collide_m(m, mrt_neq(m, rho, ux, uy), out, ux, uy, Fx, Fy, force, s_shear, s_bulk);
  }
}

// SC_psi.get_psi, EOS.py:40-47.
template <typename R>
__device__ __forceinline__ R sc_psi(R rho, R rho0) { return rho0 * (R(1) - exp(-rho / rho0)); }

// Pseudopotential.get_psi with a cubic equation of state, EOS.py:5-95:
// psi = sqrt(|6/G (p(rho, T) - rho/3)|); e = (a, b, R, G, omega, T_cr).
template <typename R>
__device__ __forceinline__ R eos_pressure(int kind, R rho, R T, const float (&e)[6]) {
  const R a = e[0], b = e[1], Rg = e[2];
  const R rep = rho * Rg * T / (R(1) - b * rho);
  if (kind == LBM_PSI_VDW) return rep - a * (rho * rho);
  if (kind == LBM_PSI_RK) return rep - a * (rho * rho) / sqrt(T) / (R(1) + b * rho);
  if (kind == LBM_PSI_PR) {
    const R w = e[4];
    const R k = R(1) + (R(0.37464) + R(1.54226) * w - R(0.269922) * (w * w)) * (R(1) - sqrt(T / (R)e[5]));
    const R br = b * rho;
    return rep - a * (k * k) * (rho * rho) / (R(1) + R(2) * br - br * br);
  }
  // Carnahan-Starling
  const R x = b * rho / R(4);
  const R om = R(1) - x;
  return rho * Rg * T * (R(1) + x + x * x - x * x * x) / (om * om * om) - a * (rho * rho);
}
template <typename R>
__device__ __forceinline__ R psi_of(int kind, R rho, R T, R rho0, const float (&e)[6]) {
  if (kind == LBM_PSI_SC) return sc_psi<R>(rho, rho0);
  return sqrt(fabs(R(6) / (R)e[3] * (eos_pressure<R>(kind, rho, T, e) - rho / R(3))));
}

// error reporting shared by every translation unit of the library
extern thread_local std::string g_err;
int fail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int hip_check(hipError_t e, const char* what);

}  // namespace lbm
