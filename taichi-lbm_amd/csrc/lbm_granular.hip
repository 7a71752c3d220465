// Granular kernels: one HIP kernel per reference @ti.kernel, over the same
// node groups, in the plane layout of lbm_abi.h.  They let a user's own
// lbm_solve() (a sequence of MacroscopicEngine / Collision / BoundaryEngine
// calls) run unchanged on the GPU, and they are the unfused baseline that the
// fused kernels are checked against at full size.
//
// A group is an int32 (count, 2) array of (i, j) in device memory; each kernel
// runs one thread per group entry, as the reference's `for m in
// range(group.count[None])` loops do.

#include "lbm_physics.h"

#include <string.h>

namespace lbm {
namespace {

struct G {
  lbm_gfields g;
  const int2* grp;
  uint32_t count;
  uint32_t nx, ny, nc, ps;
  uint32_t hs;  // psi halo plane stride (nx+2)*(ny+2)
  int32_t coll, sc;
  float visc_m, visc_nm1, tau_min, tau_max, gx, gy;
  float magic[LBM_MAX_NC];
  float s_bulk[LBM_MAX_NC];
  float g_coh[LBM_MAX_NC * LBM_MAX_NC];
  float psi_rho0[LBM_MAX_NC];
  int32_t psi_kind[LBM_MAX_NC];
  float eos[LBM_MAX_NC][6];
  float g_adh[LBM_MAX_NC];
  int32_t sc_adhesion;
  int32_t huang;       // HuangMRTCollision
  float huang_k1[LBM_MAX_NC], huang_k2[LBM_MAX_NC], huang_g[LBM_MAX_NC];
  lbm_neem_rule rule;
  int32_t kind;
  int32_t direction;   // psi boundaries
  float value;         // FixedPsi
};

__device__ __forceinline__ bool entry(const G& a, uint32_t& i, uint32_t& j, uint32_t& n) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.count) return false;
  const int2 ij = a.grp[m];
  i = (uint32_t)ij.x;
  j = (uint32_t)ij.y;
  n = i * a.ny + j;
  return true;
}

__device__ __forceinline__ float* P(float* base, uint32_t plane, uint32_t ps) { return base + (size_t)plane * ps; }

__device__ __forceinline__ uint32_t hidx(const G& a, int i, int j) {  // psi halo index of (i, j), i,j >= -1
  return (uint32_t)(i + 1) * (a.ny + 2) + (uint32_t)(j + 1);
}

__global__ void k_fill(float* p, uint32_t nplanes, uint32_t N, uint32_t ps, float v) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  for (uint32_t q = 0; q < nplanes; ++q) p[(size_t)q * ps + n] = v;
}

// LBField.init_LBM, lbm.py:119-125
__global__ void k_init(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const float ux = a.g.vel[n], uy = a.g.vel[(size_t)a.ps + n];
  const float uv = ux * ux + uy * uy;
  for (uint32_t c = 0; c < a.nc; ++c) {
    const float rho = a.g.rho[(size_t)c * a.ps + n];
    for (int k = 0; k < 9; ++k) {
      const float v = feq_k(k, ux, uy, rho, uv);
      P(a.g.f, c * 9 + k, a.ps)[n] = v;
      P(a.g.f2, c * 9 + k, a.ps)[n] = v;
    }
  }
}

// MacroscopicEngine.density, lbm.py:158-164 (after the fills)
__global__ void k_density(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  float tot = a.g.total_rho[n];
  for (uint32_t c = 0; c < a.nc; ++c) {
    float s = a.g.rho[(size_t)c * a.ps + n];
#pragma unroll
    for (int k = 0; k < 9; ++k) s += P(a.g.f, c * 9 + k, a.ps)[n];
    a.g.rho[(size_t)c * a.ps + n] = s;
    tot += s;
  }
  a.g.total_rho[n] = tot;
}

// MacroscopicEngine.pressure0/1/2, lbm.py:166-203 (after the fill)
__global__ void k_pressure(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  float* pr = a.g.pressure;
  const size_t ps = a.ps;
  if (a.sc == LBM_SC_NONE) {
    pr[n] = pr[n] + a.g.rho[n] / 3.0f;
    a.g.total_pressure[n] = pr[n];
  } else if (a.nc == 1) {
    const float psi = a.g.psi[hidx(a, (int)i, (int)j)];
    pr[n] = pr[n] + (a.g.rho[n] + 0.5f * a.g_coh[0] * (psi * psi)) / 3.0f;
    a.g.total_pressure[n] = pr[n];
  } else {
    for (uint32_t c1 = 0; c1 < a.nc; ++c1) {
      const float psi = a.g.psi[(size_t)c1 * a.hs + hidx(a, (int)i, (int)j)];
      pr[c1 * ps + n] = pr[c1 * ps + n] +
                        (a.g.rho[c1 * ps + n] + 0.5f * a.g_coh[c1 * LBM_MAX_NC + c1] * (psi * psi)) / 3.0f;
    }
    float p_ij = 0.0f, p_ii = 0.0f;
    for (uint32_t c1 = 0; c1 < a.nc; ++c1) {
      p_ii += pr[c1 * ps + n];
      for (uint32_t c2 = c1 + 1; c2 < a.nc; ++c2)
        p_ij += a.g_coh[c1 * LBM_MAX_NC + c2] * a.g.psi[(size_t)c1 * a.hs + hidx(a, (int)i, (int)j)] *
                a.g.psi[(size_t)c2 * a.hs + hidx(a, (int)i, (int)j)];
    }
    a.g.total_pressure[n] = p_ii + p_ij / 6.0f;
  }
}

// MacroscopicEngine.force_density, lbm.py:215-220
__global__ void k_force_density(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const size_t ps = a.ps;
  for (uint32_t c = 0; c < a.nc; ++c) {
    const float frac = a.g.rho[c * ps + n] / a.g.total_rho[n];
    a.g.body_force[(2 * c) * ps + n] = a.g.sc_force[(2 * c) * ps + n] + frac * a.gx;
    a.g.body_force[(2 * c + 1) * ps + n] = a.g.sc_force[(2 * c + 1) * ps + n] + frac * a.gy;
  }
}

// MacroscopicEngine.velocity, lbm.py:226-237 (after the fill)
__global__ void k_velocity(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const size_t ps = a.ps;
  float vx = a.g.vel[n], vy = a.g.vel[ps + n];
  for (uint32_t c = 0; c < a.nc; ++c) {
    float q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) q[k] = P(a.g.f, c * 9 + k, a.ps)[n];
    float tx = 0.0f, ty = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      tx += q[k] * (float)cx(k);
      ty += q[k] * (float)cy(k);
    }
    tx = tx + 0.5f * a.g.body_force[(2 * c) * ps + n];
    ty = ty + 0.5f * a.g.body_force[(2 * c + 1) * ps + n];
    vx += tx;
    vy += ty;
  }
  a.g.vel[n] = vx / a.g.total_rho[n];
  a.g.vel[ps + n] = vy / a.g.total_rho[n];
}

__device__ __forceinline__ void load9(const G& a, const float* base, uint32_t c, uint32_t n, float (&q)[9]) {
#pragma unroll
  for (int k = 0; k < 9; ++k) q[k] = base[(size_t)(c * 9 + k) * a.ps + n];
}

// MRT diag of a group node: (0, 1/tau_bulk, 1, 0, 1, 0, 1, s, s), CollisionEngine.py:301-318
__device__ __forceinline__ void mrt_rates(const G& a, uint32_t c, uint32_t n, float (&S)[9]) {
  const float s = a.g.tau[(size_t)c * a.ps + n];
  S[0] = 0.0f; S[1] = a.s_bulk[c]; S[2] = 1.0f; S[3] = 0.0f; S[4] = 1.0f;
  S[5] = 0.0f; S[6] = 1.0f; S[7] = s; S[8] = s;
}

// Collision.shear_rate: BGK CollisionEngine.py:80-92, TRT :166-178, MRT :332-377
__global__ void k_shear_rate(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const size_t ps = a.ps;
  const float ux = a.g.vel[n], uy = a.g.vel[ps + n];
  const float uv = ux * ux + uy * uy;
  for (uint32_t c = 0; c < a.nc; ++c) {
    float f[9], q[9];
    load9(a, a.g.f, c, n, f);
    const float rho = a.g.rho[c * ps + n];
#pragma unroll
    for (int k = 0; k < 9; ++k) q[k] = f[k] - feq_k(k, ux, uy, rho, uv);
    float g;
    if (a.coll == LBM_MRT) {
      float m[9], sm[9], qq[9], S[9];
      mrt_rates(a, c, n, S);
      mrt_m(q, m);
#pragma unroll
      for (int r = 0; r < 9; ++r) sm[r] = S[r] * m[r];
      mrt_minv(sm, qq);
      g = shear_from(qq, -1.5f / rho);
    } else {
      g = shear_from(q, -1.5f / rho / a.g.tau[c * ps + n]);
    }
    a.g.shear_rate[c * ps + n] = g;
  }
}

// Collision.relaxation_time: BGK :72-78, TRT :157-164, MRT :322-329
__global__ void k_relaxation_time(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const size_t ps = a.ps;
  for (uint32_t c = 0; c < a.nc; ++c) {
    const float nu = a.visc_m * powf(a.g.shear_rate[c * ps + n], a.visc_nm1);
    const float t = clamp_tau(nu * 3.0f + 0.5f, a.tau_min, a.tau_max);
    // MRT keeps diag[7] = diag[8] = 1/tau; TRT's tau_antisym follows from tau_sym
    a.g.tau[c * ps + n] = a.coll == LBM_MRT ? 1.0f / t : t;
  }
}

// Collision.apply: BGK :96-110, TRT :223-245, MRT :424-451
__global__ void k_collide(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const size_t ps = a.ps;
  const float ux = a.g.vel[n], uy = a.g.vel[ps + n];
  for (uint32_t c = 0; c < a.nc; ++c) {
    float f[9], out[9];
    load9(a, a.g.f, c, n, f);
    const float rho = a.g.rho[c * ps + n];
    const float Fx = a.g.body_force[(2 * c) * ps + n], Fy = a.g.body_force[(2 * c + 1) * ps + n];
    if (a.coll == LBM_BGK) {
      collide<LBM_BGK>(f, out, rho, ux, uy, Fx, Fy, true, a.g.tau[c * ps + n], 0.0f, 0.0f, 0.0f);
    } else if (a.coll == LBM_TRT) {
      const float ts = a.g.tau[c * ps + n];
      collide<LBM_TRT>(f, out, rho, ux, uy, Fx, Fy, true, ts, a.magic[c] / (ts - 0.5f) + 0.5f, 0.0f, 0.0f);
    } else {
      float S[9], m[9], m2[9], meq[9], mF[9];
      mrt_rates(a, c, n, S);
      mrt_m(f, m);
      const float uv = ux * ux + uy * uy;
      meq[0] = rho;
      meq[1] = rho * (-2.0f + 3.0f * uv);
      meq[2] = rho * (1.0f - 3.0f * uv);
      meq[3] = rho * ux;
      meq[4] = rho * -ux;
      meq[5] = rho * uy;
      meq[6] = rho * -uy;
      meq[7] = rho * (ux * ux - uy * uy);
      meq[8] = rho * (ux * uy);
      const float Fu = Fx * ux + Fy * uy;
      mF[0] = 0.0f;
      mF[1] = 6.0f * Fu;
      mF[2] = -6.0f * Fu;
      mF[3] = Fx;
      mF[4] = -Fx;
      mF[5] = Fy;
      mF[6] = -Fy;
      mF[7] = 2.0f * (Fx * ux - Fy * uy);
      mF[8] = Fy * ux + Fx * uy;
#pragma unroll
      for (int r = 0; r < 9; ++r) m2[r] = m[r] - S[r] * (m[r] - meq[r]) + (1.0f - S[r] / 2.0f) * mF[r];
      if (a.huang) {
        // HuangMRTCollision.apply (CollisionEngine.py:480-515): m2 + S Q with
        // the component's psi of this step (psi_field[i, j, c])
        const float psi = a.g.psi[(size_t)c * a.hs + hidx(a, (int)i, (int)j)];
        const HuangQ<float> q = huang_q<float>(Fx, Fy, psi, a.huang_k1[c], a.huang_k2[c], a.huang_g[c]);
        m2[1] = m2[1] + S[1] * q.q1;
        m2[2] = m2[2] - S[2] * q.q1;
        m2[7] = m2[7] + S[7] * q.q7;
        m2[8] = m2[8] + S[8] * q.q8;
      }
      mrt_minv(m2, out);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) P(a.g.f2, c * 9 + k, a.ps)[n] = out[k];
  }
}

// Boundary kernels, BoundaryCondition.py:154-259, :424-449
__global__ void k_boundary(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const size_t ps = a.ps;
  const uint32_t nx = a.nx, ny = a.ny;
  if (a.kind == LBM_BC_PERIODIC) {
    for (int k = 0; k < 9; ++k) {
      const uint32_t x2 = (uint32_t)(((int)i - cx(k) + (int)nx) % (int)nx);
      const uint32_t y2 = (uint32_t)(((int)j - cy(k) + (int)ny) % (int)ny);
      for (uint32_t c = 0; c < a.nc; ++c)
        P(a.g.f, c * 9 + k, a.ps)[n] = P(a.g.f2, c * 9 + k, a.ps)[x2 * ny + y2];
    }
  } else if (a.kind == LBM_BC_INSIDE || a.kind == LBM_BC_BOUNCEBACK) {
    for (int k = 0; k < 9; ++k) {
      const int ix2 = a.g.neighbor[((size_t)n * 9 + k) * 2];
      const int iy2 = a.g.neighbor[((size_t)n * 9 + k) * 2 + 1];
      if (ix2 != -1) {
        for (uint32_t c = 0; c < a.nc; ++c)
          P(a.g.f, c * 9 + k, a.ps)[n] = P(a.g.f2, c * 9 + k, a.ps)[(uint32_t)ix2 * ny + (uint32_t)iy2];
      } else if (a.kind == LBM_BC_BOUNCEBACK) {
        for (uint32_t c = 0; c < a.nc; ++c) P(a.g.f, c * 9 + k, a.ps)[n] = P(a.g.f2, c * 9 + opp(k), a.ps)[n];
      }
      // InsideBoundary with a -1 entry reads outside the reference's field
      // (no guard at :446-449); here the population is left untouched.
    }
  } else if (a.kind == LBM_BC_VELOCITY_BB) {
    // f_d += 6 w_d rho c_d . u_w on the unknown populations (BoundaryCondition.py:94-102)
    const int d = a.rule.direction;
    for (uint32_t c = 0; c < a.nc; ++c) {
      const float rhow = a.g.rho[c * ps + n];
      for (int k = 0; k < 9; ++k)
        if (neem_unknown(d, k)) {
          const float cv = (float)cx(k) * a.rule.ux + (float)cy(k) * a.rule.uy;
          float* fp = P(a.g.f, c * 9 + k, a.ps) + n;
          *fp = *fp + 6.0f * wgt(k) * rhow * cv;
        }
    }
  } else if (a.kind == LBM_BC_VELOCITY_EQ) {
    // VelocityEquilibrium, BoundaryCondition.py:135-145
    const int d = a.rule.direction;
    const float uxw = a.rule.ux, uyw = a.rule.uy, uvw = uxw * uxw + uyw * uyw;
    a.g.vel[n] = uxw;
    a.g.vel[ps + n] = uyw;
    for (uint32_t c = 0; c < a.nc; ++c) {
      float f[9];
      load9(a, a.g.f, c, n, f);
      const float rw = neem_rhow(d, uxw, uyw, f);
      a.g.rho[c * ps + n] = rw;
      for (int k = 0; k < 9; ++k) P(a.g.f, c * 9 + k, a.ps)[n] = feq_k(k, uxw, uyw, rw, uvw);
    }
  } else if (a.kind == LBM_BC_PRESSURE_ABC) {
    // PressureABC, BoundaryCondition.py:190-205 (the interior neighbour x - c_d
    // is outside the group: its velocity is not written here)
    const int d = a.rule.direction;
    const uint32_t n2 = (uint32_t)((int)i - cx(d)) * ny + (uint32_t)((int)j - cy(d));
    const float uxw = 1.5f * a.g.vel[n] - 0.5f * a.g.vel[n2];
    const float uyw = 1.5f * a.g.vel[ps + n] - 0.5f * a.g.vel[ps + n2];
    a.g.vel[n] = uxw;
    a.g.vel[ps + n] = uyw;
    a.g.total_rho[n] = a.rule.rho;
    const float uw2 = uxw * uxw + uyw * uyw;
    for (uint32_t c = 0; c < a.nc; ++c)
      for (int k = 0; k < 9; ++k)
        if (neem_unknown(d, k)) {
          const float cu = (float)cx(k) * uxw + (float)cy(k) * uyw;
          float* fp = P(a.g.f, c * 9 + k, a.ps) + n;
          *fp = -*fp + 2.0f * wgt(k) * a.rule.rho * (1.0f + 4.5f * cu * cu - 1.5f * uw2);
        }
  } else if (a.kind == LBM_BC_OPEN_EXTRAPOLATION) {
    // unknowns := 2 f(x - c_d) - f(x - 2 c_d); density and velocity untouched
    // (commented out in the reference, :412-413)
    const int d = a.rule.direction;
    const uint32_t n2 = (uint32_t)((int)i - cx(d)) * ny + (uint32_t)((int)j - cy(d));
    const uint32_t n3 = (uint32_t)((int)i - 2 * cx(d)) * ny + (uint32_t)((int)j - 2 * cy(d));
    for (uint32_t c = 0; c < a.nc; ++c)
      for (int k = 0; k < 9; ++k)
        if (neem_unknown(d, k))
          P(a.g.f, c * 9 + k, a.ps)[n] = 2.0f * P(a.g.f, c * 9 + k, a.ps)[n2] - P(a.g.f, c * 9 + k, a.ps)[n3];
  } else if (a.kind == LBM_BC_OPEN_NEUMANN || a.kind == LBM_BC_OPEN_CONVECTIVE1) {
    // OpenNeumann: copy the unknowns from x - c_d; OpenConvective1order:
    // unknowns := (f + l f(x - c_d)) / (1 + l), l = |u(x - c_d)|; then
    // OpenBoundary.density / vel at x (:274-298)
    const int d = a.rule.direction;
    const uint32_t n2 = (uint32_t)((int)i - cx(d)) * ny + (uint32_t)((int)j - cy(d));
    if (a.kind == LBM_BC_OPEN_NEUMANN) {
      for (uint32_t c = 0; c < a.nc; ++c)
        for (int k = 0; k < 9; ++k)
          if (neem_unknown(d, k)) P(a.g.f, c * 9 + k, a.ps)[n] = P(a.g.f, c * 9 + k, a.ps)[n2];
    } else {
      const float ux2 = a.g.vel[n2], uy2 = a.g.vel[ps + n2];
      const float lam = sqrtf(ux2 * ux2 + uy2 * uy2);
      for (uint32_t c = 0; c < a.nc; ++c)
        for (int k = 0; k < 9; ++k)
          if (neem_unknown(d, k)) {
            float* fp = P(a.g.f, c * 9 + k, a.ps) + n;
            *fp = (*fp + lam * P(a.g.f, c * 9 + k, a.ps)[n2]) / (1.0f + lam);
          }
    }
    float tot = 0.0f;
    for (uint32_t c = 0; c < a.nc; ++c) {
      float r = 0.0f;
      for (int k = 0; k < 9; ++k) r += P(a.g.f, c * 9 + k, a.ps)[n];
      a.g.rho[c * ps + n] = r;
      tot += r;
    }
    a.g.total_rho[n] = tot;
    float vx = 0.0f, vy = 0.0f;
    for (uint32_t c = 0; c < a.nc; ++c) {
      float tx = 0.0f, ty = 0.0f;
      for (int k = 0; k < 9; ++k) {
        const float q = P(a.g.f, c * 9 + k, a.ps)[n];
        tx += q * (float)cx(k);
        ty += q * (float)cy(k);
      }
      tx = tx + 0.5f * a.g.body_force[(2 * c) * ps + n];
      ty = ty + 0.5f * a.g.body_force[(2 * c + 1) * ps + n];
      vx += tx;
      vy += ty;
    }
    a.g.vel[n] = vx / tot;
    a.g.vel[ps + n] = vy / tot;
  } else {
    const int d = a.rule.direction;
    const uint32_t i2 = (uint32_t)((int)i - cx(d)), j2 = (uint32_t)((int)j - cy(d));
    const uint32_t n2 = i2 * ny + j2;
    const float ux2 = a.g.vel[n2], uy2 = a.g.vel[ps + n2];
    if (a.kind == LBM_BC_VELOCITY_NEEM) {
      const float uxw = a.rule.ux, uyw = a.rule.uy;
      a.g.vel[n] = uxw;
      a.g.vel[ps + n] = uyw;
      for (uint32_t c = 0; c < a.nc; ++c) {
        float f[9];
        load9(a, a.g.f, c, n, f);
        a.g.rho[c * ps + n] = neem_rhow(d, uxw, uyw, f);
      }
      const float uvw = uxw * uxw + uyw * uyw, uv2 = ux2 * ux2 + uy2 * uy2;
      for (uint32_t c = 0; c < a.nc; ++c) {
        const float rw = a.g.rho[c * ps + n], r2 = a.g.rho[c * ps + n2];
        for (int k = 0; k < 9; ++k)
          P(a.g.f, c * 9 + k, a.ps)[n] =
              feq_k(k, uxw, uyw, rw, uvw) + P(a.g.f, c * 9 + k, a.ps)[n2] - feq_k(k, ux2, uy2, r2, uv2);
      }
    } else {
      a.g.vel[n] = ux2;
      a.g.vel[ps + n] = uy2;
      a.g.total_rho[n] = a.rule.rho;
      const float uv2 = ux2 * ux2 + uy2 * uy2;
      for (uint32_t c = 0; c < a.nc; ++c) {
        const float ro = a.g.rho[c * ps + n], r2 = a.g.rho[c * ps + n2];
        for (int k = 0; k < 9; ++k)
          if (neem_unknown(d, k))
            P(a.g.f, c * 9 + k, a.ps)[n] =
                feq_k(k, ux2, uy2, ro, uv2) + P(a.g.f, c * 9 + k, a.ps)[n2] - feq_k(k, ux2, uy2, r2, uv2);
      }
    }
  }
}

// ShanChenForceC1/C2.update_psi, ExternalForce.py:70-75, 114-120
__global__ void k_update_psi(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  for (uint32_t c = 0; c < a.nc; ++c) {
    const float T = a.psi_kind[c] != LBM_PSI_SC ? a.g.temperature[n] : 0.0f;
    a.g.psi[(size_t)c * a.hs + hidx(a, (int)i, (int)j)] =
        psi_of<float>(a.psi_kind[c], a.g.rho[(size_t)c * a.ps + n], T, a.psi_rho0[c], a.eos[c]);
  }
}

// psi boundaries, SCBoundaryCondition.py:19-123, on halo-grid nodes (i, j >=
// -1).  One block: the convective kinds first average the velocity of the
// group's interior neighbours (OpenConvective1orderPsi :41-49), then every
// node is updated from psi at x - c_d (and x - 2 c_d).
__global__ __launch_bounds__(256) void k_psi_boundary(const G a) {
  __shared__ float red[2][256];
  const int d = a.direction;
  float lam = 0.0f;
  if (a.kind == LBM_PSI_BC_CONVECTIVE1 || a.kind == LBM_PSI_BC_CONVECTIVE2) {
    float ux = 0.0f, uy = 0.0f;
    for (uint32_t m = threadIdx.x; m < a.count; m += 256) {
      const int2 ij = a.grp[m];
      const uint32_t n2 = (uint32_t)(ij.x - cx(d)) * a.ny + (uint32_t)(ij.y - cy(d));
      ux += a.g.vel[n2];
      uy += a.g.vel[(size_t)a.ps + n2];
    }
    red[0][threadIdx.x] = ux;
    red[1][threadIdx.x] = uy;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) {
        red[0][threadIdx.x] += red[0][threadIdx.x + h];
        red[1][threadIdx.x] += red[1][threadIdx.x + h];
      }
      __syncthreads();
    }
    const float Ux = red[0][0] / (float)a.count, Uy = red[1][0] / (float)a.count;
    lam = 1.0f * sqrtf(Ux * Ux + Uy * Uy);
  }
  for (uint32_t m = threadIdx.x; m < a.count; m += 256) {
    const int2 ij = a.grp[m];
    const int i = ij.x, j = ij.y;
    const uint32_t x = hidx(a, i, j), x2 = hidx(a, i - cx(d), j - cy(d));
    if (a.kind == LBM_PSI_BC_FIXED) {
      a.g.psi[x] = a.value;
      continue;
    }
    for (uint32_t c = 0; c < a.nc; ++c) {
      float* p = a.g.psi + (size_t)c * a.hs;
      if (a.kind == LBM_PSI_BC_NEUMANN) {
        p[x] = p[x2];
      } else if (a.kind == LBM_PSI_BC_EXTRAPOLATION) {
        p[x] = 2.0f * p[x2] - p[hidx(a, i - 2 * cx(d), j - 2 * cy(d))];
      } else if (a.kind == LBM_PSI_BC_CONVECTIVE1) {
        p[x] = (p[x] + lam * p[x2]) / (1.0f + lam);
      } else {
        p[x] = (p[x] + 2.0f * lam * p[x2] - 0.5f * lam * p[hidx(a, i - 2 * cx(d), j - 2 * cy(d))]) / (1.0f + 1.5f * lam);
      }
    }
  }
}

// PeriodicPsi_LR (x == 0) then PeriodicPsi_BT (x == 1), SCBoundaryCondition.py:126-144
__global__ void k_psi_lr(const G a) {
  const int y = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1;
  if (y > (int)a.ny) return;
  for (uint32_t c = 0; c < a.nc; ++c) {
    float* p = a.g.psi + (size_t)c * a.hs;
    p[hidx(a, -1, y)] = p[hidx(a, (int)a.nx - 1, y)];
    p[hidx(a, (int)a.nx, y)] = p[hidx(a, 0, y)];
  }
}
__global__ void k_psi_bt(const G a) {
  const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x) - 1;
  if (x > (int)a.nx) return;
  for (uint32_t c = 0; c < a.nc; ++c) {
    float* p = a.g.psi + (size_t)c * a.hs;
    p[hidx(a, x, -1)] = p[hidx(a, x, (int)a.ny - 1)];
    p[hidx(a, x, (int)a.ny)] = p[hidx(a, x, 0)];
  }
}

// cau_SCforce, ExternalForce.py:78-91 (C1) and 124-145 (C2, every neighbour fluid)
__global__ void k_sc_force(const G a) {
  uint32_t i, j, n;
  if (!entry(a, i, j, n)) return;
  const size_t ps = a.ps;
  if (a.sc == LBM_SC_C1) {
    float Fx = 0.0f, Fy = 0.0f;
#pragma unroll
    for (int k = 1; k < 9; ++k) {
      const float p = a.g.psi[hidx(a, (int)i + cx(k), (int)j + cy(k))];
      Fx += (wgt(k) * (float)cx(k)) * p;
      Fy += (wgt(k) * (float)cy(k)) * p;
    }
    const float s = -a.g_coh[0] * a.g.psi[hidx(a, (int)i, (int)j)];
    a.g.sc_force[n] = a.g.sc_force[n] + Fx * s;
    a.g.sc_force[ps + n] = a.g.sc_force[ps + n] + Fy * s;
  } else {
    // a neighbour where fluid_strategy is false adds w_k c_k gadh[c1] (without
    // psi(x + c_k), commented out at :114) in place of the cohesion terms
    for (uint32_t c1 = 0; c1 < a.nc; ++c1) {
      float Fx = 0.0f, Fy = 0.0f, Ax = 0.0f, Ay = 0.0f;
#pragma unroll
      for (int k = 1; k < 9; ++k) {
        const uint32_t h = hidx(a, (int)i + cx(k), (int)j + cy(k));
        if (a.sc_adhesion && !a.g.sc_fluid[h]) {
          Ax += (wgt(k) * (float)cx(k)) * a.g_adh[c1];
          Ay += (wgt(k) * (float)cy(k)) * a.g_adh[c1];
          continue;
        }
        for (uint32_t c2 = 0; c2 < a.nc; ++c2) {
          if (c2 == c1) continue;
          const float p = a.g.psi[(size_t)c2 * a.hs + h];
          const float Gc = a.g_coh[c1 * LBM_MAX_NC + c2];
          Fx += (wgt(k) * (float)cx(k)) * p * Gc;
          Fy += (wgt(k) * (float)cy(k)) * p * Gc;
        }
      }
      const float s = -a.g.psi[(size_t)c1 * a.hs + hidx(a, (int)i, (int)j)];
      a.g.sc_force[(2 * c1) * ps + n] = s * (Fx + Ax);
      a.g.sc_force[(2 * c1 + 1) * ps + n] = s * (Fy + Ay);
    }
  }
}

int check_lat(const lbm_lattice* lat, const lbm_gfields* g) {
  if (!lat || !g) return fail("null lattice or fields");
  if (lat->nx < 1 || lat->ny < 1 || lat->nc < 1 || lat->nc > LBM_MAX_NC) return fail("bad lattice shape");
  if ((uint64_t)lat->nx * lat->ny >= (1ull << 31)) return fail("lattice too large");
  if ((uint64_t)lat->plane_stride < (uint64_t)lat->nx * lat->ny) return fail("plane_stride < NX*NY");
  return LBM_OK;
}

G make(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count) {
  G a;
  memset(&a, 0, sizeof a);
  a.g = *g;
  a.grp = (const int2*)group;
  a.count = count > 0 ? (uint32_t)count : 0u;
  a.nx = lat->nx;
  a.ny = lat->ny;
  a.nc = lat->nc;
  a.ps = (uint32_t)lat->plane_stride;
  a.hs = (uint32_t)((lat->nx + 2) * (lat->ny + 2));
  a.coll = lat->collision;
  a.sc = lat->shan_chen;
  a.visc_m = lat->visc_m;
  a.visc_nm1 = lat->visc_nm1;
  a.tau_min = lat->tau_min;
  a.tau_max = lat->tau_max;
  a.gx = lat->gx;
  a.gy = lat->gy;
  for (int c = 0; c < LBM_MAX_NC; ++c) {
    a.magic[c] = lat->magic[c];
    a.s_bulk[c] = lat->s_bulk[c];
    a.psi_rho0[c] = lat->psi_rho0[c];
    a.psi_kind[c] = lat->psi_kind[c];
    a.g_adh[c] = lat->g_adh[c];
    for (int q = 0; q < 6; ++q) a.eos[c][q] = lat->eos[c][q];
  }
  a.sc_adhesion = lat->sc_adhesion;
  a.huang = lat->huang;
  for (int c = 0; c < LBM_MAX_NC; ++c) {
    a.huang_k1[c] = lat->huang_k1[c];
    a.huang_k2[c] = lat->huang_k2[c];
    a.huang_g[c] = lat->huang_g[c];
  }
  for (int q = 0; q < LBM_MAX_NC * LBM_MAX_NC; ++q) a.g_coh[q] = lat->g_coh[q];
  return a;
}

int run_group(void (*fn)(const G), const G& a, hipStream_t s) {
  if (a.count == 0) return LBM_OK;
  hipLaunchKernelGGL(fn, dim3((a.count + 255) / 256), dim3(256), 0, s, a);
  return hip_check(hipGetLastError(), "granular kernel launch");
}

int fill(float* p, uint32_t nplanes, uint32_t N, uint32_t ps, float v, hipStream_t s) {
  if (!p) return fail("fill of a null field");
  hipLaunchKernelGGL(k_fill, dim3((N + 255) / 256), dim3(256), 0, s, p, nplanes, N, ps, v);
  return hip_check(hipGetLastError(), "fill launch");
}

}  // namespace
}  // namespace lbm

using namespace lbm;

#define GRANULAR_PROLOGUE                          \
  int rc = check_lat(lat, g);                      \
  if (rc) return rc;                               \
  if (count > 0 && !group) return fail("null group"); \
  hipStream_t s = (hipStream_t)stream;             \
  G a = make(lat, g, group, count);

extern "C" int lbm_g_init_lbm(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count,
                              void* stream) {
  GRANULAR_PROLOGUE
  return run_group(k_init, a, s);
}

extern "C" int lbm_g_density(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count,
                             void* stream) {
  GRANULAR_PROLOGUE
  const uint32_t N = a.nx * a.ny;
  if ((rc = fill(g->total_rho, 1, N, a.ps, 0.0f, s))) return rc;
  if ((rc = fill(g->rho, a.nc, N, a.ps, 0.0f, s))) return rc;
  return run_group(k_density, a, s);
}

extern "C" int lbm_g_pressure(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count,
                              void* stream) {
  GRANULAR_PROLOGUE
  if (a.sc != LBM_SC_NONE && !g->psi) return fail("pressure with Shan-Chen needs psi");
  if ((rc = fill(g->pressure, a.nc, a.nx * a.ny, a.ps, 0.0f, s))) return rc;
  return run_group(k_pressure, a, s);
}

extern "C" int lbm_g_force_density(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group,
                                   int32_t count, void* stream) {
  GRANULAR_PROLOGUE
  return run_group(k_force_density, a, s);
}

extern "C" int lbm_g_velocity(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count,
                              void* stream) {
  GRANULAR_PROLOGUE
  if ((rc = fill(g->vel, 2, a.nx * a.ny, a.ps, 0.0f, s))) return rc;
  return run_group(k_velocity, a, s);
}

extern "C" int lbm_g_shear_rate(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count,
                                void* stream) {
  GRANULAR_PROLOGUE
  if (!g->tau) return fail("shear_rate needs the relaxation-time field");
  return run_group(k_shear_rate, a, s);
}

extern "C" int lbm_g_relaxation_time(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group,
                                     int32_t count, void* stream) {
  GRANULAR_PROLOGUE
  if (!g->tau) return fail("relaxation_time needs the relaxation-time field");
  return run_group(k_relaxation_time, a, s);
}

extern "C" int lbm_g_collide(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count,
                             void* stream) {
  GRANULAR_PROLOGUE
  if (!g->tau) return fail("collide needs the relaxation-time field");
  if (lat->huang) {
    if (lat->collision != LBM_MRT) return fail("HuangMRTCollision needs the MRT collision");
    if (!g->psi) return fail("HuangMRTCollision.apply needs the Shan-Chen psi field");
  }
  return run_group(k_collide, a, s);
}

extern "C" int lbm_g_boundary(const lbm_lattice* lat, const lbm_gfields* g, int32_t kind, const lbm_neem_rule* rule,
                              const int32_t* group, int32_t count, void* stream) {
  GRANULAR_PROLOGUE
  if (kind < LBM_BC_PERIODIC || kind > LBM_BC_OPEN_EXTRAPOLATION) return fail("bad boundary kind %d", kind);
  if ((kind == LBM_BC_INSIDE || kind == LBM_BC_BOUNCEBACK) && !g->neighbor)
    return fail("boundary kind %d needs the neighbour table", kind);
  if (kind >= LBM_BC_VELOCITY_NEEM) {
    if (!rule) return fail("boundary kind %d needs its rule (direction, value)", kind);
    if (rule->direction < 1 || rule->direction > 4) return fail("NEEM direction %d", rule->direction);
    a.rule = *rule;
  }
  a.kind = kind;
  return run_group(k_boundary, a, s);
}

extern "C" int lbm_g_sc_psi_boundary(const lbm_lattice* lat, const lbm_gfields* g, int32_t kind, int32_t direction,
                                     float value, const int32_t* group, int32_t count, void* stream) {
  GRANULAR_PROLOGUE
  if (!g->psi) return fail("psi boundary needs psi");
  if (kind < LBM_PSI_BC_NEUMANN || kind > LBM_PSI_BC_FIXED) return fail("bad psi boundary kind %d", kind);
  if (kind != LBM_PSI_BC_FIXED && (direction < 1 || direction > 4)) return fail("psi boundary direction %d", direction);
  if ((kind == LBM_PSI_BC_CONVECTIVE1 || kind == LBM_PSI_BC_CONVECTIVE2) && !g->vel)
    return fail("convective psi boundary needs the velocity field");
  a.kind = kind;
  a.direction = direction;
  a.value = value;
  if (a.count == 0) return LBM_OK;
  hipLaunchKernelGGL(k_psi_boundary, dim3(1), dim3(256), 0, s, a);
  return hip_check(hipGetLastError(), "psi boundary launch");
}

extern "C" int lbm_g_sc_update_psi(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group,
                                   int32_t count, void* stream) {
  GRANULAR_PROLOGUE
  if (!g->psi) return fail("update_psi needs psi");
  for (int c = 0; c < lat->nc; ++c) {
    if (lat->psi_kind[c] < LBM_PSI_SC || lat->psi_kind[c] > LBM_PSI_CS) return fail("bad psi kind %d", lat->psi_kind[c]);
    if (lat->psi_kind[c] != LBM_PSI_SC && !g->temperature) return fail("a cubic-EOS psi needs the temperature field");
  }
  return run_group(k_update_psi, a, s);
}

extern "C" int lbm_g_sc_periodic_psi(const lbm_lattice* lat, const lbm_gfields* g, int32_t which, void* stream) {
  const int32_t* group = nullptr;
  int32_t count = 0;
  GRANULAR_PROLOGUE
  if (!g->psi) return fail("periodic psi needs psi");
  if (which & 1) {
    hipLaunchKernelGGL(k_psi_lr, dim3((a.ny + 2 + 255) / 256), dim3(256), 0, s, a);
    if ((rc = hip_check(hipGetLastError(), "psi LR launch"))) return rc;
  }
  if (which & 2) {
    hipLaunchKernelGGL(k_psi_bt, dim3((a.nx + 2 + 255) / 256), dim3(256), 0, s, a);
    if ((rc = hip_check(hipGetLastError(), "psi BT launch"))) return rc;
  }
  return LBM_OK;
}

extern "C" int lbm_g_sc_force(const lbm_lattice* lat, const lbm_gfields* g, const int32_t* group, int32_t count,
                              void* stream) {
  GRANULAR_PROLOGUE
  if (!g->psi || !g->sc_force) return fail("cau_SCforce needs psi and SCforce");
  if (a.sc == LBM_SC_NONE) return fail("cau_SCforce without a Shan-Chen variant");
  if (a.sc_adhesion && (a.sc != LBM_SC_C2 || !g->sc_fluid)) return fail("adhesion needs ShanChenForceC2 and sc_fluid");
  if ((rc = fill(g->sc_force, 2 * a.nc, a.nx * a.ny, a.ps, 0.0f, s))) return rc;
  return run_group(k_sc_force, a, s);
}
