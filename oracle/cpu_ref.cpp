// CPU restatement of taichi-lbm's per-step algorithm -- TEST INFRASTRUCTURE.
//
// A C++17/OpenMP restatement of the reference's *unfused* kernel sequence, in
// its AoS layout (f[i][j][c][k], i slowest, lbm.py:22) and with its node-group
// index loops, used (a) as the timed CPU baseline ("reference-algorithm CPU
// restatement", not Taichi) and (b) as a second, independent restatement that
// the tests compare with the NumPy oracle.  Only tests/, smoke() and bench.py's
// CPU-baseline leg load it.
//
// Per step (cases/Power Law Fluid.py:160-173): density (lbm.py:153-164) ->
// pressure0 (:166-173) -> [Shan-Chen: update_psi -> PeriodicPsi_LR/BT ->
// cau_SCforce (cases/Bubble or two-phase separation.py:197-212,
// ExternalForce.py:42-117, SCBoundaryCondition.py:126-144)] ->
// force_density (:215-220) -> velocity (:223-237) ->
// [shear_rate -> relaxation_time] (CollisionEngine.py:72-92, 157-178,
// 322-377) -> collide (:96-110, 223-245, 424-451) -> boundary list in order
// (BoundaryCondition.py:154-166, 215-226, 244-259, 424-449).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct cref_params {
  int32_t nx, ny, nc;
  int32_t coll;       // 0 BGK, 1 TRT, 2 MRT
  int32_t power_law;  // run shear_rate + relaxation_time
  int32_t nthreads;
  float magic;        // TRT
  float s_bulk;       // MRT diag[1]
  float visc_m, visc_nm1, tau_min, tau_max, gx, gy;
  // Shan-Chen with SC_psi on a periodic lattice whose nodes are all fluid
  // for the force (no adhesion, no psi boundaries other than periodic)
  int32_t sc;          // 0 none, 1 ShanChenForceC1 (component 0), 2 ShanChenForceC2
  float g_coh[4];      // C1: g_coh[0]; C2: the (nc, nc) matrix, row-major
  float psi_rho0[2];   // SC_psi rho0 (C1: [0]; C2: per component)
} cref_params;

typedef struct cref_bc {
  int32_t kind;  // 1 periodic, 2 inside, 3 bounceback, 4 velocity NEEM, 5 pressure NEEM, 8 pressure ABC, 9 velocity equilibrium,
                 // 6 VelocityBB, 7 OpenNeumann
  int32_t direction;
  float ux, uy, rho;
  const int32_t* nodes;  // (count, 2)
  int32_t count;
} cref_bc;

}  // extern "C"

namespace {

const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
const int OPPK[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
// unknown populations of a plane boundary with outward direction d (BoundaryCondition.py:63-72)
const int UNK[5][3] = {{0, 0, 0}, {3, 6, 7}, {4, 7, 8}, {1, 5, 8}, {2, 5, 6}};
const float W[9] = {4.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f,
                    0.25f / 9.0f, 0.25f / 9.0f, 0.25f / 9.0f, 0.25f / 9.0f};
const float MM[9][9] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1},     {-4, -1, -1, -1, -1, 2, 2, 2, 2}, {4, -2, -2, -2, -2, 1, 1, 1, 1},
    {0, 1, 0, -1, 0, 1, -1, -1, 1},  {0, -2, 0, 2, 0, 1, -1, -1, 1},   {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1},  {0, 1, -1, 1, -1, 0, 0, 0, 0},    {0, 0, 0, 0, 0, 1, -1, 1, -1}};
const float MD[9] = {9, 36, 36, 6, 12, 6, 12, 4, 4};

struct Ctx {
  cref_params p;
  float *f, *f2, *rho, *trho, *vel, *bf, *shear, *tau, *pressure, *tpressure;
  float *sc, *psi;     // SCforce (NX, NY, nc, 2); psi with its halo ((NX+2), (NY+2), nc), offset -1
  const int32_t* nb;
  const int32_t* grp;
  int32_t gcount;
  float minv[9][9];
  size_t node(int i, int j) const { return (size_t)i * p.ny + j; }
  float* F(int i, int j, int c) const { return f + (node(i, j) * p.nc + c) * 9; }
  float* F2(int i, int j, int c) const { return f2 + (node(i, j) * p.nc + c) * 9; }
};

inline void feq(float ux, float uy, float rho, float* out) {
  const float uv = ux * ux + uy * uy;
  for (int k = 0; k < 9; ++k) {
    const float eu = (float)CX[k] * ux + (float)CY[k] * uy;
    out[k] = W[k] * rho * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * uv);
  }
}

inline void fforce(float Fx, float Fy, float ux, float uy, float* out) {
  const float uF = ux * Fx + uy * Fy;
  for (int k = 0; k < 9; ++k) {
    const float cF = (float)CX[k] * Fx + (float)CY[k] * Fy;
    const float cu = (float)CX[k] * ux + (float)CY[k] * uy;
    out[k] = W[k] * (3.0f * cF + 9.0f * cF * cu - 3.0f * uF);
  }
}

inline float clampt(float t, float lo, float hi) { return t < lo ? lo : (t > hi ? hi : t); }

void mv(const float (*A)[9], const float* x, float* y) {
  for (int r = 0; r < 9; ++r) {
    float s = 0.0f;
    for (int q = 0; q < 9; ++q) s += A[r][q] * x[q];
    y[r] = s;
  }
}

void mrt_rates(const Ctx& C, int i, int j, int c, float* S) {
  const float s = C.tau[C.node(i, j) * C.p.nc + c];
  const float v[9] = {0.0f, C.p.s_bulk, 1.0f, 0.0f, 1.0f, 0.0f, 1.0f, s, s};
  memcpy(S, v, sizeof v);
}

void step(Ctx& C, const cref_bc* bcs, int nbc) {
  const cref_params& p = C.p;
  const int nc = p.nc, nx = p.nx, ny = p.ny;
  const size_t N = (size_t)nx * ny;
  // density (lbm.py:153-164) + pressure0 (:166-173)
#pragma omp parallel for schedule(static)
  for (long n = 0; n < (long)N; ++n) {
    C.trho[n] = 0.0f;
    for (int c = 0; c < nc; ++c) C.rho[n * nc + c] = 0.0f, C.pressure[n * nc + c] = 0.0f;
  }
#pragma omp parallel for schedule(static)
  for (long m = 0; m < C.gcount; ++m) {
    const int i = C.grp[2 * m], j = C.grp[2 * m + 1];
    const size_t n = C.node(i, j);
    for (int c = 0; c < nc; ++c) {
      float s = 0.0f;
      const float* q = C.F(i, j, c);
      for (int k = 0; k < 9; ++k) s += q[k];
      C.rho[n * nc + c] = s;
      C.trho[n] += s;
    }
    C.pressure[n * nc] = C.pressure[n * nc] + C.rho[n * nc] / 3.0f;
    C.tpressure[n] = C.pressure[n * nc];
  }
  if (p.sc) {
    const int py = ny + 2;
    auto P = [&](int i, int j, int c) -> float& { return C.psi[((size_t)(i + 1) * py + (j + 1)) * nc + c]; };
    // update_psi (ExternalForce.py:42-47, 86-92): SC_psi of each component's
    // density on the group (C1: one psi object for every component)
#pragma omp parallel for schedule(static)
    for (long m = 0; m < C.gcount; ++m) {
      const int i = C.grp[2 * m], j = C.grp[2 * m + 1];
      const size_t n = C.node(i, j);
      for (int c = 0; c < nc; ++c) {
        const float r0 = p.sc == 1 ? p.psi_rho0[0] : p.psi_rho0[c];
        P(i, j, c) = r0 * (1.0f - expf(-C.rho[n * nc + c] / r0));
      }
    }
    // PeriodicPsi_LR, then PeriodicPsi_BT (SCBoundaryCondition.py:126-144)
    for (int j = -1; j <= ny; ++j)
      for (int c = 0; c < nc; ++c) P(-1, j, c) = P(nx - 1, j, c), P(nx, j, c) = P(0, j, c);
    for (int i = -1; i <= nx; ++i)
      for (int c = 0; c < nc; ++c) P(i, -1, c) = P(i, ny - 1, c), P(i, ny, c) = P(i, 0, c);
    // cau_SCforce: SCforce := 0, then C1 (:78-91) F_0 = -G psi_0(x) sum_k w_k c_k psi_0(x + c_k);
    // C2 (:96-117, every neighbour fluid) F_c1 = -psi_c1(x) sum_k sum_{c2 != c1} G[c1][c2] w_k c_k psi_c2(x + c_k)
#pragma omp parallel for schedule(static)
    for (long n = 0; n < (long)N * nc * 2; ++n) C.sc[n] = 0.0f;
#pragma omp parallel for schedule(static)
    for (long m = 0; m < C.gcount; ++m) {
      const int i = C.grp[2 * m], j = C.grp[2 * m + 1];
      const size_t n = C.node(i, j);
      for (int c1 = 0; c1 < (p.sc == 1 ? 1 : nc); ++c1) {
        float Fx = 0.0f, Fy = 0.0f;
        for (int k = 1; k < 9; ++k) {
          const float wx = W[k] * (float)CX[k], wy = W[k] * (float)CY[k];
          if (p.sc == 1) {
            const float v = P(i + CX[k], j + CY[k], 0);
            Fx += wx * v, Fy += wy * v;
          } else {
            for (int c2 = 0; c2 < nc; ++c2) {
              if (c2 == c1) continue;
              const float v = P(i + CX[k], j + CY[k], c2) * p.g_coh[c1 * 2 + c2];
              Fx += wx * v, Fy += wy * v;
            }
          }
        }
        const float s = p.sc == 1 ? -p.g_coh[0] * P(i, j, 0) : -P(i, j, c1);
        C.sc[(n * nc + c1) * 2] = Fx * s;
        C.sc[(n * nc + c1) * 2 + 1] = Fy * s;
      }
    }
  }
  // force_density (lbm.py:215-220)
#pragma omp parallel for schedule(static)
  for (long m = 0; m < C.gcount; ++m) {
    const size_t n = C.node(C.grp[2 * m], C.grp[2 * m + 1]);
    for (int c = 0; c < nc; ++c) {
      const float frac = C.rho[n * nc + c] / C.trho[n];
      const float sx = p.sc ? C.sc[(n * nc + c) * 2] : 0.0f, sy = p.sc ? C.sc[(n * nc + c) * 2 + 1] : 0.0f;
      C.bf[(n * nc + c) * 2] = sx + frac * p.gx;
      C.bf[(n * nc + c) * 2 + 1] = sy + frac * p.gy;
    }
  }
  // velocity (lbm.py:223-237)
#pragma omp parallel for schedule(static)
  for (long n = 0; n < (long)N; ++n) C.vel[2 * n] = C.vel[2 * n + 1] = 0.0f;
#pragma omp parallel for schedule(static)
  for (long m = 0; m < C.gcount; ++m) {
    const int i = C.grp[2 * m], j = C.grp[2 * m + 1];
    const size_t n = C.node(i, j);
    for (int c = 0; c < nc; ++c) {
      const float* q = C.F(i, j, c);
      float tx = 0.0f, ty = 0.0f;
      for (int k = 0; k < 9; ++k) tx += q[k] * (float)CX[k], ty += q[k] * (float)CY[k];
      tx += 0.5f * C.bf[(n * nc + c) * 2];
      ty += 0.5f * C.bf[(n * nc + c) * 2 + 1];
      C.vel[2 * n] += tx;
      C.vel[2 * n + 1] += ty;
    }
    C.vel[2 * n] /= C.trho[n];
    C.vel[2 * n + 1] /= C.trho[n];
  }
  if (p.power_law) {
    // shear_rate
#pragma omp parallel for schedule(static)
    for (long m = 0; m < C.gcount; ++m) {
      const int i = C.grp[2 * m], j = C.grp[2 * m + 1];
      const size_t n = C.node(i, j);
      for (int c = 0; c < nc; ++c) {
        float fe[9], q[9];
        const float rho = C.rho[n * nc + c];
        feq(C.vel[2 * n], C.vel[2 * n + 1], rho, fe);
        const float* fp = C.F(i, j, c);
        for (int k = 0; k < 9; ++k) q[k] = fp[k] - fe[k];
        float cof;
        if (p.coll == 2) {
          float mq[9], S[9], sm[9], qq[9];
          mv(MM, q, mq);
          mrt_rates(C, i, j, c, S);
          for (int r = 0; r < 9; ++r) sm[r] = S[r] * mq[r];
          mv(C.minv, sm, qq);
          memcpy(q, qq, sizeof q);
          cof = -1.5f / rho;
        } else {
          cof = -1.5f / rho / C.tau[n * nc + c];
        }
        const float Sxx = (q[1] + q[3] + q[5] + q[6] + q[7] + q[8]) * cof;
        const float Syy = (q[2] + q[4] + q[5] + q[6] + q[7] + q[8]) * cof;
        const float Sxy = (q[5] - q[6] + q[7] - q[8]) * cof;
        C.shear[n * nc + c] = sqrtf(2.0f * (Sxx * Sxx + Syy * Syy + 2.0f * Sxy * Sxy));
      }
    }
    // relaxation_time
#pragma omp parallel for schedule(static)
    for (long m = 0; m < C.gcount; ++m) {
      const size_t n = C.node(C.grp[2 * m], C.grp[2 * m + 1]);
      for (int c = 0; c < nc; ++c) {
        const float nu = p.visc_m * powf(C.shear[n * nc + c], p.visc_nm1);
        const float t = clampt(nu * 3.0f + 0.5f, p.tau_min, p.tau_max);
        C.tau[n * nc + c] = p.coll == 2 ? 1.0f / t : t;
      }
    }
  }
  // collide
#pragma omp parallel for schedule(static)
  for (long m = 0; m < C.gcount; ++m) {
    const int i = C.grp[2 * m], j = C.grp[2 * m + 1];
    const size_t n = C.node(i, j);
    const float ux = C.vel[2 * n], uy = C.vel[2 * n + 1];
    for (int c = 0; c < nc; ++c) {
      const float* fp = C.F(i, j, c);
      float* out = C.F2(i, j, c);
      const float rho = C.rho[n * nc + c];
      const float Fx = C.bf[(n * nc + c) * 2], Fy = C.bf[(n * nc + c) * 2 + 1];
      float fe[9], ff[9];
      feq(ux, uy, rho, fe);
      fforce(Fx, Fy, ux, uy, ff);
      if (p.coll == 0) {
        const float om = 1.0f / C.tau[n * nc + c];
        for (int k = 0; k < 9; ++k) out[k] = fp[k] + (-om * (fp[k] - fe[k])) + (1.0f - 0.5f * om) * ff[k];
      } else if (p.coll == 1) {
        const float ts = C.tau[n * nc + c];
        const float ws = 1.0f / ts, wa = 1.0f / (p.magic / (ts - 0.5f) + 0.5f);
        for (int k = 0; k < 9; ++k) {
          const int o = OPPK[k];
          const float fs = k ? (fp[k] + fp[o]) / 2 : fp[0], es = k ? (fe[k] + fe[o]) / 2 : fe[0];
          const float Fs = k ? (ff[k] + ff[o]) / 2 : ff[0];
          const float coll = -ws * (fs - es) - wa * ((fp[k] - fs) - (fe[k] - es));
          out[k] = fp[k] + coll + ((1.0f - 0.5f * ws) * Fs + (1.0f - 0.5f * wa) * (ff[k] - Fs));
        }
      } else {
        float mq[9], meq[9], mF[9], S[9], m2[9];
        mv(MM, fp, mq);
        const float uv = ux * ux + uy * uy;
        const float e[9] = {1.0f, -2.0f + 3.0f * uv, 1.0f - 3.0f * uv, ux, -ux, uy, -uy, ux * ux - uy * uy, ux * uy};
        for (int r = 0; r < 9; ++r) meq[r] = e[r] * rho;
        const float Fu = Fx * ux + Fy * uy;
        const float g[9] = {0.0f, 6 * Fu, -6 * Fu, Fx, -Fx, Fy, -Fy, 2 * (Fx * ux - Fy * uy), Fy * ux + Fx * uy};
        memcpy(mF, g, sizeof g);
        mrt_rates(C, i, j, c, S);
        for (int r = 0; r < 9; ++r) m2[r] = mq[r] - S[r] * (mq[r] - meq[r]) + (1.0f - S[r] / 2.0f) * mF[r];
        mv(C.minv, m2, out);
      }
    }
  }
  // boundary list in order
  for (int b = 0; b < nbc; ++b) {
    const cref_bc& bc = bcs[b];
    const int d = bc.direction;
#pragma omp parallel for schedule(static)
    for (long m = 0; m < bc.count; ++m) {
      const int i = bc.nodes[2 * m], j = bc.nodes[2 * m + 1];
      const size_t n = C.node(i, j);
      if (bc.kind == 1) {
        for (int k = 0; k < 9; ++k) {
          const int x2 = (i - CX[k] + nx) % nx, y2 = (j - CY[k] + ny) % ny;
          for (int c = 0; c < nc; ++c) C.F(i, j, c)[k] = C.F2(x2, y2, c)[k];
        }
      } else if (bc.kind == 2 || bc.kind == 3) {
        for (int k = 0; k < 9; ++k) {
          const int x2 = C.nb[(n * 9 + k) * 2], y2 = C.nb[(n * 9 + k) * 2 + 1];
          for (int c = 0; c < nc; ++c) {
            if (x2 != -1) C.F(i, j, c)[k] = C.F2(x2, y2, c)[k];
            else if (bc.kind == 3) C.F(i, j, c)[k] = C.F2(i, j, c)[OPPK[k]];
          }
        }
      } else if (bc.kind == 6) {
        // VelocityBB, BoundaryCondition.py:95-102
        for (int c = 0; c < nc; ++c) {
          const float rw = C.rho[n * nc + c];
          for (int u = 0; u < 3; ++u) {
            const int k = UNK[d][u];
            const float cv = (float)CX[k] * bc.ux + (float)CY[k] * bc.uy;
            C.F(i, j, c)[k] += 6 * W[k] * rw * cv;
          }
        }
      } else if (bc.kind == 9) {
        // VelocityEquilibrium, BoundaryCondition.py:132-145
        C.vel[2 * n] = bc.ux;
        C.vel[2 * n + 1] = bc.uy;
        for (int c = 0; c < nc; ++c) {
          float* q = C.F(i, j, c);
          float rw = 0.0f;
          if (d == 2) rw = 1.0f / (1.0f + bc.uy) * (q[0] + q[1] + q[3] + 2 * (q[2] + q[5] + q[6]));
          if (d == 4) rw = 1.0f / (1.0f - bc.uy) * (q[0] + q[1] + q[3] + 2 * (q[4] + q[7] + q[8]));
          if (d == 1) rw = 1.0f / (1.0f + bc.ux) * (q[0] + q[2] + q[4] + 2 * (q[1] + q[5] + q[8]));
          if (d == 3) rw = 1.0f / (1.0f - bc.ux) * (q[0] + q[2] + q[4] + 2 * (q[3] + q[6] + q[7]));
          C.rho[n * nc + c] = rw;
          feq(bc.ux, bc.uy, rw, q);
        }
      } else {
        const int i2 = i - CX[d], j2 = j - CY[d];
        const size_t n2 = C.node(i2, j2);
        if (bc.kind == 8) {
          // PressureABC, BoundaryCondition.py:177-205: u_w extrapolated from the
          // velocity pass, unknowns := -f + 2 w rho_value (1 + 4.5 (c.u_w)^2 - 1.5 u_w^2)
          const float uxw = 1.5f * C.vel[2 * n] - 0.5f * C.vel[2 * n2];
          const float uyw = 1.5f * C.vel[2 * n + 1] - 0.5f * C.vel[2 * n2 + 1];
          C.vel[2 * n] = uxw;
          C.vel[2 * n + 1] = uyw;
          C.trho[n] = bc.rho;
          const float uw2 = uxw * uxw + uyw * uyw;
          for (int c = 0; c < nc; ++c) {
            float* q = C.F(i, j, c);
            for (int u = 0; u < 3; ++u) {
              const int k = UNK[d][u];
              const float cu = (float)CX[k] * uxw + (float)CY[k] * uyw;
              q[k] = -q[k] + 2 * W[k] * bc.rho * (1.0f + 4.5f * cu * cu - 1.5f * uw2);
            }
          }
        } else if (bc.kind == 7) {
          // OpenNeumann, :327-334, then OpenBoundary.density / vel, :274-298
          float t = 0.0f, vx = 0.0f, vy = 0.0f;
          for (int c = 0; c < nc; ++c) {
            float* q = C.F(i, j, c);
            for (int u = 0; u < 3; ++u) q[UNK[d][u]] = C.F(i2, j2, c)[UNK[d][u]];
            float r = 0.0f, ax = 0.0f, ay = 0.0f;
            for (int k = 0; k < 9; ++k) {
              r += q[k];
              ax += q[k] * CX[k];
              ay += q[k] * CY[k];
            }
            C.rho[n * nc + c] = r;
            t += r;
            vx += ax + 0.5f * C.bf[(n * nc + c) * 2];
            vy += ay + 0.5f * C.bf[(n * nc + c) * 2 + 1];
          }
          C.trho[n] = t;
          C.vel[2 * n] = vx / t;
          C.vel[2 * n + 1] = vy / t;
        } else if (bc.kind == 4) {
          C.vel[2 * n] = bc.ux;
          C.vel[2 * n + 1] = bc.uy;
          for (int c = 0; c < nc; ++c) {
            const float* q = C.F(i, j, c);
            float rw = 0.0f;
            if (d == 2) rw = 1.0f / (1.0f + bc.uy) * (q[0] + q[1] + q[3] + 2 * (q[2] + q[5] + q[6]));
            if (d == 4) rw = 1.0f / (1.0f - bc.uy) * (q[0] + q[1] + q[3] + 2 * (q[4] + q[7] + q[8]));
            if (d == 1) rw = 1.0f / (1.0f + bc.ux) * (q[0] + q[2] + q[4] + 2 * (q[1] + q[5] + q[8]));
            if (d == 3) rw = 1.0f / (1.0f - bc.ux) * (q[0] + q[2] + q[4] + 2 * (q[3] + q[6] + q[7]));
            C.rho[n * nc + c] = rw;
          }
          for (int c = 0; c < nc; ++c) {
            float a[9], b2[9];
            feq(bc.ux, bc.uy, C.rho[n * nc + c], a);
            feq(C.vel[2 * n2], C.vel[2 * n2 + 1], C.rho[n2 * nc + c], b2);
            for (int k = 0; k < 9; ++k) C.F(i, j, c)[k] = a[k] + C.F(i2, j2, c)[k] - b2[k];
          }
        } else {
          C.vel[2 * n] = C.vel[2 * n2];
          C.vel[2 * n + 1] = C.vel[2 * n2 + 1];
          C.trho[n] = bc.rho;
          for (int c = 0; c < nc; ++c) {
            float a[9], b2[9];
            feq(C.vel[2 * n], C.vel[2 * n + 1], C.rho[n * nc + c], a);
            feq(C.vel[2 * n2], C.vel[2 * n2 + 1], C.rho[n2 * nc + c], b2);
            for (int u = 0; u < 3; ++u) {
              const int k = UNK[d][u];
              C.F(i, j, c)[k] = a[k] + C.F(i2, j2, c)[k] - b2[k];
            }
          }
        }
      }
    }
  }
}

}  // namespace

// sc_force (NX, NY, nc, 2) and psi ((NX+2), (NY+2), nc) are read and written
// when p->sc is set (may be NULL otherwise).
extern "C" int cref_run(const cref_params* p, float* f, float* f2, float* rho, float* total_rho, float* vel,
                        float* body_force, float* shear_rate, float* tau_state, float* pressure,
                        float* total_pressure, const int32_t* neighbor, const int32_t* group, int32_t count,
                        const cref_bc* bcs, int32_t nbc, int32_t nsteps, float* sc_force, float* psi) {
  if (p->sc && (!sc_force || !psi || p->nc > 2 || (p->sc == 2 && p->nc != 2))) return 1;
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
  Ctx C;
  C.p = *p;
  C.f = f, C.f2 = f2, C.rho = rho, C.trho = total_rho, C.vel = vel, C.bf = body_force;
  C.shear = shear_rate, C.tau = tau_state, C.pressure = pressure, C.tpressure = total_pressure;
  C.nb = neighbor, C.grp = group, C.gcount = count;
  C.sc = sc_force, C.psi = psi;
  for (int k = 0; k < 9; ++k)
    for (int r = 0; r < 9; ++r) C.minv[k][r] = MM[r][k] * (1.0f / MD[r]);
  for (int s = 0; s < nsteps; ++s) step(C, bcs, nbc);
  return 0;
}

extern "C" int cref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
