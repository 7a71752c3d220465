// Fused pull-scheme stream + collide for D2Q9 on MI355X (gfx950).
//
// One launch per time step.  A launch reads the post-collision populations f2
// of step n-1 and produces those of step n, doing in between, per node:
//   S7(n-1)  the reference's ordered boundary pass (BoundaryCondition.py:
//            481-483): periodic / table pull, halfway bounce-back (:244-259),
//            VelocityNEEM (:154-166) and PressureNEEM (:215-226), compiled on
//            the host into one uint16 code per node;
//   S1..S6(n) density (lbm.py:153-164), force_density (:215-220), velocity
//            (:223-237), power-law shear_rate/relaxation_time
//            (CollisionEngine.py:72-92, 157-178, 322-377) and the BGK / TRT /
//            MRT collision (:96-110, 223-245, 424-451).
// Shan-Chen runs add a first pass per step that produces psi from rho
// (ExternalForce.py:70-75, 114-120); the collide pass then reads psi at the
// 8 neighbours (ExternalForce.py:78-91, 124-145).
//
// Memory layout: SoA planes, node (i, j) at i*NY + j, population plane
// c*9 + k.  One thread owns one node; a wavefront covers 64 consecutive nodes
// of a row, so every population load is a coalesced (at most one-line
// misaligned) 256 B access and every store is aligned and non-temporal.
// Solid nodes read only their code (2 B).  No LDS and no MFMA: the step is a
// bandwidth-bound stencil (about 2-5 flop per byte).

#include "lbm_physics.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>

namespace lbm {

struct StepArgs {
  const void* src;     // populations read this launch
  void* dst;           // populations written this launch
  const float* mac_in;  // macro planes of the previous step (NEEM inputs)
  float* mac_out;       // macro planes of this step
  float* tau;           // per-node relaxation state (in place for power law)
  const float* psi;     // Shan-Chen psi of this step (nc planes)
  float* body_force;
  float* sc_force;
  float* shear_rate;
  const uint16_t* code;
  uint64_t mdiv;
  uint32_t nx, ny, N;
  uint32_t ps;          // plane stride (elements)
  int32_t local_read;   // populations are post-stream f: read in place
  int32_t write_macro;  // write every node's macro fields (last step)
  int32_t tau_mode;
  int32_t force;        // gravity non-zero
  int32_t sc;           // LBM_SC_*
  int32_t n_neem;
  int32_t xcd_remap;
  uint32_t nblocks;
  float tau0[LBM_MAX_NC], tau_anti0[LBM_MAX_NC], magic[LBM_MAX_NC];
  float s_shear[LBM_MAX_NC], s_bulk[LBM_MAX_NC];
  float visc_m, visc_nm1, tau_min, tau_max, gx, gy;
  float g_coh[LBM_MAX_NC * LBM_MAX_NC];
  float psi_rho0[LBM_MAX_NC];
  lbm_neem_rule neem[LBM_MAX_NEEM];
};

// Block index with XCD-aware remapping: the hardware dispatches consecutive
// workgroups round-robin over the 8 XCDs; remapping gives each XCD one
// contiguous range of the lattice so the lines shared by neighbouring blocks
// stay in one L2.
__device__ __forceinline__ uint32_t logical_block(const StepArgs& a) {
  const uint32_t b = blockIdx.x;
  if (!a.xcd_remap) return b;
  const uint32_t per = a.nblocks >> 3;  // host enables remap only if nblocks % 8 == 0
  return (b & 7u) * per + (b >> 3);
}

// Post-stream populations of node nb (all components): pull f2(x - c_k, k),
// or f2(x, opp k) where the code's bounce bit is set.
template <int NC, int S>
__device__ __forceinline__ void gather(const typename Pop<S>::T* __restrict__ src, uint32_t ps,
                                       const Nbr& nb, uint32_t bmask, float (&f)[NC][9]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const bool bounce = (k > 0) && ((bmask >> (k - 1)) & 1u);
      const uint32_t plane = (uint32_t)(c * 9 + (bounce ? opp(k) : k));
      const uint32_t node = bounce ? nb.n : nb.src(k);
      f[c][k] = Pop<S>::ld(src + (size_t)plane * ps, node, bounce ? opp(k) : k);
    }
  }
}

// Apply NEEM rule `r` at node (i, j) whose pre-rule populations are f.
// Reads the interior neighbour's post-stream populations (re-gathered) and
// its rho/u of the previous step; on `finish` also writes the reference's
// vel / rho / total_rho overrides.
template <int NC, int S>
__device__ void apply_neem(const StepArgs& a, const lbm_neem_rule& r, uint32_t i, uint32_t j,
                           uint32_t n, float (&f)[NC][9], float* mac_override) {
  using T = typename Pop<S>::T;
  const int d = r.direction;
  int di = 0, dj = 0;
  switch (d) {
    case 1: di = 1; break;
    case 2: dj = 1; break;
    case 3: di = -1; break;
    case 4: dj = -1; break;
  }
  const uint32_t inb = (uint32_t)((int)i - di), jnb = (uint32_t)((int)j - dj);
  const Nbr nb(inb, jnb, a.nx, a.ny);
  const uint32_t code_nb = a.code[nb.n];
  float fn[NC][9];
  gather<NC, S>((const T*)a.src, a.ps, nb, code_nb & LBM_CODE_BOUNCE_MASK, fn);
  const float* mi = a.mac_in;
  const uint32_t ps = a.ps;
  const float ux_nb = mi[(size_t)(NC + 1) * ps + nb.n];
  const float uy_nb = mi[(size_t)(NC + 2) * ps + nb.n];
  const float uv_nb = ux_nb * ux_nb + uy_nb * uy_nb;
  if (r.kind == LBM_NEEM_VELOCITY) {
    const float uxw = r.ux, uyw = r.uy;
    const float uvw = uxw * uxw + uyw * uyw;
    float rw[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) rw[c] = neem_rhow(d, uxw, uyw, f[c]);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float rho_nb = mi[(size_t)c * ps + nb.n];
#pragma unroll
      for (int k = 0; k < 9; ++k)
        f[c][k] = feq_k(k, uxw, uyw, rw[c], uvw) + fn[c][k] - feq_k(k, ux_nb, uy_nb, rho_nb, uv_nb);
    }
    if (mac_override) {
#pragma unroll
      for (int c = 0; c < NC; ++c) mac_override[(size_t)c * ps + n] = rw[c];
      mac_override[(size_t)(NC + 1) * ps + n] = uxw;
      mac_override[(size_t)(NC + 2) * ps + n] = uyw;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float rho_own = mi[(size_t)c * ps + n];
      const float rho_nb = mi[(size_t)c * ps + nb.n];
#pragma unroll
      for (int k = 0; k < 9; ++k)
        if (neem_unknown(d, k))
          f[c][k] = feq_k(k, ux_nb, uy_nb, rho_own, uv_nb) + fn[c][k] - feq_k(k, ux_nb, uy_nb, rho_nb, uv_nb);
    }
    if (mac_override) {
      mac_override[(size_t)NC * ps + n] = r.rho;
      mac_override[(size_t)(NC + 1) * ps + n] = ux_nb;
      mac_override[(size_t)(NC + 2) * ps + n] = uy_nb;
    }
  }
}

enum { MODE_STEP = 0, MODE_FINISH = 1, MODE_PSI = 2 };

template <int COLL, int NC, int S, int MODE>
__global__ __launch_bounds__(kBlock) void k_fused(const StepArgs a) {
  using T = typename Pop<S>::T;
  const uint32_t n = logical_block(a) * kBlock + threadIdx.x;
  if (n >= a.N) return;
  const uint32_t ps = a.ps;
  const uint32_t code = a.code[n];
  if (!(code & LBM_CODE_ACTIVE)) {
    if (MODE == MODE_STEP && a.write_macro) {
#pragma unroll
      for (int p = 0; p < NC + 3; ++p) a.mac_out[(size_t)p * ps + n] = 0.0f;
    }
    return;
  }
  const uint32_t i = div_ny(n, a.mdiv);
  const uint32_t j = n - i * a.ny;
  const T* __restrict__ src = (const T*)a.src;

  float f[NC][9];
  if (MODE == MODE_STEP && a.local_read) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 9; ++k) f[c][k] = Pop<S>::ld(src + (size_t)(c * 9 + k) * ps, n, k);
  } else {
    const Nbr nb(i, j, a.nx, a.ny);
    gather<NC, S>(src, ps, nb, code & LBM_CODE_BOUNCE_MASK, f);
    const uint32_t e = (code & LBM_CODE_NEEM_MASK) >> LBM_CODE_NEEM_SHIFT;
    if (e) apply_neem<NC, S>(a, a.neem[e - 1], i, j, n, f, MODE == MODE_FINISH ? a.mac_out : nullptr);
  }

  if constexpr (MODE == MODE_FINISH) {
    T* dst = (T*)a.dst;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 9; ++k) Pop<S>::st_cached(dst + (size_t)(c * 9 + k) * ps, n, k, f[c][k]);
    return;
  }

  // S1 density (lbm.py:153-164)
  float rho[NC];
  float rtot = 0.0f;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) s += f[c][k];
    rho[c] = s;
    rtot += s;
  }

  if constexpr (MODE == MODE_PSI) {
    float* psi_out = (float*)a.dst;
#pragma unroll
    for (int c = 0; c < NC; ++c) psi_out[(size_t)c * ps + n] = sc_psi(rho[c], a.psi_rho0[c]);
    return;
  }

  // Shan-Chen force (ExternalForce.py:78-91 / 124-145); psi read at x + c_k.
  float scx[NC], scy[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) scx[c] = scy[c] = 0.0f;
  if (a.sc != LBM_SC_NONE) {
    const Nbr nb(i, j, a.nx, a.ny);
    const float* psi = a.psi;
    if (a.sc == LBM_SC_C1) {
      float Fx = 0.0f, Fy = 0.0f;
#pragma unroll
      for (int k = 1; k < 9; ++k) {
        const float p = psi[nb.src(opp(k))];  // x - c_opp(k) = x + c_k
        Fx += (wgt(k) * (float)cx(k)) * p;
        Fy += (wgt(k) * (float)cy(k)) * p;
      }
      const float s = -a.g_coh[0] * psi[n];
      scx[0] = Fx * s;
      scy[0] = Fy * s;
    } else {
#pragma unroll
      for (int c1 = 0; c1 < NC; ++c1) {
        float Fx = 0.0f, Fy = 0.0f;
#pragma unroll
        for (int k = 1; k < 9; ++k) {
          const uint32_t xk = nb.src(opp(k));
#pragma unroll
          for (int c2 = 0; c2 < NC; ++c2) {
            if (c2 == c1) continue;
            const float p = psi[(size_t)c2 * ps + xk];
            const float G = a.g_coh[c1 * LBM_MAX_NC + c2];
            Fx += (wgt(k) * (float)cx(k)) * p * G;
            Fy += (wgt(k) * (float)cy(k)) * p * G;
          }
        }
        const float s = -psi[(size_t)c1 * ps + n];
        scx[c1] = s * Fx;
        scy[c1] = s * Fy;
      }
    }
  }

  // S3 force_density (lbm.py:215-220)
  const bool force = a.force || a.sc != LBM_SC_NONE;
  float Fx[NC], Fy[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const float frac = rho[c] / rtot;
    Fx[c] = scx[c] + frac * a.gx;
    Fy[c] = scy[c] + frac * a.gy;
  }

  // S4 velocity (lbm.py:223-237)
  float vx = 0.0f, vy = 0.0f;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const float* q = f[c];
    float tx = q[1] - q[3] + q[5] - q[6] - q[7] + q[8];
    float ty = q[2] - q[4] + q[5] + q[6] - q[7] - q[8];
    if (force) {
      tx = tx + 0.5f * Fx[c];
      ty = ty + 0.5f * Fy[c];
    }
    vx += tx;
    vy += ty;
  }
  const float ux = vx / rtot, uy = vy / rtot;

  // S5 power law + S6 collide, per component
  T* __restrict__ dst = (T*)a.dst;
  float gam[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    float tau = a.tau0[c], tau_anti = a.tau_anti0[c], s_shear = a.s_shear[c];
    gam[c] = 0.0f;
    if (a.tau_mode == LBM_TAU_FIELD) {
      const float state = a.tau[(size_t)c * ps + n];
      if constexpr (COLL == LBM_MRT) {
        s_shear = state;
      } else {
        tau = state;
        if constexpr (COLL == LBM_TRT) tau_anti = a.magic[c] / (state - 0.5f) + 0.5f;
      }
    } else if (a.tau_mode == LBM_TAU_POWER_LAW) {
      float* tp = a.tau + (size_t)c * ps + n;
      const float state = *tp;
      const float g = shear_rate_of<COLL>(f[c], rho[c], ux, uy, state, a.s_bulk[c]);
      gam[c] = g;
      const float nu = a.visc_m * powf(g, a.visc_nm1);
      const float t = clamp_tau(nu * 3.0f + 0.5f, a.tau_min, a.tau_max);
      if constexpr (COLL == LBM_MRT) {
        s_shear = 1.0f / t;
        *tp = s_shear;
      } else {
        tau = t;
        *tp = t;
        if constexpr (COLL == LBM_TRT) tau_anti = a.magic[c] / (t - 0.5f) + 0.5f;
      }
    }
    float out[9];
    collide<COLL>(f[c], out, rho[c], ux, uy, Fx[c], Fy[c], force, tau, tau_anti, s_shear, a.s_bulk[c]);
#pragma unroll
    for (int k = 0; k < 9; ++k) Pop<S>::st(dst + (size_t)(c * 9 + k) * ps, n, k, out[k]);
  }

  if (a.write_macro || (code & LBM_CODE_KEEP_MACRO)) {
    float* mo = a.mac_out;
#pragma unroll
    for (int c = 0; c < NC; ++c) mo[(size_t)c * ps + n] = rho[c];
    mo[(size_t)NC * ps + n] = rtot;
    mo[(size_t)(NC + 1) * ps + n] = ux;
    mo[(size_t)(NC + 2) * ps + n] = uy;
  }
  if (a.write_macro) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (force) {
        a.body_force[(size_t)(2 * c) * ps + n] = Fx[c];
        a.body_force[(size_t)(2 * c + 1) * ps + n] = Fy[c];
      }
      if (a.sc != LBM_SC_NONE) {
        a.sc_force[(size_t)(2 * c) * ps + n] = scx[c];
        a.sc_force[(size_t)(2 * c + 1) * ps + n] = scy[c];
      }
      if (a.tau_mode == LBM_TAU_POWER_LAW) a.shear_rate[(size_t)c * ps + n] = gam[c];
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

}  // namespace lbm

extern "C" const char* lbm_last_error(void) { return lbm::g_err.c_str(); }
extern "C" int lbm_abi_version(void) { return LBM_ABI_VERSION; }

namespace lbm {

int fail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return LBM_E_INVALID;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return LBM_OK;
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return LBM_E_HIP;
}

int validate(const lbm_lattice* lat, const lbm_fields* fld) {
  if (!lat || !fld) return fail("null lattice or fields");
  if (lat->nx < 1 || lat->ny < 1) return fail("bad lattice size %d x %d", lat->nx, lat->ny);
  if (lat->nc < 1 || lat->nc > LBM_MAX_NC) return fail("num_components %d unsupported (1..%d)", lat->nc, LBM_MAX_NC);
  const uint64_t N = (uint64_t)lat->nx * (uint64_t)lat->ny;
  if (N >= (1ull << 31)) return fail("lattice of %llu nodes exceeds 2^31", (unsigned long long)N);
  if ((uint64_t)lat->plane_stride < N) return fail("plane_stride %lld < NX*NY", (long long)lat->plane_stride);
  if ((uint64_t)lat->plane_stride * (uint64_t)(lat->nc * 9) >= (1ull << 32))
    return fail("population buffer exceeds 2^32 elements");
  if (lat->collision < LBM_BGK || lat->collision > LBM_MRT) return fail("bad collision kind %d", lat->collision);
  if (lat->storage < LBM_STORE_F32 || lat->storage > LBM_STORE_BF16_SHIFT) return fail("bad storage %d", lat->storage);
  if (lat->shan_chen < LBM_SC_NONE || lat->shan_chen > LBM_SC_C2) return fail("bad shan_chen %d", lat->shan_chen);
  if (lat->shan_chen == LBM_SC_C2 && lat->nc != 2) return fail("ShanChenForceC2 needs 2 components");
  if (lat->n_neem < 0 || lat->n_neem > LBM_MAX_NEEM) return fail("n_neem %d out of range", lat->n_neem);
  if (!fld->f[0] || !fld->f[1] || !fld->macro[0] || !fld->macro[1] || !fld->code)
    return fail("missing population / macro / code buffer");
  if (lat->tau_mode < LBM_TAU_CONST || lat->tau_mode > LBM_TAU_POWER_LAW) return fail("bad tau_mode %d", lat->tau_mode);
  if (lat->tau_mode != LBM_TAU_CONST && !fld->tau) return fail("tau_mode %d needs the tau buffer", lat->tau_mode);
  if (lat->tau_mode == LBM_TAU_POWER_LAW && !fld->shear_rate) return fail("power law needs the shear_rate buffer");
  if (lat->shan_chen && (!fld->psi[0] || !fld->psi[1] || !fld->sc_force))
    return fail("Shan-Chen needs psi and sc_force buffers");
  if (!fld->body_force) return fail("missing body_force buffer");
  return LBM_OK;
}

StepArgs make_args(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem) {
  StepArgs a;
  memset(&a, 0, sizeof a);
  a.nx = (uint32_t)lat->nx;
  a.ny = (uint32_t)lat->ny;
  a.N = a.nx * a.ny;
  a.ps = (uint32_t)lat->plane_stride;
  a.mdiv = lat->ny == 1 ? 0ull : (~0ull) / (uint64_t)lat->ny + 1ull;
  a.code = fld->code;
  a.tau = fld->tau;
  a.body_force = fld->body_force;
  a.sc_force = fld->sc_force;
  a.shear_rate = fld->shear_rate;
  a.tau_mode = lat->tau_mode;
  a.force = (lat->gx != 0.0f || lat->gy != 0.0f);
  a.sc = lat->shan_chen;
  a.n_neem = lat->n_neem;
  for (int c = 0; c < LBM_MAX_NC; ++c) {
    a.tau0[c] = lat->tau[c];
    a.tau_anti0[c] = lat->tau_anti[c];
    a.magic[c] = lat->magic[c];
    a.s_shear[c] = lat->s_shear[c];
    a.s_bulk[c] = lat->s_bulk[c];
    a.psi_rho0[c] = lat->psi_rho0[c];
  }
  for (int q = 0; q < LBM_MAX_NC * LBM_MAX_NC; ++q) a.g_coh[q] = lat->g_coh[q];
  a.visc_m = lat->visc_m;
  a.visc_nm1 = lat->visc_nm1;
  a.tau_min = lat->tau_min;
  a.tau_max = lat->tau_max;
  a.gx = lat->gx;
  a.gy = lat->gy;
  for (int r = 0; r < lat->n_neem && neem; ++r) a.neem[r] = neem[r];
  a.nblocks = (a.N + kBlock - 1) / kBlock;
  a.xcd_remap = (a.nblocks % 8 == 0 && a.nblocks >= 64) ? 1 : 0;
  return a;
}

using KernelFn = void (*)(const StepArgs);

template <int COLL, int NC, int S>
KernelFn pick_mode(int mode) {
  switch (mode) {
    case MODE_STEP: return k_fused<COLL, NC, S, MODE_STEP>;
    case MODE_FINISH: return k_fused<COLL, NC, S, MODE_FINISH>;
    default: return k_fused<COLL, NC, S, MODE_PSI>;
  }
}
template <int NC, int S>
KernelFn pick_coll(int coll, int mode) {
  // FINISH and PSI do not depend on the collision model.
  if (mode != MODE_STEP) return pick_mode<LBM_BGK, NC, S>(mode);
  switch (coll) {
    case LBM_BGK: return pick_mode<LBM_BGK, NC, S>(mode);
    case LBM_TRT: return pick_mode<LBM_TRT, NC, S>(mode);
    default: return pick_mode<LBM_MRT, NC, S>(mode);
  }
}
template <int NC>
KernelFn pick_store(int store, int coll, int mode) {
  switch (store) {
    case LBM_STORE_F32: return pick_coll<NC, LBM_STORE_F32>(coll, mode);
    case LBM_STORE_F16_SHIFT: return pick_coll<NC, LBM_STORE_F16_SHIFT>(coll, mode);
    default: return pick_coll<NC, LBM_STORE_BF16_SHIFT>(coll, mode);
  }
}
KernelFn pick(int nc, int store, int coll, int mode) {
  return nc == 1 ? pick_store<1>(store, coll, mode) : pick_store<2>(store, coll, mode);
}

int launch(KernelFn fn, const StepArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(fn, dim3(a.nblocks), dim3(kBlock), 0, s, a);
  return hip_check(hipGetLastError(), "kernel launch");
}

}  // namespace lbm

extern "C" int lbm_run(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
                       int32_t nsteps, int32_t parity, int32_t first_local, void* stream) {
  using namespace lbm;
  int rc = validate(lat, fld);
  if (rc) return rc;
  if (nsteps < 0) return fail("nsteps %d < 0", nsteps);
  if (parity != 0 && parity != 1) return fail("parity must be 0 or 1");
  if (lat->n_neem > 0 && !neem) return fail("missing NEEM rule table");
  hipStream_t s = (hipStream_t)stream;
  StepArgs a = make_args(lat, fld, neem);
  KernelFn step = pick(lat->nc, lat->storage, lat->collision, MODE_STEP);
  KernelFn psik = pick(lat->nc, lat->storage, lat->collision, MODE_PSI);
  int p = parity;
  for (int t = 0; t < nsteps; ++t) {
    a.local_read = (t == 0 && first_local) ? 1 : 0;
    a.src = fld->f[p];
    a.mac_in = fld->macro[p];
    a.mac_out = fld->macro[1 - p];
    a.write_macro = (t == nsteps - 1) ? 1 : 0;
    if (lat->shan_chen) {
      StepArgs b = a;
      b.dst = fld->psi[1 - p];
      if ((rc = launch(psik, b, s))) return rc;
      a.psi = fld->psi[1 - p];
    }
    a.dst = fld->f[1 - p];
    if ((rc = launch(step, a, s))) return rc;
    p = 1 - p;
  }
  return LBM_OK;
}

extern "C" int lbm_finish(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
                          int32_t parity, void* stream) {
  using namespace lbm;
  int rc = validate(lat, fld);
  if (rc) return rc;
  if (parity != 0 && parity != 1) return fail("parity must be 0 or 1");
  if (lat->n_neem > 0 && !neem) return fail("missing NEEM rule table");
  StepArgs a = make_args(lat, fld, neem);
  a.src = fld->f[parity];
  a.dst = fld->f[1 - parity];
  a.mac_in = fld->macro[parity];
  a.mac_out = fld->macro[parity];
  return launch(pick(lat->nc, lat->storage, lat->collision, MODE_FINISH), a, (hipStream_t)stream);
}
