// Fused pull-scheme stream + collide for D2Q9 on MI355X (gfx950).
//
// One launch per time step.  A launch reads the post-collision populations f2
// of step n-1 and produces those of step n, doing in between, per node:
//   S7(n-1)  the reference's ordered boundary pass (BoundaryCondition.py:
//            481-483): periodic / table pull, halfway bounce-back (:244-259),
//            VelocityNEEM (:154-166), PressureNEEM (:215-226), VelocityBB
//            (:85-102) and OpenNeumann (:321-334), compiled on the host into
//            one uint16 code per node;
//   S1..S6(n) density (lbm.py:153-164), force_density (:215-220), velocity
//            (:223-237), power-law shear_rate/relaxation_time
//            (CollisionEngine.py:72-92, 157-178, 322-377) and the BGK / TRT /
//            MRT collision (:96-110, 223-245, 424-451).
// Shan-Chen steps (ExternalForce.py:70-145) run as one tiled launch that
// stages psi of the tile and a one-node ring in LDS (k_sc below); FP64
// storage keeps a psi pass of its own before the step.
//
// Memory layout: SoA planes, node (i, j) at i*NY + j, population plane
// c*9 + k.  One thread owns one node; a wavefront covers 64 consecutive nodes
// of a row, so every population load is a coalesced (at most one-line
// misaligned) 256 B access and every store is aligned and non-temporal.
// Solid nodes read only their code (2 B).  No LDS and no MFMA: the step is a
// bandwidth-bound stencil (about 2-5 flop per byte).
//
// The per-node body (node_update) is shared by the single-lattice kernel
// (k_fused: one block per 256 nodes) and the batched kernel (k_batch: a
// persistent grid over the 256-node tiles of many independent lattices,
// each described by a device-resident Lat record, replayable from a graph).

#include "lbm_physics.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

namespace lbm {

// One lattice as the kernels see it.  Single-lattice launches pass it by
// value; batched launches read it from a device array.
struct Lat {
  void* pop[2];          // population ping-pong (nc*9 planes each)
  void* mac[2];          // macro ping-pong: rho_0..rho_{nc-1}, total_rho, ux, uy (element type R)
  void* tau;             // per-node relaxation state (R), nc planes
  void* psi[2];          // Shan-Chen psi (R), nc planes each
  float* body_force;
  float* sc_force;
  float* shear_rate;
  float* vel_old;        // 2 planes: velocity of the last residual check (batches)
  const uint16_t* code;
  uint64_t mdiv;
  uint32_t nx, ny, N, ps;
  int32_t coll, tau_mode, force, sc, n_neem;
  int32_t start_step;    // batched runs: launch index whose populations are read in place
  uint32_t block0;       // batched runs: first 256-node tile of this lattice
  uint32_t active;
  void* tau2;                  // batched two-step runs of power-law lattices: second state buffer
  uint32_t tile_oi, tile_oj;   // batched two-step runs: shift of the tile grid (tile (ti, tj) starts at
                               // row ti*TI - tile_oi, column tj*W - tile_oj)
  float tau0[LBM_MAX_NC], tau_anti0[LBM_MAX_NC], magic[LBM_MAX_NC];
  float s_shear[LBM_MAX_NC], s_bulk[LBM_MAX_NC];
  float visc_m, visc_nm1, tau_min, tau_max, gx, gy;
  float g_coh[LBM_MAX_NC * LBM_MAX_NC];
  float psi_rho0[LBM_MAX_NC];
  int32_t psi_kind[LBM_MAX_NC];
  float eos[LBM_MAX_NC][6];
  float g_adh[LBM_MAX_NC];
  int32_t sc_adhesion;         // some node carries LBM_CODE_SC_SOLID
  const float* temperature;    // lb_field.T (cubic-EOS psi)
  int32_t huang;               // HuangMRTCollision: MRT with the source S Q (huang_q)
  float huang_k1[LBM_MAX_NC], huang_k2[LBM_MAX_NC], huang_g[LBM_MAX_NC];
  uint32_t parity_xor;         // batched Shan-Chen runs: this lattice's ping-pong phase against the launch's
  uint32_t stochastic;         // 16-bit storage: stochastic rounding of the steps' stores
  uint32_t round_base;         // steps the lattice has taken before this call (16-bit stochastic rounding)
  const uint32_t* round_ctr;   // batches: the same count in device memory (advanced by k_round_tick)
  lbm_neem_rule neem[LBM_MAX_NEEM];
};

// psi of component c from the node's density (ShanChenForceC1/C2.update_psi,
// ExternalForce.py:42-47, 86-92, with the psi objects of EOS.py)
template <typename R>
__device__ __forceinline__ R node_psi(const Lat& a, int c, R rho, uint32_t n) {
  if (a.psi_kind[c] == LBM_PSI_SC) return sc_psi<R>(rho, (R)a.psi_rho0[c]);
  return psi_of<R>(a.psi_kind[c], rho, (R)a.temperature[n], (R)a.psi_rho0[c], a.eos[c]);
}

// Per-launch switches.
struct Launch {
  uint32_t parity;       // pop[parity] / mac[parity] are read
  int32_t step;          // launch index within the run (batches: in-place read when == start_step)
  int32_t local_read;    // single lattice: populations are post-stream f, read in place
  int32_t write_macro;   // write every node's macro fields (last step of a run)
  int32_t xcd_remap;
  uint32_t nblocks;
  uint32_t round_step;   // steps of the call up to and including the one this launch stores
};

enum { MODE_STEP = 0, MODE_FINISH = 1, MODE_PSI = 2 };

// Block index with XCD-aware remapping: the hardware dispatches consecutive
// workgroups round-robin over the 8 XCDs; remapping gives each XCD one
// contiguous range of the lattice so the lines shared by neighbouring blocks
// stay in one L2.
__device__ __forceinline__ uint32_t logical_block(const Launch& q) {
  const uint32_t b = blockIdx.x;
  if (!q.xcd_remap) return b;
  const uint32_t per = q.nblocks >> 3;  // host enables remap only if nblocks % 8 == 0
  return (b & 7u) * per + (b >> 3);
}

// Post-stream populations of node nb (all components): pull f2(x - c_k, k),
// or f2(x, opp k) where the code's bounce bit is set.
template <int NC, int S>
__device__ __forceinline__ void gather(const typename Pop<S>::T* __restrict__ src, uint32_t ps, const Nbr& nb,
                                       uint32_t bmask, typename Pop<S>::R (&f)[NC][9]) {
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

// Where a node rule reads the previous step's state: post-collision
// populations f2 and macroscopic planes in HBM.  (The two-step kernel has an
// LDS counterpart with the same interface.)
template <int S>
struct GlobalView {
  using T = typename Pop<S>::T;
  using R = typename Pop<S>::R;
  struct Node {
    uint32_t i, j;
    Nbr nb;
  };
  const T* src;
  const R* mac;
  const uint16_t* code;
  uint32_t ps, nx, ny;
  __device__ __forceinline__ Node node(uint32_t i, uint32_t j) const { return Node{i, j, Nbr(i, j, nx, ny)}; }
  // x - (di, dj): rule neighbours lie inside the lattice (checked on the host)
  __device__ __forceinline__ Node back(const Node& x, int di, int dj) const {
    return node((uint32_t)((int)x.i - di), (uint32_t)((int)x.j - dj));
  }
  __device__ __forceinline__ uint32_t code_of(const Node& x) const { return code[x.nb.n]; }
  __device__ __forceinline__ R macro(const Node& x, int plane) const { return mac[(size_t)plane * ps + x.nb.n]; }
  // post-stream population k of component c at x: pull f2(x - c_k, k), or
  // f2(x, opp k) where x's code has the bounce bit of k
  __device__ __forceinline__ R pop(const Node& x, uint32_t bm, int c, int k) const {
    const bool bounce = (k > 0) && ((bm >> (k - 1)) & 1u);
    const int kk = bounce ? opp(k) : k;
    return Pop<S>::ld(src + (size_t)(c * 9 + kk) * ps, bounce ? x.nb.n : x.nb.src(k), kk);
  }
};

// Apply node rule `r` at node x (global index n) whose pre-rule populations
// are f.  NEEM rules read the interior neighbour's post-stream populations
// and its rho/u of the previous step; VelocityBB reads the node's own rho of
// the previous step; OpenNeumann copies the unknowns from the neighbour.  On
// `finish` the reference's vel / rho / total_rho overrides are written into
// mac_override.
template <int NC, int S, class V>
__device__ void apply_rule(const Lat& a, const V& v, const lbm_neem_rule& r, const typename V::Node& x, uint32_t n,
                           typename Pop<S>::R (&f)[NC][9], typename Pop<S>::R* mac_override) {
  using R = typename Pop<S>::R;
  const int d = r.direction;
  const uint32_t ps = a.ps;
  if (r.kind == LBM_RULE_VELOCITY_BB) {
    // BoundaryCondition.py:95-102: f_k += 6 w_k rho_c(x) (c_k . u_w) on the unknowns
    const R uxw = r.ux, uyw = r.uy;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const R rw = v.macro(x, c);
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        // a select per population, not a branch: branches on d let the
        // compiler merge the updates of different k into one store through
        // a pointer, which moves f to scratch memory
        const R cv = (R)cx(k) * uxw + (R)cy(k) * uyw;
        const R upd = f[c][k] + R(6) * wgt<R>(k) * rw * cv;
        f[c][k] = neem_unknown(d, k) ? upd : f[c][k];
      }
    }
    return;
  }
  int di = 0, dj = 0;
  switch (d) {
    case 1: di = 1; break;
    case 2: dj = 1; break;
    case 3: di = -1; break;
    case 4: dj = -1; break;
  }
  if (r.kind == LBM_RULE_VELOCITY_EQ) {
    // BoundaryCondition.py:135-145: rho_w from the node's populations (Zou-He
    // form), then all nine := feq(u_w, rho_w)
    const R uxw = r.ux, uyw = r.uy;
    const R uvw = uxw * uxw + uyw * uyw;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const R rw = neem_rhow<R>(d, uxw, uyw, f[c]);
#pragma unroll
      for (int k = 0; k < 9; ++k) f[c][k] = feq_k<R>(k, uxw, uyw, rw, uvw);
      if (mac_override) mac_override[(size_t)c * ps + n] = rw;
    }
    if (mac_override) {
      mac_override[(size_t)(NC + 1) * ps + n] = uxw;
      mac_override[(size_t)(NC + 2) * ps + n] = uyw;
    }
    return;
  }
  const typename V::Node y = v.back(x, di, dj);
  const uint32_t bm_y = v.code_of(y) & LBM_CODE_BOUNCE_MASK;
  if (r.kind == LBM_RULE_PRESSURE_ABC) {
    // BoundaryCondition.py:190-205: u_w = 1.5 u(x) - 0.5 u(x - c_d) from the
    // velocity pass of the step, then on the unknowns (written by an earlier
    // bounce-back) f_k := -f_k + 2 w_k rho_value (1 + 4.5 (c_k . u_w)^2 - 1.5 u_w^2)
    const R uxw = R(1.5) * v.macro(x, NC + 1) - R(0.5) * v.macro(y, NC + 1);
    const R uyw = R(1.5) * v.macro(x, NC + 2) - R(0.5) * v.macro(y, NC + 2);
    const R uw2 = uxw * uxw + uyw * uyw;
    const R rw = r.rho;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const R cu = (R)cx(k) * uxw + (R)cy(k) * uyw;
        const R upd = -f[c][k] + R(2) * wgt<R>(k) * rw * (R(1) + R(4.5) * cu * cu - R(1.5) * uw2);
        f[c][k] = neem_unknown(d, k) ? upd : f[c][k];
      }
    }
    if (mac_override) {
      mac_override[(size_t)NC * ps + n] = rw;
      mac_override[(size_t)(NC + 1) * ps + n] = uxw;
      mac_override[(size_t)(NC + 2) * ps + n] = uyw;
    }
    return;
  }
  if (r.kind == LBM_RULE_OPEN_NEUMANN || r.kind == LBM_RULE_OPEN_CONVECTIVE1) {
    // BoundaryCondition.py:327-334: unknowns := f(x - c_d); :349-359:
    // unknowns := (f + l f(x - c_d)) / (1 + l), l = |u(x - c_d)| of the
    // step's velocity pass; then density and velocity of the node
    // (OpenBoundary.density / vel, :274-298)
    if (r.kind == LBM_RULE_OPEN_NEUMANN) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const R g = neem_unknown(d, k) ? v.pop(y, bm_y, c, k) : f[c][k];
          f[c][k] = g;
        }
    } else {
      const R ux2 = v.macro(y, NC + 1), uy2 = v.macro(y, NC + 2);
      const R lam = sqrt(ux2 * ux2 + uy2 * uy2);
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const R g = neem_unknown(d, k) ? (f[c][k] + lam * v.pop(y, bm_y, c, k)) / (R(1) + lam) : f[c][k];
          f[c][k] = g;
        }
    }
    if (mac_override) {
      R rtot = R(0), vx = R(0), vy = R(0);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R s = R(0);
#pragma unroll
        for (int k = 0; k < 9; ++k) s += f[c][k];
        mac_override[(size_t)c * ps + n] = s;
        rtot += s;
        const R* g = f[c];
        R tx = g[1] - g[3] + g[5] - g[6] - g[7] + g[8];
        R ty = g[2] - g[4] + g[5] + g[6] - g[7] - g[8];
        if ((a.force || a.sc != LBM_SC_NONE) && a.body_force) {
          tx = tx + R(0.5) * (R)a.body_force[(size_t)(2 * c) * ps + n];
          ty = ty + R(0.5) * (R)a.body_force[(size_t)(2 * c + 1) * ps + n];
        }
        vx += tx;
        vy += ty;
      }
      mac_override[(size_t)NC * ps + n] = rtot;
      mac_override[(size_t)(NC + 1) * ps + n] = vx / rtot;
      mac_override[(size_t)(NC + 2) * ps + n] = vy / rtot;
    }
    return;
  }
  R fn[NC][9];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int k = 0; k < 9; ++k) fn[c][k] = v.pop(y, bm_y, c, k);
  const R ux_nb = v.macro(y, NC + 1);
  const R uy_nb = v.macro(y, NC + 2);
  const R uv_nb = ux_nb * ux_nb + uy_nb * uy_nb;
  if (r.kind == LBM_NEEM_VELOCITY) {
    const R uxw = r.ux, uyw = r.uy;
    const R uvw = uxw * uxw + uyw * uyw;
    R rw[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) rw[c] = neem_rhow<R>(d, uxw, uyw, f[c]);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const R rho_nb = v.macro(y, c);
#pragma unroll
      for (int k = 0; k < 9; ++k)
        f[c][k] = feq_k<R>(k, uxw, uyw, rw[c], uvw) + fn[c][k] - feq_k<R>(k, ux_nb, uy_nb, rho_nb, uv_nb);
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
      const R rho_own = v.macro(x, c);
      const R rho_nb = v.macro(y, c);
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const R upd = feq_k<R>(k, ux_nb, uy_nb, rho_own, uv_nb) + fn[c][k] - feq_k<R>(k, ux_nb, uy_nb, rho_nb, uv_nb);
        f[c][k] = neem_unknown(d, k) ? upd : f[c][k];
      }
    }
    if (mac_override) {
      mac_override[(size_t)NC * ps + n] = (R)r.rho;
      mac_override[(size_t)(NC + 1) * ps + n] = ux_nb;
      mac_override[(size_t)(NC + 2) * ps + n] = uy_nb;
    }
  }
}

// Everything one step computes at a node once its post-boundary populations
// f are known (S1 density .. S6 collide), kept in registers.
template <int NC, typename R>
struct StepOut {
  R f[NC][9];                 // post-collision populations
  R rho[NC], rtot, ux, uy;    // macroscopic fields of the step
  R Fx[NC], Fy[NC];           // body force (force_density)
  R scx[NC], scy[NC];         // Shan-Chen force
  R gam[NC];                  // squared shear rate (power law)
  R state[NC];                // new relaxation state (power law)
};

// A node's new populations to pop planes (BUF: buffer stores).  With 16-bit
// storage and a.stochastic the step's stores round stochastically, seeded by
// the node's row and column and the lattice's step count at the stored step
// (round_step: steps of this call up to it); else to nearest even.
template <int NC, int S, bool BUF>
__device__ __forceinline__ void store_pops(const Lat& a, uint32_t round_step, typename Pop<S>::T* __restrict__ dst,
                                           Rsrc rdst, uint32_t ps, uint32_t n, uint32_t gi, uint32_t gj,
                                           const typename Pop<S>::R (&f)[NC][9]) {
  using T = typename Pop<S>::T;
  constexpr bool sr = S == LBM_STORE_F16_SHIFT || S == LBM_STORE_BF16_SHIFT;
  if (sr && a.stochastic) {
    const uint32_t base = a.round_ctr ? *a.round_ctr : a.round_base;
    const SrSeed sd(gi, gj, round_key(base + round_step));
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const uint32_t seed = sd.bits(k) ^ (0x9e3779b9u * (uint32_t)c);
        if constexpr (BUF)
          Pop<S>::stb_sr(rdst, n * (uint32_t)sizeof(T), (uint32_t)(c * 9 + k) * ps * (uint32_t)sizeof(T), k, f[c][k],
                         seed);
        else
          Pop<S>::st_sr(dst + (size_t)(c * 9 + k) * ps, n, k, f[c][k], seed);
      }
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      if constexpr (BUF)
        Pop<S>::stb(rdst, n * (uint32_t)sizeof(T), (uint32_t)(c * 9 + k) * ps * (uint32_t)sizeof(T), k, f[c][k]);
      else
        Pop<S>::st(dst + (size_t)(c * 9 + k) * ps, n, k, f[c][k]);
    }
  }
}

// Where a step reads Shan-Chen psi of the step at x + c_k (k = 0: x itself):
// psi.at(c, k) for component c.  PsiPlanes: the psi plane set `set` in HBM,
// periodic neighbours; PsiNone: lattices without Shan-Chen forces (never
// read).  The LDS tile of the fused Shan-Chen kernel is a third source.
template <typename R>
struct PsiPlanes {
  const Lat& a;
  uint32_t set, i, j;
  __device__ __forceinline__ R at(int c, int k) const {
    const Nbr nb(i, j, a.nx, a.ny);
    return ((const R*)a.psi[set])[(size_t)c * a.ps + (k == 0 ? nb.n : nb.src(opp(k)))];   // x - c_opp(k) = x + c_k
  }
  // x + c_k is not fluid for the Shan-Chen force (fluid_strategy false)
  __device__ __forceinline__ bool solid(int k) const {
    const Nbr nb(i, j, a.nx, a.ny);
    return (a.code[nb.src(opp(k))] & LBM_CODE_SC_SOLID) != 0;
  }
};
template <typename R>
struct PsiNone {
  __device__ __forceinline__ R at(int, int) const { return R(0); }
  __device__ __forceinline__ bool solid(int) const { return false; }
};

// Immersed-boundary correction of a node's force (no Shan-Chen): the body
// force that multi-direct forcing spread onto the node (nc x 2 planes), added
// to the node's own force before its velocity is taken (ibm.py:229-248: the
// spread adds F/(2 rho_tot) to u and F rho_c/rho_tot to each component's
// force, which is the velocity's half-force term of the added force).
struct IbForce {
  const float* df;
  uint32_t ps, n;
  __device__ __forceinline__ float at(int, int) const { return 0.0f; }
  __device__ __forceinline__ bool solid(int) const { return false; }
  __device__ __forceinline__ float fx(int c) const { return df[(size_t)(2 * c) * ps + n]; }
  __device__ __forceinline__ float fy(int c) const { return df[(size_t)(2 * c + 1) * ps + n]; }
};
template <class P, class = void>
struct has_extra_force : std::false_type {};
template <class P>
struct has_extra_force<P, std::void_t<decltype(&P::fx)>> : std::true_type {};
// Psi sources whose steps always carry a force (the Shan-Chen tiles: their
// lattices have a Shan-Chen force; the IB fix-up: the spread force); the
// force terms are then compiled in unconditionally.
template <class P>
struct always_forced : has_extra_force<P> {};

// S1..S6 of one step at node n = (i, j).  state_in is the per-node relaxation
// state (BGK tau, TRT tau_sym, MRT 1/tau) for LBM_TAU_FIELD / _POWER_LAW.
// Shan-Chen reads psi of this step through `psi` (PsiPlanes and the like).
// FMODE: 0 reads the force switch at run time; 1 (kernels that run only
// lattices with a body force) compiles the force terms in, 2 (lattices
// without one) leaves them out -- no run-time switch either way.
template <int COLL, int NC, typename R, class PS, int FMODE = 0>
__device__ __forceinline__ void step_math(const Lat& a, const PS& psi, uint32_t n, uint32_t i, uint32_t j,
                                          const R (&f)[NC][9], const R (&state_in)[NC], StepOut<NC, R>& o) {
  const uint32_t ps = a.ps;
  // S1 density (lbm.py:153-164) with the first moments, from the MRT moment
  // sums: m0 = rho, m3 = jx, m5 = jy (the other moments feed the MRT
  // collision; BGK / TRT do not compute them)
  R mom[NC][9];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    mrt_m(f[c], mom[c]);
    o.rho[c] = mom[c][0];
  }
  o.rtot = o.rho[0];
#pragma unroll
  for (int c = 1; c < NC; ++c) o.rtot += o.rho[c];

  // Shan-Chen force (ExternalForce.py:78-91 / 124-145); psi read at x + c_k.
#pragma unroll
  for (int c = 0; c < NC; ++c) o.scx[c] = o.scy[c] = R(0);
  if (a.sc != LBM_SC_NONE) {
    if (a.sc == LBM_SC_C1) {
      R Fx = R(0), Fy = R(0);
#pragma unroll
      for (int k = 1; k < 9; ++k) {
        const R v = psi.at(0, k);
        Fx += (wgt<R>(k) * (R)cx(k)) * v;
        Fy += (wgt<R>(k) * (R)cy(k)) * v;
      }
      const R s = -(R)a.g_coh[0] * psi.at(0, 0);
      o.scx[0] = Fx * s;
      o.scy[0] = Fy * s;
    } else if (!a.sc_adhesion) {
#pragma unroll
      for (int c1 = 0; c1 < NC; ++c1) {
        R Fx = R(0), Fy = R(0);
#pragma unroll
        for (int k = 1; k < 9; ++k) {
#pragma unroll
          for (int c2 = 0; c2 < NC; ++c2) {
            if (c2 == c1) continue;
            const R v = psi.at(c2, k);
            const R G = a.g_coh[c1 * LBM_MAX_NC + c2];
            Fx += (wgt<R>(k) * (R)cx(k)) * v * G;
            Fy += (wgt<R>(k) * (R)cy(k)) * v * G;
          }
        }
        const R s = -psi.at(c1, 0);
        o.scx[c1] = s * Fx;
        o.scy[c1] = s * Fy;
      }
    } else {
      // ExternalForce.py:105-116: a neighbour where fluid_strategy is false
      // adds w_k c_k gadh[c1] (without psi(x + c_k), commented out at :114)
      // in place of its cohesion terms
#pragma unroll
      for (int c1 = 0; c1 < NC; ++c1) {
        R Fx = R(0), Fy = R(0), Ax = R(0), Ay = R(0);
#pragma unroll
        for (int k = 1; k < 9; ++k) {
          const bool sol = psi.solid(k);
          const R ga = (R)a.g_adh[c1];
          Ax = sol ? Ax + (wgt<R>(k) * (R)cx(k)) * ga : Ax;
          Ay = sol ? Ay + (wgt<R>(k) * (R)cy(k)) * ga : Ay;
#pragma unroll
          for (int c2 = 0; c2 < NC; ++c2) {
            if (c2 == c1) continue;
            const R v = psi.at(c2, k);
            const R G = a.g_coh[c1 * LBM_MAX_NC + c2];
            Fx = sol ? Fx : Fx + (wgt<R>(k) * (R)cx(k)) * v * G;
            Fy = sol ? Fy : Fy + (wgt<R>(k) * (R)cy(k)) * v * G;
          }
        }
        const R s = -psi.at(c1, 0);
        o.scx[c1] = s * (Fx + Ax);
        o.scy[c1] = s * (Fy + Ay);
      }
    }
  }

  // S3 force_density (lbm.py:215-220)
  const bool force = FMODE == 1 || always_forced<PS>::value || (FMODE == 0 && (a.force || a.sc != LBM_SC_NONE));
  const R inv_rtot = rcp(o.rtot);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const R frac = o.rho[c] * inv_rtot;
    o.Fx[c] = o.scx[c] + frac * (R)a.gx;
    o.Fy[c] = o.scy[c] + frac * (R)a.gy;
    if constexpr (has_extra_force<PS>::value) {
      o.Fx[c] += (R)psi.fx(c);
      o.Fy[c] += (R)psi.fy(c);
    }
  }

  // S4 velocity (lbm.py:223-237)
  R vx = R(0), vy = R(0);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    R tx = mom[c][3], ty = mom[c][5];
    if (force) {
      tx = tx + R(0.5) * o.Fx[c];
      ty = ty + R(0.5) * o.Fy[c];
    }
    vx = c == 0 ? tx : vx + tx;
    vy = c == 0 ? ty : vy + ty;
  }
  o.ux = vx * inv_rtot;
  o.uy = vy * inv_rtot;

  // S5 power law + S6 collide, per component.  COLL < 0: the collision model
  // is read from the descriptor (batches mixing models share one kernel; the
  // branches are wave-uniform).
  const int coll = COLL >= 0 ? COLL : a.coll;
  const bool is_mrt = coll == LBM_MRT, is_trt = coll == LBM_TRT;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    R tau = a.tau0[c], tau_anti = a.tau_anti0[c], s_shear = a.s_shear[c];
    const R s_bulk = a.s_bulk[c];
    o.gam[c] = R(0);
    o.state[c] = state_in[c];
    const MrtNeq<R> nq = mrt_neq(mom[c], o.rho[c], o.ux, o.uy);   // MRT, and the power-law shear rate
    if (a.tau_mode == LBM_TAU_FIELD) {
      const R state = state_in[c];
      if (is_mrt) {
        s_shear = state;
      } else {
        tau = state;
        if (is_trt) tau_anti = (R)a.magic[c] * rcp(state - R(0.5)) + R(0.5);
      }
    } else if (a.tau_mode == LBM_TAU_POWER_LAW) {
      const R state = state_in[c];
      const R inv_tau = is_mrt ? R(0) : rcp(state);
      const R g2 = is_mrt ? mrt_shear2(nq, rcp(o.rho[c]), state, s_bulk)
                          : mrt_shear2(nq, rcp(o.rho[c]), inv_tau, inv_tau);
      o.gam[c] = g2;
      const R nu = (R)a.visc_m * pow_shear2(g2, (R)a.visc_nm1);
      const R t = clamp_tau<R>(nu * R(3) + R(0.5), (R)a.tau_min, (R)a.tau_max);
      if (is_mrt) {
        s_shear = rcp(t);
        o.state[c] = s_shear;
      } else {
        tau = t;
        o.state[c] = t;
        if (is_trt) tau_anti = (R)a.magic[c] * rcp(t - R(0.5)) + R(0.5);
      }
    }
    if (is_mrt) {
      // HuangMRTCollision only with a Shan-Chen force (its psi): kernels
      // without one carry no trace of it
      constexpr bool may_huang = !std::is_same<PS, PsiNone<R>>::value;
      if (may_huang && a.huang) {
        const HuangQ<R> hq =
            huang_q<R>(o.Fx[c], o.Fy[c], psi.at(c, 0), (R)a.huang_k1[c], (R)a.huang_k2[c], (R)a.huang_g[c]);
        mrt_collide_m(mom[c], nq, o.f[c], o.ux, o.uy, o.Fx[c], o.Fy[c], force, s_shear, s_bulk, &hq);
      } else {
        mrt_collide_m(mom[c], nq, o.f[c], o.ux, o.uy, o.Fx[c], o.Fy[c], force, s_shear, s_bulk);
      }
    } else if constexpr (COLL >= 0) {
      collide<COLL, R>(f[c], o.f[c], o.rho[c], o.ux, o.uy, o.Fx[c], o.Fy[c], force, tau, tau_anti, s_shear, s_bulk);
    } else if (is_trt) {
      collide<LBM_TRT, R>(f[c], o.f[c], o.rho[c], o.ux, o.uy, o.Fx[c], o.Fy[c], force, tau, tau_anti, s_shear, s_bulk);
    } else {
      collide<LBM_BGK, R>(f[c], o.f[c], o.rho[c], o.ux, o.uy, o.Fx[c], o.Fy[c], force, tau, tau_anti, s_shear, s_bulk);
    }
  }
}

// Writes a step's macroscopic fields (macro plane set `mo`, and on the last
// step of a run the body force, Shan-Chen force and shear rate).
template <int NC, typename R>
__device__ __forceinline__ void store_macro(const Lat& a, R* mo, uint32_t n, const StepOut<NC, R>& o,
                                            bool keep, bool last) {
  const uint32_t ps = a.ps;
  if (last || keep) {
#pragma unroll
    for (int c = 0; c < NC; ++c) mo[(size_t)c * ps + n] = o.rho[c];
    mo[(size_t)NC * ps + n] = o.rtot;
    mo[(size_t)(NC + 1) * ps + n] = o.ux;
    mo[(size_t)(NC + 2) * ps + n] = o.uy;
  }
  if (last) {
    const bool force = a.force || a.sc != LBM_SC_NONE;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (force && a.body_force) {
        a.body_force[(size_t)(2 * c) * ps + n] = (float)o.Fx[c];
        a.body_force[(size_t)(2 * c + 1) * ps + n] = (float)o.Fy[c];
      }
      if (a.sc != LBM_SC_NONE) {
        a.sc_force[(size_t)(2 * c) * ps + n] = (float)o.scx[c];
        a.sc_force[(size_t)(2 * c + 1) * ps + n] = (float)o.scy[c];
      }
      if (a.tau_mode == LBM_TAU_POWER_LAW && a.shear_rate) a.shear_rate[(size_t)c * ps + n] = (float)sqrt(o.gam[c]);
    }
  }
}

template <int COLL, int NC, int S, int MODE>
__device__ __forceinline__ void node_update(const Lat& a, const Launch& q, uint32_t n, bool local) {
  using T = typename Pop<S>::T;
  using R = typename Pop<S>::R;
  const uint32_t ps = a.ps;
  const uint32_t p = q.parity;
  const uint32_t code = a.code[n];
  if (!(code & LBM_CODE_ACTIVE)) {
    if (MODE == MODE_STEP && q.write_macro) {
      R* mo = (R*)a.mac[p ^ 1u];
#pragma unroll
      for (int c = 0; c < NC + 3; ++c) mo[(size_t)c * ps + n] = R(0);
    }
    return;
  }
  const uint32_t i = div_ny(n, a.mdiv);
  const uint32_t j = n - i * a.ny;
  const T* __restrict__ src = (const T*)a.pop[p];

  R f[NC][9];
  // The first launch of a run may read post-stream f in place (no boundary
  // pass): the psi pass and the step of that launch must both do so.
  if (MODE != MODE_FINISH && local) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 9; ++k) f[c][k] = Pop<S>::ld(src + (size_t)(c * 9 + k) * ps, n, k);
  } else {
    const Nbr nb(i, j, a.nx, a.ny);
    gather<NC, S>(src, ps, nb, code & LBM_CODE_BOUNCE_MASK, f);
    const uint32_t e = (code & LBM_CODE_NEEM_MASK) >> LBM_CODE_NEEM_SHIFT;
    if (e) {
      const GlobalView<S> v{src, (const R*)a.mac[p], a.code, ps, a.nx, a.ny};
      apply_rule<NC, S>(a, v, a.neem[e - 1], v.node(i, j), n, f, MODE == MODE_FINISH ? (R*)a.mac[p] : nullptr);
    }
  }

  if constexpr (MODE == MODE_FINISH) {
    T* dst = (T*)a.pop[p ^ 1u];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 9; ++k) Pop<S>::st_cached(dst + (size_t)(c * 9 + k) * ps, n, k, f[c][k]);
    return;
  }

  if constexpr (MODE == MODE_PSI) {
    R* psi_out = (R*)a.psi[p ^ 1u];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      R s = R(0);
#pragma unroll
      for (int k = 0; k < 9; ++k) s += f[c][k];
      psi_out[(size_t)c * ps + n] = node_psi<R>(a, c, s, n);
    }
    return;
  }

  R st_in[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) st_in[c] = a.tau_mode != LBM_TAU_CONST ? ((const R*)a.tau)[(size_t)c * ps + n] : R(0);
  StepOut<NC, R> o;
  step_math<COLL, NC, R>(a, PsiPlanes<R>{a, p ^ 1u, i, j}, n, i, j, f, st_in, o);
  T* __restrict__ dst = (T*)a.pop[p ^ 1u];
  store_pops<NC, S, false>(a, q.round_step, dst, make_rsrc(dst, 0u), ps, n, i, j, o.f);
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (a.tau_mode == LBM_TAU_POWER_LAW) ((R*)a.tau)[(size_t)c * ps + n] = o.state[c];
  store_macro<NC, R>(a, (R*)a.mac[p ^ 1u], n, o, (code & LBM_CODE_KEEP_MACRO) != 0, q.write_macro != 0);
}

template <int COLL, int NC, int S, int MODE>
__global__ __launch_bounds__(kBlock) void k_fused(const Lat a, const Launch q) {
  const uint32_t n = logical_block(q) * kBlock + threadIdx.x;
  if (n >= a.N) return;
  node_update<COLL, NC, S, MODE>(a, q, n, q.local_read != 0);
}

// ---------------------------------------------------------------------------
// Immersed-boundary coupling in the fused step.  The reference runs multi-
// direct forcing between MacroscopicEngine.velocity and the collision
// (ibm.py:248-253), i.e. inside one step.  Here a step is: (0) at every
// stencil node of the marker list, the step's density, force and velocity
// from the pulled populations (step_math: what the fused step computes
// there) into scratch planes, the spread force cleared; (1) multi-direct
// forcing on the scratch planes (lbm_ib_mdf); (2) the fused step on every
// node; (3) at the stencil nodes again, the step recomputed with the spread
// force added, over the fused step's result.  Nodes touched by several
// markers are computed by each with the same inputs and write the same
// values.  Lattices with node rules, Shan-Chen or a power law do not take
// this path (their step at a node needs more than its own pulls).
// ---------------------------------------------------------------------------
struct IbStencil {
  const int2* group;
  const int32_t* count;
  const float* node_pos;   // 2 planes of cap
  uint32_t cap, K;
  int32_t max_number;
  float* vel;              // scratch: 2 planes (u of the step at stencil nodes)
  float* rho;              // nc planes
  float* total_rho;        // 1 plane
  float* body_force;       // nc * 2 planes (the spread force)
};

template <int COLL, int NC, int MODE>   // MODE 0: macros into the scratch planes; 1: fix-up
__global__ __launch_bounds__(256) void k_ib_stencil(const Lat a, const IbStencil ib, uint32_t parity, int32_t local,
                                                    int32_t write_macro) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ib.cap || (int32_t)t >= *ib.count) return;
  const int2 mn = ib.group[t];
  if (mn.x < 0 || mn.x >= ib.max_number || mn.y < 0 || (uint32_t)mn.y >= ib.K) return;
  const uint32_t q = (uint32_t)mn.x * ib.K + (uint32_t)mn.y;
  const float px = fminf(fmaxf(ib.node_pos[q], -2.0f), (float)a.nx + 1.0f);
  const float py = fminf(fmaxf(ib.node_pos[ib.cap + q], -2.0f), (float)a.ny + 1.0f);
  const int x0 = (int)floorf(px), x1 = (int)ceilf(px), y0 = (int)floorf(py), y1 = (int)ceilf(py);
  const uint32_t ps = a.ps;
  const float* __restrict__ src = (const float*)a.pop[parity];
  for (int ix = x0; ix <= x1; ++ix) {
    if (ix < 0 || ix >= (int)a.nx) continue;
    for (int iy = y0; iy <= y1; ++iy) {
      if (iy < 0 || iy >= (int)a.ny) continue;
      const uint32_t i = (uint32_t)ix, j = (uint32_t)iy, n = i * a.ny + j;
      const uint32_t code = a.code[n];
      if (!(code & LBM_CODE_ACTIVE)) {
        if (MODE == 0) {   // as the granular fields outside the fluid group
          ib.vel[n] = 0.0f;
          ib.vel[ps + n] = 0.0f;
          ib.total_rho[n] = 0.0f;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            ib.rho[(size_t)c * ps + n] = 0.0f;
            ib.body_force[(size_t)(2 * c) * ps + n] = 0.0f;
            ib.body_force[(size_t)(2 * c + 1) * ps + n] = 0.0f;
          }
        }
        continue;
      }
      float f[NC][9];
      if (local) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int k = 0; k < 9; ++k) f[c][k] = src[(size_t)(c * 9 + k) * ps + n];
      } else {
        gather<NC, LBM_STORE_F32>(src, ps, Nbr(i, j, a.nx, a.ny), code & LBM_CODE_BOUNCE_MASK, f);
      }
      float st_in[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c)
        st_in[c] = a.tau_mode == LBM_TAU_FIELD ? ((const float*)a.tau)[(size_t)c * ps + n] : 0.0f;
      StepOut<NC, float> o;
      if constexpr (MODE == 0) {
        step_math<COLL, NC, float>(a, PsiNone<float>{}, n, i, j, f, st_in, o);
        ib.vel[n] = o.ux;
        ib.vel[ps + n] = o.uy;
        ib.total_rho[n] = o.rtot;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          ib.rho[(size_t)c * ps + n] = o.rho[c];
          ib.body_force[(size_t)(2 * c) * ps + n] = 0.0f;
          ib.body_force[(size_t)(2 * c + 1) * ps + n] = 0.0f;
        }
      } else {
        step_math<COLL, NC, float>(a, IbForce{ib.body_force, ps, n}, n, i, j, f, st_in, o);
        float* dst = (float*)a.pop[parity ^ 1u];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int k = 0; k < 9; ++k) dst[(size_t)(c * 9 + k) * ps + n] = o.f[c][k];
        store_macro<NC, float>(a, (float*)a.mac[parity ^ 1u], n, o, (code & LBM_CODE_KEEP_MACRO) != 0,
                               write_macro != 0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Two steps per launch (temporal blocking through LDS).
//
// A block owns a tile of TI rows x W columns.  Phase A runs step n on the
// tile and a ring of H nodes around it (pull gathers and node rules read HBM,
// exactly as the one-step kernel) and keeps the post-collision populations,
// relaxation state and -- with node rules -- the macroscopic fields in LDS;
// phase B runs step n+1 on the tile, reading step n from LDS, and writes only
// its result to HBM.  Per pair of steps HBM sees the ring's re-reads plus one
// write, instead of two reads and two writes.  The per-node arithmetic
// (step_math, apply_rule) is shared with the one-step kernel, so both give
// bit-identical fields.
//
// Phase B's pulls are plain LDS reads at fixed offsets.  Population k of node
// x streams from slot x - c_k of plane k, and only x reads that slot of that
// plane; so phase A stores there whatever x must receive:
//   * the post-collision value of node x - c_k (its own write);
//   * for a node outside the fluid group that an active node pulls
//     (LBM_CODE_PULLED), its unchanged f2, read from HBM;
//   * where x bounces population k, x's own post-collision population opp(k)
//     (a "reflection", written after a barrier so that it wins over the
//     first two).
// The LDS tile carries one pad node around the ring, where the reflections
// of ring-edge nodes land (no node of the tile reads them).
//
// Ring width: H = 1 streams; H = 2 when the lattice has node rules (a NEEM or
// OpenNeumann node reads its interior neighbour's post-stream populations,
// two nodes away).  Shan-Chen (k_sc: one step per launch) and FP64 storage run
// one step per launch.  The result of a launch goes to pop[parity^1] like a
// one-step launch: a two-step launch flips the buffers once.
// ---------------------------------------------------------------------------
constexpr int kBlock2 = 1024;
// Two-step mode M: 1 = one-node ring, no node rules; 2 = two-node ring with
// node rules (a rule node's interior neighbour may lie in the ring); 3 =
// one-node ring with node rules (every such neighbour lies in its rule
// node's tile: batches shift a lattice's tile grid to make it so).
// Mode of a two-step tile kernel: 1 = one-node ring without node rules, 2 =
// two-node ring with rules, 3 = one-node ring with rules (batches, shifted
// tile grids); plus kModeForced / kModeUnforced for kernels built for
// lattices with / without a body force (step_math's FMODE 1 / 2).
constexpr int kModeForced = 8, kModeUnforced = 16;
constexpr int ring_of(int m) { return (m & 7) == 2 ? 2 : 1; }
constexpr bool rules_of(int m) { return (m & 7) != 1; }
constexpr int fmode_of(int m) { return (m & kModeForced) ? 1 : ((m & kModeUnforced) ? 2 : 0); }
// Threads of a two-step block (one per node of the ring-extended tile).  A
// tile's two steps are a chain of dependent phases (code, gathers,
// arithmetic, barrier, LDS, arithmetic, stores), so several smaller blocks
// per CU overlap better than two large ones; the ring's share of the tile
// sets the lower limit.  Ring 1: 256 threads for 16 columns, 512 for 32 and
// 64; ring 2: 512 for 16 and 32 columns, 1024 for 64.
constexpr int block2_threads(int h, int w) { return h == 1 ? (w == 16 ? 256 : 512) : (w == 64 ? 1024 : 512); }
constexpr int kDefaultTileCols = 32;

// Tile of W = 16, 32 or 64 columns; the rows are as many as let the
// ring-extended tile fit one round of the block (block2_threads).  Narrow
// tiles leave fewer lanes idle at a lattice's last tile column; wide ones
// read longer rows.
template <int H, int W, int NT_ = block2_threads(H, W)>
struct TileShape {
  static constexpr int NT = NT_;            // threads of the block
  static constexpr int TI = NT / (W + 2 * H) - 2 * H;
  static constexpr int RI = TI + 2 * H;     // ring-extended rows
  static constexpr int RJ = W + 2 * H;      // ring-extended columns
  static constexpr int RN = RI * RJ;
  static constexpr int PJ = RJ + 2;         // LDS row pitch (one pad column each side)
  static constexpr int PN = (RI + 2) * PJ;  // LDS slots (one pad row above and below)
  static_assert(RN <= NT, "ring-extended tile larger than the block");
  // LDS slot of ring node (rr, cc), rr < RI, cc < RJ
  static __device__ __forceinline__ uint32_t slot(uint32_t rr, uint32_t cc) { return (rr + 1) * PJ + cc + 1; }
  // slot offset of x - c_k from x
  static constexpr int back(int k) { return cx(k) * PJ + cy(k); }
};

template <int NC, int S, int M, int W>
struct TileLds {
  using R = typename Pop<S>::R;
  static constexpr int H = ring_of(M);
  static constexpr int PN = TileShape<H, W>::PN;
  R f[NC * 9][PN];                      // what each node's pull of step n+1 receives (see above)
  R mac[rules_of(M) ? NC + 3 : 1][rules_of(M) ? PN : 1];   // rho_c, total_rho, ux, uy of step n (node rules)
  R st[NC][PN];                         // relaxation state after step n
  uint16_t code[PN];
};

// Step n as phase B reads it: the LDS tile, nodes addressed by their slot.
template <int NC, int S, int M, int W>
struct LdsView {
  static constexpr int H = ring_of(M);
  using R = typename Pop<S>::R;
  struct Node {
    uint32_t q;
  };
  const TileLds<NC, S, M, W>* L;
  __device__ __forceinline__ Node back(const Node& x, int di, int dj) const {
    return Node{(uint32_t)((int)x.q - di * TileShape<H, W>::PJ - dj)};
  }
  __device__ __forceinline__ uint32_t code_of(const Node& x) const { return L->code[x.q]; }
  __device__ __forceinline__ R macro(const Node& x, int plane) const { return L->mac[plane][x.q]; }
  // post-stream population k (bounces were resolved by phase A's reflections)
  __device__ __forceinline__ R pop(const Node& x, uint32_t, int c, int k) const {
    return L->f[c * 9 + k][x.q - TileShape<H, W>::back(k)];
  }
};

struct Launch2 {
  uint32_t parity;            // pop[parity]: f2 of step n-1 (post-stream f if local_read)
  int32_t local_read;         // single lattice: the first launch reads f in place
  int32_t step;               // batches: launch index (in-place read where == start_step)
  int32_t write_macro;        // step n+1 ends the run
  int32_t xcd_remap;
  uint32_t tiles_i, tiles_j;  // single lattice
  const void* tau_in;         // single lattice, power law: state before step n
  void* tau_out;              // single lattice, power law: state after step n+1
  uint32_t tau_flip;          // batches, power law: read tau2, write tau (else the reverse)
  int32_t write_psi;          // Shan-Chen: keep this step's psi in HBM (the last two steps of a run)
  uint32_t round_step;        // steps of the call up to and including the one this launch stores
};

// Phase A inputs of one ring node: gathered populations and relaxation state.
template <int NC, int S>
struct RingIn {
  using R = typename Pop<S>::R;
  R f[NC][9];
  R st[NC];
};

// Ring node (rr, cc) of the tile whose first row / column is (i0, j0): global
// coordinates, wrapped periodically.
template <int H>
__device__ __forceinline__ void ring_coords(uint32_t rr, uint32_t cc, uint32_t i0, uint32_t j0, uint32_t nx,
                                            uint32_t ny, uint32_t& gi, uint32_t& gj) {
  gi = i0 + rr + nx - H;
  gj = j0 + cc + ny - H;
  if (gi >= nx) gi -= nx;
  if (gj >= ny) gj -= ny;
  // lattices smaller than a tile wrap more than once
  if (gi >= nx) gi %= nx;
  if (gj >= ny) gj %= ny;
}

// Pull gather of one node with buffer addressing (the population buffer spans
// fewer than 2^32 bytes; nx, ny < 2^24).  Byte offsets: plane p at node m is
// p ps B + m B.  Population k pulls node x - c_k of plane k, or, bounced,
// node x of plane opp(k); both are addressed from the lower of the two planes
// (a scalar offset), the per-lane part selected by the bounce bit.
template <int NC, int S>
__device__ __forceinline__ void gather_buf(const Rsrc& rs, uint32_t ps, uint32_t nx, uint32_t ny, uint32_t gi,
                                           uint32_t gj, uint32_t code, typename Pop<S>::R (&f)[NC][9]) {
  constexpr uint32_t B = sizeof(typename Pop<S>::T);
  const uint32_t two = 2u * ps * B;   // planes k and opp(k) are two apart
  const uint32_t rc = __umul24(gi, ny);
  const uint32_t rm = gi == 0 ? rc + (nx - 1u) * ny : rc - ny;
  const uint32_t rp = gi + 1u == nx ? 0u : rc + ny;
  const uint32_t jm = gj == 0 ? ny - 1u : gj - 1u;
  const uint32_t jp = gj + 1u == ny ? 0u : gj + 1u;
  const uint32_t nB = (rc + gj) * B, nB2 = nB + two;
  const uint32_t row[3] = {rc * B, rm * B, rp * B};   // source row of cx = 0, +1, -1
  const uint32_t col[3] = {gj * B, jm * B, jp * B};   // source column of cy = 0, +1, -1
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    f[c][0] = Pop<S>::ldb(rs, nB, (uint32_t)(c * 9) * ps * B, 0);
#pragma unroll
    for (int k = 1; k < 9; ++k) {
      const bool low = opp(k) > k;   // k = 1, 2, 5, 6: plane k is the lower one
      const uint32_t src = row[cx(k) == 0 ? 0 : (cx(k) == 1 ? 1 : 2)] + col[cy(k) == 0 ? 0 : (cy(k) == 1 ? 1 : 2)] +
                           (low ? 0u : two);
      const uint32_t own = low ? nB2 : nB;
      const uint32_t m = (uint32_t)((int32_t)(code << (32 - k)) >> 31);   // all ones where bounced
      const uint32_t off = (m & own) | (~m & src);
      f[c][k] = Pop<S>::ldb(rs, off, (uint32_t)(c * 9 + (low ? k : opp(k))) * ps * B, k);
    }
  }
}

// Phase A loads of one active ring node.  BUF: buffer addressing (unsigned
// 32-bit offsets; see gather_buf); otherwise 64-bit addresses.
template <int NC, int S, bool BUF>
__device__ __forceinline__ void ring_load(const Lat& a, bool local, const typename Pop<S>::T* __restrict__ src,
                                          const Rsrc& rsrc, const typename Pop<S>::R* tau_a, uint32_t gi,
                                          uint32_t gj, uint32_t code, RingIn<NC, S>& in) {
  using R = typename Pop<S>::R;
  const uint32_t ps = a.ps, n = gi * a.ny + gj;
  if (local) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 9; ++k) in.f[c][k] = Pop<S>::ld(src + (size_t)(c * 9 + k) * ps, n, k);
  } else if constexpr (BUF) {
    gather_buf<NC, S>(rsrc, ps, a.nx, a.ny, gi, gj, code, in.f);
  } else {
    gather<NC, S>(src, ps, Nbr(gi, gj, a.nx, a.ny), code & LBM_CODE_BOUNCE_MASK, in.f);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) in.st[c] = a.tau_mode != LBM_TAU_CONST ? tau_a[(size_t)c * ps + n] : R(0);
}

// Phase A arithmetic of one active ring node: node rules (reading step n-1's
// state in HBM), step n, results into LDS slot q.  The populations that the
// node bounces are returned in `refl` for the reflection writes.
template <int COLL, int NC, int S, int M, int W>
__device__ __forceinline__ void ring_step(const Lat& a, bool local, const typename Pop<S>::T* __restrict__ src,
                                          uint32_t parity, uint32_t q, uint32_t gi, uint32_t gj, uint32_t code,
                                          RingIn<NC, S>& in, TileLds<NC, S, M, W>& L,
                                          typename Pop<S>::R (&refl)[NC][9]) {
  constexpr int H = ring_of(M);
  using R = typename Pop<S>::R;
  const uint32_t n = gi * a.ny + gj;
  if constexpr (rules_of(M)) {
    const uint32_t e = (code & LBM_CODE_NEEM_MASK) >> LBM_CODE_NEEM_SHIFT;
    if (e && !local) {
      const GlobalView<S> v{src, (const R*)a.mac[parity], a.code, a.ps, a.nx, a.ny};
      apply_rule<NC, S>(a, v, a.neem[e - 1], v.node(gi, gj), n, in.f, nullptr);
    }
  }
  StepOut<NC, R> o;
  step_math<COLL, NC, R, PsiNone<R>, fmode_of(M)>(a, PsiNone<R>{}, n, gi, gj, in.f, in.st, o);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      L.f[c * 9 + k][q] = o.f[c][k];
      refl[c][k] = o.f[c][opp(k)];
    }
    L.st[c][q] = o.state[c];
  }
  if constexpr (rules_of(M)) {
#pragma unroll
    for (int c = 0; c < NC; ++c) L.mac[c][q] = o.rho[c];
    L.mac[NC][q] = o.rtot;
    L.mac[NC + 1][q] = o.ux;
    L.mac[NC + 2][q] = o.uy;
  }
}

// Reflections of one active ring node (after the barrier that follows the
// own writes): population k that the node bounces goes to slot q - c_k of
// plane k, the slot its pull reads.
template <int NC, int S, int M, int W>
__device__ __forceinline__ void ring_reflect(uint32_t q, uint32_t code, const typename Pop<S>::R (&refl)[NC][9],
                                             TileLds<NC, S, M, W>& L) {
  constexpr int H = ring_of(M);
#pragma unroll
  for (int k = 1; k < 9; ++k) {
    if (code & (1u << (k - 1))) {
#pragma unroll
      for (int c = 0; c < NC; ++c) L.f[c * 9 + k][q - TileShape<H, W>::back(k)] = refl[c][k];
    }
  }
}

// A node outside the fluid group that an active node pulls: its unchanged f2
// (the stored value in HBM, as the one-step kernel reads it) into its slot.
template <int NC, int S, int M, int W>
__device__ __forceinline__ void ring_pulled(const Lat& a, const typename Pop<S>::T* __restrict__ stale, uint32_t q,
                                            uint32_t n, TileLds<NC, S, M, W>& L) {
  constexpr int H = ring_of(M);
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int k = 0; k < 9; ++k) L.f[c * 9 + k][q] = Pop<S>::ld(stale + (size_t)(c * 9 + k) * a.ps, n, k);
}

// Phase B: step n+1 on the tile from LDS; results to pop[parity^1] (and the
// relaxation state to tau_out under the power law, macroscopic fields when
// the run ends).
template <int COLL, int NC, int S, int M, int W, bool BUF>
__device__ __forceinline__ void tile_second_step(const Lat& a, const Launch2& q, uint32_t i0, uint32_t j0,
                                                 void* tau_out, const TileLds<NC, S, M, W>& L) {
  constexpr int H = ring_of(M);
  using T = typename Pop<S>::T;
  using R = typename Pop<S>::R;
  using Shape = TileShape<H, W>;
  const uint32_t ps = a.ps, nx = a.nx, ny = a.ny;
  const bool pl = a.tau_mode == LBM_TAU_POWER_LAW;
  T* __restrict__ dst = (T*)a.pop[q.parity ^ 1u];
  const Rsrc rdst = make_rsrc(dst, BUF ? (uint32_t)(NC * 9) * ps * (uint32_t)sizeof(T) : 0u);
  R* mo = (R*)a.mac[q.parity ^ 1u];
  const LdsView<NC, S, M, W> v{&L};
  for (uint32_t t = threadIdx.x; t < (uint32_t)(Shape::TI * W); t += Shape::NT) {
    const uint32_t rr = t / W, cc = t - rr * W;
    const uint32_t gi = i0 + rr, gj = j0 + cc;
    if (gi >= nx || gj >= ny) continue;
    const uint32_t n = gi * ny + gj;
    const typename LdsView<NC, S, M, W>::Node x{Shape::slot(rr + H, cc + H)};
    const uint32_t code = L.code[x.q];
    if (!(code & LBM_CODE_ACTIVE)) {
      if (q.write_macro) {
#pragma unroll
        for (int c = 0; c < NC + 3; ++c) mo[(size_t)c * ps + n] = R(0);
      }
      continue;
    }
    R f[NC][9];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int k = 0; k < 9; ++k) f[c][k] = v.pop(x, 0u, c, k);
    if constexpr (rules_of(M)) {
      const uint32_t e = (code & LBM_CODE_NEEM_MASK) >> LBM_CODE_NEEM_SHIFT;
      if (e) apply_rule<NC, S>(a, v, a.neem[e - 1], x, n, f, nullptr);
    }
    R st_in[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      st_in[c] = pl ? L.st[c][x.q] : (a.tau_mode == LBM_TAU_FIELD ? ((const R*)a.tau)[(size_t)c * ps + n] : R(0));
    StepOut<NC, R> o;
    step_math<COLL, NC, R, PsiNone<R>, fmode_of(M)>(a, PsiNone<R>{}, n, gi, gj, f, st_in, o);
    store_pops<NC, S, BUF>(a, q.round_step, dst, rdst, ps, n, gi, gj, o.f);
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (pl) ((R*)tau_out)[(size_t)c * ps + n] = o.state[c];
    store_macro<NC, R>(a, mo, n, o, (code & LBM_CODE_KEEP_MACRO) != 0, q.write_macro != 0);
  }
}

template <int COLL, int NC, int S, int M, int W, bool BUF>
// codes / cnx / cny: the lattice's node codes and sides for the tile's first
// loads (the same as a.code, a.nx, a.ny; a batch passes them from the tile
// record so those loads do not wait for the lattice's descriptor).
__device__ __forceinline__ void tile_two_steps(const Lat& a, const Launch2& q, bool local, uint32_t i0, uint32_t j0,
                                               const void* tau_in, void* tau_out, TileLds<NC, S, M, W>& L,
                                               const uint16_t* __restrict__ codes, uint32_t cnx, uint32_t cny) {
  constexpr int H = ring_of(M);
  using T = typename Pop<S>::T;
  using R = typename Pop<S>::R;
  using Shape = TileShape<H, W>;
  const T* __restrict__ src = (const T*)a.pop[q.parity];
  const Rsrc rsrc = make_rsrc(src, BUF ? (uint32_t)(NC * 9) * a.ps * (uint32_t)sizeof(T) : 0u);
  const R* tau_a = (const R*)(a.tau_mode == LBM_TAU_POWER_LAW ? tau_in : a.tau);
  const uint32_t r1 = threadIdx.x;
  const uint32_t rr = r1 / (uint32_t)Shape::RJ, cc = r1 - rr * (uint32_t)Shape::RJ;
  const uint32_t qs = Shape::slot(rr, cc);
  uint32_t code = 0;
  R refl[NC][9];
  // the tile's gathers issue at raised wave priority (the waves of other
  // tiles on the CU are computing or storing meanwhile); measured: 0.7 % on
  // the 8192^2 MRT power-law lattice
  __builtin_amdgcn_s_setprio(3);
  if (r1 < (uint32_t)Shape::RN) {
    uint32_t gi, gj;
    ring_coords<H>(rr, cc, i0, j0, cnx, cny, gi, gj);
    code = codes[gi * cny + gj];
    L.code[qs] = (uint16_t)code;
    if (code & LBM_CODE_ACTIVE) {
      RingIn<NC, S> in;
      ring_load<NC, S, BUF>(a, local, src, rsrc, tau_a, gi, gj, code, in);
      __builtin_amdgcn_s_setprio(0);
      ring_step<COLL, NC, S, M, W>(a, local, src, q.parity, qs, gi, gj, code, in, L, refl);
    } else if (code & LBM_CODE_PULLED) {
      ring_pulled<NC, S, M, W>(a, (const T*)a.pop[q.parity ^ 1u], qs, gi * a.ny + gj, L);
    }
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  if ((code & (LBM_CODE_ACTIVE | LBM_CODE_BOUNCE_MASK)) > LBM_CODE_ACTIVE) ring_reflect<NC, S, M, W>(qs, code, refl, L);
  __syncthreads();
  tile_second_step<COLL, NC, S, M, W, BUF>(a, q, i0, j0, tau_out, L);
}

template <int COLL, int NC, int S, int M, int W, bool BUF>
__global__ __launch_bounds__(block2_threads(ring_of(M), W)) void k_fused2(const Lat a, const Launch2 q) {
  constexpr int H = ring_of(M);
  __shared__ TileLds<NC, S, M, W> L;
  const uint32_t nt = q.tiles_i * q.tiles_j;
  uint32_t b = blockIdx.x;
  if (q.xcd_remap) b = (b & 7u) * (nt >> 3) + (b >> 3);
  const uint32_t ti = b / q.tiles_j, tj = b - ti * q.tiles_j;
  tile_two_steps<COLL, NC, S, M, W, BUF>(a, q, q.local_read != 0, ti * TileShape<H, W>::TI, tj * W, q.tau_in,
                                         q.tau_out, L, a.code, a.nx, a.ny);
}

// ---------------------------------------------------------------------------
// Shan-Chen step with the psi halo in LDS (one step per launch).
//
// The force at x needs psi of this step at the 8 neighbours, and psi is a
// function of the post-stream density (ExternalForce.py:114-120): a separate
// pass would read every population twice per step.  Here a block owns a tile
// of TI rows x 32 columns and one thread per node of the tile and a one-node
// ring around it: every thread gathers its node (pull, bounce-back and node
// rules, as the one-step kernel), sums the density and writes psi into LDS;
// after one barrier the tile's threads take the force from LDS and finish
// the step from the populations still in their registers.  The ring's
// gathers are the only re-reads (mostly L2 hits: the ring is the next tile's
// edge).  psi is written to HBM only on the last two steps of a run (the
// host reads it, and the pressure observable uses the older one).  Nodes
// outside the fluid group keep the psi the planes hold (their force terms
// read it as the two-pass step does).  The arithmetic is the one-step
// kernel's (same psi sum order, same force expressions): fields are
// bit-identical to the two-pass path.
// ---------------------------------------------------------------------------
constexpr int kScCols = 32;
constexpr int kScThreads = 512;   // 13 x 32 nodes per tile with the ring: 82 % of the threads own a tile node

template <int NT>
using ScShape = TileShape<1, kScCols, NT>;

// psi of component c at x + c_k from the LDS tile (slot q of x)
template <typename R, class Shape>
struct PsiTile {
  const R* L;         // psi of component c at L[c * Shape::PN + slot]
  const uint8_t* S;   // 1 where the slot's node is not fluid for the force (LBM_CODE_SC_SOLID)
  uint32_t q;
  __device__ __forceinline__ R at(int c, int k) const { return L[c * Shape::PN + q + Shape::back(k)]; }
  // x + c_k = slot q - back(opp(k)) = q + back(k)
  __device__ __forceinline__ bool solid(int k) const { return S[q + Shape::back(k)] != 0; }
};
template <typename R, class Shape>
struct always_forced<PsiTile<R, Shape>> : std::true_type {};

// Eight waves per SIMD where the registers allow it without spilling (one
// component, or two with MRT: 62 VGPRs; two-component BGK / TRT would spill)
constexpr int sc_min_waves(int coll, int nc) { return (nc == 1 || coll == LBM_MRT) ? 8 : 1; }

// One Shan-Chen tile (rows i0.., columns j0..) of lattice `a` for one step;
// psiL / solL: the block's LDS (psi of the ring-extended tile per component,
// adhesion flags).
template <int COLL, int NC, int S, bool BUF, int NT>
__device__ __forceinline__ void sc_tile(const Lat& a, const Launch2& q, bool local, uint32_t i0, uint32_t j0,
                                        typename Pop<S>::R* psiL, uint8_t* solL) {
  using Shape = ScShape<NT>;
  using T = typename Pop<S>::T;
  using R = typename Pop<S>::R;
  const uint32_t ps = a.ps, nx = a.nx, ny = a.ny;
  const uint32_t r1 = threadIdx.x;
  const uint32_t rr = r1 / (uint32_t)Shape::RJ, cc = r1 - rr * (uint32_t)Shape::RJ;
  const uint32_t qs = Shape::slot(rr, cc);
  const uint32_t pset = q.parity ^ 1u;   // the psi plane set of this step
  const T* __restrict__ src = (const T*)a.pop[q.parity];
  const Rsrc rsrc = make_rsrc(src, BUF ? (uint32_t)(NC * 9) * ps * (uint32_t)sizeof(T) : 0u);
  R f[NC][9];
  uint32_t code = 0, gi = 0, gj = 0;
  __builtin_amdgcn_s_setprio(3);   // gathers at raised wave priority (3 % on 2048^2 x 2)
  if (r1 < (uint32_t)Shape::RN) {
    ring_coords<1>(rr, cc, i0, j0, nx, ny, gi, gj);
    const uint32_t n = gi * ny + gj;
    code = a.code[n];
    if (a.sc_adhesion) solL[qs] = (uint8_t)(code >> 15);
    // the pulls issue without waiting for the code; the populations the code
    // bounces are loaded again from the node itself afterwards
    if (local) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < 9; ++k) f[c][k] = Pop<S>::ld(src + (size_t)(c * 9 + k) * ps, n, k);
    } else if constexpr (BUF) {
      gather_buf<NC, S>(rsrc, ps, nx, ny, gi, gj, 0u, f);
    } else {
      gather<NC, S>(src, ps, Nbr(gi, gj, nx, ny), 0u, f);
    }
    if (code & LBM_CODE_ACTIVE) {
      if (!local) {
        if (code & LBM_CODE_BOUNCE_MASK) {
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int k = 1; k < 9; ++k)
              if ((code >> (k - 1)) & 1u) f[c][k] = Pop<S>::ld(src + (size_t)(c * 9 + opp(k)) * ps, n, opp(k));
        }
        const uint32_t e = (code & LBM_CODE_NEEM_MASK) >> LBM_CODE_NEEM_SHIFT;
        if (e) {
          const GlobalView<S> v{src, (const R*)a.mac[q.parity], a.code, ps, nx, ny};
          apply_rule<NC, S>(a, v, a.neem[e - 1], v.node(gi, gj), n, f, nullptr);
        }
      }
      __builtin_amdgcn_s_setprio(0);
      // ShanChenForceC1/C2.update_psi: psi = psi(rho_c), rho_c summed k = 0..8
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R s = R(0);
#pragma unroll
        for (int k = 0; k < 9; ++k) s += f[c][k];
        psiL[c * Shape::PN + qs] = node_psi<R>(a, c, s, n);
      }
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) psiL[c * Shape::PN + qs] = ((const R*)a.psi[pset])[(size_t)c * ps + n];
    }
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  // the tile's own nodes (ring threads are done)
  if (rr < 1 || rr > (uint32_t)Shape::TI || cc < 1 || cc > (uint32_t)kScCols) return;
  if (i0 + rr - 1 >= nx || j0 + cc - 1 >= ny) return;   // past the lattice's last tile row / column
  const uint32_t n = gi * ny + gj;                        // (inside the lattice gi, gj are unwrapped)
  if (!(code & LBM_CODE_ACTIVE)) {
    if (q.write_macro) {
      R* mo = (R*)a.mac[pset];
#pragma unroll
      for (int c = 0; c < NC + 3; ++c) mo[(size_t)c * ps + n] = R(0);
    }
    return;
  }
  R st_in[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) st_in[c] = a.tau_mode != LBM_TAU_CONST ? ((const R*)a.tau)[(size_t)c * ps + n] : R(0);
  StepOut<NC, R> o;
  step_math<COLL, NC, R>(a, PsiTile<R, Shape>{psiL, solL, qs}, n, gi, gj, f, st_in, o);
  T* __restrict__ dst = (T*)a.pop[pset];
  const Rsrc rdst = make_rsrc(dst, BUF ? (uint32_t)(NC * 9) * ps * (uint32_t)sizeof(T) : 0u);
  store_pops<NC, S, BUF>(a, q.round_step, dst, rdst, ps, n, gi, gj, o.f);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (a.tau_mode == LBM_TAU_POWER_LAW) ((R*)a.tau)[(size_t)c * ps + n] = o.state[c];
    if (q.write_psi) ((R*)a.psi[pset])[(size_t)c * ps + n] = psiL[c * Shape::PN + qs];
  }
  store_macro<NC, R>(a, (R*)a.mac[pset], n, o, (code & LBM_CODE_KEEP_MACRO) != 0, q.write_macro != 0);
}

template <int COLL, int NC, int S, bool BUF, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(sc_min_waves(COLL, NC)))) void k_sc(
    const Lat a, const Launch2 q) {
  using Shape = ScShape<NT>;
  using R = typename Pop<S>::R;
  __shared__ R psiL[NC * Shape::PN];
  __shared__ uint8_t solL[Shape::PN];   // adhesion flags of the ring-extended tile
  const uint32_t nt = q.tiles_i * q.tiles_j;
  uint32_t b = blockIdx.x;
  if (q.xcd_remap) b = (b & 7u) * (nt >> 3) + (b >> 3);
  const uint32_t ti = b / q.tiles_j, tj = b - ti * q.tiles_j;
  sc_tile<COLL, NC, S, BUF, NT>(a, q, q.local_read != 0, ti * Shape::TI, tj * kScCols, psiL, solL);
}

// ---------------------------------------------------------------------------
// Shan-Chen, two steps per launch.
//
// Step n+1 on a tile needs psi of step n+1 on the tile and a one-node ring
// (R1), i.e. the post-stream populations of step n+1 there, i.e. step n
// computed on a two-node ring (R2), whose force needs psi of step n on a
// three-node ring (R3), i.e. the post-stream populations of step n gathered
// on R3.  A block owns one thread per node of R3 and runs:
//   1. (R3) gather f_n from HBM (pull, bounce-back), psi_n -> LDS;
//   2. (R2) step n (force from LDS psi_n), its post-collision populations
//      into the LDS slots their pulls read, then the reflections of bounced
//      populations (as the two-step kernel above);
//   3. (R1) pull f_{n+1} from LDS, psi_{n+1} -> LDS (over psi_n, now dead);
//   4. (tile) step n+1 (force from LDS psi_{n+1}) -> HBM.
// Per pair of steps HBM sees the gathers of R3 (the rings are mostly L2
// hits) and one write, instead of two reads and two writes.  The per-node
// arithmetic is step_math's, with psi summed in the same order as k_sc: the
// fields are bit-identical to two k_sc launches.  No node rules, no power
// law (their lattices keep one step per launch).
// ---------------------------------------------------------------------------
template <int NC, int NT>
struct ScTwoLds {
  using Shape = TileShape<3, kScCols, NT>;
  float f[NC * 9][Shape::PN];   // post-collision populations of step n, at the slots their pulls read
  float psi[NC][Shape::PN];     // psi of step n (phases 1-2), then of step n+1 (phases 3-4)
  uint8_t sol[Shape::PN];       // adhesion flags (LBM_CODE_SC_SOLID)
};

// Two Shan-Chen steps on one tile (rows i0.., columns j0..) of lattice `a`.
template <int COLL, int NC, bool BUF, int NT>
__device__ __forceinline__ void sc2_tile(const Lat& a, const Launch2& q, bool local, uint32_t i0, uint32_t j0,
                                         ScTwoLds<NC, NT>& L) {
  using Shape = TileShape<3, kScCols, NT>;
  using R = float;
  using T = typename Pop<LBM_STORE_F32>::T;
  constexpr int S = LBM_STORE_F32;
  const uint32_t ps = a.ps, nx = a.nx, ny = a.ny;
  const uint32_t r1 = threadIdx.x;
  const uint32_t rr = r1 / (uint32_t)Shape::RJ, cc = r1 - rr * (uint32_t)Shape::RJ;
  const uint32_t qs = Shape::slot(rr, cc);
  // distance of this node from the tile (0: a tile node; 1..3: ring)
  const int dr = rr < 3u ? 3 - (int)rr : ((int)rr >= 3 + Shape::TI ? (int)rr - (2 + Shape::TI) : 0);
  const int dc = cc < 3u ? 3 - (int)cc : ((int)cc >= 3 + kScCols ? (int)cc - (2 + kScCols) : 0);
  const int ring = dr > dc ? dr : dc;
  const bool has = r1 < (uint32_t)Shape::RN;
  // psi plane sets: step n's psi goes to set `parity`, step n+1's to the set
  // of the launch's result (parity ^ 1), so that the newest psi and the
  // newest populations share an index and the psi of the step before is in
  // the other set, as after one-step launches; inactive nodes hold equal sets
  const uint32_t pa = q.parity, pb = q.parity ^ 1u;
  const T* __restrict__ src = (const T*)a.pop[q.parity];
  const Rsrc rsrc = make_rsrc(src, BUF ? (uint32_t)(NC * 9) * ps * (uint32_t)sizeof(T) : 0u);
  R f[NC][9];
  uint32_t code = 0, gi = 0, gj = 0, n = 0;
  // 1. gather f_n on R3, psi_n
  __builtin_amdgcn_s_setprio(3);
  if (has) {
    ring_coords<3>(rr, cc, i0, j0, nx, ny, gi, gj);
    n = gi * ny + gj;
    code = a.code[n];
    if (a.sc_adhesion) L.sol[qs] = (uint8_t)(code >> 15);
    if (code & LBM_CODE_ACTIVE) {
      if (local) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int k = 0; k < 9; ++k) f[c][k] = Pop<S>::ld(src + (size_t)(c * 9 + k) * ps, n, k);
      } else if constexpr (BUF) {
        gather_buf<NC, S>(rsrc, ps, nx, ny, gi, gj, code, f);
      } else {
        gather<NC, S>(src, ps, Nbr(gi, gj, nx, ny), code & LBM_CODE_BOUNCE_MASK, f);
      }
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R s = R(0);
#pragma unroll
        for (int k = 0; k < 9; ++k) s += f[c][k];
        L.psi[c][qs] = node_psi<R>(a, c, s, n);
      }
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) L.psi[c][qs] = ((const R*)a.psi[pa])[(size_t)c * ps + n];
    }
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  // 2. step n on R2
  R refl[NC][9];
  const bool act = (code & LBM_CODE_ACTIVE) != 0;
  if (has && ring <= 2) {
    if (act) {
      R st_in[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c)
        st_in[c] = a.tau_mode == LBM_TAU_FIELD ? ((const R*)a.tau)[(size_t)c * ps + n] : R(0);
      StepOut<NC, R> o;
      step_math<COLL, NC, R>(a, PsiTile<R, Shape>{&L.psi[0][0], L.sol, qs}, n, gi, gj, f, st_in, o);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          L.f[c * 9 + k][qs] = o.f[c][k];
          refl[c][k] = o.f[c][opp(k)];
        }
      }
      // psi of step n of the tile's nodes (the run's second-last step); tile
      // slots past the lattice's last row / column wrap onto other tiles' nodes
      if (ring == 0 && q.write_psi && i0 + rr - 3 < nx && j0 + cc - 3 < ny) {
#pragma unroll
        for (int c = 0; c < NC; ++c) ((R*)a.psi[pa])[(size_t)c * ps + n] = L.psi[c][qs];
      }
    } else if (code & LBM_CODE_PULLED) {
      // an inactive node that an active node pulls: its unchanged f2
      const T* stale = (const T*)a.pop[q.parity ^ 1u];
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int k = 0; k < 9; ++k) L.f[c * 9 + k][qs] = Pop<S>::ld(stale + (size_t)(c * 9 + k) * ps, n, k);
    }
  }
  __syncthreads();
  if (has && ring <= 2 && act && (code & LBM_CODE_BOUNCE_MASK)) {
#pragma unroll
    for (int k = 1; k < 9; ++k) {
      if (code & (1u << (k - 1))) {
#pragma unroll
        for (int c = 0; c < NC; ++c) L.f[c * 9 + k][qs - Shape::back(k)] = refl[c][k];
      }
    }
  }
  __syncthreads();
  // 3. pull f_{n+1} on R1, psi_{n+1}
  if (has && ring <= 1) {
    if (act) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        R s = R(0);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          f[c][k] = L.f[c * 9 + k][qs - Shape::back(k)];
          s += f[c][k];
        }
        L.psi[c][qs] = node_psi<R>(a, c, s, n);
      }
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) L.psi[c][qs] = ((const R*)a.psi[pb])[(size_t)c * ps + n];
    }
  }
  __syncthreads();
  // 4. step n+1 on the tile
  if (!has || ring != 0) return;
  if (i0 + rr - 3 >= nx || j0 + cc - 3 >= ny) return;   // past the lattice's last tile row / column
  if (!act) {
    if (q.write_macro) {
      R* mo = (R*)a.mac[pb];
#pragma unroll
      for (int c = 0; c < NC + 3; ++c) mo[(size_t)c * ps + n] = R(0);
    }
    return;
  }
  R st_in[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) st_in[c] = a.tau_mode == LBM_TAU_FIELD ? ((const R*)a.tau)[(size_t)c * ps + n] : R(0);
  StepOut<NC, R> o;
  step_math<COLL, NC, R>(a, PsiTile<R, Shape>{&L.psi[0][0], L.sol, qs}, n, gi, gj, f, st_in, o);
  T* __restrict__ dst = (T*)a.pop[pb];
  const Rsrc rdst = make_rsrc(dst, BUF ? (uint32_t)(NC * 9) * ps * (uint32_t)sizeof(T) : 0u);
  store_pops<NC, S, BUF>(a, q.round_step, dst, rdst, ps, n, gi, gj, o.f);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (q.write_psi) ((R*)a.psi[pb])[(size_t)c * ps + n] = L.psi[c][qs];
  }
  store_macro<NC, R>(a, (R*)a.mac[pb], n, o, (code & LBM_CODE_KEEP_MACRO) != 0, q.write_macro != 0);
}

template <int COLL, int NC, bool BUF, int NT>
__global__ __launch_bounds__(NT) void k_sc2(const Lat a, const Launch2 q) {
  using Shape = TileShape<3, kScCols, NT>;
  __shared__ ScTwoLds<NC, NT> L;
  const uint32_t nt = q.tiles_i * q.tiles_j;
  uint32_t b = blockIdx.x;
  if (q.xcd_remap) b = (b & 7u) * (nt >> 3) + (b >> 3);
  const uint32_t ti = b / q.tiles_j, tj = b - ti * q.tiles_j;
  sc2_tile<COLL, NC, BUF, NT>(a, q, q.local_read != 0, ti * Shape::TI, tj * kScCols, L);
}

// One tile of a batched two-step launch (32 bytes): what the tile's first
// loads need, so that the node codes are fetched together with the
// lattice's descriptor instead of after it.
struct Tile2 {
  const uint16_t* code;   // the lattice's node codes
  uint32_t l;             // lattice index (kEmptyTile: no tile)
  uint32_t i0, j0;        // first row / column (shifted grids wrap below zero)
  uint32_t nx, ny;
  uint32_t pad;
};
constexpr uint32_t kEmptyTile = 0xFFFFFFFFu;

// Batched two-step launch: one block per tile of many lattices (tile ->
// lattice and tile -> index tables and the tile count in device memory; a
// captured graph launches one block per entry of the table's capacity and
// blocks past the current count exit, so the graph follows admissions and
// retirements).  Each XCD takes one contiguous eighth of the current tile
// list, so the tiles spread over all XCDs whatever the capacity.  Power-law
// lattices alternate their relaxation state between tau and tau2.  Each entry carries the tile's origin and the lattice's code
// pointer and sides; the table is striped by XCD with empty entries, so no
// tile count has to be read before the tile's own loads.
template <int NC, int S, int M, int W, bool BUF>
__global__ __launch_bounds__(block2_threads(ring_of(M), W)) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch2(
    const Lat* __restrict__ items, const Tile2* __restrict__ tiles, uint32_t stride, const Launch2 q) {
  __shared__ TileLds<NC, S, M, W> L;
  // the table holds 8 stripes of `stride` entries, stripe x the x-th eighth
  // of the tile list (the rest empty); blocks b with b % 8 == x run on XCD x
  const Tile2 tr = tiles[(blockIdx.x & 7u) * stride + (blockIdx.x >> 3)];
  if (tr.l == kEmptyTile) return;
  const Lat& a = items[tr.l];
  const bool local = a.start_step == q.step;
  // power law: the state alternates between tau and tau2 from launch to launch
  const void* tin = q.tau_flip ? a.tau2 : a.tau;
  void* tout = q.tau_flip ? a.tau : a.tau2;
  // the collision model is the lattice's (uniform over the block): one
  // specialised tile path per model
  switch (a.coll) {
    case LBM_BGK:
      tile_two_steps<LBM_BGK, NC, S, M, W, BUF>(a, q, local, tr.i0, tr.j0, tin, tout, L, tr.code, tr.nx, tr.ny);
      break;
    case LBM_TRT:
      tile_two_steps<LBM_TRT, NC, S, M, W, BUF>(a, q, local, tr.i0, tr.j0, tin, tout, L, tr.code, tr.nx, tr.ny);
      break;
    default:
      tile_two_steps<LBM_MRT, NC, S, M, W, BUF>(a, q, local, tr.i0, tr.j0, tin, tout, L, tr.code, tr.nx, tr.ny);
      break;
  }
}

// Many Shan-Chen lattices, one step per launch: one block per tile of any of
// them, from a table striped by XCD as the two-step batch's (Tile2 records);
// one table and launch per collision model present, so that each kernel
// keeps k_sc's register budget.  Each lattice's buffers sit at its own
// ping-pong phase: the launch parity is offset by the lattice's parity_xor.
// Per tile the arithmetic is k_sc's, so every lattice's fields equal its run
// alone bit for bit.
template <int COLL, int NC, bool BUF>
__global__ __launch_bounds__(kScThreads) __attribute__((amdgpu_waves_per_eu(sc_min_waves(COLL, NC)))) void k_batch_sc(
    const Lat* __restrict__ items, const Tile2* __restrict__ tiles, uint32_t stride, const Launch2 q) {
  using Shape = ScShape<kScThreads>;
  __shared__ float psiL[NC * Shape::PN];
  __shared__ uint8_t solL[Shape::PN];
  const Tile2 tr = tiles[(blockIdx.x & 7u) * stride + (blockIdx.x >> 3)];
  if (tr.l == kEmptyTile) return;
  const Lat& a = items[tr.l];
  Launch2 ql = q;
  ql.parity ^= a.parity_xor;
  sc_tile<COLL, NC, LBM_STORE_F32, BUF, kScThreads>(a, ql, q.local_read != 0, tr.i0, tr.j0, psiL, solL);
}

// Two steps per launch for batched Shan-Chen lattices (every lattice without
// node rules or power law): k_sc2's tile over the same kind of striped table
// (tiles of the three-node ring's shape).
template <int COLL, int NC, bool BUF, int NT>
__global__ __launch_bounds__(NT) void k_batch_sc2(const Lat* __restrict__ items, const Tile2* __restrict__ tiles,
                                                  uint32_t stride, const Launch2 q) {
  __shared__ ScTwoLds<NC, NT> L;
  const Tile2 tr = tiles[(blockIdx.x & 7u) * stride + (blockIdx.x >> 3)];
  if (tr.l == kEmptyTile) return;
  const Lat& a = items[tr.l];
  Launch2 ql = q;
  ql.parity ^= a.parity_xor;
  sc2_tile<COLL, NC, BUF, NT>(a, ql, q.local_read != 0, tr.i0, tr.j0, L);
}

// Tile-grid shifts that keep every rule node's interior neighbour in its own
// tile (two-step mode 3).  For each active rule node that reads x - c_d
// (VelocityNEEM, PressureNEEM, OpenNeumann, PressureABC), the one row shift (d = 1, 3;
// tiles of ti rows) or column shift (d = 2, 4; w columns) that would put x
// on the tile edge facing its neighbour is marked forbidden in
// masks[3l] (rows) and masks[3l + 1..2] (columns, 64 bits).
__global__ void k_rule_shifts(const Lat* __restrict__ items, uint32_t ti, uint32_t w, uint32_t* masks) {
  const uint32_t l = blockIdx.y;
  const Lat& a = items[l];
  if (!a.active || a.n_neem == 0) return;
  for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += gridDim.x * blockDim.x) {
    const uint32_t code = a.code[n];
    const uint32_t e = (code & LBM_CODE_NEEM_MASK) >> LBM_CODE_NEEM_SHIFT;
    if (!e || !(code & LBM_CODE_ACTIVE)) continue;
    const int kind = a.neem[e - 1].kind, d = a.neem[e - 1].direction;
    if (kind == LBM_RULE_VELOCITY_BB || kind == LBM_RULE_VELOCITY_EQ) continue;   // read their own node only
    const uint32_t i = n / a.ny, j = n - i * a.ny;
    if (d == 1 || d == 3) {
      const uint32_t o = d == 1 ? (ti - i % ti) % ti : (2 * ti - 1 - i % ti) % ti;   // x first / last row
      atomicOr(&masks[3 * l], 1u << o);
    } else {
      const uint32_t o = d == 2 ? (w - j % w) % w : (2 * w - 1 - j % w) % w;        // x first / last column
      atomicOr(&masks[3 * l + 1 + (o >> 5)], 1u << (o & 31u));
    }
  }
}

// After an odd number of batched two-step launches the power-law state of the
// active nodes is in tau2: copy it back to tau (the one-step, finish and
// readback paths read tau), over the batch's 256-node tiles.
__global__ __launch_bounds__(kBlock) void k_batch_tau_back(const Lat* __restrict__ items,
                                                           const uint32_t* __restrict__ tile_item,
                                                           const uint32_t* __restrict__ tile_count, uint32_t nc) {
  const uint32_t total = *tile_count;
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const uint32_t l = __builtin_amdgcn_readfirstlane(tile_item[w]);
    const Lat& a = items[l];
    if (a.tau_mode != LBM_TAU_POWER_LAW || !a.tau2) continue;
    const uint32_t n = (w - a.block0) * kBlock + threadIdx.x;
    // active nodes only: the others keep their state, as in one-step runs
    if (n >= a.N || !(a.code[n] & LBM_CODE_ACTIVE)) continue;
    for (uint32_t c = 0; c < nc; ++c) ((float*)a.tau)[(size_t)c * a.ps + n] = ((const float*)a.tau2)[(size_t)c * a.ps + n];
  }
}

// Step counts of a batch's lattices (the keys of their stochastic rounding):
// the last launch of every batched run and captured graph adds the steps it
// took to each active lattice's count.
__global__ __launch_bounds__(kBlock) void k_round_tick(const Lat* __restrict__ items, uint32_t count,
                                                       uint32_t steps, uint32_t* __restrict__ ctr) {
  for (uint32_t l = threadIdx.x; l < count; l += kBlock)
    if (items[l].active) ctr[l] += steps;
}
__global__ void k_round_set(uint32_t* __restrict__ ctr, uint32_t value) {
  if (threadIdx.x == 0) *ctr = value;
}

// The same for one lattice (lbm_run_blocked): a.tau2 holds the state to copy.
__global__ __launch_bounds__(kBlock) void k_tau_back(const Lat a, uint32_t nc) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= a.N || !(a.code[n] & LBM_CODE_ACTIVE)) return;
  for (uint32_t c = 0; c < nc; ++c) ((float*)a.tau)[(size_t)c * a.ps + n] = ((const float*)a.tau2)[(size_t)c * a.ps + n];
}

// Persistent batched step over the 256-node tiles of many lattices.  Each XCD
// takes one contiguous range of the tile list; its blocks stride through it.
// The tile -> lattice table and the tile count live in device memory so a
// captured graph keeps working after the host changes the batch.
template <int NC, int S, int MODE>
__global__ __launch_bounds__(kBlock) void k_batch(const Lat* __restrict__ items,
                                                  const uint32_t* __restrict__ tile_item,
                                                  const uint32_t* __restrict__ tile_count, const Launch q) {
  const uint32_t total = *tile_count;
  const uint32_t xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
  const uint32_t lo = (uint32_t)(((uint64_t)total * xcd) >> 3);
  const uint32_t hi = (uint32_t)(((uint64_t)total * (xcd + 1)) >> 3);
  for (uint32_t w = lo + slot; w < hi; w += per) {
    const uint32_t l = __builtin_amdgcn_readfirstlane(tile_item[w]);
    const Lat& a = items[l];
    const uint32_t n = (w - a.block0) * kBlock + threadIdx.x;
    if (n >= a.N) continue;
    const bool local = (a.start_step == q.step);
    if constexpr (MODE == MODE_STEP) {
      switch (a.coll) {
        case LBM_BGK: node_update<LBM_BGK, NC, S, MODE>(a, q, n, local); break;
        case LBM_TRT: node_update<LBM_TRT, NC, S, MODE>(a, q, n, local); break;
        default: node_update<LBM_MRT, NC, S, MODE>(a, q, n, local); break;
      }
    } else {
      node_update<LBM_BGK, NC, S, MODE>(a, q, n, local);
    }
  }
}

// Velocity residual of every lattice of a batch: acc[2l] += sum |u - u_old|^2,
// acc[2l+1] += sum |u|^2 over its active nodes, then u_old = u.
template <int NC, typename M>
__global__ __launch_bounds__(kBlock) void k_batch_residual(const Lat* __restrict__ items,
                                                           const uint32_t* __restrict__ tile_item,
                                                           const uint32_t* __restrict__ tile_count,
                                                           uint32_t parity, double* acc) {
  const uint32_t total = *tile_count;
  __shared__ double part[kBlock / 64][2];
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const uint32_t l = __builtin_amdgcn_readfirstlane(tile_item[w]);
    const Lat& a = items[l];
    const uint32_t n = (w - a.block0) * kBlock + threadIdx.x;
    double d2 = 0.0, u2 = 0.0;
    if (n < a.N && (a.code[n] & LBM_CODE_ACTIVE) && a.vel_old) {
      const M* m = (const M*)a.mac[parity];
      const double ux = (double)m[(size_t)(NC + 1) * a.ps + n], uy = (double)m[(size_t)(NC + 2) * a.ps + n];
      const double ox = (double)a.vel_old[n], oy = (double)a.vel_old[(size_t)a.ps + n];
      d2 = (ux - ox) * (ux - ox) + (uy - oy) * (uy - oy);
      u2 = ux * ux + uy * uy;
      a.vel_old[n] = (float)ux;
      a.vel_old[(size_t)a.ps + n] = (float)uy;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      d2 += __shfl_down(d2, off, 64);
      u2 += __shfl_down(u2, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      part[threadIdx.x >> 6][0] = d2;
      part[threadIdx.x >> 6][1] = u2;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      double s = 0.0;
#pragma unroll
      for (int wv = 0; wv < kBlock / 64; ++wv) s += part[wv][threadIdx.x];
      atomicAdd(acc + 2 * l + threadIdx.x, s);
    }
    __syncthreads();
  }
}

__global__ void k_zero_f64(double* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
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

int validate(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem) {
  if (!lat || !fld) return fail("null lattice or fields");
  if (lat->nx < 1 || lat->ny < 1) return fail("bad lattice size %d x %d", lat->nx, lat->ny);
  if (lat->nc < 1 || lat->nc > LBM_MAX_NC) return fail("num_components %d unsupported (1..%d)", lat->nc, LBM_MAX_NC);
  const uint64_t N = (uint64_t)lat->nx * (uint64_t)lat->ny;
  if (N >= (1ull << 31)) return fail("lattice of %llu nodes exceeds 2^31", (unsigned long long)N);
  if ((uint64_t)lat->plane_stride < N) return fail("plane_stride %lld < NX*NY", (long long)lat->plane_stride);
  if ((uint64_t)lat->plane_stride * (uint64_t)(lat->nc * 9) >= (1ull << 32))
    return fail("population buffer exceeds 2^32 elements");
  if (lat->collision < LBM_BGK || lat->collision > LBM_MRT) return fail("bad collision kind %d", lat->collision);
  if (lat->storage < LBM_STORE_F32 || lat->storage > LBM_STORE_F64) return fail("bad storage %d", lat->storage);
  if (lat->rounding != LBM_ROUND_NEAREST && lat->rounding != LBM_ROUND_STOCHASTIC)
    return fail("bad rounding %d", lat->rounding);
  if (lat->shan_chen < LBM_SC_NONE || lat->shan_chen > LBM_SC_C2) return fail("bad shan_chen %d", lat->shan_chen);
  if (lat->shan_chen == LBM_SC_C2 && lat->nc != 2) return fail("ShanChenForceC2 needs 2 components");
  if (lat->n_neem < 0 || lat->n_neem > LBM_MAX_NEEM) return fail("n_neem %d out of range", lat->n_neem);
  if (lat->n_neem > 0 && !neem) return fail("missing NEEM rule table");
  for (int r = 0; r < lat->n_neem; ++r) {
    if (neem[r].direction < 1 || neem[r].direction > 4) return fail("NEEM rule %d: direction %d", r, neem[r].direction);
    if (neem[r].kind < LBM_NEEM_VELOCITY || neem[r].kind > LBM_RULE_OPEN_CONVECTIVE1)
      return fail("node rule %d: kind %d", r, neem[r].kind);
  }
  if (!fld->f[0] || !fld->f[1] || !fld->macro[0] || !fld->macro[1] || !fld->code)
    return fail("missing population / macro / code buffer");
  if (lat->tau_mode < LBM_TAU_CONST || lat->tau_mode > LBM_TAU_POWER_LAW)
    return fail("bad tau_mode %d", lat->tau_mode);
  if (lat->tau_mode != LBM_TAU_CONST && !fld->tau) return fail("tau_mode %d needs the tau buffer", lat->tau_mode);
  if (lat->shan_chen && (!fld->psi[0] || !fld->psi[1] || !fld->sc_force))
    return fail("Shan-Chen needs psi and sc_force buffers");
  if (lat->shan_chen) {
    for (int c = 0; c < lat->nc; ++c) {
      if (lat->psi_kind[c] < LBM_PSI_SC || lat->psi_kind[c] > LBM_PSI_CS) return fail("bad psi kind %d", lat->psi_kind[c]);
      if (lat->psi_kind[c] != LBM_PSI_SC && !fld->temperature) return fail("a cubic-EOS psi needs the temperature plane");
    }
    if (lat->sc_adhesion && lat->shan_chen != LBM_SC_C2) return fail("adhesion needs ShanChenForceC2");
  }
  if (lat->huang) {
    if (lat->collision != LBM_MRT) return fail("HuangMRTCollision needs the MRT collision");
    if (lat->shan_chen == LBM_SC_NONE) return fail("HuangMRTCollision needs a Shan-Chen force (its psi)");
  }
  return LBM_OK;
}

Lat make_lat(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem) {
  Lat a;
  memset(&a, 0, sizeof a);
  a.round_base = lat->round_step;
  a.stochastic = lat->rounding == LBM_ROUND_STOCHASTIC ? 1u : 0u;
  for (int q = 0; q < 2; ++q) {
    a.pop[q] = fld->f[q];
    a.mac[q] = fld->macro[q];
    a.psi[q] = fld->psi[q];
  }
  a.tau = fld->tau;
  a.tau2 = fld->tau2;
  a.body_force = fld->body_force;
  a.sc_force = fld->sc_force;
  a.shear_rate = fld->shear_rate;
  a.code = fld->code;
  a.nx = (uint32_t)lat->nx;
  a.ny = (uint32_t)lat->ny;
  a.N = a.nx * a.ny;
  a.ps = (uint32_t)lat->plane_stride;
  a.mdiv = lat->ny == 1 ? 0ull : (~0ull) / (uint64_t)lat->ny + 1ull;
  a.coll = lat->collision;
  a.tau_mode = lat->tau_mode;
  a.force = (lat->gx != 0.0f || lat->gy != 0.0f);
  a.sc = lat->shan_chen;
  a.n_neem = lat->n_neem;
  a.start_step = -1;
  a.active = 1;
  for (int c = 0; c < LBM_MAX_NC; ++c) {
    a.tau0[c] = lat->tau[c];
    a.tau_anti0[c] = lat->tau_anti[c];
    a.magic[c] = lat->magic[c];
    a.s_shear[c] = lat->s_shear[c];
    a.s_bulk[c] = lat->s_bulk[c];
    a.psi_rho0[c] = lat->psi_rho0[c];
    a.psi_kind[c] = lat->psi_kind[c];
    a.g_adh[c] = lat->g_adh[c];
    for (int q = 0; q < 6; ++q) a.eos[c][q] = lat->eos[c][q];
  }
  a.sc_adhesion = lat->sc_adhesion;
  a.temperature = fld->temperature;
  a.huang = lat->huang;
  for (int c = 0; c < LBM_MAX_NC; ++c) {
    a.huang_k1[c] = lat->huang_k1[c];
    a.huang_k2[c] = lat->huang_k2[c];
    a.huang_g[c] = lat->huang_g[c];
  }
  for (int q = 0; q < LBM_MAX_NC * LBM_MAX_NC; ++q) a.g_coh[q] = lat->g_coh[q];
  a.visc_m = lat->visc_m;
  a.visc_nm1 = lat->visc_nm1;
  a.tau_min = lat->tau_min;
  a.tau_max = lat->tau_max;
  a.gx = lat->gx;
  a.gy = lat->gy;
  for (int r = 0; r < lat->n_neem && neem; ++r) a.neem[r] = neem[r];
  return a;
}

Launch make_launch(uint32_t N) {
  Launch q;
  memset(&q, 0, sizeof q);
  q.nblocks = (N + kBlock - 1) / kBlock;
  q.xcd_remap = (q.nblocks % 8 == 0 && q.nblocks >= 64) ? 1 : 0;
  return q;
}

using KernelFn = void (*)(const Lat, const Launch);

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
    case LBM_STORE_BF16_SHIFT: return pick_coll<NC, LBM_STORE_BF16_SHIFT>(coll, mode);
    default: return pick_coll<NC, LBM_STORE_F64>(coll, mode);
  }
}
KernelFn pick(int nc, int store, int coll, int mode) {
  return nc == 1 ? pick_store<1>(store, coll, mode) : pick_store<2>(store, coll, mode);
}

int launch(KernelFn fn, const Lat& a, const Launch& q, hipStream_t s) {
  hipLaunchKernelGGL(fn, dim3(q.nblocks), dim3(kBlock), 0, s, a, q);
  return hip_check(hipGetLastError(), "kernel launch");
}

// ---- batches ------------------------------------------------------------------
using BatchFn = void (*)(const Lat*, const uint32_t*, const uint32_t*, const Launch);

template <int NC, int S>
BatchFn pick_batch_mode(int mode) {
  return mode == MODE_STEP ? k_batch<NC, S, MODE_STEP> : k_batch<NC, S, MODE_FINISH>;
}
template <int NC>
BatchFn pick_batch_store(int store, int mode) {
  switch (store) {
    case LBM_STORE_F32: return pick_batch_mode<NC, LBM_STORE_F32>(mode);
    case LBM_STORE_F16_SHIFT: return pick_batch_mode<NC, LBM_STORE_F16_SHIFT>(mode);
    case LBM_STORE_BF16_SHIFT: return pick_batch_mode<NC, LBM_STORE_BF16_SHIFT>(mode);
    default: return pick_batch_mode<NC, LBM_STORE_F64>(mode);
  }
}
BatchFn pick_batch(int nc, int store, int mode) {
  return nc == 1 ? pick_batch_store<1>(store, mode) : pick_batch_store<2>(store, mode);
}

}  // namespace lbm

struct lbm_batch {
  int nc = 1, storage = 0, count = 0;
  std::vector<lbm::Lat> host;        // descriptors (host copy)
  lbm::Lat* items = nullptr;         // device copy
  uint32_t* tile_item = nullptr;     // device: tile -> lattice
  uint32_t* tile_count = nullptr;    // device: number of tiles
  double* acc = nullptr;             // device: residual accumulators, 2 per lattice
  size_t tile_cap = 0;
  uint32_t tiles = 0;
  std::vector<uint32_t> host_tiles;
  // two-step launches: (lattice, tile) tables over the ring-extended tiling
  lbm::Tile2* tile2 = nullptr;       // device: 8 stripes of tile2_stride entries
  size_t tile2_stride = 0;
  size_t tile2_cap = 0;              // 8 * tile2_stride
  uint32_t tiles2 = 0;
  int mode2 = 0;                     // 0: one step per launch, else the ring width H
  bool unforced = false;             // no active lattice has a body force (kernels without force terms)
  int tile_cols = lbm::kDefaultTileCols;   // two-step tile columns (16, 32 or 64)
  int ring_pref = 0;                 // 0: one-node ring with rules where the tile grids allow it; 2: two-node
  bool power_law = false;            // an active lattice has a power-law state (two-step: tau / tau2)
  uint32_t* shift_masks = nullptr;   // device: 3 words per lattice (k_rule_shifts)
  int shift_cap = 0;
  std::vector<lbm::Tile2> host_t2, host_t2lay;   // tile list; its striped layout (uploaded)
  bool captured = false;             // a graph holds the table addresses: no reallocation
  uint32_t* round_ctr = nullptr;     // device: each lattice's step count (k_round_tick)
};

namespace {
// two-step table with room for at least `tiles` tiles (8 stripes); the
// contents are not kept
int tiles2_alloc(lbm_batch* b, size_t tiles) {
  using namespace lbm;
  const size_t stride = std::max<size_t>(1, (tiles + 7) / 8);
  Tile2* t = nullptr;
  int rc = hip_check(hipMalloc((void**)&t, sizeof(Tile2) * 8 * stride), "hipMalloc(tile2 table)");
  if (rc) return rc;
  (void)hipFree(b->tile2);
  b->tile2 = t;
  b->tile2_stride = stride;
  b->tile2_cap = 8 * stride;
  return LBM_OK;
}
// lays the tile list out in the striped table (stripe x: the x-th eighth of
// the list, then empty entries) and uploads it, asynchronously on `s`:
// host_t2lay stays untouched until the stream has passed the copy
int tiles2_upload(lbm_batch* b, hipStream_t s) {
  using namespace lbm;
  const size_t st = b->tile2_stride, n = b->tiles2;
  Tile2 empty{};
  empty.l = kEmptyTile;
  b->host_t2lay.assign(8 * st, empty);
  for (size_t x = 0; x < 8; ++x) {
    const size_t lo = n * x / 8, hi = n * (x + 1) / 8;
    for (size_t k = lo; k < hi; ++k) b->host_t2lay[x * st + (k - lo)] = b->host_t2[k];
  }
  return hip_check(hipMemcpyAsync(b->tile2, b->host_t2lay.data(), sizeof(Tile2) * 8 * st, hipMemcpyHostToDevice, s),
                   "upload tile2 table");
}
}  // namespace

namespace lbm {
// Shan-Chen runs through k_sc (defined below); false where that kernel is
// not built for the lattice (then the two-pass path runs).
bool run_sc(const lbm_lattice* lat, const Lat& a, int32_t nsteps, uint32_t parity, int32_t first_local,
            hipStream_t s, int* rc, uint32_t step0 = 0);
}  // namespace lbm

extern "C" int lbm_run(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
                       int32_t nsteps, int32_t parity, int32_t first_local, void* stream) {
  using namespace lbm;
  int rc = validate(lat, fld, neem);
  if (rc) return rc;
  if (nsteps < 0) return fail("nsteps %d < 0", nsteps);
  if (parity != 0 && parity != 1) return fail("parity must be 0 or 1");
  hipStream_t s = (hipStream_t)stream;
  const Lat a = make_lat(lat, fld, neem);
  if (lat->shan_chen && run_sc(lat, a, nsteps, (uint32_t)parity, first_local, s, &rc)) return rc;
  Launch q = make_launch(a.N);
  KernelFn step = pick(lat->nc, lat->storage, lat->collision, MODE_STEP);
  KernelFn psik = pick(lat->nc, lat->storage, lat->collision, MODE_PSI);
  uint32_t p = (uint32_t)parity;
  for (int t = 0; t < nsteps; ++t) {
    q.parity = p;
    q.step = t;
    q.local_read = (t == 0 && first_local) ? 1 : 0;
    q.write_macro = (t == nsteps - 1) ? 1 : 0;
    q.round_step = (uint32_t)t + 1u;
    if (lat->shan_chen && (rc = launch(psik, a, q, s))) return rc;
    if ((rc = launch(step, a, q, s))) return rc;
    p ^= 1u;
  }
  return LBM_OK;
}

namespace lbm {
using Kernel2Fn = void (*)(const Lat, const Launch2);

template <int NC, int S, int H, int W, bool BUF>
Kernel2Fn pick2_coll(int coll) {
  switch (coll) {
    case LBM_BGK: return k_fused2<LBM_BGK, NC, S, H, W, BUF>;
    case LBM_TRT: return k_fused2<LBM_TRT, NC, S, H, W, BUF>;
    default: return k_fused2<LBM_MRT, NC, S, H, W, BUF>;
  }
}
template <int H, int W, bool BUF>
Kernel2Fn pick2_h(int nc, int store, int coll) {
  switch (store) {
    case LBM_STORE_F32:
      return nc == 1 ? pick2_coll<1, LBM_STORE_F32, H, W, BUF>(coll) : pick2_coll<2, LBM_STORE_F32, H, W, BUF>(coll);
    case LBM_STORE_F16_SHIFT:
      return nc == 1 ? pick2_coll<1, LBM_STORE_F16_SHIFT, H, W, BUF>(coll)
                     : pick2_coll<2, LBM_STORE_F16_SHIFT, H, W, BUF>(coll);
    case LBM_STORE_BF16_SHIFT:
      return nc == 1 ? pick2_coll<1, LBM_STORE_BF16_SHIFT, H, W, BUF>(coll)
                     : pick2_coll<2, LBM_STORE_BF16_SHIFT, H, W, BUF>(coll);
    default: return nullptr;
  }
}
// buffer addressing (unsigned 32-bit offsets and record count, 24-bit row
// products) when the population buffer spans fewer than 2^32 bytes and both
// sides are below 2^24
bool fits_buffer(uint64_t nc, uint64_t nx, uint64_t ny, uint64_t ps, int storage) {
  const uint64_t t = storage == LBM_STORE_F32 ? 4 : (storage == LBM_STORE_F64 ? 8 : 2);
  return nc * 9ull * ps * t < (1ull << 32) && nx < (1ull << 24) && ny < (1ull << 24);
}
bool fits_buffer(const lbm_lattice* lat) {
  return fits_buffer(lat->nc, lat->nx, lat->ny, lat->plane_stride, lat->storage);
}
// nullptr where the two-step kernel is not built: FP64 storage (the
// high-precision reference path; in a 1024-thread block its registers would
// spill) runs one step per launch.  Lattices beyond buffer addressing use
// 64-column tiles.
// The one-component kernels of 32 and 64 columns with buffer addressing come
// in builds for lattices with and without a body force (F = kModeForced /
// kModeUnforced); fmode 1 / 2 picks them.
template <int H, int W, int F>
Kernel2Fn pick2_fm(int store, int coll) {
  switch (store) {
    case LBM_STORE_F32: return pick2_coll<1, LBM_STORE_F32, H | F, W, true>(coll);
    case LBM_STORE_F16_SHIFT: return pick2_coll<1, LBM_STORE_F16_SHIFT, H | F, W, true>(coll);
    case LBM_STORE_BF16_SHIFT: return pick2_coll<1, LBM_STORE_BF16_SHIFT, H | F, W, true>(coll);
    default: return nullptr;
  }
}
template <int F>
Kernel2Fn pick2_f(int store, int coll, int h, int w) {
  if (w == 32) return h == 1 ? pick2_fm<1, 32, F>(store, coll) : pick2_fm<2, 32, F>(store, coll);
  return h == 1 ? pick2_fm<1, 64, F>(store, coll) : pick2_fm<2, 64, F>(store, coll);
}
Kernel2Fn pick2(int nc, int store, int coll, int h, int w, bool buf, int fmode = 0) {
  if (fmode && buf && nc == 1 && (w == 32 || w == 64))
    return fmode == 1 ? pick2_f<kModeForced>(store, coll, h, w) : pick2_f<kModeUnforced>(store, coll, h, w);
  if (!buf) return h == 1 ? pick2_h<1, 64, false>(nc, store, coll) : pick2_h<2, 64, false>(nc, store, coll);
  if (w == 16) return h == 1 ? pick2_h<1, 16, true>(nc, store, coll) : pick2_h<2, 16, true>(nc, store, coll);
  if (w == 32) return h == 1 ? pick2_h<1, 32, true>(nc, store, coll) : pick2_h<2, 32, true>(nc, store, coll);
  return h == 1 ? pick2_h<1, 64, true>(nc, store, coll) : pick2_h<2, 64, true>(nc, store, coll);
}
bool valid_tile_cols(int32_t w) { return w == 0 || w == 16 || w == 32 || w == 64; }
// tile columns: 16, 32 or 64 as asked, else the default
int tile_cols_of(int32_t asked) { return asked == 16 || asked == 32 || asked == 64 ? asked : kDefaultTileCols; }
uint32_t tile_rows(int h, int w) {
  return (uint32_t)(block2_threads(h, w) / (w + 2 * h) - 2 * h);
}
int ring_width(const lbm_lattice* lat) { return lat->n_neem > 0 ? 2 : 1; }

// ---- Shan-Chen ----------------------------------------------------------------
using ScFn = void (*)(const Lat, const Launch2);
template <int NC, int S, bool BUF>
ScFn pick_sc_coll(int coll) {
  switch (coll) {
    case LBM_BGK: return k_sc<LBM_BGK, NC, S, BUF, kScThreads>;
    case LBM_TRT: return k_sc<LBM_TRT, NC, S, BUF, kScThreads>;
    default: return k_sc<LBM_MRT, NC, S, BUF, kScThreads>;
  }
}
// nullptr where k_sc is not built: FP64 storage (the high-precision path) and
// shifted FP16 / BF16 storage (single lattices of the Python API store FP32)
// run the two-pass step
ScFn pick_sc(int nc, int store, int coll, bool buf) {
  if (store != LBM_STORE_F32) return nullptr;
  if (buf) return nc == 1 ? pick_sc_coll<1, LBM_STORE_F32, true>(coll) : pick_sc_coll<2, LBM_STORE_F32, true>(coll);
  return nc == 1 ? pick_sc_coll<1, LBM_STORE_F32, false>(coll) : pick_sc_coll<2, LBM_STORE_F32, false>(coll);
}

// Two Shan-Chen steps per launch (k_sc2): FP32 storage, no node rules, no
// power law.  Threads per block: one per node of the three-node-ring tile.
// Two components: 768 (20 x 38 nodes, a 14 x 32 tile, 71 KB of LDS: two
// blocks per CU); measured on 2048^2 x 2 (MRT): 0.101 ms/step, against
// 0.140 with 1024 threads (26 x 38 nodes, 90 KB: one block per CU).  One
// component: 1024 (46 KB of LDS; two blocks per CU either way, and the
// larger tile re-reads less ring).
constexpr int sc2_threads(int nc) { return nc == 2 ? 768 : 1024; }
template <int NC, bool BUF>
ScFn pick_sc2_coll(int coll) {
  switch (coll) {
    case LBM_BGK: return k_sc2<LBM_BGK, NC, BUF, sc2_threads(NC)>;
    case LBM_TRT: return k_sc2<LBM_TRT, NC, BUF, sc2_threads(NC)>;
    default: return k_sc2<LBM_MRT, NC, BUF, sc2_threads(NC)>;
  }
}
ScFn pick_sc2(const lbm_lattice* lat) {
  if (lat->storage != LBM_STORE_F32 || lat->n_neem != 0 || lat->tau_mode == LBM_TAU_POWER_LAW) return nullptr;
  const bool buf = fits_buffer(lat);
  if (buf) return lat->nc == 1 ? pick_sc2_coll<1, true>(lat->collision) : pick_sc2_coll<2, true>(lat->collision);
  return lat->nc == 1 ? pick_sc2_coll<1, false>(lat->collision) : pick_sc2_coll<2, false>(lat->collision);
}

// Shan-Chen steps two per launch where k_sc2 takes the lattice (a last odd
// step, and lattices it does not take, through run_sc); false where neither
// kernel is built.  *out_parity: the buffer holding the newest f2.
bool run_sc2(const lbm_lattice* lat, const Lat& a, int32_t nsteps, uint32_t parity, int32_t first_local,
             hipStream_t s, int* rc, int32_t* out_parity) {
  ScFn fn2 = pick_sc2(lat);
  const int pairs = fn2 ? nsteps / 2 : 0;
  uint32_t p = parity;
  *rc = LBM_OK;
  if (pairs > 0) {
    const int threads = sc2_threads(lat->nc);
    const uint32_t ti = lat->nc == 2 ? (uint32_t)TileShape<3, kScCols, sc2_threads(2)>::TI
                                     : (uint32_t)TileShape<3, kScCols, sc2_threads(1)>::TI;
    Launch2 q;
    memset(&q, 0, sizeof q);
    q.tiles_i = (a.nx + ti - 1) / ti;
    q.tiles_j = (a.ny + kScCols - 1) / kScCols;
    const uint64_t nt = (uint64_t)q.tiles_i * q.tiles_j;
    if (nt >= (1ull << 31)) {
      *rc = fail("lattice too large for the Shan-Chen tiling");
      return true;
    }
    q.xcd_remap = (nt % 8 == 0 && nt >= 64) ? 1 : 0;
    for (int t = 0; t < pairs; ++t) {
      q.parity = p;
      q.local_read = (t == 0 && first_local) ? 1 : 0;
      q.write_macro = (t == pairs - 1 && (nsteps & 1) == 0) ? 1 : 0;
      q.write_psi = (t == pairs - 1) ? 1 : 0;
      q.round_step = 2u * (uint32_t)t + 2u;
      hipLaunchKernelGGL(fn2, dim3((unsigned)nt), dim3(threads), 0, s, a, q);
      if ((*rc = hip_check(hipGetLastError(), "Shan-Chen two-step launch"))) return true;
      p ^= 1u;
    }
  }
  const int rest = nsteps - 2 * pairs;
  if (rest > 0) {
    if (!run_sc(lat, a, rest, p, (pairs == 0 && first_local) ? 1 : 0, s, rc, 2u * (uint32_t)pairs)) return false;
    if (*rc) return true;
    p ^= (uint32_t)(rest & 1);
  }
  *out_parity = (int32_t)p;
  return true;
}

bool run_sc(const lbm_lattice* lat, const Lat& a, int32_t nsteps, uint32_t parity, int32_t first_local,
            hipStream_t s, int* rc, uint32_t step0) {
  ScFn fn = pick_sc(lat->nc, lat->storage, lat->collision, fits_buffer(lat));
  if (!fn) return false;
  using Shape = ScShape<kScThreads>;
  Launch2 q;
  memset(&q, 0, sizeof q);
  q.tiles_i = (a.nx + Shape::TI - 1) / Shape::TI;
  q.tiles_j = (a.ny + kScCols - 1) / kScCols;
  const uint64_t nt = (uint64_t)q.tiles_i * q.tiles_j;
  if (nt >= (1ull << 31)) {
    *rc = fail("lattice too large for the Shan-Chen tiling");
    return true;
  }
  q.xcd_remap = (nt % 8 == 0 && nt >= 64) ? 1 : 0;
  uint32_t p = parity;
  *rc = LBM_OK;
  for (int t = 0; t < nsteps; ++t) {
    q.parity = p;
    q.step = t;
    q.local_read = (t == 0 && first_local) ? 1 : 0;
    q.write_macro = (t == nsteps - 1) ? 1 : 0;
    q.write_psi = (t >= nsteps - 2) ? 1 : 0;
    q.round_step = step0 + (uint32_t)t + 1u;
    hipLaunchKernelGGL(fn, dim3((unsigned)nt), dim3(kScThreads), 0, s, a, q);
    if ((*rc = hip_check(hipGetLastError(), "Shan-Chen launch"))) return true;
    p ^= 1u;
  }
  return true;
}
}  // namespace lbm

namespace lbm {
// lbm_run_blocked with the second relaxation-state buffer given (a caller's
// buffer, or the fixed workspace of lbm_run_blocked_noalloc).  Launches only.
int run_blocked(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem, int32_t nsteps,
                int32_t parity, int32_t first_local, int32_t tile_cols, void* tau_next, int32_t* out_parity,
                void* stream) {
  int rc = validate(lat, fld, neem);
  if (rc) return rc;
  if (nsteps < 0) return fail("nsteps %d < 0", nsteps);
  if (parity != 0 && parity != 1) return fail("parity must be 0 or 1");
  if (!out_parity) return fail("null out_parity");
  hipStream_t s = (hipStream_t)stream;
  const bool pl = lat->tau_mode == LBM_TAU_POWER_LAW;
  const int h = ring_width(lat);
  if (!valid_tile_cols(tile_cols)) return fail("tile_cols must be 0, 16, 32 or 64");
  const bool buf = fits_buffer(lat);
  const int w = buf ? tile_cols_of(tile_cols) : 64;
  // lattices in this path have no Shan-Chen force: gravity is their only one
  Kernel2Fn two = pick2(lat->nc, lat->storage, lat->collision, h, w, buf, (lat->gx != 0.0f || lat->gy != 0.0f) ? 1 : 2);
  if (lat->shan_chen != LBM_SC_NONE) {
    const Lat a = make_lat(lat, fld, neem);
    if (run_sc2(lat, a, nsteps, (uint32_t)parity, first_local, s, &rc, out_parity)) return rc;
  }
  const bool ok2 = two && lat->shan_chen == LBM_SC_NONE && (!pl || tau_next);
  if (!ok2) {
    rc = lbm_run(lat, fld, neem, nsteps, parity, first_local, stream);
    if (!rc) *out_parity = parity ^ (nsteps & 1);
    return rc;
  }
  Lat a = make_lat(lat, fld, neem);
  const void* tcur = fld->tau;
  void* tnext = tau_next;
  Launch2 q2;
  memset(&q2, 0, sizeof q2);
  q2.tiles_i = (a.nx + tile_rows(h, w) - 1) / tile_rows(h, w);
  q2.tiles_j = (a.ny + w - 1) / w;
  const uint64_t nt = (uint64_t)q2.tiles_i * q2.tiles_j;
  if (nt >= (1ull << 31)) return fail("lattice too large for the two-step tiling");
  q2.xcd_remap = (nt % 8 == 0 && nt >= 64) ? 1 : 0;
  uint32_t p = (uint32_t)parity;
  const int pairs = nsteps / 2;
  for (int t = 0; t < pairs; ++t) {
    q2.parity = p;
    q2.local_read = (t == 0 && first_local) ? 1 : 0;
    q2.write_macro = (t == pairs - 1 && (nsteps & 1) == 0) ? 1 : 0;
    q2.tau_in = tcur;
    q2.tau_out = tnext;
    q2.round_step = 2u * (uint32_t)t + 2u;
    hipLaunchKernelGGL(two, dim3((unsigned)nt), dim3(block2_threads(h, w)), 0, s, a, q2);
    if ((rc = hip_check(hipGetLastError(), "two-step launch"))) return rc;
    if (pl) {
      void* w = tnext;
      tnext = const_cast<void*>(tcur);
      tcur = w;
    }
    p ^= 1u;
  }
  if (nsteps & 1) {
    // the odd step: a one-step launch on the current relaxation-state buffer
    a.tau = const_cast<void*>(tcur);
    Launch q = make_launch(a.N);
    q.parity = p;
    q.step = 0;
    q.local_read = (pairs == 0 && first_local) ? 1 : 0;
    q.write_macro = 1;
    q.round_step = 2u * (uint32_t)pairs + 1u;
    if ((rc = launch(pick(lat->nc, lat->storage, lat->collision, MODE_STEP), a, q, s))) return rc;
    p ^= 1u;
  }
  if (pl && tcur != fld->tau) {
    // the state of the active nodes is in tau_next: copy it back (the other
    // nodes keep theirs; tau_next holds nothing defined there)
    Lat t = a;
    t.tau = fld->tau;   // (the odd step above pointed a.tau at the current buffer)
    t.tau2 = const_cast<void*>(tcur);
    hipLaunchKernelGGL(k_tau_back, dim3((a.N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, t, (uint32_t)lat->nc);
    if ((rc = hip_check(hipGetLastError(), "copy tau state"))) return rc;
  }
  *out_parity = (int32_t)p;
  return LBM_OK;
}

// IB-coupled fused steps (k_ib_stencil around one-step launches).
using IbStencilFn = void (*)(const Lat, const IbStencil, uint32_t, int32_t, int32_t);
template <int MODE>
IbStencilFn pick_ib_stencil(int nc, int coll) {
  if (nc == 1) {
    switch (coll) {
      case LBM_BGK: return k_ib_stencil<LBM_BGK, 1, MODE>;
      case LBM_TRT: return k_ib_stencil<LBM_TRT, 1, MODE>;
      default: return k_ib_stencil<LBM_MRT, 1, MODE>;
    }
  }
  switch (coll) {
    case LBM_BGK: return k_ib_stencil<LBM_BGK, 2, MODE>;
    case LBM_TRT: return k_ib_stencil<LBM_TRT, 2, MODE>;
    default: return k_ib_stencil<LBM_MRT, 2, MODE>;
  }
}

// Workspace of lbm_run_blocked_noalloc: the second relaxation-state buffer of
// two-step power-law runs (nc planes of plane_stride floats), a fixed device
// array instead of a caller's buffer.  Every GPU holds its own copy, at its
// own address, until the process ends; the name is a host-side stand-in, so
// host code uses only the per-device address from hipGetSymbolAddress.
// Nothing in it has to start at zero: a launch writes the state of every
// active node of its tiles before the next launch reads it, and the final
// copy-back reads active nodes only.
constexpr size_t kNoAllocStateMax = size_t(1) << 22;   // floats (16 MiB)
static __device__ __align__(256) float scratch[kNoAllocStateMax];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

// Device address of the current device's workspace, looked up once per
// device (with that device current).
int scratch_of_current_device(void** out) {
  int dev = -1;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) return fail("device %d beyond the %d the workspace table holds", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if ((rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))) return rc;
    if (!ptr) return fail("hipGetSymbolAddress returned no address for the workspace");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return LBM_OK;
}
}  // namespace lbm

extern "C" int lbm_run_blocked(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
                               int32_t nsteps, int32_t parity, int32_t first_local, int32_t tile_cols,
                               void* tau_next, int32_t* out_parity, void* stream) {
  return lbm::run_blocked(lat, fld, neem, nsteps, parity, first_local, tile_cols, tau_next, out_parity, stream);
}

extern "C" int lbm_ib_run(const lbm_lattice* lat, const lbm_fields* fld, const lbm_ib_particles* p,
                          const lbm_gfields* scratch, int32_t iterations, int32_t nsteps, int32_t parity,
                          int32_t first_local, int32_t* out_parity, void* stream) {
  using namespace lbm;
  int rc = validate(lat, fld, nullptr);
  if (rc) return rc;
  if (!p || !scratch || !out_parity) return fail("null particle set, scratch fields or out_parity");
  if (lat->n_neem != 0 || lat->shan_chen != LBM_SC_NONE || lat->tau_mode == LBM_TAU_POWER_LAW ||
      lat->storage != LBM_STORE_F32)
    return fail("IB-coupled fused steps need FP32 storage without node rules, Shan-Chen or a power law");
  if (nsteps < 0) return fail("nsteps %d < 0", nsteps);
  if (iterations < 0) return fail("MDFIteration = %d", iterations);
  if (parity != 0 && parity != 1) return fail("parity must be 0 or 1");
  if (p->max_number < 1 || p->max_num_ib < 1 || !p->group || !p->count || !p->node_pos)
    return fail("bad particle set");
  const uint64_t cap = (uint64_t)p->max_number * (uint64_t)p->max_num_ib;
  if (cap >= (1ull << 31)) return fail("particle capacity too large");
  if (!scratch->vel || !scratch->rho || !scratch->total_rho || !scratch->body_force)
    return fail("IB scratch planes missing (vel, rho, total_rho, body_force)");
  hipStream_t s = (hipStream_t)stream;
  const Lat a = make_lat(lat, fld, nullptr);
  IbStencil ib;
  ib.group = (const int2*)p->group;
  ib.count = p->count;
  ib.node_pos = p->node_pos;
  ib.cap = (uint32_t)cap;
  ib.K = (uint32_t)p->max_num_ib;
  ib.max_number = p->max_number;
  ib.vel = scratch->vel;
  ib.rho = scratch->rho;
  ib.total_rho = scratch->total_rho;
  ib.body_force = scratch->body_force;
  const IbStencilFn macro = pick_ib_stencil<0>(lat->nc, lat->collision);
  const IbStencilFn fixup = pick_ib_stencil<1>(lat->nc, lat->collision);
  const KernelFn step = pick(lat->nc, lat->storage, lat->collision, MODE_STEP);
  const dim3 gib((unsigned)((cap + 255) / 256)), bib(256);
  Launch q = make_launch(a.N);
  uint32_t par = (uint32_t)parity;
  for (int t = 0; t < nsteps; ++t) {
    const int32_t local = (t == 0 && first_local) ? 1 : 0;
    const int32_t last = (t == nsteps - 1) ? 1 : 0;
    hipLaunchKernelGGL(macro, gib, bib, 0, s, a, ib, par, local, last);
    if ((rc = hip_check(hipGetLastError(), "IB stencil launch"))) return rc;
    if ((rc = lbm_ib_mdf(lat, scratch, p, iterations, stream))) return rc;
    q.parity = par;
    q.step = t;
    q.local_read = local;
    q.write_macro = last;
    if ((rc = launch(step, a, q, s))) return rc;
    hipLaunchKernelGGL(fixup, gib, bib, 0, s, a, ib, par, local, last);
    if ((rc = hip_check(hipGetLastError(), "IB fix-up launch"))) return rc;
    par ^= 1u;
  }
  *out_parity = (int32_t)par;
  return LBM_OK;
}

extern "C" int lbm_run_blocked_noalloc(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
                                       int32_t nsteps, int32_t parity, int32_t first_local, int32_t tile_cols,
                                       int32_t* out_parity, void* stream) {
  using namespace lbm;
  int rc = validate(lat, fld, neem);
  if (rc) return rc;
  void* ws = nullptr;
  // the workspace is used where lbm_run_blocked would use tau_next: two-step
  // power-law runs (FP64 storage and Shan-Chen run one step per launch)
  if (lat->tau_mode == LBM_TAU_POWER_LAW && lat->storage != LBM_STORE_F64 && lat->shan_chen == LBM_SC_NONE) {
    const uint64_t need = (uint64_t)lat->nc * (uint64_t)lat->plane_stride;
    if (need > kNoAllocStateMax)
      return fail("lbm_run_blocked_noalloc: the power-law state of %d component(s) x %lld nodes exceeds the "
                  "workspace of %llu values; use lbm_run_blocked with a tau_next buffer",
                  lat->nc, (long long)lat->plane_stride, (unsigned long long)kNoAllocStateMax);
    if ((rc = scratch_of_current_device(&ws))) return rc;
  }
  return run_blocked(lat, fld, neem, nsteps, parity, first_local, tile_cols, ws, out_parity, stream);
}

extern "C" int lbm_finish(const lbm_lattice* lat, const lbm_fields* fld, const lbm_neem_rule* neem,
                          int32_t parity, void* stream) {
  using namespace lbm;
  int rc = validate(lat, fld, neem);
  if (rc) return rc;
  if (parity != 0 && parity != 1) return fail("parity must be 0 or 1");
  const Lat a = make_lat(lat, fld, neem);
  Launch q = make_launch(a.N);
  q.parity = (uint32_t)parity;
  return launch(pick(lat->nc, lat->storage, lat->collision, MODE_FINISH), a, q, (hipStream_t)stream);
}

// ---- batch entry points ---------------------------------------------------------
extern "C" int lbm_batch_create(const lbm_lattice* lats, const lbm_fields* flds, const lbm_neem_rule* neem,
                                int32_t count, lbm_batch** out) {
  using namespace lbm;
  if (!lats || !flds || !out || count < 1) return fail("bad lbm_batch_create arguments");
  for (int l = 0; l < count; ++l) {
    int rc = validate(&lats[l], &flds[l], neem ? neem + (size_t)l * LBM_MAX_NEEM : nullptr);
    if (rc) return rc;
    if (lats[l].shan_chen) return fail("batched lattices cannot use Shan-Chen forces");
    if (lats[l].nc != lats[0].nc || lats[l].storage != lats[0].storage)
      return fail("every lattice of a batch needs the same num_components and storage");
  }
  lbm_batch* b = new lbm_batch();
  b->nc = lats[0].nc;
  b->storage = lats[0].storage;
  b->count = count;
  b->host.resize(count);
  for (int l = 0; l < count; ++l) b->host[l] = make_lat(&lats[l], &flds[l], neem ? neem + (size_t)l * LBM_MAX_NEEM : nullptr);
  int rc = hip_check(hipMalloc((void**)&b->items, sizeof(Lat) * count), "hipMalloc(batch items)");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->tile_count, sizeof(uint32_t)), "hipMalloc(tile count)");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->acc, sizeof(double) * 2 * count), "hipMalloc(residuals)");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->round_ctr, sizeof(uint32_t) * count), "hipMalloc(step counts)");
  if (!rc) rc = hip_check(hipMemset(b->round_ctr, 0, sizeof(uint32_t) * count), "hipMemset(step counts)");
  if (rc) {
    (void)hipFree(b->items);
    (void)hipFree(b->tile_count);
    (void)hipFree(b->acc);
    (void)hipFree(b->round_ctr);
    delete b;
    return rc;
  }
  for (int l = 0; l < count; ++l) b->host[l].round_ctr = b->round_ctr + l;
  *out = b;
  return LBM_OK;
}

extern "C" int lbm_batch_set(lbm_batch* b, int32_t index, int32_t active, int32_t start_step, float* vel_old) {
  using namespace lbm;
  if (!b || index < 0 || index >= b->count) return fail("bad batch index %d", index);
  b->host[index].active = active ? 1u : 0u;
  b->host[index].start_step = start_step;
  b->host[index].vel_old = vel_old;
  return LBM_OK;
}

extern "C" int lbm_batch_update(lbm_batch* b, int32_t index, const lbm_lattice* lat, const lbm_fields* fld,
                                const lbm_neem_rule* neem) {
  using namespace lbm;
  if (!b || index < 0 || index >= b->count) return fail("bad batch index %d", index);
  int rc = validate(lat, fld, neem);
  if (rc) return rc;
  if (lat->nc != b->nc || lat->storage != b->storage || lat->shan_chen)
    return fail("lattice does not match the batch (components, storage, no Shan-Chen)");
  const Lat keep = b->host[index];
  b->host[index] = make_lat(lat, fld, neem);
  b->host[index].round_ctr = keep.round_ctr;
  b->host[index].active = keep.active;
  b->host[index].start_step = keep.start_step;
  b->host[index].vel_old = keep.vel_old;
  return LBM_OK;
}

extern "C" int lbm_batch_commit(lbm_batch* b, void* stream) {
  using namespace lbm;
  if (!b) return fail("null batch");
  hipStream_t s = (hipStream_t)stream;
  b->host_tiles.clear();
  for (int l = 0; l < b->count; ++l) {
    Lat& a = b->host[l];
    a.block0 = (uint32_t)b->host_tiles.size();
    if (!a.active) continue;
    const uint32_t nt = (a.N + kBlock - 1) / kBlock;
    for (uint32_t t = 0; t < nt; ++t) b->host_tiles.push_back((uint32_t)l);
  }
  b->tiles = (uint32_t)b->host_tiles.size();
  // two-step tiling: every active lattice without power law or fp64 storage
  bool ok2 = b->storage != LBM_STORE_F64, rules = false;
  b->power_law = false;
  for (int l = 0; l < b->count; ++l) {
    Lat& a = b->host[l];
    a.tile_oi = a.tile_oj = 0;
    if (!a.active) continue;
    if (a.tau_mode == LBM_TAU_POWER_LAW) {
      if (!a.tau2) ok2 = false;   // no second state buffer
      b->power_law = true;
    }
    if (!fits_buffer(b->nc, a.nx, a.ny, a.ps, b->storage)) ok2 = false;   // buffer addressing
    if (a.n_neem > 0) rules = true;
  }
  int mode = ok2 ? (rules ? 2 : 1) : 0;
  if (b->tiles > b->tile_cap) {
    // a captured graph holds the old table's address: grow only before capture
    if (b->captured)
      return fail("tile table full (%u > %zu) after a graph capture: reserve before capturing", b->tiles,
                  b->tile_cap);
    if (b->tile_item) (void)hipFree(b->tile_item);
    b->tile_cap = b->tiles + b->tiles / 4 + 1024;
    int rc = hip_check(hipMalloc((void**)&b->tile_item, sizeof(uint32_t) * b->tile_cap), "hipMalloc(tile table)");
    if (rc) {
      b->tile_item = nullptr;
      b->tile_cap = 0;
      return rc;
    }
  }
  int rc = hip_check(hipMemcpyAsync(b->items, b->host.data(), sizeof(Lat) * b->count, hipMemcpyHostToDevice, s),
                     "upload batch items");
  if (!rc && b->tiles)
    rc = hip_check(hipMemcpyAsync(b->tile_item, b->host_tiles.data(), sizeof(uint32_t) * b->tiles,
                                  hipMemcpyHostToDevice, s), "upload tile table");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(b->tile_count, &b->tiles, sizeof(uint32_t), hipMemcpyHostToDevice, s),
                   "upload tile count");
  if (rc) return rc;
  const uint32_t w = (uint32_t)b->tile_cols;
  if (mode == 2 && b->ring_pref == 0) {
    // node rules with a one-node ring (mode 3) where every lattice's tile grid
    // can be shifted so that each rule node's interior neighbour shares its tile
    const uint32_t ti1 = tile_rows(1, (int)w);
    if (b->shift_cap < b->count) {
      (void)hipFree(b->shift_masks);
      b->shift_masks = nullptr;
      b->shift_cap = 0;
      if ((rc = hip_check(hipMalloc((void**)&b->shift_masks, sizeof(uint32_t) * 3 * (size_t)b->count),
                          "hipMalloc(shift masks)")))
        return rc;
      b->shift_cap = b->count;
    }
    std::vector<uint32_t> masks(3 * (size_t)b->count, 0u);
    rc = hip_check(hipMemsetAsync(b->shift_masks, 0, sizeof(uint32_t) * masks.size(), s), "clear shift masks");
    if (!rc) {
      hipLaunchKernelGGL(k_rule_shifts, dim3(64, (unsigned)b->count), dim3(256), 0, s, b->items, ti1, w,
                         b->shift_masks);
      rc = hip_check(hipGetLastError(), "rule-shift scan");
    }
    if (!rc)
      rc = hip_check(hipMemcpyAsync(masks.data(), b->shift_masks, sizeof(uint32_t) * masks.size(),
                                    hipMemcpyDeviceToHost, s), "read shift masks");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (rc) return rc;
    bool ok3 = true;
    for (int l = 0; l < b->count && ok3; ++l) {
      Lat& a = b->host[l];
      if (!a.active || a.n_neem == 0) continue;
      const uint64_t cols = (uint64_t)masks[3 * l + 1] | ((uint64_t)masks[3 * l + 2] << 32);
      uint32_t oi = 0, oj = 0;
      while (oi < ti1 && ((masks[3 * l] >> oi) & 1u)) ++oi;
      while (oj < w && ((cols >> oj) & 1u)) ++oj;
      // no free shift, or a shifted grid over a lattice narrower than a tile
      if (oi == ti1 || oj == w || (oi && a.nx < ti1) || (oj && a.ny < w)) ok3 = false;
      a.tile_oi = oi;
      a.tile_oj = oj;
    }
    if (ok3) {
      mode = 3;
    } else {
      for (int l = 0; l < b->count; ++l) b->host[l].tile_oi = b->host[l].tile_oj = 0;
    }
  }
  b->host_t2.clear();
  if (mode) {
    const uint32_t ti = tile_rows(ring_of(mode), (int)w);
    for (int l = 0; l < b->count; ++l) {
      const Lat& a = b->host[l];
      if (!a.active) continue;
      const uint32_t nti = (a.nx + a.tile_oi + ti - 1) / ti, ntj = (a.ny + a.tile_oj + w - 1) / w;
      for (uint32_t r = 0; r < nti; ++r)
        for (uint32_t c = 0; c < ntj; ++c)
          b->host_t2.push_back(Tile2{a.code, (uint32_t)l, r * ti - a.tile_oi, c * w - a.tile_oj, a.nx, a.ny, 0u});
    }
  }
  b->mode2 = mode;
  b->unforced = true;
  for (int l = 0; l < b->count; ++l)
    if (b->host[l].active && b->host[l].force) b->unforced = false;
  b->tiles2 = (uint32_t)b->host_t2.size();
  if ((b->tiles2 + 7) / 8 > b->tile2_stride || !b->tile2) {
    if (b->captured)
      return fail("two-step tile table full (%u > %zu) after a graph capture: reserve before capturing",
                  b->tiles2, b->tile2_cap);
    if ((rc = tiles2_alloc(b, (size_t)b->tiles2 + b->tiles2 / 16 + 64))) {
      b->mode2 = 0;
      return rc;
    }
  }
  if (mode == 3)   // the descriptors now carry the grid shifts
    rc = hip_check(hipMemcpyAsync(b->items, b->host.data(), sizeof(Lat) * b->count, hipMemcpyHostToDevice, s),
                   "upload batch items");
  if (!rc) rc = tiles2_upload(b, s);
  // the host vectors are read by the copies above: wait before they can change
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return rc;
}

extern "C" int lbm_batch_tiles(const lbm_batch* b) { return b ? (int)b->tiles : -1; }
extern "C" int lbm_batch_capacity(const lbm_batch* b) { return b ? (int)b->tile_cap : -1; }

extern "C" int lbm_batch_reserve(lbm_batch* b, int32_t tiles) {
  using namespace lbm;
  if (!b || tiles < 0) return fail("bad lbm_batch_reserve arguments");
  if ((size_t)tiles <= b->tile_cap) return LBM_OK;
  uint32_t* t = nullptr;
  int rc = hip_check(hipMalloc((void**)&t, sizeof(uint32_t) * (size_t)tiles), "hipMalloc(tile table)");
  if (rc) return rc;
  if (b->tile_item) {
    rc = hip_check(hipMemcpy(t, b->tile_item, sizeof(uint32_t) * b->tiles, hipMemcpyDeviceToDevice), "copy tiles");
    (void)hipFree(b->tile_item);
  }
  b->tile_item = t;
  b->tile_cap = (size_t)tiles;
  return rc;
}

extern "C" int lbm_batch_reserve2(lbm_batch* b, int32_t tiles) {
  using namespace lbm;
  if (!b || tiles < 0) return fail("bad lbm_batch_reserve2 arguments");
  if ((uint32_t)tiles < b->tiles2) tiles = (int32_t)b->tiles2;   // never below the current count
  const size_t stride = std::max<size_t>(1, ((size_t)tiles + 7) / 8);
  if (stride == b->tile2_stride && b->tile2) return LBM_OK;
  if (b->captured) {
    if (stride < b->tile2_stride) return LBM_OK;   // a graph holds the table: keep it
    return fail("cannot grow the two-step tile table after a graph capture");
  }
  int rc = tiles2_alloc(b, (size_t)tiles);
  if (!rc) rc = tiles2_upload(b, nullptr);
  if (!rc) rc = hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  return rc;
}

namespace {
int batch_launch(lbm_batch* b, int mode, uint32_t parity, int32_t step, int32_t write_macro, int32_t grid,
                 hipStream_t s, uint32_t round_step = 0) {
  using namespace lbm;
  if (!b->tile_item) return fail("commit the batch before stepping it");
  if (grid <= 0) grid = 256 * 8;
  if (grid % 8) return fail("grid_blocks must be a multiple of 8 (one range per XCD)");
  Launch q;
  memset(&q, 0, sizeof q);
  q.parity = parity;
  q.step = step;
  q.write_macro = write_macro;
  q.round_step = round_step;
  hipLaunchKernelGGL(pick_batch(b->nc, b->storage, mode), dim3(grid), dim3(kBlock), 0, s, b->items, b->tile_item,
                     b->tile_count, q);
  return hip_check(hipGetLastError(), "batch launch");
}

// After a batched run (or at the end of a captured one): the lattices' step
// counts advance by the run's steps.
int round_tick(lbm_batch* b, int32_t nsteps, hipStream_t s) {
  using namespace lbm;
  if (nsteps <= 0) return LBM_OK;
  hipLaunchKernelGGL(k_round_tick, dim3(1), dim3(kBlock), 0, s, b->items, (uint32_t)b->count, (uint32_t)nsteps,
                     b->round_ctr);
  return hip_check(hipGetLastError(), "step-count launch");
}
}  // namespace

namespace {
using Batch2Fn = void (*)(const lbm::Lat*, const lbm::Tile2*, uint32_t, const lbm::Launch2);
template <int NC, int S, int W, int F = 0>
Batch2Fn pick_batch2_w(int m) {
  return m == 1 ? lbm::k_batch2<NC, S, 1 | F, W, true>
                : (m == 2 ? lbm::k_batch2<NC, S, 2 | F, W, true> : lbm::k_batch2<NC, S, 3 | F, W, true>);
}
template <int NC, int S>
Batch2Fn pick_batch2_s(int m, int w) {
  // batched lattices are small: buffer addressing (each lattice's segment is
  // checked against 2^32 bytes at commit)
  if (w == 16) return pick_batch2_w<NC, S, 16>(m);
  if (w == 32) return pick_batch2_w<NC, S, 32>(m);
  return pick_batch2_w<NC, S, 64>(m);
}
// Batches of one component whose lattices have no body force, in 32- or
// 64-column tiles: kernels built without the force terms.
template <int S>
Batch2Fn pick_batch2_unforced(int m, int w) {
  return w == 32 ? pick_batch2_w<1, S, 32, lbm::kModeUnforced>(m) : pick_batch2_w<1, S, 64, lbm::kModeUnforced>(m);
}
Batch2Fn pick_batch2(int nc, int store, int h, int w, bool unforced = false) {
  if (unforced && nc == 1 && (w == 32 || w == 64)) {
    switch (store) {
      case LBM_STORE_F32: return pick_batch2_unforced<LBM_STORE_F32>(h, w);
      case LBM_STORE_F16_SHIFT: return pick_batch2_unforced<LBM_STORE_F16_SHIFT>(h, w);
      case LBM_STORE_BF16_SHIFT: return pick_batch2_unforced<LBM_STORE_BF16_SHIFT>(h, w);
      default: return nullptr;
    }
  }
  switch (store) {
    case LBM_STORE_F32:
      return nc == 1 ? pick_batch2_s<1, LBM_STORE_F32>(h, w) : pick_batch2_s<2, LBM_STORE_F32>(h, w);
    case LBM_STORE_F16_SHIFT:
      return nc == 1 ? pick_batch2_s<1, LBM_STORE_F16_SHIFT>(h, w) : pick_batch2_s<2, LBM_STORE_F16_SHIFT>(h, w);
    case LBM_STORE_BF16_SHIFT:
      return nc == 1 ? pick_batch2_s<1, LBM_STORE_BF16_SHIFT>(h, w) : pick_batch2_s<2, LBM_STORE_BF16_SHIFT>(h, w);
    default: return nullptr;
  }
}

// The launch sequence of nsteps batched steps, two per launch where the
// batch allows it; returns the parity after it.
int batch_sequence(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid, hipStream_t s, bool capture,
                   int32_t* out) {
  using namespace lbm;
  if (!b->tile_item) return fail("commit the batch before stepping it");
  Batch2Fn two = b->mode2 ? pick_batch2(b->nc, b->storage, b->mode2, b->tile_cols, b->unforced) : nullptr;
  uint32_t p = (uint32_t)parity;
  const int total = nsteps;
  int t0 = 0;
  if (two) {
    // one block per entry of the striped table (empty entries exit): a graph
    // follows later commits while the tiles fit the stripes
    (void)capture;
    const int g2 = (int)(8 * b->tile2_stride);
    const int pairs = nsteps / 2;
    for (int t = 0; t < pairs; ++t) {
      Launch2 q2;
      memset(&q2, 0, sizeof q2);
      q2.parity = p;
      q2.step = t;
      q2.write_macro = (t == pairs - 1 && (nsteps & 1) == 0) ? 1 : 0;
      q2.tau_flip = (uint32_t)(t & 1);
      q2.round_step = 2u * (uint32_t)t + 2u;
      hipLaunchKernelGGL(two, dim3(g2), dim3(block2_threads(ring_of(b->mode2), b->tile_cols)), 0, s, b->items,
                         b->tile2, (uint32_t)b->tile2_stride, q2);
      int rc = hip_check(hipGetLastError(), "batched two-step launch");
      if (rc) return rc;
      p ^= 1u;
    }
    if ((pairs & 1) && b->power_law) {
      hipLaunchKernelGGL(k_batch_tau_back, dim3(2048), dim3(kBlock), 0, s, b->items, b->tile_item, b->tile_count,
                         (uint32_t)b->nc);
      int rc = hip_check(hipGetLastError(), "relaxation-state copy-back");
      if (rc) return rc;
    }
    t0 = pairs;
    nsteps -= 2 * pairs;
  }
  for (int t = 0; t < nsteps; ++t) {
    int rc = batch_launch(b, MODE_STEP, p, t0 + t, t == nsteps - 1, grid, s, 2u * (uint32_t)t0 + (uint32_t)t + 1u);
    if (rc) return rc;
    p ^= 1u;
  }
  int rc = round_tick(b, total, s);
  if (rc) return rc;
  *out = (int32_t)p;
  return LBM_OK;
}
}  // namespace

extern "C" int lbm_batch_mode(const lbm_batch* b) {
  if (!b) return -1;
  const bool unforced = b->mode2 && b->unforced && b->nc == 1 && (b->tile_cols == 32 || b->tile_cols == 64);
  return b->mode2 | (unforced ? lbm::kModeUnforced : 0);
}

extern "C" int lbm_batch_tile_cols(lbm_batch* b, int32_t tile_cols) {
  using namespace lbm;
  if (!b) return fail("null batch");
  if (!valid_tile_cols(tile_cols)) return fail("tile_cols must be 0, 16, 32 or 64");
  const int w = tile_cols_of(tile_cols);
  if (w == b->tile_cols) return LBM_OK;
  if (b->captured) return fail("cannot change the tile shape after a graph capture");
  b->tile_cols = w;
  return LBM_OK;
}

extern "C" int lbm_batch_ring(lbm_batch* b, int32_t ring) {
  using namespace lbm;
  if (!b) return fail("null batch");
  if (ring != 0 && ring != 2) return fail("ring must be 0 (automatic) or 2");
  if (ring == b->ring_pref) return LBM_OK;
  if (b->captured) return fail("cannot change the ring after a graph capture");
  b->ring_pref = ring;
  return LBM_OK;
}

extern "C" int lbm_batch_tiles2(const lbm_batch* b) { return b ? (int)b->tiles2 : -1; }

extern "C" int lbm_batch_release(lbm_batch* b) {
  using namespace lbm;
  if (!b) return fail("null batch");
  b->captured = false;
  return LBM_OK;
}

extern "C" int lbm_tile_rows(int32_t ring, int32_t tile_cols) {
  using namespace lbm;
  if ((ring != 1 && ring != 2) || !valid_tile_cols(tile_cols))
    return fail("lbm_tile_rows: ring 1 or 2, tile_cols 0, 16, 32 or 64");
  return (int)tile_rows(ring, tile_cols_of(tile_cols));
}

extern "C" int lbm_batch_run_blocked(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks,
                                     int32_t* out_parity, void* stream) {
  using namespace lbm;
  if (!b || nsteps < 0 || (parity != 0 && parity != 1) || !out_parity)
    return fail("bad lbm_batch_run_blocked arguments");
  return batch_sequence(b, nsteps, parity, grid_blocks, (hipStream_t)stream, false, out_parity);
}

extern "C" int lbm_batch_graph_blocked(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks,
                                       void* stream, void** graph_exec, int32_t* out_parity) {
  using namespace lbm;
  if (!b || nsteps < 1 || !graph_exec || !out_parity || (parity != 0 && parity != 1))
    return fail("bad lbm_batch_graph_blocked arguments");
  hipStream_t s = (hipStream_t)stream;
  if (!s) return fail("graph capture needs a non-null stream");
  int rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (rc) return rc;
  b->captured = true;
  rc = batch_sequence(b, nsteps, parity, grid_blocks, s, true, out_parity);
  hipGraph_t g = nullptr;
  const int rc2 = hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  if (rc || rc2) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : rc2;
  }
  hipGraphExec_t ge = nullptr;
  rc = hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  if (rc) return rc;
  *graph_exec = (void*)ge;
  return LBM_OK;
}

extern "C" int lbm_batch_run(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks, void* stream) {
  using namespace lbm;
  if (!b || nsteps < 0 || (parity != 0 && parity != 1)) return fail("bad lbm_batch_run arguments");
  hipStream_t s = (hipStream_t)stream;
  for (int t = 0; t < nsteps; ++t) {
    int rc = batch_launch(b, MODE_STEP, (uint32_t)((parity + t) & 1), t, t == nsteps - 1, grid_blocks, s,
                          (uint32_t)t + 1u);
    if (rc) return rc;
  }
  return round_tick(b, nsteps, s);
}

extern "C" int lbm_batch_finish(lbm_batch* b, int32_t parity, int32_t grid_blocks, void* stream) {
  using namespace lbm;
  if (!b || (parity != 0 && parity != 1)) return fail("bad lbm_batch_finish arguments");
  return batch_launch(b, MODE_FINISH, (uint32_t)parity, -2, 0, grid_blocks, (hipStream_t)stream);
}

extern "C" int lbm_batch_graph(lbm_batch* b, int32_t nsteps, int32_t parity, int32_t grid_blocks, void* stream,
                               void** graph_exec) {
  using namespace lbm;
  if (!b || nsteps < 1 || !graph_exec || (parity != 0 && parity != 1)) return fail("bad lbm_batch_graph arguments");
  hipStream_t s = (hipStream_t)stream;
  if (!s) return fail("graph capture needs a non-null stream");
  int rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (rc) return rc;
  b->captured = true;
  for (int t = 0; t < nsteps && !rc; ++t)
    rc = batch_launch(b, MODE_STEP, (uint32_t)((parity + t) & 1), t, t == nsteps - 1, grid_blocks, s,
                      (uint32_t)t + 1u);
  if (!rc) rc = round_tick(b, nsteps, s);
  hipGraph_t g = nullptr;
  const int rc2 = hip_check(hipStreamEndCapture(s, &g), "hipStreamEndCapture");
  if (rc || rc2) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : rc2;
  }
  hipGraphExec_t ge = nullptr;
  rc = hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  if (rc) return rc;
  *graph_exec = (void*)ge;
  return LBM_OK;
}

extern "C" int lbm_graph_launch(void* graph_exec, void* stream) {
  using namespace lbm;
  if (!graph_exec) return fail("null graph");
  return hip_check(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream), "hipGraphLaunch");
}

extern "C" int lbm_graph_destroy(void* graph_exec) {
  using namespace lbm;
  if (!graph_exec) return LBM_OK;
  return hip_check(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
}

extern "C" int lbm_batch_residual(lbm_batch* b, int32_t parity, double* out, void* stream) {
  using namespace lbm;
  if (!b || !out || (parity != 0 && parity != 1)) return fail("bad lbm_batch_residual arguments");
  if (!b->tile_item) return fail("commit the batch first");
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = 2u * (uint32_t)b->count;
  hipLaunchKernelGGL(k_zero_f64, dim3((n + 255) / 256), dim3(256), 0, s, b->acc, n);
  int rc = hip_check(hipGetLastError(), "zero launch");
  if (rc) return rc;
  const uint32_t grid = b->tiles < 4096u ? (b->tiles ? b->tiles : 1u) : 4096u;
  using ResFn = void (*)(const Lat*, const uint32_t*, const uint32_t*, uint32_t, double*);
  const bool f64 = b->storage == LBM_STORE_F64;
  const ResFn fn = b->nc == 1 ? (f64 ? k_batch_residual<1, double> : k_batch_residual<1, float>)
                              : (f64 ? k_batch_residual<2, double> : k_batch_residual<2, float>);
  hipLaunchKernelGGL(fn, dim3(grid), dim3(kBlock), 0, s, b->items, b->tile_item, b->tile_count, (uint32_t)parity,
                     b->acc);
  if ((rc = hip_check(hipGetLastError(), "residual launch"))) return rc;
  if ((rc = hip_check(hipMemcpyAsync(out, b->acc, sizeof(double) * n, hipMemcpyDeviceToHost, s), "copy out")))
    return rc;
  return hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

extern "C" int lbm_batch_destroy(lbm_batch* b) {
  if (!b) return LBM_OK;
  (void)hipFree(b->items);
  (void)hipFree(b->tile_item);
  (void)hipFree(b->tile_count);
  (void)hipFree(b->acc);
  (void)hipFree(b->tile2);
  (void)hipFree(b->shift_masks);
  (void)hipFree(b->round_ctr);
  delete b;
  return LBM_OK;
}

extern "C" int lbm_batch_set_round_step(lbm_batch* b, int32_t index, uint32_t step, void* stream) {
  using namespace lbm;
  if (!b || index < 0 || index >= b->count) return fail("bad batch index %d", index);
  hipLaunchKernelGGL(k_round_set, dim3(1), dim3(64), 0, (hipStream_t)stream, b->round_ctr + index, step);
  return hip_check(hipGetLastError(), "step-count set");
}

// ---- batched Shan-Chen lattices ------------------------------------------------
// Descriptors and the striped tile tables (one per collision model) are built
// once (lbm_sc_batch_create); each lattice keeps its own buffers.  A run
// launches k_batch_sc once per step and model present.
struct lbm_sc_batch {
  int nc = 1;
  int count = 0;
  bool buf = true;               // every lattice fits 32-bit buffer offsets
  lbm::Lat* items = nullptr;     // device descriptors
  lbm::Tile2* tiles[3] = {nullptr, nullptr, nullptr};  // per model: 8 stripes of stride[m] entries
  size_t stride[3] = {0, 0, 0};
  size_t ntiles = 0;
  bool two = false;              // two steps per launch (no lattice has node rules or power law)
  lbm::Tile2* tiles2[3] = {nullptr, nullptr, nullptr};  // the same for the three-node-ring tiles
  size_t stride2[3] = {0, 0, 0};
  size_t ntiles2 = 0;
};

namespace {
// striped upload of one tile list (8 stripes, stripe x the x-th eighth)
int sc_tiles_upload(const std::vector<lbm::Tile2>& list, lbm::Tile2** dev, size_t* stride_out) {
  using namespace lbm;
  const size_t n = list.size();
  if (n == 0) return LBM_OK;
  const size_t stride = (n + 7) / 8;
  if (8 * stride >= (1ull << 31)) return fail("Shan-Chen batch of %zu tiles too large", n);
  Tile2 empty{};
  empty.l = kEmptyTile;
  std::vector<Tile2> lay(8 * stride, empty);
  for (size_t x = 0; x < 8; ++x) {
    const size_t lo = n * x / 8, hi = n * (x + 1) / 8;
    for (size_t k = lo; k < hi; ++k) lay[x * stride + (k - lo)] = list[k];
  }
  *stride_out = stride;
  int rc = hip_check(hipMalloc((void**)dev, sizeof(Tile2) * lay.size()), "hipMalloc(Shan-Chen batch tiles)");
  if (!rc)
    rc = hip_check(hipMemcpy(*dev, lay.data(), sizeof(Tile2) * lay.size(), hipMemcpyHostToDevice),
                   "upload Shan-Chen batch tiles");
  return rc;
}
}  // namespace

extern "C" int lbm_sc_batch_destroy(lbm_sc_batch* b);

extern "C" int lbm_sc_batch_create(const lbm_lattice* lats, const lbm_fields* flds, const lbm_neem_rule* neem,
                                   const int32_t* parity_offset, int32_t count, lbm_sc_batch** out) {
  using namespace lbm;
  using Shape = ScShape<kScThreads>;
  if (!out) return fail("null out");
  *out = nullptr;
  if (count < 1 || !lats || !flds) return fail("a Shan-Chen batch needs at least one lattice");
  std::vector<Lat> host((size_t)count);
  std::vector<Tile2> list[3], list2[3];
  bool buf = true, two = true;
  const uint32_t TI2 = lats[0].nc == 2 ? (uint32_t)TileShape<3, kScCols, sc2_threads(2)>::TI
                                       : (uint32_t)TileShape<3, kScCols, sc2_threads(1)>::TI;
  for (int l = 0; l < count; ++l) {
    const lbm_lattice* L = &lats[l];
    const lbm_neem_rule* rules = neem ? neem + (size_t)l * LBM_MAX_NEEM : nullptr;
    int rc = validate(L, &flds[l], rules);
    if (rc) return rc;
    if (L->shan_chen == LBM_SC_NONE) return fail("lattice %d of a Shan-Chen batch has no Shan-Chen force", l);
    if (L->storage != LBM_STORE_F32) return fail("lattice %d: Shan-Chen batches store FP32 populations", l);
    if (L->nc != lats[0].nc) return fail("every lattice of a batch needs the same num_components");
    host[l] = make_lat(L, &flds[l], rules);
    host[l].parity_xor = parity_offset ? (uint32_t)(parity_offset[l] & 1) : 0u;
    buf = buf && fits_buffer(L);
    two = two && L->n_neem == 0 && L->tau_mode != LBM_TAU_POWER_LAW;
    Tile2 t{};
    t.code = host[l].code;
    t.l = (uint32_t)l;
    t.nx = (uint32_t)L->nx;
    t.ny = (uint32_t)L->ny;
    const uint32_t tj = ((uint32_t)L->ny + kScCols - 1) / kScCols;
    for (uint32_t i = 0; i < ((uint32_t)L->nx + Shape::TI - 1) / Shape::TI; ++i)
      for (uint32_t j = 0; j < tj; ++j) {
        t.i0 = i * Shape::TI;
        t.j0 = j * kScCols;
        list[L->collision].push_back(t);
      }
    for (uint32_t i = 0; i < ((uint32_t)L->nx + TI2 - 1) / TI2; ++i)
      for (uint32_t j = 0; j < tj; ++j) {
        t.i0 = i * TI2;
        t.j0 = j * kScCols;
        list2[L->collision].push_back(t);
      }
  }
  lbm_sc_batch* b = new lbm_sc_batch;
  b->nc = lats[0].nc;
  b->count = count;
  b->buf = buf;
  b->two = two;
  int rc = hip_check(hipMalloc((void**)&b->items, sizeof(Lat) * host.size()), "hipMalloc(Shan-Chen batch items)");
  if (!rc)
    rc = hip_check(hipMemcpy(b->items, host.data(), sizeof(Lat) * host.size(), hipMemcpyHostToDevice),
                   "upload Shan-Chen batch items");
  for (int m = 0; m < 3 && !rc; ++m) {
    b->ntiles += list[m].size();
    rc = sc_tiles_upload(list[m], &b->tiles[m], &b->stride[m]);
    if (!rc && two) {
      b->ntiles2 += list2[m].size();
      rc = sc_tiles_upload(list2[m], &b->tiles2[m], &b->stride2[m]);
    }
  }
  if (rc) {
    lbm_sc_batch_destroy(b);
    return rc;
  }
  *out = b;
  return LBM_OK;
}

namespace lbm {
using ScBatchFn = void (*)(const Lat*, const Tile2*, uint32_t, const Launch2);
template <int COLL>
ScBatchFn pick_sc_batch(int nc, bool buf) {
  return nc == 1 ? (buf ? k_batch_sc<COLL, 1, true> : k_batch_sc<COLL, 1, false>)
                 : (buf ? k_batch_sc<COLL, 2, true> : k_batch_sc<COLL, 2, false>);
}
template <int COLL>
ScBatchFn pick_sc_batch2(int nc, bool buf) {
  return nc == 1 ? (buf ? k_batch_sc2<COLL, 1, true, sc2_threads(1)> : k_batch_sc2<COLL, 1, false, sc2_threads(1)>)
                 : (buf ? k_batch_sc2<COLL, 2, true, sc2_threads(2)> : k_batch_sc2<COLL, 2, false, sc2_threads(2)>);
}
}  // namespace lbm

extern "C" int lbm_sc_batch_run(lbm_sc_batch* b, int32_t nsteps, int32_t parity, int32_t first_local,
                                int32_t* out_parity, void* stream) {
  using namespace lbm;
  if (!b) return fail("null Shan-Chen batch");
  if (!out_parity) return fail("null out_parity");
  if (nsteps < 0) return fail("nsteps %d < 0", nsteps);
  if (parity != 0 && parity != 1) return fail("parity must be 0 or 1");
  const ScBatchFn one[3] = {pick_sc_batch<LBM_BGK>(b->nc, b->buf), pick_sc_batch<LBM_TRT>(b->nc, b->buf),
                            pick_sc_batch<LBM_MRT>(b->nc, b->buf)};
  const ScBatchFn two[3] = {pick_sc_batch2<LBM_BGK>(b->nc, b->buf), pick_sc_batch2<LBM_TRT>(b->nc, b->buf),
                            pick_sc_batch2<LBM_MRT>(b->nc, b->buf)};
  hipStream_t s = (hipStream_t)stream;
  Launch2 q;
  memset(&q, 0, sizeof q);
  uint32_t p = (uint32_t)parity;
  // as lbm_run with Shan-Chen: pairs of steps first, an odd last step alone
  const int pairs = b->two ? nsteps / 2 : 0;
  for (int t = 0; t < pairs; ++t) {
    q.parity = p;
    q.local_read = (t == 0 && first_local) ? 1 : 0;
    q.write_macro = (t == pairs - 1 && (nsteps & 1) == 0) ? 1 : 0;
    q.write_psi = (t == pairs - 1) ? 1 : 0;
    for (int m = 0; m < 3; ++m) {
      if (!b->tiles2[m]) continue;
      hipLaunchKernelGGL(two[m], dim3((unsigned)(8 * b->stride2[m])), dim3(sc2_threads(b->nc)), 0, s, b->items,
                         b->tiles2[m], (uint32_t)b->stride2[m], q);
      int rc = hip_check(hipGetLastError(), "Shan-Chen batch two-step launch");
      if (rc) return rc;
    }
    p ^= 1u;
  }
  const int rest = nsteps - 2 * pairs;
  for (int t = 0; t < rest; ++t) {
    q.parity = p;
    q.step = t;
    q.local_read = (t == 0 && pairs == 0 && first_local) ? 1 : 0;
    q.write_macro = (t == rest - 1) ? 1 : 0;
    q.write_psi = (t >= rest - 2) ? 1 : 0;
    for (int m = 0; m < 3; ++m) {
      if (!b->tiles[m]) continue;
      hipLaunchKernelGGL(one[m], dim3((unsigned)(8 * b->stride[m])), dim3(kScThreads), 0, s, b->items, b->tiles[m],
                         (uint32_t)b->stride[m], q);
      int rc = hip_check(hipGetLastError(), "Shan-Chen batch launch");
      if (rc) return rc;
    }
    p ^= 1u;
  }
  *out_parity = (int32_t)p;
  return LBM_OK;
}

extern "C" int lbm_sc_batch_tiles(const lbm_sc_batch* b) {
  return b ? (int)b->ntiles : lbm::fail("null Shan-Chen batch");
}

extern "C" int lbm_sc_batch_destroy(lbm_sc_batch* b) {
  if (!b) return LBM_OK;
  (void)hipFree(b->items);
  for (int m = 0; m < 3; ++m) {
    (void)hipFree(b->tiles[m]);
    (void)hipFree(b->tiles2[m]);
  }
  delete b;
  return LBM_OK;
}
