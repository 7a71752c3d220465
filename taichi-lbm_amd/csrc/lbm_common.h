// Shared device-side helpers for the D2Q9 kernels (gfx950).
//
// D2Q9 numbering, weights and opposite map follow the reference:
// lbm.py:43-45 (c, w) and BoundaryCondition.py:242 (bounce map).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "lbm_abi.h"

namespace lbm {

constexpr int kBlock = 256;

__host__ __device__ constexpr int cx(int k) {
  return (k == 1 || k == 5 || k == 8) ? 1 : ((k == 3 || k == 6 || k == 7) ? -1 : 0);
}
__host__ __device__ constexpr int cy(int k) {
  return (k == 2 || k == 5 || k == 6) ? 1 : ((k == 4 || k == 7 || k == 8) ? -1 : 0);
}
__host__ __device__ constexpr int opp(int k) {
  return k == 0 ? 0 : (k <= 4 ? ((k + 1) % 4) + 1 : ((k - 5 + 2) % 4) + 5);
}
static_assert(opp(1) == 3 && opp(2) == 4 && opp(3) == 1 && opp(4) == 2, "opp");
static_assert(opp(5) == 7 && opp(6) == 8 && opp(7) == 5 && opp(8) == 6, "opp");

// lbm.py:44: an fp32 vector (4,1,1,1,1,1/4,1/4,1/4,1/4) divided by 9.0 in fp32.
__host__ __device__ constexpr float wgt(int k) {
  return k == 0 ? 4.0f / 9.0f : (k <= 4 ? 1.0f / 9.0f : 0.25f / 9.0f);
}

// Collision.f_eq, CollisionEngine.py:14-18: w*rho*(1+3eu+4.5eu^2-1.5uu).
__device__ __forceinline__ float feq_k(int k, float ux, float uy, float rho, float uv) {
  const float eu = (float)cx(k) * ux + (float)cy(k) * uy;
  return wgt(k) * rho * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * uv);
}

// Collision.f_force (Guo), CollisionEngine.py:20-25: w*(3cF + 9 cF cu - 3 uF).
__device__ __forceinline__ float fforce_k(int k, float Fx, float Fy, float ux, float uy, float uF) {
  const float cF = (float)cx(k) * Fx + (float)cy(k) * Fy;
  const float cu = (float)cx(k) * ux + (float)cy(k) * uy;
  return wgt(k) * (3.0f * cF + 9.0f * cF * cu - 3.0f * uF);
}

// Collision.tau_max_min, CollisionEngine.py:28-34.
__device__ __forceinline__ float clamp_tau(float t, float tmin, float tmax) {
  if (t < tmin) t = tmin;
  else if (t > tmax) t = tmax;
  return t;
}

// ---- population storage ---------------------------------------------------
// Shifted formats store g = f - w_k so that the stored magnitude is the
// deviation from rest; arithmetic is always fp32.
template <int S> struct Pop;
template <> struct Pop<LBM_STORE_F32> {
  using T = float;
  static __device__ __forceinline__ float ld(const T* p, uint32_t i, int) { return p[i]; }
  static __device__ __forceinline__ void st(T* p, uint32_t i, int, float v) {
    __builtin_nontemporal_store(v, p + i);
  }
  static __device__ __forceinline__ void st_cached(T* p, uint32_t i, int, float v) { p[i] = v; }
};
template <> struct Pop<LBM_STORE_F16_SHIFT> {
  using T = _Float16;
  static __device__ __forceinline__ float ld(const T* p, uint32_t i, int k) {
    return (float)p[i] + wgt(k);
  }
  static __device__ __forceinline__ void st(T* p, uint32_t i, int k, float v) {
    __builtin_nontemporal_store((T)(v - wgt(k)), p + i);
  }
  static __device__ __forceinline__ void st_cached(T* p, uint32_t i, int k, float v) {
    p[i] = (T)(v - wgt(k));
  }
};
template <> struct Pop<LBM_STORE_BF16_SHIFT> {
  using T = uint16_t;  // raw bfloat16 bits
  static __device__ __forceinline__ float ld(const T* p, uint32_t i, int k) {
    return __uint_as_float(((uint32_t)p[i]) << 16) + wgt(k);
  }
  static __device__ __forceinline__ uint16_t enc(float x) {
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
  static __device__ __forceinline__ void st(T* p, uint32_t i, int k, float v) {
    __builtin_nontemporal_store(enc(v - wgt(k)), p + i);
  }
  static __device__ __forceinline__ void st_cached(T* p, uint32_t i, int k, float v) {
    p[i] = enc(v - wgt(k));
  }
};

// n / ny via a 64-bit multiply-high: mdiv = floor((2^64-1)/ny) + 1 is exact for
// every n < 2^32 (ny == 1 is passed as mdiv == 0).
__device__ __forceinline__ uint32_t div_ny(uint32_t n, uint64_t mdiv) {
  return mdiv ? (uint32_t)__umul64hi((uint64_t)n, mdiv) : n;
}

// Row/column neighbours with periodic wrap, shared by the pull gather.
struct Nbr {
  uint32_t n;      // this node
  uint32_t rc, rm, rp;  // row offsets of i, i-1, i+1 (wrapped) times ny
  uint32_t j, jm, jp;   // columns j, j-1, j+1 (wrapped)
  __device__ __forceinline__ Nbr(uint32_t i, uint32_t j_, uint32_t nx, uint32_t ny) {
    j = j_;
    rc = i * ny;
    n = rc + j;
    rm = (i == 0 ? nx - 1 : i - 1) * ny;
    rp = (i + 1 == nx ? 0 : i + 1) * ny;
    jm = (j == 0 ? ny - 1 : j - 1);
    jp = (j + 1 == ny ? 0 : j + 1);
  }
  // source node of population k for a pull: x - c_k (wrapped)
  __device__ __forceinline__ uint32_t src(int k) const {
    const uint32_t row = cx(k) == 1 ? rm : (cx(k) == -1 ? rp : rc);
    const uint32_t col = cy(k) == 1 ? jm : (cy(k) == -1 ? jp : j);
    return row + col;
  }
};

}  // namespace lbm
