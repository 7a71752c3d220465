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

// lbm.py:44: a vector (4,1,1,1,1,1/4,1/4,1/4,1/4) divided by 9.0 in the
// working precision (fp32 as the reference's `float`, or fp64).
template <typename R = float>
__host__ __device__ constexpr R wgt(int k) {
  return k == 0 ? R(4) / R(9) : (k <= 4 ? R(1) / R(9) : R(0.25) / R(9));
}

// Collision.f_eq, CollisionEngine.py:14-18: w*rho*(1+3eu+4.5eu^2-1.5uu).
template <typename R>
__device__ __forceinline__ R feq_k(int k, R ux, R uy, R rho, R uv) {
  const R eu = (R)cx(k) * ux + (R)cy(k) * uy;
  return wgt<R>(k) * rho * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * uv);
}

// Collision.f_force (Guo), CollisionEngine.py:20-25: w*(3cF + 9 cF cu - 3 uF).
template <typename R>
__device__ __forceinline__ R fforce_k(int k, R Fx, R Fy, R ux, R uy, R uF) {
  const R cF = (R)cx(k) * Fx + (R)cy(k) * Fy;
  const R cu = (R)cx(k) * ux + (R)cy(k) * uy;
  return wgt<R>(k) * (R(3) * cF + R(9) * cF * cu - R(3) * uF);
}

// 1/x: the hardware reciprocal (1 ulp) in FP32, exact division in FP64 (the
// high-precision path).  The hot path divides by a handful of per-node
// quantities (rho, tau, tau - 1/2); an IEEE-exact FP32 division costs about
// ten instructions, the reciprocal one.
template <typename R>
__device__ __forceinline__ R rcp(R x) {
  if constexpr (sizeof(R) == 4) {
    return __builtin_amdgcn_rcpf(x);
  } else {
    return R(1) / x;
  }
}

// g^e for the power law nu = m g^(n-1) (FluidModels.py:26-29), from the
// squared rate g2 = g^2: exp2((e/2) log2 g2) with the hardware log2 / exp2 in
// FP32 (a few ulp against pow's one; the parity tolerance on tau is 1e-5),
// pow(sqrt g2, e) in FP64.  g = 0 gives +inf for e < 0 and 0 for e > 0 as pow
// does; e = 0 (n = 1) gives 1.
template <typename R>
__device__ __forceinline__ R pow_shear2(R g2, R e) {
  if constexpr (sizeof(R) == 4) {
    if (e == R(0)) return R(1);
    return __builtin_amdgcn_exp2f((e * R(0.5)) * __builtin_amdgcn_logf(g2));
  } else {
    return pow(sqrt(g2), e);
  }
}

// Collision.tau_max_min, CollisionEngine.py:28-34 (as selects: NaN passes
// through, as in the reference).
template <typename R>
__device__ __forceinline__ R clamp_tau(R t, R tmin, R tmax) {
  return t < tmin ? tmin : (t > tmax ? tmax : t);
}

// ---- buffer addressing ------------------------------------------------------
// A raw buffer resource over a range of fewer than 2^32 bytes: loads and
// stores then take a 32-bit per-lane offset plus a wave-uniform (scalar)
// offset instead of a 64-bit address per lane, and accesses past the range
// read 0 / are dropped by the hardware.
using Rsrc = __amdgpu_buffer_rsrc_t;
// raw buffer of `bytes` bytes (< 2^32; the record count is an unsigned
// 32-bit field, passed as its bit pattern)
__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
constexpr int kNonTemporal = 2;   // cache-policy bits of a streaming store

// ---- stochastic rounding of 16-bit stores ----------------------------------
// A step's post-collision populations are rounded to FP16 / BF16 with
// gfx950's stochastic conversions (v_cvt_sr_f16_f32 adds seed >> 19 to the
// FP32 bits before truncating to FP16, v_cvt_sr_bf16_f32 adds seed >> 16
// before truncating to BF16): the expected stored value is the FP32 value, so
// a steady flow's rounding does not push the same way every step.  Round to
// nearest even at every store left errors of 2e-5 (FP16) and 4e-4 (BF16) in
// the steady velocity of a force-driven channel, against 2e-6 and 2e-5 with
// these seeds.  The seed is a hash of the node's row and column and of the
// lattice's step count, so a lattice rounds the same alone, in a batch and
// across split runs.
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
// The step key of the populations a launch stores: `step` counts the steps
// of the lattice from its initial state up to and including the stored one.
__host__ __device__ __forceinline__ uint32_t round_key(uint32_t step) {
  return mix32(step * 0x9e3779b9u + 0x632be5abu);
}
// 64 random bits per node and step; direction k takes the 32-bit window at
// bit 4k (the conversions read a window's top bits).
struct SrSeed {
  uint32_t lo, hi;
  __device__ __forceinline__ SrSeed(uint32_t i, uint32_t j, uint32_t key) {
    lo = mix32(((i << 16) + j) ^ key);
    hi = mix32(lo ^ 0x85ebca6bu);
  }
  __device__ __forceinline__ uint32_t bits(int k) const {
    return k < 8 ? __builtin_amdgcn_alignbit(hi, lo, 4 * k) : hi;
  }
};
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf162_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f16_sr(float x, uint32_t seed) {
  return __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_f16_f32((half2_t){0, 0}, x, seed, false)[0]);
}
__device__ __forceinline__ uint16_t bf16_sr(float x, uint32_t seed) {
  return __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_bf16_f32((bf162_t){0, 0}, x, seed, false)[0]);
}

// ---- population storage ---------------------------------------------------
// Pop<S>::T is the stored type, Pop<S>::R the arithmetic type.  Shifted
// formats store g = f - w_k so that the stored magnitude is the deviation
// from rest; their arithmetic is fp32.  st / stb round to nearest even (a
// re-store of loaded values: exact); st_sr / stb_sr store a step's new
// populations with stochastic rounding (seed from SrSeed::bits; FP32 and
// FP64 ignore it).
template <int S> struct Pop;
template <> struct Pop<LBM_STORE_F32> {
  using T = float;
  using R = float;
  static __device__ __forceinline__ R ld(const T* p, uint32_t i, int) { return p[i]; }
  static __device__ __forceinline__ void st(T* p, uint32_t i, int, R v) { __builtin_nontemporal_store(v, p + i); }
  static __device__ __forceinline__ void st_cached(T* p, uint32_t i, int, R v) { p[i] = v; }
  // byte offsets: voff per lane, soff wave-uniform
  static __device__ __forceinline__ R ldb(Rsrc r, uint32_t voff, uint32_t soff, int) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
  }
  static __device__ __forceinline__ void stb(Rsrc r, uint32_t voff, uint32_t soff, int, R v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)voff, (int)soff, kNonTemporal);
  }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, int k, R v, uint32_t) { st(p, i, k, v); }
  static __device__ __forceinline__ void stb_sr(Rsrc r, uint32_t voff, uint32_t soff, int k, R v, uint32_t) {
    stb(r, voff, soff, k, v);
  }
};
template <> struct Pop<LBM_STORE_F64> {
  using T = double;
  using R = double;
  static __device__ __forceinline__ R ld(const T* p, uint32_t i, int) { return p[i]; }
  static __device__ __forceinline__ void st(T* p, uint32_t i, int, R v) { __builtin_nontemporal_store(v, p + i); }
  static __device__ __forceinline__ void st_cached(T* p, uint32_t i, int, R v) { p[i] = v; }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, int k, R v, uint32_t) { st(p, i, k, v); }
};
template <> struct Pop<LBM_STORE_F16_SHIFT> {
  using T = _Float16;
  using R = float;
  static __device__ __forceinline__ R ld(const T* p, uint32_t i, int k) { return (float)p[i] + wgt<float>(k); }
  static __device__ __forceinline__ void st(T* p, uint32_t i, int k, R v) {
    __builtin_nontemporal_store((T)(v - wgt<float>(k)), p + i);
  }
  static __device__ __forceinline__ void st_cached(T* p, uint32_t i, int k, R v) { p[i] = (T)(v - wgt<float>(k)); }
  static __device__ __forceinline__ R ldb(Rsrc r, uint32_t voff, uint32_t soff, int k) {
    const uint16_t b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
    return (float)__builtin_bit_cast(_Float16, b) + wgt<float>(k);
  }
  static __device__ __forceinline__ void stb(Rsrc r, uint32_t voff, uint32_t soff, int k, R v) {
    const uint16_t b = __builtin_bit_cast(uint16_t, (T)(v - wgt<float>(k)));
    __builtin_amdgcn_raw_buffer_store_b16(b, r, (int)voff, (int)soff, kNonTemporal);
  }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, int k, R v, uint32_t seed) {
    __builtin_nontemporal_store(__builtin_bit_cast(T, f16_sr(v - wgt<float>(k), seed)), p + i);
  }
  static __device__ __forceinline__ void stb_sr(Rsrc r, uint32_t voff, uint32_t soff, int k, R v, uint32_t seed) {
    __builtin_amdgcn_raw_buffer_store_b16(f16_sr(v - wgt<float>(k), seed), r, (int)voff, (int)soff, kNonTemporal);
  }
};
template <> struct Pop<LBM_STORE_BF16_SHIFT> {
  using T = uint16_t;  // raw bfloat16 bits
  using R = float;
  static __device__ __forceinline__ R ld(const T* p, uint32_t i, int k) {
    return __uint_as_float(((uint32_t)p[i]) << 16) + wgt<float>(k);
  }
  static __device__ __forceinline__ uint16_t enc(float x) {
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
  static __device__ __forceinline__ void st(T* p, uint32_t i, int k, R v) {
    __builtin_nontemporal_store(enc(v - wgt<float>(k)), p + i);
  }
  static __device__ __forceinline__ void st_cached(T* p, uint32_t i, int k, R v) { p[i] = enc(v - wgt<float>(k)); }
  static __device__ __forceinline__ R ldb(Rsrc r, uint32_t voff, uint32_t soff, int k) {
    const uint16_t b = __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
    return __uint_as_float(((uint32_t)b) << 16) + wgt<float>(k);
  }
  static __device__ __forceinline__ void stb(Rsrc r, uint32_t voff, uint32_t soff, int k, R v) {
    __builtin_amdgcn_raw_buffer_store_b16(enc(v - wgt<float>(k)), r, (int)voff, (int)soff, kNonTemporal);
  }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, int k, R v, uint32_t seed) {
    __builtin_nontemporal_store(bf16_sr(v - wgt<float>(k), seed), p + i);
  }
  static __device__ __forceinline__ void stb_sr(Rsrc r, uint32_t voff, uint32_t soff, int k, R v, uint32_t seed) {
    __builtin_amdgcn_raw_buffer_store_b16(bf16_sr(v - wgt<float>(k), seed), r, (int)voff, (int)soff, kNonTemporal);
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
