// FP64 reductions over the lattice (mass per component, momentum, NaN count,
// velocity residual) and the plain copy kernel used as the bandwidth
// reference.  Wavefront reduction with DPP-backed shuffles, one FP64 atomic
// per workgroup.

#include "lbm_physics.h"

namespace lbm {
namespace {

constexpr int kRedBlock = 256;
constexpr int kMaxRed = LBM_MAX_NC + 3;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int NV>
__device__ __forceinline__ void block_sum_atomic(double (&v)[NV], double* out) {
  __shared__ double part[kRedBlock / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = wave_sum(v[q]);
    if (lane == 0) part[wave][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kRedBlock / 64; ++w) s += part[w][threadIdx.x];
    atomicAdd(out + threadIdx.x, s);
  }
}

// out[c] = sum rho_c, out[nc..nc+1] = sum total_rho*u, out[nc+2] = non-finite count
template <int NC>
__global__ __launch_bounds__(kRedBlock) void k_reduce_macro(const float* __restrict__ mac,
                                                             const uint16_t* __restrict__ code, uint32_t N,
                                                             uint32_t ps, double* out) {
  double v[NC + 3];
#pragma unroll
  for (int q = 0; q < NC + 3; ++q) v[q] = 0.0;
  for (uint32_t n = blockIdx.x * kRedBlock + threadIdx.x; n < N; n += gridDim.x * kRedBlock) {
    if (!(code[n] & LBM_CODE_ACTIVE)) continue;
    float bad = 0.0f;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float r = mac[(size_t)c * ps + n];
      v[c] += (double)r;
      bad += isfinite(r) ? 0.0f : 1.0f;
    }
    const float rt = mac[(size_t)NC * ps + n];
    const float ux = mac[(size_t)(NC + 1) * ps + n], uy = mac[(size_t)(NC + 2) * ps + n];
    v[NC] += (double)rt * (double)ux;
    v[NC + 1] += (double)rt * (double)uy;
    bad += (isfinite(ux) && isfinite(uy) && isfinite(rt)) ? 0.0f : 1.0f;
    v[NC + 2] += (double)bad;
  }
  block_sum_atomic<NC + 3>(v, out);
}

__global__ __launch_bounds__(kRedBlock) void k_residual(const float* __restrict__ vel, float* __restrict__ old,
                                                         uint32_t N, uint32_t ps, double* out) {
  double v[1] = {0.0};
  for (uint32_t n = blockIdx.x * kRedBlock + threadIdx.x; n < N; n += gridDim.x * kRedBlock) {
    const float ux = vel[n], uy = vel[(size_t)ps + n];
    const float dx = ux - old[n], dy = uy - old[(size_t)ps + n];
    v[0] += (double)(dx * dx + dy * dy);
    old[n] = ux;
    old[(size_t)ps + n] = uy;
  }
  block_sum_atomic<1>(v, out);
}

// Clears the accumulators with a kernel (not a memset node) so the
// reductions stay correct when a caller captures them in a graph.
__global__ void k_zero(double* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

typedef float v4f __attribute__((ext_vector_type(4)));

// One 16-byte element per thread, streamed (non-temporal) in and out: the
// fastest plain copy measured on MI355X (a grid-stride loop over a capped
// grid reaches about 80 % of its bandwidth).
__global__ __launch_bounds__(256) void k_copy(const v4f* __restrict__ src, v4f* __restrict__ dst, size_t n4) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + q), dst + q);
}

__global__ void k_copy_tail(const float* __restrict__ src, float* __restrict__ dst, size_t start, size_t n) {
  const size_t q = start + threadIdx.x;
  if (q < n) dst[q] = src[q];
}

// Scratch for the reductions: device accumulators and a host copy.  Owned by
// the library per calling thread (one allocation, reused).
struct RedScratch {
  double* dev = nullptr;
  int device = -1;
};
thread_local RedScratch g_red;

int red_buffer(double** p) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (!g_red.dev || g_red.device != dev) {
    if ((rc = hip_check(hipMalloc((void**)&g_red.dev, sizeof(double) * kMaxRed), "hipMalloc(reduction)"))) return rc;
    g_red.device = dev;
  }
  *p = g_red.dev;
  return LBM_OK;
}

uint32_t red_grid(uint32_t N) {
  uint32_t b = (N + kRedBlock - 1) / kRedBlock;
  return b > 2048 ? 2048 : (b ? b : 1);
}

}  // namespace
}  // namespace lbm

using namespace lbm;

extern "C" int lbm_reduce_macro(const lbm_lattice* lat, const float* macro, const uint16_t* code, double* out,
                                void* stream) {
  if (!lat || !macro || !code || !out) return fail("null argument to lbm_reduce_macro");
  if (lat->nc < 1 || lat->nc > LBM_MAX_NC) return fail("bad nc");
  hipStream_t s = (hipStream_t)stream;
  double* d = nullptr;
  int rc = red_buffer(&d);
  if (rc) return rc;
  const int nv = lat->nc + 3;
  const uint32_t N = (uint32_t)lat->nx * (uint32_t)lat->ny;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kMaxRed), 0, s, d, nv);
  if ((rc = hip_check(hipGetLastError(), "zero launch"))) return rc;
  if (lat->nc == 1)
    hipLaunchKernelGGL(k_reduce_macro<1>, dim3(red_grid(N)), dim3(kRedBlock), 0, s, macro, code, N,
                       (uint32_t)lat->plane_stride, d);
  else
    hipLaunchKernelGGL(k_reduce_macro<2>, dim3(red_grid(N)), dim3(kRedBlock), 0, s, macro, code, N,
                       (uint32_t)lat->plane_stride, d);
  if ((rc = hip_check(hipGetLastError(), "reduce launch"))) return rc;
  if ((rc = hip_check(hipMemcpyAsync(out, d, sizeof(double) * nv, hipMemcpyDeviceToHost, s), "copy out")))
    return rc;
  return hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

extern "C" int lbm_residual(int64_t n_nodes, int64_t plane_stride, const float* vel, float* vel_old, double* out,
                            void* stream) {
  if (!vel || !vel_old || !out) return fail("null argument to lbm_residual");
  if (n_nodes < 0 || n_nodes >= (1ll << 31) || plane_stride < n_nodes) return fail("bad residual sizes");
  hipStream_t s = (hipStream_t)stream;
  double* d = nullptr;
  int rc = red_buffer(&d);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(kMaxRed), 0, s, d, 1);
  if ((rc = hip_check(hipGetLastError(), "zero launch"))) return rc;
  hipLaunchKernelGGL(k_residual, dim3(red_grid((uint32_t)n_nodes)), dim3(kRedBlock), 0, s, vel, vel_old,
                     (uint32_t)n_nodes, (uint32_t)plane_stride, d);
  if ((rc = hip_check(hipGetLastError(), "residual launch"))) return rc;
  if ((rc = hip_check(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, s), "copy out"))) return rc;
  return hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

extern "C" int lbm_copy_f32(const float* src, float* dst, int64_t n, void* stream) {
  if (!src || !dst || n < 0) return fail("bad lbm_copy_f32 arguments");
  if (((uintptr_t)src | (uintptr_t)dst) & 15) return fail("lbm_copy_f32 needs 16-byte aligned buffers");
  hipStream_t s = (hipStream_t)stream;
  const size_t n4 = (size_t)n / 4;
  if (n4) {
    const size_t blocks = (n4 + 255) / 256;
    if (blocks >= (1ull << 31)) return fail("lbm_copy_f32: copy too large for one launch");
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const v4f*)src, (v4f*)dst, n4);
    int rc = hip_check(hipGetLastError(), "copy launch");
    if (rc) return rc;
  }
  if ((size_t)n > n4 * 4) {
    hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(4), 0, s, src, dst, n4 * 4, (size_t)n);
    return hip_check(hipGetLastError(), "copy tail launch");
  }
  return LBM_OK;
}
