// Immersed-boundary particles (taichi_lbm/ibm.py): circular particles whose
// boundary is a ring of Lagrangian markers ("IB nodes") coupled to the
// lattice by multi-direct forcing -- bilinear interpolation of the lattice
// velocity at the markers, a correction force 2 (u_desired - u) per marker,
// and its spreading back onto the lattice's velocity and body force.
//
// Layout (lbm_ib_particles in lbm_abi.h): sphere fields are arrays over the
// Max_Number capacity, marker fields arrays over Max_Number * Max_NumIB with
// marker (m, n) at m * Max_NumIB + n; vectors are two consecutive planes
// (x then y).  Lattice fields use the plane layout of the other kernels.
//
// The marker list (IBNodeInsiderGroup) is compacted in (m, n) order with
// per-sphere block scans, so it does not depend on scheduling; the spread is
// a scatter onto at most four lattice nodes per marker, accumulated with
// float atomics as the reference's parallel `+=` is.  Stencil nodes outside
// the lattice are skipped (the reference reads them unchecked; the marker
// selection keeps them inside for every finite sphere state).

#include "lbm_physics.h"

#include <math.h>

namespace lbm {
namespace {

constexpr int kIbBlock = 256;
constexpr int kScanBlock = 1024;

struct IB {
  lbm_ib_particles p;
  uint32_t cap;  // Max_Number * Max_NumIB
  uint32_t K;    // Max_NumIB
};

struct Lat {
  float* rho;        // nc planes
  float* total_rho;  // 1 plane
  float* vel;        // 2 planes
  float* body_force; // nc * 2 planes
  int32_t nx, ny, nc;
  uint32_t ps;
};

// SphereIB: NumIB = round(pi * 2 * radius / NUMIB_cof), ibm.py:61
__device__ __forceinline__ int marker_count(float radius, float cof) {
  const float x = 6.28318530717958647692f * radius / cof;
  // halves round up; NaN and values beyond the int range give no markers
  if (!(x > -1.0f && x < 2147483520.0f)) return 0;
  return (int)floorf(x + 0.5f);
}

// ibm.py:63 / :72: angle = 2.0 * math.pi * n / NumIB
__device__ __forceinline__ float marker_angle(int n, int N) {
  return 6.28318530717958647692f * (float)n / (float)N;
}

// SphereIB.init_IB_nodes, ibm.py:57-67: one block per sphere
__global__ void __launch_bounds__(kIbBlock) k_ib_init(const IB a) {
  const int m = blockIdx.x;
  const lbm_ib_particles& p = a.p;
  const int N = marker_count(p.radius[m], p.numib_cof);
  if (threadIdx.x == 0) p.num_ib[m] = N;
  const int nmax = min(max(N, 0), (int)a.K);
  const float cx = p.pos[m], cy = p.pos[p.max_number + m], r = p.radius[m];
  for (int n = threadIdx.x; n < nmax; n += blockDim.x) {
    const float ang = marker_angle(n, N);
    const uint32_t q = (uint32_t)m * a.K + n;
    p.node_pos[q] = cx + r * cosf(ang);
    p.node_pos[a.cap + q] = cy + r * sinf(ang);
  }
}

// IBDEMCouplerEngine.CheckPos, ibm.py:92-108
__device__ __forceinline__ int check_pos(float x, float y, float d, int nx, int ny) {
  const float x_min = x - d, x_max = x + d, y_min = y - d, y_max = y + d;
  const float fx = (float)(nx - 1), fy = (float)(ny - 1);
  if (x_min >= 0.0f && x_max <= fx && y_min >= 0.0f && y_max <= fy) return 1;
  if (x_max < 0.0f || x_min > fx || y_max < 0.0f || y_min > fy) return 0;
  return 2;
}

// IBDEMCouplerEngine.InDomain, ibm.py:111-116
__device__ __forceinline__ bool in_domain(float x, float y, int nx, int ny) {
  return x > 0.0f && x < (float)(nx - 1) && y > 0.0f && y < (float)(ny - 1);
}

// Block-wide exclusive scan of one flag per thread; returns the prefix and
// writes the block total.
__device__ __forceinline__ int block_scan_flag(bool flag, int* lds_waves, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned long long b = __ballot(flag);
  const int below = __popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) lds_waves[wave] = __popcll(b);
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    const int c = lds_waves[w];
    base += w < wave ? c : 0;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return base + below;
}

// IBDEMCouplerEngine.SphereIBUpdate, ibm.py:120-141, first pass: one block
// per sphere moves its markers (computer_IB_nodes_pos, ibm.py:69-78), sets
// their desired velocity and ranks the selected ones within the sphere.
__global__ void __launch_bounds__(kIbBlock) k_ib_update(const IB a, int32_t nx, int32_t ny) {
  __shared__ int lds[kIbBlock / 64];
  const int m = blockIdx.x;
  const lbm_ib_particles& p = a.p;
  int32_t* rank = p.work;
  int32_t* sphere_count = p.work + a.cap;
  const float cx = p.pos[m], cy = p.pos[p.max_number + m], r = p.radius[m];
  const int flag = check_pos(cx, cy, 2.0f * r, nx, ny);
  const int N = p.num_ib[m];
  const int nmax = flag == 0 ? 0 : min(max(N, 0), (int)a.K);
  const float th = p.angle[m];
  const float ct = cosf(th), st = sinf(th);
  const float vx = p.vel[m], vy = p.vel[p.max_number + m];
  int carry = 0;
  for (int n0 = 0; n0 < (int)a.K; n0 += blockDim.x) {
    const int n = n0 + threadIdx.x;
    const uint32_t q = (uint32_t)m * a.K + n;
    bool sel = false;
    if (n < nmax) {
      const float ang = marker_angle(n, N);
      const float xref = r * cosf(ang), yref = r * sinf(ang);
      const float x = cx + xref * ct - yref * st;
      const float y = cy + xref * st + yref * ct;
      p.node_pos[q] = x;
      p.node_pos[a.cap + q] = y;
      p.node_vel_desired[q] = vx;
      p.node_vel_desired[a.cap + q] = vy;
      sel = flag == 1 || in_domain(x, y, nx, ny);
    }
    int tot;
    const int pre = block_scan_flag(sel, lds, tot);
    if (n < (int)a.K) rank[q] = sel ? carry + pre : -1;
    carry += tot;
  }
  if (threadIdx.x == 0) sphere_count[m] = carry;
}

// second pass: exclusive scan of the per-sphere counts (one block)
__global__ void __launch_bounds__(kScanBlock) k_ib_scan(const IB a) {
  __shared__ int lds[kScanBlock / 64];
  const lbm_ib_particles& p = a.p;
  const int32_t* sphere_count = p.work + a.cap;
  int32_t* offset = p.work + a.cap + p.max_number;
  int carry = 0;
  for (int m0 = 0; m0 < p.num; m0 += blockDim.x) {
    const int m = m0 + threadIdx.x;
    const int c = m < p.num ? sphere_count[m] : 0;
    // inclusive wave scan, then wave totals through LDS
    int v = c;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(v, d, 64);
      if (lane >= d) v += t;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
      base += w < wave ? lds[w] : 0;
      tot += lds[w];
    }
    __syncthreads();
    if (m < p.num) offset[m] = carry + base + v - c;
    carry += tot;
  }
  if (threadIdx.x == 0) *p.count = carry;
}

// third pass: scatter the selected markers into the group, clear the rest
// (the reference fills the whole group with (0, 0) first, ibm.py:123)
__global__ void __launch_bounds__(kIbBlock) k_ib_compact(const IB a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.cap) return;
  const lbm_ib_particles& p = a.p;
  const int32_t count = *p.count;
  int2* group = (int2*)p.group;
  if ((int32_t)t >= count) group[t] = make_int2(0, 0);
  const uint32_t m = t / a.K, n = t - m * a.K;
  if ((int32_t)m >= p.num) return;
  const int32_t rk = p.work[t];
  if (rk >= 0) group[p.work[a.cap + p.max_number + m] + rk] = make_int2((int)m, (int)n);
}

// SphereIB.reset_force, ibm.py:80-83
__global__ void __launch_bounds__(kIbBlock) k_ib_reset_force(const IB a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.cap) return;
  a.p.node_force[t] = 0.0f;
  a.p.node_force[a.cap + t] = 0.0f;
}

// The marker of group entry t, or false for an entry past the count or
// naming no marker of the capacity.
__device__ __forceinline__ bool group_marker(const IB& a, uint32_t t, uint32_t& q) {
  if ((int32_t)t >= *a.p.count) return false;
  const int2 mn = ((const int2*)a.p.group)[t];
  if (mn.x < 0 || mn.x >= a.p.max_number || mn.y < 0 || (uint32_t)mn.y >= a.K) return false;
  q = (uint32_t)mn.x * a.K + (uint32_t)mn.y;
  return true;
}

// IBEngine.BottomAndTop, ibm.py:187-194, with coordinates clamped before the
// conversion (NaN and huge values give stencils outside the lattice)
__device__ __forceinline__ void stencil(float x, float y, int nx, int ny, int& x0, int& x1, int& y0, int& y1) {
  const float cx = fminf(fmaxf(x, -2.0f), (float)nx + 1.0f);
  const float cy = fminf(fmaxf(y, -2.0f), (float)ny + 1.0f);
  x0 = (int)floorf(cx);
  x1 = (int)ceilf(cx);
  y0 = (int)floorf(cy);
  y1 = (int)ceilf(cy);
}

// IBEngine.weight, ibm.py:196-198
__device__ __forceinline__ float ib_weight(float x, float y, int ix, int iy) {
  return (1.0f - fabsf(x - (float)ix)) * (1.0f - fabsf(y - (float)iy));
}

// IBEngine.InterpolateBodyVelocities, ibm.py:201-223
__global__ void __launch_bounds__(kIbBlock) k_ib_interpolate(const IB a, const Lat L) {
  uint32_t q;
  if (!group_marker(a, blockIdx.x * blockDim.x + threadIdx.x, q)) return;
  const lbm_ib_particles& p = a.p;
  const float x = p.node_pos[q], y = p.node_pos[a.cap + q];
  int x0, x1, y0, y1;
  stencil(x, y, L.nx, L.ny, x0, x1, y0, y1);
  float vx = 0.0f, vy = 0.0f;
  for (int ix = x0; ix <= x1; ++ix) {
    if (ix < 0 || ix >= L.nx) continue;
    for (int iy = y0; iy <= y1; ++iy) {
      if (iy < 0 || iy >= L.ny) continue;
      const float w = ib_weight(x, y, ix, iy);
      const uint32_t n = (uint32_t)ix * L.ny + iy;
      vx += L.vel[n] * w;
      vy += L.vel[L.ps + n] * w;
    }
  }
  p.node_vel[q] = vx;
  p.node_vel[a.cap + q] = vy;
  const float fx = 2.0f * (p.node_vel_desired[q] - vx);
  const float fy = 2.0f * (p.node_vel_desired[a.cap + q] - vy);
  p.node_force_temp[q] = fx;
  p.node_force_temp[a.cap + q] = fy;
  p.node_force[q] += fx;
  p.node_force[a.cap + q] += fy;
}

// IBEngine.SpreadBodyForces, ibm.py:225-244
__global__ void __launch_bounds__(kIbBlock) k_ib_spread(const IB a, const Lat L) {
  uint32_t q;
  if (!group_marker(a, blockIdx.x * blockDim.x + threadIdx.x, q)) return;
  const lbm_ib_particles& p = a.p;
  const float x = p.node_pos[q], y = p.node_pos[a.cap + q];
  const float Fx = p.node_force_temp[q], Fy = p.node_force_temp[a.cap + q];
  int x0, x1, y0, y1;
  stencil(x, y, L.nx, L.ny, x0, x1, y0, y1);
  for (int ix = x0; ix <= x1; ++ix) {
    if (ix < 0 || ix >= L.nx) continue;
    for (int iy = y0; iy <= y1; ++iy) {
      if (iy < 0 || iy >= L.ny) continue;
      const float w = ib_weight(x, y, ix, iy);
      const uint32_t n = (uint32_t)ix * L.ny + iy;
      const float tx = Fx * w, ty = Fy * w;
      const float rt = L.total_rho[n];
      const float r2 = rt * 2.0f;
      unsafeAtomicAdd(&L.vel[n], tx / r2);
      unsafeAtomicAdd(&L.vel[L.ps + n], ty / r2);
      for (int c = 0; c < L.nc; ++c) {
        const float rc = L.rho[(size_t)c * L.ps + n];
        unsafeAtomicAdd(&L.body_force[(size_t)(2 * c) * L.ps + n], tx * rc / rt);
        unsafeAtomicAdd(&L.body_force[(size_t)(2 * c + 1) * L.ps + n], ty * rc / rt);
      }
    }
  }
}

// IBEngine.ComputeParticleForce, ibm.py:171-185: one block per sphere sums
// force and torque (F_x r_y - F_y r_x) over the sphere's group entries,
// found through the offsets of the last SphereIBUpdate; a sphere without
// entries keeps its previous force and torque.
__global__ void __launch_bounds__(kIbBlock) k_ib_particle_force(const IB a) {
  __shared__ float lds[3][kIbBlock];
  const int m = blockIdx.x;
  const lbm_ib_particles& p = a.p;
  const int cnt = p.work[a.cap + m];
  const int off = p.work[a.cap + p.max_number + m];
  if (cnt <= 0 || off < 0 || (uint64_t)off + (uint64_t)cnt > a.cap) return;
  const float cx = p.pos[m], cy = p.pos[p.max_number + m];
  const int2* group = (const int2*)p.group;
  float tq = 0.0f, sx = 0.0f, sy = 0.0f;
  for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
    const int2 mn = group[off + e];
    if (mn.x != m || mn.y < 0 || (uint32_t)mn.y >= a.K) continue;
    const uint32_t q = (uint32_t)m * a.K + (uint32_t)mn.y;
    const float fx = p.node_force[q], fy = p.node_force[a.cap + q];
    const float rx = p.node_pos[q] - cx, ry = p.node_pos[a.cap + q] - cy;
    tq += fx * ry - fy * rx;
    sx += fx;
    sy += fy;
  }
  lds[0][threadIdx.x] = tq;
  lds[1][threadIdx.x] = sx;
  lds[2][threadIdx.x] = sy;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + h];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + h];
      lds[2][threadIdx.x] += lds[2][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    p.torque[m] = lds[0][0];
    p.force[m] = lds[1][0];
    p.force[p.max_number + m] = lds[2][0];
  }
}

int check_ib(const lbm_ib_particles* p, IB& a) {
  if (!p) return fail("null particle set");
  if (p->max_number < 1 || p->max_num_ib < 1) return fail("Max_Number and Max_NumIB must be positive");
  const uint64_t cap = (uint64_t)p->max_number * (uint64_t)p->max_num_ib;
  if (cap + 2ull * (uint64_t)p->max_number >= (1ull << 31)) return fail("particle capacity too large");
  if (p->num < 0 || p->num > p->max_number) return fail("num = %d outside [0, Max_Number = %d]", p->num, p->max_number);
  if (!p->pos || !p->vel || !p->force || !p->torque || !p->radius || !p->angle || !p->num_ib || !p->node_pos ||
      !p->node_vel || !p->node_vel_desired || !p->node_force || !p->node_force_temp || !p->group || !p->count ||
      !p->work)
    return fail("a particle buffer is missing");
  a.p = *p;
  a.cap = (uint32_t)cap;
  a.K = (uint32_t)p->max_num_ib;
  return LBM_OK;
}

int check_lattice(const lbm_lattice* lat, const lbm_gfields* g, Lat& L) {
  if (!lat || !g) return fail("null lattice or fields");
  if (lat->nx < 1 || lat->ny < 1 || lat->nc < 1 || lat->nc > LBM_MAX_NC) return fail("bad lattice shape");
  if ((uint64_t)lat->nx * lat->ny >= (1ull << 31)) return fail("lattice too large");
  if ((uint64_t)lat->plane_stride < (uint64_t)lat->nx * lat->ny || lat->plane_stride >= (1ll << 31))
    return fail("bad plane_stride");
  if (!g->rho || !g->total_rho || !g->vel || !g->body_force) return fail("IB coupling needs rho, total_rho, vel, body_force");
  L.rho = g->rho;
  L.total_rho = g->total_rho;
  L.vel = g->vel;
  L.body_force = g->body_force;
  L.nx = lat->nx;
  L.ny = lat->ny;
  L.nc = lat->nc;
  L.ps = (uint32_t)lat->plane_stride;
  return LBM_OK;
}

int grid_of(uint32_t n) { return (int)((n + kIbBlock - 1) / kIbBlock); }

int interp_spread(const IB& a, const Lat& L, int iterations, hipStream_t s) {
  for (int it = 0; it < iterations; ++it) {
    hipLaunchKernelGGL(k_ib_interpolate, dim3(grid_of(a.cap)), dim3(kIbBlock), 0, s, a, L);
    hipLaunchKernelGGL(k_ib_spread, dim3(grid_of(a.cap)), dim3(kIbBlock), 0, s, a, L);
  }
  return hip_check(hipGetLastError(), "IB interpolate/spread launch");
}

}  // namespace
}  // namespace lbm

using namespace lbm;

extern "C" int64_t lbm_ib_work_ints(int32_t max_number, int32_t max_num_ib) {
  if (max_number < 1 || max_num_ib < 1) return 0;
  return (int64_t)max_number * max_num_ib + 2 * (int64_t)max_number;
}

extern "C" int lbm_ib_init_nodes(const lbm_ib_particles* p, void* stream) {
  IB a;
  int rc = check_ib(p, a);
  if (rc) return rc;
  if (a.p.num == 0) return LBM_OK;
  hipLaunchKernelGGL(k_ib_init, dim3(a.p.num), dim3(kIbBlock), 0, (hipStream_t)stream, a);
  return hip_check(hipGetLastError(), "IB init launch");
}

extern "C" int lbm_ib_update(const lbm_lattice* lat, const lbm_ib_particles* p, void* stream) {
  IB a;
  int rc = check_ib(p, a);
  if (rc) return rc;
  if (!lat || lat->nx < 1 || lat->ny < 1) return fail("bad lattice shape");
  hipStream_t s = (hipStream_t)stream;
  if (a.p.num > 0) hipLaunchKernelGGL(k_ib_update, dim3(a.p.num), dim3(kIbBlock), 0, s, a, lat->nx, lat->ny);
  hipLaunchKernelGGL(k_ib_scan, dim3(1), dim3(kScanBlock), 0, s, a);
  hipLaunchKernelGGL(k_ib_compact, dim3(grid_of(a.cap)), dim3(kIbBlock), 0, s, a);
  return hip_check(hipGetLastError(), "IB update launch");
}

extern "C" int lbm_ib_reset_force(const lbm_ib_particles* p, void* stream) {
  IB a;
  int rc = check_ib(p, a);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ib_reset_force, dim3(grid_of(a.cap)), dim3(kIbBlock), 0, (hipStream_t)stream, a);
  return hip_check(hipGetLastError(), "IB reset_force launch");
}

extern "C" int lbm_ib_interpolate(const lbm_lattice* lat, const lbm_gfields* g, const lbm_ib_particles* p,
                                  void* stream) {
  IB a;
  Lat L;
  int rc = check_ib(p, a);
  if (rc || (rc = check_lattice(lat, g, L))) return rc;
  hipLaunchKernelGGL(k_ib_interpolate, dim3(grid_of(a.cap)), dim3(kIbBlock), 0, (hipStream_t)stream, a, L);
  return hip_check(hipGetLastError(), "IB interpolate launch");
}

extern "C" int lbm_ib_spread(const lbm_lattice* lat, const lbm_gfields* g, const lbm_ib_particles* p, void* stream) {
  IB a;
  Lat L;
  int rc = check_ib(p, a);
  if (rc || (rc = check_lattice(lat, g, L))) return rc;
  hipLaunchKernelGGL(k_ib_spread, dim3(grid_of(a.cap)), dim3(kIbBlock), 0, (hipStream_t)stream, a, L);
  return hip_check(hipGetLastError(), "IB spread launch");
}

extern "C" int lbm_ib_mdf(const lbm_lattice* lat, const lbm_gfields* g, const lbm_ib_particles* p, int32_t iterations,
                          void* stream) {
  IB a;
  Lat L;
  int rc = check_ib(p, a);
  if (rc || (rc = check_lattice(lat, g, L))) return rc;
  if (iterations < 0) return fail("MDFIteration = %d", iterations);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ib_reset_force, dim3(grid_of(a.cap)), dim3(kIbBlock), 0, s, a);
  if ((rc = hip_check(hipGetLastError(), "IB reset_force launch"))) return rc;
  return interp_spread(a, L, iterations, s);
}

extern "C" int lbm_ib_particle_force(const lbm_ib_particles* p, void* stream) {
  IB a;
  int rc = check_ib(p, a);
  if (rc) return rc;
  if (a.p.num == 0) return LBM_OK;
  hipLaunchKernelGGL(k_ib_particle_force, dim3(a.p.num), dim3(kIbBlock), 0, (hipStream_t)stream, a);
  return hip_check(hipGetLastError(), "IB particle force launch");
}
