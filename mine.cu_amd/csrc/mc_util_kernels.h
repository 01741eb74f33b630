// mc_util_kernels.h - the small one-purpose kernels of the library: episode
// ring set-up, standalone world generation and reset ops, the camera basis,
// the graph-mode slot words and the test probes of the device arithmetic.
// Included by mc_capi.hip only (the step and render kernels are instantiated
// by mc_kernels.hip).
#pragma once
#include "mc_device.h"

namespace mc {

// Episode-ring set-up, one workgroup of 1024 threads.
//   mode 0: ring[slot] = any(step_count >= L), ring[(slot+1)%3] = 0, and when
//           a deferred-env list is given (p.fb), its slot-0 header words
//           (count, finished workgroups) = 0
//   mode 1 (before a reset): ring[3] = any(mask) (1 when mask is NULL),
//           ring[slot] = ring[(slot+1)%3] = 0
__global__ __launch_bounds__(1024) void k_ring_setup(Params p, int mode) {
  uint32_t any = 0;
  const int slot = p.slot_dev ? (int)*p.slot_dev : p.slot;
  if (mode == 0) {
    if (p.episode_length >= 0)
      for (int i = threadIdx.x; i < p.B; i += 1024) any |= (p.step_count[i] >= p.episode_length);
  } else {
    if (!p.mask) any = 1;
    else
      for (int i = threadIdx.x; i < p.B; i += 1024) any |= (p.mask[i] != 0);
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      p.ring[slot] = any ? 1u : 0u;
      if (p.fb) {
        p.fb[0] = 0u;
        p.fb[1] = 0u;
      }
    } else {
      p.ring[3] = any ? 1u : 0u;
      p.ring[slot] = 0u;
    }
    p.ring[(slot + 1) % 3] = 0u;
  }
}

// Masked regeneration of whole worlds (mask NULL = all); one workgroup per env.
__global__ __launch_bounds__(256) void k_regen(Params p) {
  const int b = blockIdx.x;
  if (p.mask && !p.mask[b]) return;
  const int nvox = p.X * p.Y * p.Z;
  store_fresh_world(p, b, p.voxels + (int64_t)b * nvox, nullptr, nvox, threadIdx.x, 256);
}

// k_regen for the seeded terrain world.
__global__ __launch_bounds__(256) void k_regen_terrain(Params p) {
  const int b = blockIdx.x;
  if (p.mask && !p.mask[b]) return;
  const int nvox = p.X * p.Y * p.Z;
  store_fresh_world<true>(p, b, p.voxels + (int64_t)b * nvox, nullptr, nvox, threadIdx.x, 256);
}

// generate_world_kernel (src/kernels.cu:1793-1828), 16 voxels per thread.
__global__ void k_generate_world(int8_t* voxels, int64_t total, int XZ, int Y, int gh) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  const int64_t wxyz = (int64_t)XZ * Y;
  if ((wxyz & 15) == 0) {
    const int64_t i0 = i << 4;
    if (i0 >= total) return;
    int8_t v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = layer_block((int)(((i0 + k) % wxyz) / XZ), gh);
    uint4 r;
    r.x = pack4(v[0], v[1], v[2], v[3]);
    r.y = pack4(v[4], v[5], v[6], v[7]);
    r.z = pack4(v[8], v[9], v[10], v[11]);
    r.w = pack4(v[12], v[13], v[14], v[15]);
    reinterpret_cast<uint4*>(voxels)[i] = r;
  } else {
    if (i >= total) return;
    voxels[i] = layer_block((int)((i % wxyz) / XZ), gh);
  }
}

// place_tree_kernel (src/kernels.cu:1831-1865); x/z outside the world are skipped.
__global__ void k_place_tree(int8_t* voxels, int B, int X, int Y, int Z, int tx, int tz, int gh) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || tx < 0 || tx >= X || tz < 0 || tz >= Z) return;
  const int64_t base = (int64_t)b * X * Y * Z;
  for (int dy = 0; dy < 4; ++dy) {
    const int y = gh + dy;
    if (y >= 0 && y < Y) voxels[base + (int64_t)y * X * Z + tx * Z + tz] = (dy < 3) ? kOakLog : kLeaves;
  }
}

// place_house_kernel (src/kernels.cu:1868-1946) via the closed form above.
__global__ void k_place_house(int8_t* voxels, int B, int X, int Y, int Z, int hx, int hz, int gh) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int64_t base = (int64_t)b * X * Y * Z;
  for (int dy = 0; dy <= 4; ++dy)
    for (int dx = 0; dx < 5; ++dx)
      for (int dz = 0; dz < 4; ++dz) {
        const int vx = hx + dx, vy = gh + dy, vz = hz + dz;
        if (vx < 0 || vx >= X || vy < 0 || vy >= Y || vz < 0 || vz >= Z) continue;
        const int8_t h = house_block(dy, dx, dz);
        if (h != kNoWrite) voxels[base + (int64_t)vy * X * Z + vx * Z + vz] = h;
      }
}

// reset_kernel (src/kernels.cu:1760-1786).
__global__ void k_reset(float* pos, float* vel, float* yaw, float* pitch, uint8_t* og,
                        const uint8_t* do_reset, int B, float sx, float sy, float sz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B || !do_reset[i]) return;
  pos[i * 3 + 0] = sx; pos[i * 3 + 1] = sy; pos[i * 3 + 2] = sz;
  vel[i * 3 + 0] = 0.0f; vel[i * 3 + 1] = 0.0f; vel[i * 3 + 2] = 0.0f;
  yaw[i] = 0.0f; pitch[i] = 0.0f; og[i] = 0;
}

// Camera basis of every env from its [B, 5] camera (src/kernels.cu:511-561),
// one thread per env.
__global__ void k_camera_basis(Params p) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= p.B) return;
  const float* cam = p.cameras + (int64_t)b * 5;
  float sin_yaw, cos_yaw, sin_pitch, cos_pitch;
  mc_sincosf(cam[3], &sin_yaw, &cos_yaw);
  mc_sincosf(cam[4], &sin_pitch, &cos_pitch);
  Frame f;
  f.ex = cam[0];
  f.ey = cam[1] + 1.6f;
  f.ez = cam[2];
  f.fx = sin_yaw * cos_pitch;
  f.fy = sin_pitch;
  f.fz = cos_yaw * cos_pitch;
  f.rx = cos_yaw;
  f.rz = -sin_yaw;
  f.ux = -sin_yaw * sin_pitch;
  f.uy = cos_pitch;
  f.uz = -cos_yaw * sin_pitch;
  store_basis(p, b, f);
}

// binary16 conversion, sum and order of n float pairs (test hook for the
// half-precision DDA arithmetic).
__global__ void k_debug_f16(const float* a, const float* b, uint16_t* ha, uint16_t* sum,
                            uint8_t* lt, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const _Float16 x = (_Float16)a[i], y = (_Float16)b[i];
  const _Float16 z = x + y;
  ha[i] = __builtin_bit_cast(uint16_t, x);
  sum[i] = __builtin_bit_cast(uint16_t, z);
  lt[i] = (x < y) ? 1 : 0;
}

// Scan of the ray-setup arithmetic cores against the full IEEE operations
// (test hook).  mode 0: 1 / x for the float bit patterns begin .. begin +
// count - 1 with |x| in [2^-90, 2^90]; mode 1: sqrt(x) for the patterns with x
// in [2^-60, 2^60]; mode 2: n / d for count hashed pairs, |d| in [2^-30, 2^30]
// and n = 0 or |n| in [2^-60, 2^30].  out[0] += inputs checked, out[1] +=
// mismatches, out[2] = one mismatching input (bits; mode 2: the denominator).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}
__global__ void k_debug_fast_math(int mode, uint32_t begin, uint32_t count, uint32_t* out) {
  uint32_t checked = 0, bad = 0, bad_in = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
    const uint32_t bits = begin + (uint32_t)k;
    if (mode == 0) {
      const float x = __uint_as_float(bits);
      const float ax = fabsf(x);
      if (!(ax >= 0x1p-90f && ax <= 0x1p90f)) continue;
      const float f = rcp_cr(x);
      const float g = 1.0f / x;
      ++checked;
      if (__float_as_uint(f) != __float_as_uint(g)) { ++bad; bad_in = bits; }
    } else if (mode == 1) {
      const float x = __uint_as_float(bits);
      if (!(x >= 0x1p-60f && x <= 0x1p60f)) continue;
      const float f = sqrt_core(x);
      const float g = sqrtf(x);
      ++checked;
      if (__float_as_uint(f) != __float_as_uint(g)) { ++bad; bad_in = bits; }
    } else {
      const uint32_t h1 = mix32(bits * 2u + 1u), h2 = mix32(bits * 2u + 2u);
      // |d| in [2^-30, 2^30): exponent field 97..156; n: 67..156, or 0
      const float d = __uint_as_float((h1 & 0x807fffffu) | ((97u + (h1 >> 23) % 60u) << 23));
      const float n = ((h2 & 255u) == 0u) ? 0.0f
                      : __uint_as_float((h2 & 0x807fffffu) | ((67u + (h2 >> 23) % 90u) << 23));
      const float f = div_core(n, d, recip_core(d));
      const float g = n / d;
      ++checked;
      // The cores may return -0 for a zero quotient where IEEE gives +0.
      if (__float_as_uint(f) != __float_as_uint(g) && !(f == 0.0f && g == 0.0f)) {
        ++bad;
        bad_in = __float_as_uint(d);
      }
    }
  }
  if (checked) atomicAdd(&out[0], checked);
  if (bad) {
    atomicAdd(&out[1], bad);
    out[2] = bad_in;
  }
}

// Graph-replayable steps: advance the device-resident ring slot (one thread,
// launched after the step on the same stream).
__global__ void k_slot_advance(uint32_t* slot) { slot[0] = (slot[0] + 1u) % 3u; }
// Set the device-resident slot (graph mode switch-on).
__global__ void k_slot_set(uint32_t* slot, uint32_t v) { slot[0] = v; }

__global__ void k_debug_sincos(const float* x, float* s, float* c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mc_sincosf(x[i], &s[i], &c[i]);
}

}  // namespace mc
