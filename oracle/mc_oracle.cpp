// mc_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Scalar CPU restatement of the reference environment step (Infatoshi/mine.cu,
// src/kernels.cu + minecu/__init__.py).  It is the checker for the gfx950 HIP
// kernels: only tests, the smoke check and the CPU-baseline leg of the
// benchmark load it.  No product code path links or calls it.
//
// Every function follows one reference kernel, in the reference's operation
// order (left-to-right products, no FP contraction -- built with
// -ffp-contract=off -fno-fast-math).  Numerics deliberately differ from the
// reference binary in one place: cos/sin/tan come from mc_math.h (IEEE,
// Payne-Hanek + Taylor in binary64, rounded once) instead of NVIDIA's
// fast-math intrinsics, and the per-pixel tanf(fov/2) is evaluated once on the
// host with the same code.  Division and sqrt are IEEE correctly rounded.
//
// Non-cubic worlds: the reference's uint8/float32 step paths index every axis
// with world_size = voxels.size(1) (src/bindings.cpp:546, kernels.cu:2619-2649);
// this restatement takes X, Y, Z separately (as the prebasis path does,
// src/bindings.cpp:759-761).  For cubic worlds the two are identical.
#include <math.h>

#include <cmath>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "mc_math.h"

// ORC_MUTANT (0 = none) compiles one planted slip into the shading, only for
// the test that checks the fog pins catch it (tests/test_oracle_known_answers.py);
// the oracle built by the Makefile has none.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

namespace {

// Block ids (src/kernels.cu:11-24).
constexpr int8_t AIR = -1, GRASS = 0, DIRT = 1, OAKLOG = 3, LEAVES = 4, BRICK = 8,
                 COBBLESTONE = 9, PLANKS = 10, BEDROCK = 12;

// Shaded colour table (src/kernels.cu:27-44).
const float BLOCK_COLORS[13][3] = {
    {0.30f, 0.65f, 0.20f}, {0.55f, 0.35f, 0.20f}, {0.50f, 0.50f, 0.50f},
    {0.55f, 0.40f, 0.25f}, {0.20f, 0.55f, 0.15f}, {0.90f, 0.85f, 0.60f},
    {0.20f, 0.40f, 0.80f}, {0.85f, 0.90f, 0.95f}, {0.70f, 0.35f, 0.30f},
    {0.40f, 0.40f, 0.40f}, {0.75f, 0.60f, 0.40f}, {0.95f, 0.95f, 0.98f},
    {0.15f, 0.15f, 0.15f},
};
const float SKY_COLOR[3] = {0.53f, 0.81f, 0.92f};

// Slab test against [min, max]^3 (src/kernels.cu:51-79).
bool ray_intersects_aabb(float ox, float oy, float oz, float dx, float dy, float dz,
                         float minx, float miny, float minz, float maxx, float maxy,
                         float maxz, float* t_entry, float* t_exit) {
  float inv_dx = (dx != 0.0f) ? 1.0f / dx : 1e30f;
  float inv_dy = (dy != 0.0f) ? 1.0f / dy : 1e30f;
  float inv_dz = (dz != 0.0f) ? 1.0f / dz : 1e30f;
  float t1 = (minx - ox) * inv_dx;
  float t2 = (maxx - ox) * inv_dx;
  float t3 = (miny - oy) * inv_dy;
  float t4 = (maxy - oy) * inv_dy;
  float t5 = (minz - oz) * inv_dz;
  float t6 = (maxz - oz) * inv_dz;
  float tmin = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
  float tmax = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
  *t_entry = tmin;
  *t_exit = tmax;
  if (tmax < 0 || tmin > tmax) return false;
  return true;
}

struct PixelResult {
  float rgb[3];
  int32_t hit_index;  // y*X*Z + x*Z + z of the hit voxel, -1 for sky
};

// One pixel of render_kernel / render_kernel_uint8 (src/kernels.cu:87-293,
// 300-470): camera basis, ray, AABB cull, DDA, face shade + fog.
PixelResult render_pixel(const int8_t* vox, int X, int Y, int Z, const float* cam, int px,
                         int py, int width, int height, int max_steps, float view_distance,
                         float half_fov) {
  PixelResult res;
  float cam_x = cam[0], cam_y = cam[1], cam_z = cam[2], yaw = cam[3], pitch = cam[4];
  float eye_x = cam_x;
  float eye_y = cam_y + 1.6f;
  float eye_z = cam_z;
  float sin_yaw, cos_yaw, sin_pitch, cos_pitch;
  mc_sincosf(yaw, &sin_yaw, &cos_yaw);
  mc_sincosf(pitch, &sin_pitch, &cos_pitch);
  float forward_x = sin_yaw * cos_pitch;
  float forward_y = sin_pitch;
  float forward_z = cos_yaw * cos_pitch;
  float right_x = cos_yaw;
  float right_z = -sin_yaw;
  float up_x = -sin_yaw * sin_pitch;
  float up_y = cos_pitch;
  float up_z = -cos_yaw * sin_pitch;
  float aspect = (float)width / (float)height;
  float u = (2.0f * (px + 0.5f) / width - 1.0f) * half_fov * aspect;
  float v = (1.0f - 2.0f * (py + 0.5f) / height) * half_fov;
  float ray_x = forward_x + u * right_x + v * up_x;
  float ray_y = forward_y + v * up_y;
  float ray_z = forward_z + u * right_z + v * up_z;
  float ray_len = sqrtf(ray_x * ray_x + ray_y * ray_y + ray_z * ray_z);
  ray_x /= ray_len;
  ray_y /= ray_len;
  ray_z /= ray_len;

  float t_entry, t_exit;
  bool hits_world = ray_intersects_aabb(eye_x, eye_y, eye_z, ray_x, ray_y, ray_z, 0.0f, 0.0f,
                                        0.0f, (float)X, (float)Y, (float)Z, &t_entry, &t_exit);
  res.hit_index = -1;
  if (!hits_world || t_entry > view_distance) {
    for (int c = 0; c < 3; ++c) res.rgb[c] = SKY_COLOR[c];
    return res;
  }
  int voxel_x = (int)floorf(eye_x);
  int voxel_y = (int)floorf(eye_y);
  int voxel_z = (int)floorf(eye_z);
  int step_x = (ray_x >= 0) ? 1 : -1;
  int step_y = (ray_y >= 0) ? 1 : -1;
  int step_z = (ray_z >= 0) ? 1 : -1;
  float inv_ray_x = (ray_x != 0.0f) ? 1.0f / ray_x : 1e30f;
  float inv_ray_y = (ray_y != 0.0f) ? 1.0f / ray_y : 1e30f;
  float inv_ray_z = (ray_z != 0.0f) ? 1.0f / ray_z : 1e30f;
  float t_delta_x = fabsf(inv_ray_x);
  float t_delta_y = fabsf(inv_ray_y);
  float t_delta_z = fabsf(inv_ray_z);
  float t_max_x = (ray_x >= 0) ? ((float)(voxel_x + 1) - eye_x) * inv_ray_x
                               : ((float)voxel_x - eye_x) * inv_ray_x;
  float t_max_y = (ray_y >= 0) ? ((float)(voxel_y + 1) - eye_y) * inv_ray_y
                               : ((float)voxel_y - eye_y) * inv_ray_y;
  float t_max_z = (ray_z >= 0) ? ((float)(voxel_z + 1) - eye_z) * inv_ray_z
                               : ((float)voxel_z - eye_z) * inv_ray_z;
  float t = 0.0f;
  int hit_face = -1;
  int8_t hit_block = AIR;
  // The reference unrolls this loop 4x with a goto exit (kernels.cu:413-442);
  // every unrolled body re-checks the same exit conditions first, so the plain
  // loop below visits exactly the same voxels.
  for (int step = 0;; ++step) {
    if (step >= max_steps || t >= view_distance) break;
    if (voxel_x >= 0 && voxel_x < X && voxel_y >= 0 && voxel_y < Y && voxel_z >= 0 &&
        voxel_z < Z) {
      int idx = voxel_y * X * Z + voxel_x * Z + voxel_z;
      int8_t block = vox[idx];
      if (block >= 0) {
        hit_block = block;
        res.hit_index = idx;
        break;
      }
    } else if (t > t_exit) {
      break;
    }
    if (t_max_x < t_max_y && t_max_x < t_max_z) {
      t = t_max_x; t_max_x += t_delta_x; voxel_x += step_x; hit_face = 0;
    } else if (t_max_y < t_max_z) {
      t = t_max_y; t_max_y += t_delta_y; voxel_y += step_y; hit_face = 1;
    } else {
      t = t_max_z; t_max_z += t_delta_z; voxel_z += step_z; hit_face = 2;
    }
  }
  // Ids outside the 13-entry table (never produced by the environment) show
  // the sky; the reference would index past BLOCK_COLORS.
  if (hit_block >= 0 && hit_block < 13) {
    float r = BLOCK_COLORS[hit_block][0];
    float g = BLOCK_COLORS[hit_block][1];
    float b = BLOCK_COLORS[hit_block][2];
    float shade = 1.0f;
    if (hit_face == 0) shade = 0.8f;
    else if (hit_face == 2) shade = 0.9f;
    else if (hit_face == 1 && step_y < 0) shade = 0.6f;
    float fog = fminf(t / view_distance, 1.0f);
#if ORC_MUTANT == 1  // planted slip for the pin tests: fog not squared
#elif ORC_MUTANT == 2  // planted slip: (1 - fog) and fog swapped
    fog = 1.0f - fog * fog;
#elif ORC_MUTANT == 3  // planted slip: face shade dropped
    fog = fog * fog;
    shade = 1.0f;
#else
    fog = fog * fog;
#endif
    res.rgb[0] = r * shade * (1.0f - fog) + SKY_COLOR[0] * fog;
    res.rgb[1] = g * shade * (1.0f - fog) + SKY_COLOR[1] * fog;
    res.rgb[2] = b * shade * (1.0f - fog) + SKY_COLOR[2] * fog;
  } else {
    if (hit_block < 0) res.hit_index = -1;
    for (int c = 0; c < 3; ++c) res.rgb[c] = SKY_COLOR[c];
  }
  return res;
}

// Flat uint8 colour table and sky (src/kernels.cu:479-495).
const uint8_t BLOCK_COLORS_U8[13][3] = {
    {76, 166, 51},   {140, 89, 51},   {127, 127, 127}, {140, 102, 64}, {51, 140, 38},
    {229, 217, 153}, {51, 102, 204},  {217, 229, 242}, {178, 89, 76},  {102, 102, 102},
    {191, 153, 102}, {242, 242, 250}, {38, 38, 38},
};
const uint8_t SKY_COLOR_U8[3] = {135, 206, 235};

// Camera basis of one camera (precompute_camera_basis_kernel,
// src/kernels.cu:511-561): [0-2] eye, [3-5] forward, [6-7] right x/z,
// [8-10] up, [11] half_fov * aspect, [12] half_fov; [13] is not written.
void camera_basis_row(const float* cam, int width, int height, float half_fov_flat,
                      float* basis) {
  float sin_yaw, cos_yaw, sin_pitch, cos_pitch;
  mc_sincosf(cam[3], &sin_yaw, &cos_yaw);
  mc_sincosf(cam[4], &sin_pitch, &cos_pitch);
  float aspect = (float)width / (float)height;
  basis[0] = cam[0];
  basis[1] = cam[1] + 1.6f;
  basis[2] = cam[2];
  basis[3] = sin_yaw * cos_pitch;
  basis[4] = sin_pitch;
  basis[5] = cos_yaw * cos_pitch;
  basis[6] = cos_yaw;
  basis[7] = -sin_yaw;
  basis[8] = -sin_yaw * sin_pitch;
  basis[9] = cos_pitch;
  basis[10] = -cos_yaw * sin_pitch;
  basis[11] = half_fov_flat * aspect;
  basis[12] = half_fov_flat;
}

// IEEE binary16 in software (gcc has no _Float16 on this host), round to
// nearest even, subnormals kept, overflow to infinity -- the semantics of
// CUDA's __float2half / __hadd / __hlt used by render_kernel_uint8_fp16
// (src/kernels.cu:1190-1225).  A binary16 sum is exact in binary64 (at most
// 40 exponent steps + 11 significant bits), so add = one rounding of the
// exact sum.
uint16_t f16_from_double(double x) {
  const uint16_t sign = std::signbit(x) ? 0x8000 : 0;
  const double ax = fabs(x);
  if (std::isnan(x)) return 0x7e00;
  if (ax >= 65520.0) return sign | 0x7c00;          // rounds past 65504: infinity
  if (ax < 6.103515625e-05) {                       // below 2^-14: subnormal (or zero)
    const double q = nearbyint(ax * 16777216.0);    // units of 2^-24, ties to even
    return sign | (uint16_t)q;                      // q == 1024 is the smallest normal
  }
  int e;
  const double m = frexp(ax, &e);                   // ax = m 2^e, m in [0.5, 1)
  double q = nearbyint(m * 2048.0);                 // 11 significant bits
  if (q == 2048.0) { q = 1024.0; ++e; }
  const int biased = e - 1 + 15;
  if (biased >= 31) return sign | 0x7c00;
  return sign | (uint16_t)((biased << 10) | ((int)q - 1024));
}

double f16_to_double(uint16_t h) {
  const int e = (h >> 10) & 0x1f, f = h & 0x3ff;
  double v;
  if (e == 0) v = ldexp((double)f, -24);
  else if (e == 31) v = f ? NAN : INFINITY;
  else v = ldexp((double)(f | 0x400), e - 25);
  return (h & 0x8000) ? -v : v;
}

struct Half {
  uint16_t bits;
  static Half from_float(float f) { return Half{f16_from_double((double)f)}; }
  double value() const { return f16_to_double(bits); }
  Half operator+(Half o) const { return Half{f16_from_double(value() + o.value())}; }
  bool operator<(Half o) const { return value() < o.value(); }
};

// One pixel of the flat-colour kernels (render_kernel_uint8_minimal,
// src/kernels.cu:985-1119; render_kernel_uint8_prebasis, 568-700): ray from
// the basis and image-plane coordinates (u, v), normalised with the
// reciprocal square root (restated as 1 / sqrtf, correctly rounded), DDA from
// the eye voxel for at most max_steps lookups, out-of-bounds voxels skipped,
// no view-distance limit.  Returns the block hit (-1: none); *hit_index gets
// its in-world index or -1.
struct F32 {
  float x;
  static F32 from_float(float f) { return F32{f}; }
  F32 operator+(F32 o) const { return F32{x + o.x}; }
  bool operator<(F32 o) const { return x < o.x; }
};

// T = F32: the binary32 kernels; T = Half: render_kernel_uint8_fp16
// (src/kernels.cu:1127-1262), whose 1/0 stand-in is 65504 and whose t_max /
// t_delta are binary16.
template <class T>
int flat_pixel(const int8_t* vox, int X, int Y, int Z, const float* bs, float u, float v,
               int max_steps, int32_t* hit_index) {
  const float kBig = sizeof(T) == 2 ? 65504.0f : 1e30f;
  float eye_x = bs[0], eye_y = bs[1], eye_z = bs[2];
  float ray_x = bs[3] + u * bs[6] + v * bs[8];
  float ray_y = bs[4] + v * bs[9];
  float ray_z = bs[5] + u * bs[7] + v * bs[10];
  float ray_len = 1.0f / sqrtf(ray_x * ray_x + ray_y * ray_y + ray_z * ray_z);
  ray_x *= ray_len;
  ray_y *= ray_len;
  ray_z *= ray_len;
  int voxel_x = (int)floorf(eye_x);
  int voxel_y = (int)floorf(eye_y);
  int voxel_z = (int)floorf(eye_z);
  int step_x = (ray_x >= 0) ? 1 : -1;
  int step_y = (ray_y >= 0) ? 1 : -1;
  int step_z = (ray_z >= 0) ? 1 : -1;
  float inv_ray_x = (ray_x != 0.0f) ? 1.0f / ray_x : kBig;
  float inv_ray_y = (ray_y != 0.0f) ? 1.0f / ray_y : kBig;
  float inv_ray_z = (ray_z != 0.0f) ? 1.0f / ray_z : kBig;
  T t_delta_x = T::from_float(fabsf(inv_ray_x));
  T t_delta_y = T::from_float(fabsf(inv_ray_y));
  T t_delta_z = T::from_float(fabsf(inv_ray_z));
  T t_max_x = T::from_float((ray_x >= 0) ? ((float)(voxel_x + 1) - eye_x) * inv_ray_x
                                         : ((float)voxel_x - eye_x) * inv_ray_x);
  T t_max_y = T::from_float((ray_y >= 0) ? ((float)(voxel_y + 1) - eye_y) * inv_ray_y
                                         : ((float)voxel_y - eye_y) * inv_ray_y);
  T t_max_z = T::from_float((ray_z >= 0) ? ((float)(voxel_z + 1) - eye_z) * inv_ray_z
                                         : ((float)voxel_z - eye_z) * inv_ray_z);
  *hit_index = -1;
  // The reference unrolls by 4 with a limit check in front of every body
  // (kernels.cu:1088-1110); this loop visits the same voxels.
  for (int step = 0; step < max_steps; ++step) {
    if (voxel_x >= 0 && voxel_x < X && voxel_y >= 0 && voxel_y < Y && voxel_z >= 0 &&
        voxel_z < Z) {
      int idx = voxel_y * X * Z + voxel_x * Z + voxel_z;
      int8_t block = vox[idx];
      if (block >= 0) {
        *hit_index = idx;
        return block;
      }
    }
    if (t_max_x < t_max_y && t_max_x < t_max_z) {
      t_max_x = t_max_x + t_delta_x; voxel_x += step_x;
    } else if (t_max_y < t_max_z) {
      t_max_y = t_max_y + t_delta_y; voxel_y += step_y;
    } else {
      t_max_z = t_max_z + t_delta_z; voxel_z += step_z;
    }
  }
  return -1;
}

template <class F>
void parallel_for(int n, int nthreads, F&& f) {
  if (nthreads <= 1 || n <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  nthreads = std::min(nthreads, n);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([&, t] {
      int lo = (int)((int64_t)n * t / nthreads), hi = (int)((int64_t)n * (t + 1) / nthreads);
      for (int i = lo; i < hi; ++i) f(i);
    });
  }
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// ---- numerics --------------------------------------------------------------
void orc_sincosf(const float* x, float* s, float* c, int n) {
  for (int i = 0; i < n; ++i) mc_sincosf(x[i], &s[i], &c[i]);
}

float orc_half_fov_tan(float fov_degrees) { return mc_half_fov_tan(fov_degrees); }

// ---- world generation ------------------------------------------------------
// generate_world_kernel (src/kernels.cu:1793-1828): layers by y; seed unused.
void orc_generate_world(int8_t* voxels, int B, int X, int Y, int Z, int ground_height,
                        unsigned int seed) {
  (void)seed;
  const int64_t wxyz = (int64_t)X * Y * Z;
  for (int64_t idx = 0; idx < (int64_t)B * wxyz; ++idx) {
    int voxel_idx = (int)(idx % wxyz);
    int y = voxel_idx / (X * Z);
    int8_t block = AIR;
    if (y == 0) block = BEDROCK;
    else if (y < ground_height - 1) block = DIRT;
    else if (y == ground_height - 1) block = GRASS;
    voxels[idx] = block;
  }
}

// place_tree_kernel (src/kernels.cu:1831-1865).  The reference bounds-checks y
// only; a tree outside the x/z range writes outside the env's slab there.  This
// restatement skips such voxels (callers keep the tree inside the world).
void orc_place_tree(int8_t* voxels, int B, int X, int Y, int Z, int tree_x, int tree_z,
                    int ground_height) {
  const int64_t wxyz = (int64_t)X * Y * Z;
  const bool xz_ok = tree_x >= 0 && tree_x < X && tree_z >= 0 && tree_z < Z;
  for (int b = 0; b < B; ++b) {
    if (!xz_ok) continue;
    for (int dy = 0; dy < 3; ++dy) {
      int y = ground_height + dy;
      if (y < Y) voxels[b * wxyz + (int64_t)y * X * Z + tree_x * Z + tree_z] = OAKLOG;
    }
    int leaf_y = ground_height + 3;
    if (leaf_y < Y) voxels[b * wxyz + (int64_t)leaf_y * X * Z + tree_x * Z + tree_z] = LEAVES;
  }
}

// place_house_kernel (src/kernels.cu:1868-1946): 5x4 footprint, floor, walls
// with a door, corner logs, flat brick roof; every write bounds-checked.
void orc_place_house(int8_t* voxels, int B, int X, int Y, int Z, int house_x, int house_z,
                     int ground_height) {
  const int64_t wxyz = (int64_t)X * Y * Z;
  const int W = 5, D = 4, H = 4;
  for (int b = 0; b < B; ++b) {
    auto set = [&](int dy, int dx, int dz, int8_t block) {
      int vx = house_x + dx, vy = ground_height + dy, vz = house_z + dz;
      if (vx >= 0 && vx < X && vy >= 0 && vy < Y && vz >= 0 && vz < Z)
        voxels[b * wxyz + (int64_t)vy * X * Z + vx * Z + vz] = block;
    };
    for (int dx = 0; dx < W; ++dx)
      for (int dz = 0; dz < D; ++dz) set(0, dx, dz, COBBLESTONE);
    for (int dy = 1; dy <= 3; ++dy) {
      for (int dx = 0; dx < W; ++dx) {
        if (dx == 2 && dy <= 2) continue;
        set(dy, dx, 0, PLANKS);
      }
      for (int dx = 0; dx < W; ++dx) set(dy, dx, D - 1, PLANKS);
      for (int dz = 1; dz < D - 1; ++dz) set(dy, 0, dz, PLANKS);
      for (int dz = 1; dz < D - 1; ++dz) set(dy, W - 1, dz, PLANKS);
    }
    for (int dy = 1; dy <= 3; ++dy) {
      set(dy, 0, 0, OAKLOG);
      set(dy, W - 1, 0, OAKLOG);
      set(dy, 0, D - 1, OAKLOG);
      set(dy, W - 1, D - 1, OAKLOG);
    }
    for (int dx = 0; dx < W; ++dx)
      for (int dz = 0; dz < D; ++dz) set(H, dx, dz, BRICK);
  }
}

// ---- agent -----------------------------------------------------------------
// reset_kernel (src/kernels.cu:1760-1786).
void orc_reset(float* positions, float* velocities, float* yaws, float* pitches,
               uint8_t* on_ground, const uint8_t* do_reset, int B, float spawn_x, float spawn_y,
               float spawn_z) {
  for (int i = 0; i < B; ++i) {
    if (!do_reset[i]) continue;
    positions[i * 3 + 0] = spawn_x;
    positions[i * 3 + 1] = spawn_y;
    positions[i * 3 + 2] = spawn_z;
    velocities[i * 3 + 0] = 0.0f;
    velocities[i * 3 + 1] = 0.0f;
    velocities[i * 3 + 2] = 0.0f;
    yaws[i] = 0.0f;
    pitches[i] = 0.0f;
    on_ground[i] = 0;
  }
}

// Discrete action ids -> multi-hot buttons + look deltas.  The reference's
// current MineEnv.step takes buttons/look only (minecu/__init__.py:304-331); the
// 11-action table it documents (docs/environment.md:3-17) is mapped here:
// 0 no-op, 1 forward, 2 back, 3 strafe left, 4 strafe right, 5 jump, 6 break,
// 7/8 yaw +/-look_delta (+ = look left), 9/10 pitch +/-look_delta (+ = up).
// Ids outside [0, 11) act as no-op.  look_delta is a project parameter: the
// reference never defines its magnitude.
void orc_discrete_to_buttons(const int64_t* actions, int B, float look_delta, int8_t* buttons,
                             float* look) {
  for (int i = 0; i < B; ++i) {
    int64_t a = actions[i];
    int8_t* bt = buttons + i * 8;
    memset(bt, 0, 8);
    float dyaw = 0.0f, dpitch = 0.0f;
    switch (a) {
      case 1: bt[0] = 1; break;
      case 2: bt[1] = 1; break;
      case 3: bt[2] = 1; break;
      case 4: bt[3] = 1; break;
      case 5: bt[4] = 1; break;
      case 6: bt[5] = 1; break;
      case 7: dyaw = look_delta; break;
      case 8: dyaw = -look_delta; break;
      case 9: dpitch = look_delta; break;
      case 10: dpitch = -look_delta; break;
      default: break;
    }
    look[i * 2 + 0] = dyaw;
    look[i * 2 + 1] = dpitch;
  }
}

// decode_actions_kernel (src/kernels.cu:1459-1504).
void orc_decode_actions(const int8_t* buttons, const float* look, float* forward_in,
                        float* strafe_in, float* delta_yaw_in, float* delta_pitch_in,
                        uint8_t* jump_in, uint8_t* do_break, uint8_t* do_place,
                        float* speed_mult, int B, float sprint_multiplier) {
  for (int i = 0; i < B; ++i) {
    int8_t fwd = buttons[i * 8 + 0], back = buttons[i * 8 + 1];
    int8_t left = buttons[i * 8 + 2], right = buttons[i * 8 + 3];
    int8_t jump = buttons[i * 8 + 4], brk = buttons[i * 8 + 5];
    int8_t place = buttons[i * 8 + 6], sprint = buttons[i * 8 + 7];
    forward_in[i] = (float)fwd - (float)back;
    strafe_in[i] = (float)right - (float)left;
    delta_yaw_in[i] = look[i * 2 + 0];
    delta_pitch_in[i] = look[i * 2 + 1];
    jump_in[i] = (jump != 0);
    do_break[i] = (brk != 0);
    do_place[i] = (place != 0);
    speed_mult[i] = (fwd && sprint) ? sprint_multiplier : 1.0f;
  }
}

// physics_kernel (src/kernels.cu:1511-1584).
void orc_physics(float* positions, float* velocities, float* yaws, float* pitches,
                 uint8_t* on_ground, const float* forward_in, const float* strafe_in,
                 const float* delta_yaw_in, const float* delta_pitch_in, const uint8_t* jump_in,
                 const float* speed_mult, int B, int world_x, int world_y, int world_z, float dt,
                 float gravity, float walk_speed, float jump_vel) {
  for (int i = 0; i < B; ++i) {
    yaws[i] += delta_yaw_in[i];
    float new_pitch = pitches[i] + delta_pitch_in[i];
    pitches[i] = fminf(fmaxf(new_pitch, -1.56f), 1.56f);
    float sin_yaw, cos_yaw;
    mc_sincosf(yaws[i], &sin_yaw, &cos_yaw);
    float fwd = forward_in[i];
    float str = strafe_in[i];
    float speed = walk_speed * speed_mult[i];
    float move_x = sin_yaw * fwd * speed * dt + cos_yaw * str * walk_speed * dt;
    float move_z = cos_yaw * fwd * speed * dt - sin_yaw * str * walk_speed * dt;
    float vel_y = velocities[i * 3 + 1];
    if (jump_in[i] && on_ground[i]) vel_y = jump_vel;
    vel_y += gravity * dt;
    float new_x = positions[i * 3 + 0] + move_x;
    float new_y = positions[i * 3 + 1] + vel_y * dt;
    float new_z = positions[i * 3 + 2] + move_z;
    float ground_y = (float)(world_y / 4);
    if (new_y <= ground_y) {
      new_y = ground_y;
      vel_y = 0.0f;
      on_ground[i] = 1;
    } else {
      on_ground[i] = 0;
    }
    float margin = 1.0f;
    new_x = fminf(fmaxf(new_x, margin), (float)world_x - margin);
    new_z = fminf(fmaxf(new_z, margin), (float)world_z - margin);
    positions[i * 3 + 0] = new_x;
    positions[i * 3 + 1] = new_y;
    positions[i * 3 + 2] = new_z;
    velocities[i * 3 + 1] = vel_y;
  }
}

// Voxel collision (the library's MineEnv(collision=True); the reference's
// only ground is the plane y = world_y / 4, src/kernels.cu:1566).  Same
// decode and integration as orc_physics, then, on env i's world as it is
// before this step's break/place:
//   clamp x, z to [1, dim - 1];
//   a move into a column with a solid block at layer floor(y) or floor(y) + 1
//   (y: the position before the step) is undone in x and z together;
//   rising with a solid block at layer floor(new_y + 1.8) keeps y, vy = 0;
//   the ground is 1 + the highest solid layer <= floor(y) of the column (0
//   when none); new_y <= ground lands on it.  Solid: a block id >= 0 inside
//   the world.
void orc_physics_collide(float* positions, float* velocities, float* yaws, float* pitches,
                         uint8_t* on_ground, const float* forward_in, const float* strafe_in,
                         const float* delta_yaw_in, const float* delta_pitch_in,
                         const uint8_t* jump_in, const float* speed_mult, const int8_t* voxels,
                         int B, int world_x, int world_y, int world_z, float dt, float gravity,
                         float walk_speed, float jump_vel) {
  const int64_t wxyz = (int64_t)world_x * world_y * world_z;
  for (int i = 0; i < B; ++i) {
    const int8_t* w = voxels + i * wxyz;
    auto is_solid = [&](int x, int y, int z) {
      if (x < 0 || x >= world_x || y < 0 || y >= world_y || z < 0 || z >= world_z) return false;
      return w[((int64_t)y * world_x + x) * world_z + z] >= 0;
    };
    yaws[i] += delta_yaw_in[i];
    float new_pitch = pitches[i] + delta_pitch_in[i];
    pitches[i] = fminf(fmaxf(new_pitch, -1.56f), 1.56f);
    float sin_yaw, cos_yaw;
    mc_sincosf(yaws[i], &sin_yaw, &cos_yaw);
    const float fwd = forward_in[i], str = strafe_in[i];
    const float speed = walk_speed * speed_mult[i];
    const float move_x = sin_yaw * fwd * speed * dt + cos_yaw * str * walk_speed * dt;
    const float move_z = cos_yaw * fwd * speed * dt - sin_yaw * str * walk_speed * dt;
    float vel_y = velocities[i * 3 + 1];
    if (jump_in[i] && on_ground[i]) vel_y = jump_vel;
    vel_y += gravity * dt;
    const float old_x = positions[i * 3 + 0], old_y = positions[i * 3 + 1],
                old_z = positions[i * 3 + 2];
    float new_x = fminf(fmaxf(old_x + move_x, 1.0f), (float)world_x - 1.0f);
    float new_y = old_y + vel_y * dt;
    float new_z = fminf(fmaxf(old_z + move_z, 1.0f), (float)world_z - 1.0f);
    const int feet = (int)std::floor(old_y);
    if (is_solid((int)std::floor(new_x), feet, (int)std::floor(new_z)) ||
        is_solid((int)std::floor(new_x), feet + 1, (int)std::floor(new_z))) {
      new_x = old_x;
      new_z = old_z;
    }
    const int col_x = (int)std::floor(new_x), col_z = (int)std::floor(new_z);
    if (vel_y > 0.0f && is_solid(col_x, (int)std::floor(new_y + 1.8f), col_z)) {
      new_y = old_y;
      vel_y = 0.0f;
    }
    int ground = 0;
    for (int y = std::min(feet, world_y - 1); y >= 0 && ground == 0; --y)
      if (is_solid(col_x, y, col_z)) ground = y + 1;
    if (new_y <= (float)ground) {
      new_y = (float)ground;
      vel_y = 0.0f;
      on_ground[i] = 1;
    } else {
      on_ground[i] = 0;
    }
    positions[i * 3 + 0] = new_x;
    positions[i * 3 + 1] = new_y;
    positions[i * 3 + 2] = new_z;
    velocities[i * 3 + 1] = vel_y;
  }
}

// raycast_interact_kernel (src/kernels.cu:1591-1680): 32 fixed samples along
// the view ray; first solid in-bounds voxel is broken (reward if it matches the
// target) or, for place, DIRT goes into the previous in-bounds air sample.
void orc_raycast_interact(int8_t* voxels, const float* positions, const float* yaws,
                          const float* pitches, const uint8_t* do_break,
                          const uint8_t* do_place, float* rewards, int B, int world_x,
                          int world_y, int world_z, int target_block, float reward_value) {
  const int wx = world_x, wy = world_y, wz = world_z;
  const int wxz = wx * wz;
  const int64_t wxyz = (int64_t)wx * wy * wz;
  for (int i = 0; i < B; ++i) {
    rewards[i] = 0.0f;
    bool wants_break = do_break[i], wants_place = do_place[i];
    if (!wants_break && !wants_place) continue;
    float sin_yaw, cos_yaw, sin_pitch, cos_pitch;
    mc_sincosf(yaws[i], &sin_yaw, &cos_yaw);
    mc_sincosf(pitches[i], &sin_pitch, &cos_pitch);
    float dir_x = sin_yaw * cos_pitch;
    float dir_y = sin_pitch;
    float dir_z = cos_yaw * cos_pitch;
    float eye_x = positions[i * 3 + 0];
    float eye_y = positions[i * 3 + 1] + 1.6f;
    float eye_z = positions[i * 3 + 2];
    float reach = 4.0f;
    int prev_vx = -1, prev_vy = -1, prev_vz = -1;
    for (int step = 0; step < 32; ++step) {
      float t = 0.1f + step * (reach / 32.0f);
      float px = eye_x + dir_x * t;
      float py = eye_y + dir_y * t;
      float pz = eye_z + dir_z * t;
      int vx = (int)floorf(px), vy = (int)floorf(py), vz = (int)floorf(pz);
      if (vx >= 0 && vx < wx && vy >= 0 && vy < wy && vz >= 0 && vz < wz) {
        int64_t idx = i * wxyz + vy * wxz + vx * wz + vz;
        int8_t block = voxels[idx];
        if (block >= 0) {
          if (wants_break) {
            if (target_block < 0 || block == target_block) rewards[i] = reward_value;
            voxels[idx] = AIR;
          } else if (wants_place && prev_vx >= 0) {
            voxels[i * wxyz + prev_vy * wxz + prev_vx * wz + prev_vz] = DIRT;
          }
          break;
        }
        prev_vx = vx; prev_vy = vy; prev_vz = vz;
      }
    }
  }
}

// update_cameras_kernel (src/kernels.cu:2320-2335).
void orc_update_cameras(float* cameras, const float* positions, const float* yaws,
                        const float* pitches, int B) {
  for (int i = 0; i < B; ++i) {
    cameras[i * 5 + 0] = positions[i * 3 + 0];
    cameras[i * 5 + 1] = positions[i * 3 + 1];
    cameras[i * 5 + 2] = positions[i * 3 + 2];
    cameras[i * 5 + 3] = yaws[i];
    cameras[i * 5 + 4] = pitches[i];
  }
}

// render_kernel_uint8 (out_u8 != 0; src/kernels.cu:300-470) or render_kernel
// (float32; src/kernels.cu:87-293).  hit_index (optional, [B,H,W] int32)
// receives the hit voxel's in-world index or -1.
void orc_render(const int8_t* voxels, const float* cameras, void* out, int out_u8, int B, int X,
                int Y, int Z, int width, int height, int max_steps, float view_distance,
                float fov_degrees, int32_t* hit_index, int nthreads) {
  const float half_fov = mc_half_fov_tan(fov_degrees);
  const int64_t wxyz = (int64_t)X * Y * Z;
  const int64_t npix = (int64_t)width * height;
  parallel_for(B, nthreads, [&](int b) {
    const int8_t* vox = voxels + b * wxyz;
    const float* cam = cameras + b * 5;
    for (int py = 0; py < height; ++py) {
      for (int px = 0; px < width; ++px) {
        PixelResult r = render_pixel(vox, X, Y, Z, cam, px, py, width, height, max_steps,
                                     view_distance, half_fov);
        int64_t o = (b * npix + (int64_t)py * width + px) * 3;
        if (out_u8) {
          uint8_t* ou = (uint8_t*)out;
          for (int c = 0; c < 3; ++c) ou[o + c] = (uint8_t)(r.rgb[c] * 255.0f);
        } else {
          float* of = (float*)out;
          for (int c = 0; c < 3; ++c) of[o + c] = r.rgb[c];
        }
        if (hit_index) hit_index[b * npix + (int64_t)py * width + px] = r.hit_index;
      }
    }
  });
}

float orc_half_fov_tan_flat(float fov_degrees) { return mc_half_fov_tan_flat(fov_degrees); }

// precompute_camera_basis_kernel (src/kernels.cu:511-561) over B cameras.
void orc_camera_basis(const float* cameras, float* basis, int B, int width, int height,
                      float fov_degrees) {
  const float hf = mc_half_fov_tan_flat(fov_degrees);
  for (int b = 0; b < B; ++b) camera_basis_row(cameras + b * 5, width, height, hf, basis + b * 14);
}

// Flat-colour render.  mode 0: render_kernel_uint8_minimal from cameras [B,5]
// (src/kernels.cu:985-1119, u = (a * half_fov) * aspect); mode 1:
// render_kernel_uint8_prebasis from basis [B,14] (src/kernels.cu:568-700,
// u = a * scale_u); the smem kernel (kernels.cu:835-982) computes the same
// image as mode 1; mode 2: render_kernel_uint8_fp16 from cameras (mode 0's
// image plane, binary16 DDA, src/kernels.cu:1127-1262).  Ids >= 13 show the
// sky (the reference would index past BLOCK_COLORS_U8).
void orc_render_flat(const int8_t* voxels, const float* cams, uint8_t* out, int mode, int B,
                     int X, int Y, int Z, int width, int height, int max_steps,
                     float fov_degrees, int32_t* hit_index, int nthreads) {
  const float hf = mc_half_fov_tan_flat(fov_degrees);
  const float aspect = (float)width / (float)height;
  const int64_t wxyz = (int64_t)X * Y * Z;
  const int64_t npix = (int64_t)width * height;
  parallel_for(B, nthreads, [&](int b) {
    const int8_t* vox = voxels + b * wxyz;
    float bs[14];
    if (mode != 1) camera_basis_row(cams + b * 5, width, height, hf, bs);
    else for (int k = 0; k < 13; ++k) bs[k] = cams[b * 14 + k];
    for (int py = 0; py < height; ++py) {
      for (int px = 0; px < width; ++px) {
        float a = 2.0f * (px + 0.5f) / width - 1.0f;
        float c = 1.0f - 2.0f * (py + 0.5f) / height;
        float u = (mode != 1) ? a * hf * aspect : a * bs[11];
        float v = (mode != 1) ? c * hf : c * bs[12];
        int32_t hi;
        int blk = (mode == 2) ? flat_pixel<Half>(vox, X, Y, Z, bs, u, v, max_steps, &hi)
                              : flat_pixel<F32>(vox, X, Y, Z, bs, u, v, max_steps, &hi);
        const uint8_t* col = (blk >= 0 && blk < 13) ? BLOCK_COLORS_U8[blk] : SKY_COLOR_U8;
        int64_t o = (b * npix + (int64_t)py * width + px) * 3;
        for (int k = 0; k < 3; ++k) out[o + k] = col[k];
        if (hit_index) hit_index[b * npix + (int64_t)py * width + px] = hi;
      }
    }
  });
}

// binary16 conversion, sum and order of n float pairs (the software binary16
// used by mode 2 of orc_render_flat).
void orc_f16(const float* a, const float* b, uint16_t* ha, uint16_t* sum, uint8_t* lt, int n) {
  for (int i = 0; i < n; ++i) {
    const Half x = Half::from_float(a[i]), y = Half::from_float(b[i]);
    ha[i] = x.bits;
    sum[i] = (x + y).bits;
    lt[i] = (x < y) ? 1 : 0;
  }
}

// Largest sample length a caller may pass (documentation aid for ctypes users).
int orc_version() { return 1; }

// ---- seeded terrain ---------------------------------------------------------
// No reference algorithm: the reference's generator ignores its seed
// (src/kernels.cu:1800-1828) and lists noise terrain as a TODO (README.md:71).
// This restates the library's MC_TERRAIN_NOISE definition from its statement:
//   corner(b, i, j) = F(seed ^ F(b * 0x9e3779b9 ^ F(i * 0x632be5ab + j))) mod (2A + 1),
//   F = MurmurHash3's fmix32 (shifts 16, 13, 16; multipliers 0x85ebca6b, 0xc2b2ae35);
//   height(x, z) = ground - A + floor((bilinear weight sum over 16 + 8) / 16);
// then generate_world's layers per column, the tree and the house on the
// surface of their own corner column.  Built here as: the corner lattice of
// each env first, then its height map, then columns, then the oracle's own
// place_tree / place_house on that env's slab.
uint32_t orc_fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// One column (x, z), any integers: lattice cell by floor division.
int orc_terrain_column(int b, int x, int z, int ground_height, int amp, unsigned int seed) {
  auto fdiv4 = [](int a) { return (a - (((a % 4) + 4) % 4)) / 4; };
  const int i = fdiv4(x), j = fdiv4(z), u = x - 4 * i, w = z - 4 * j;
  auto corner = [&](int ci, int cj) {
    const uint32_t inner = orc_fmix32((uint32_t)ci * 0x632be5abu + (uint32_t)cj);
    return (int)(orc_fmix32(seed ^ orc_fmix32((uint32_t)b * 0x9e3779b9u ^ inner)) %
                 (uint32_t)(2 * amp + 1));
  };
  const int sum = corner(i, j) * (4 - u) * (4 - w) + corner(i + 1, j) * u * (4 - w) +
                  corner(i, j + 1) * (4 - u) * w + corner(i + 1, j + 1) * u * w;
  return ground_height - amp + (sum + 8) / 16;
}

void orc_terrain_heights(int* heights, int b, int X, int Z, int ground_height, int amp,
                         unsigned int seed) {
  const int CX = X / 4 + 2, CZ = Z / 4 + 2;
  std::vector<int> corner((size_t)CX * CZ);
  for (int i = 0; i < CX; ++i)
    for (int j = 0; j < CZ; ++j) {
      const uint32_t inner = orc_fmix32((uint32_t)i * 0x632be5abu + (uint32_t)j);
      const uint32_t mid = orc_fmix32((uint32_t)b * 0x9e3779b9u ^ inner);
      corner[(size_t)i * CZ + j] = (int)(orc_fmix32(seed ^ mid) % (uint32_t)(2 * amp + 1));
    }
  for (int x = 0; x < X; ++x)
    for (int z = 0; z < Z; ++z) {
      const int i = x / 4, j = z / 4, u = x % 4, w = z % 4;
      const int c00 = corner[(size_t)i * CZ + j], c10 = corner[(size_t)(i + 1) * CZ + j];
      const int c01 = corner[(size_t)i * CZ + j + 1], c11 = corner[(size_t)(i + 1) * CZ + j + 1];
      const int sum = c00 * (4 - u) * (4 - w) + c10 * u * (4 - w) + c01 * (4 - u) * w + c11 * u * w;
      heights[x * Z + z] = ground_height - amp + (sum + 8) / 16;
    }
}

// ids: the env index of each of the B worlds (the terrain is seeded by it), or
// null for 0 .. B-1.
void orc_terrain_world(int8_t* voxels, int B, int X, int Y, int Z, int ground_height, int amp,
                       unsigned int seed, int tree_x, int tree_z, int house_x, int house_z,
                       const int* ids) {
  const int64_t wxyz = (int64_t)X * Y * Z;
  std::vector<int> h((size_t)X * Z);
  for (int k = 0; k < B; ++k) {
    int8_t* v = voxels + k * wxyz;
    const int b = ids ? ids[k] : k;
    orc_terrain_heights(h.data(), b, X, Z, ground_height, amp, seed);
    for (int y = 0; y < Y; ++y)
      for (int x = 0; x < X; ++x)
        for (int z = 0; z < Z; ++z) {
          const int g = h[x * Z + z];
          int8_t block = AIR;
          if (y == 0) block = BEDROCK;
          else if (y < g - 1) block = DIRT;
          else if (y == g - 1) block = GRASS;
          v[(int64_t)y * X * Z + x * Z + z] = block;
        }
    // The tree and house corner columns may lie outside the world (small
    // worlds): their surface follows the same definition, evaluated alone.
    auto at = [&](int x, int z) {
      if (x >= 0 && x < X && z >= 0 && z < Z) return h[x * Z + z];
      return orc_terrain_column(b, x, z, ground_height, amp, seed);
    };
    orc_place_tree(v, 1, X, Y, Z, tree_x, tree_z, at(tree_x, tree_z));
    orc_place_house(v, 1, X, Y, Z, house_x, house_z, at(house_x, house_z));
  }
}

}  // extern "C"
