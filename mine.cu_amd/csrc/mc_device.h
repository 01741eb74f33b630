// mc_device.h - gfx950 device code of the batched voxel-environment step.
//
// One workgroup owns one environment (its 64-lane waves share the env's world
// staged in LDS).  Wave 0 runs the per-env logic -- episode check, action
// decode, physics, the 32-sample break/place raycast (one sample per lane,
// resolved with a wave ballot), masked respawn -- and publishes the camera
// frame; then every lane traces rays through the LDS copy of the world and
// writes the observation with wide coalesced stores.  World regeneration, when
// any env of the group is done, is fused after the render.
//
// Arithmetic follows the reference operation by operation (src/kernels.cu,
// cited per function) with FP contraction off and IEEE division/sqrt, so the
// results are bit-identical to the scalar host restatement.
#pragma once
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "mc_math.h"

namespace mc {

constexpr int kWave = 64;

// Global-address-space views: pointers loaded from memory (the batched
// kernel's per-group descriptors) would otherwise be accessed as flat.
#define MC_GLOBAL __attribute__((address_space(1)))
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef MC_GLOBAL v4u guint4;
typedef MC_GLOBAL uint32_t guint32;
typedef MC_GLOBAL v4f gfloat4;
// LDS of one workgroup: gfx950 lets a workgroup allocate up to 160 KiB.
// kStaticLdsReserve bounds the kernels' static LDS (palette, frames,
// work-item bookkeeping; the global-world kernel's tables); the rest holds
// the items' LDS (item_lds_bytes).  The host
// also caps the world bytes by the device's reported per-workgroup maximum.
constexpr int kLdsPerWorkgroup = 160 * 1024;
constexpr int kStaticLdsReserve = 8192;
constexpr int kMaxLdsWorld = kLdsPerWorkgroup - kStaticLdsReserve;

// Block ids (src/kernels.cu:11-24).
constexpr int8_t kAir = -1, kGrass = 0, kDirt = 1, kOakLog = 3, kLeaves = 4, kBrick = 8,
                 kCobble = 9, kPlanks = 10, kBedrock = 12;

// Shaded palette, 13 blocks x RGB then the sky (src/kernels.cu:27-44).  Copied
// into LDS at kernel start (one constant-memory load per thread) so the
// per-pixel lookup is a ds_read.
constexpr int kPaletteLen = 42;
constexpr int kSkyBase = 39;
static __device__ __constant__ float kPaletteTab[kPaletteLen] = {
    0.30f, 0.65f, 0.20f,  // grass
    0.55f, 0.35f, 0.20f,  // dirt
    0.50f, 0.50f, 0.50f,  // stone
    0.55f, 0.40f, 0.25f,  // oak log
    0.20f, 0.55f, 0.15f,  // leaves
    0.90f, 0.85f, 0.60f,  // sand
    0.20f, 0.40f, 0.80f,  // water
    0.85f, 0.90f, 0.95f,  // glass
    0.70f, 0.35f, 0.30f,  // brick
    0.40f, 0.40f, 0.40f,  // cobblestone
    0.75f, 0.60f, 0.40f,  // planks
    0.95f, 0.95f, 0.98f,  // snow
    0.15f, 0.15f, 0.15f,  // bedrock
    0.53f, 0.81f, 0.92f,  // sky
};

// Flat uint8 colours, 13 blocks then the sky (src/kernels.cu:479-495), packed
// 0x00BBGGRR.
static __device__ __constant__ uint32_t kFlatTab[14] = {
    76u | (166u << 8) | (51u << 16),     // grass
    140u | (89u << 8) | (51u << 16),     // dirt
    127u | (127u << 8) | (127u << 16),   // stone
    140u | (102u << 8) | (64u << 16),    // oak log
    51u | (140u << 8) | (38u << 16),     // leaves
    229u | (217u << 8) | (153u << 16),   // sand
    51u | (102u << 8) | (204u << 16),    // water
    217u | (229u << 8) | (242u << 16),   // glass
    178u | (89u << 8) | (76u << 16),     // brick
    102u | (102u << 8) | (102u << 16),   // cobblestone
    191u | (153u << 8) | (102u << 16),   // planks
    242u | (242u << 8) | (250u << 16),   // snow
    38u | (38u << 8) | (38u << 16),      // bedrock
    135u | (206u << 8) | (235u << 16),   // sky
};
constexpr int kFlatLen = 14;
constexpr int kFlatSky = 13;

// Everything one launch needs; passed by value as the kernel argument.
struct Params {
  int B, X, Y, Z, W, H;
  int max_steps;
  float view_distance, half_fov, aspect;
  float dt, gravity, walk_speed, sprint_mult, jump_vel;
  int target_block;
  float reward_value;
  float spawn_x, spawn_y, spawn_z;
  int ground_height, tree_x, tree_z, house_x, house_z;
  int episode_length;          // < 0: none
  float look_delta;
  int action_kind;             // 0 buttons+look, 1 discrete int64
  const void* actions;
  const float* look;
  int8_t* voxels;
  float* positions;
  float* velocities;
  float* yaws;
  float* pitches;
  uint8_t* on_ground;
  int32_t* step_count;
  uint8_t* do_reset;
  uint8_t* dones;
  float* cameras;
  void* obs;
  float* rewards;
  const uint8_t* mask;         // reset: NULL = all
  // Episode flag ring (4 device words, NULL: no episode logic).  A step
  // launch reads ring[slot] ("some env is done this step", which decides the
  // regeneration of every world), sets ring[(slot+1)%3] when some env will be
  // done on the next step, and clears ring[(slot+2)%3] for the launch after.
  // ring[3] is the reset path's "regenerate" flag.
  uint32_t* ring;
  int slot;
  // Graph-replayable launches read the slot from device memory instead
  // (NULL: use `slot`); the host advances it with k_slot_advance after each
  // step launch, so a captured step replays with the right slot.
  const uint32_t* slot_dev;
  // done_mode 0: done = step_count >= episode_length, with the MineEnv
  // bookkeeping (step_count, do_reset, dones, regeneration).  done_mode 1:
  // done = do_reset[b] as given, nothing else touched (launch_step_uint8).
  int done_mode;
  // Optional decoded-control outputs ([B] each; NULL = not written), the
  // reference's decode buffers (src/kernels.cu:1459-1504).
  float* forward_in;
  float* strafe_in;
  float* delta_yaw_in;
  float* delta_pitch_in;
  uint8_t* jump_in;
  uint8_t* do_break;
  uint8_t* do_place;
  float* speed_mult_in;
  // Flat-colour variants: tan of the half field of view with the flat
  // kernels' degree conversion, and that times the aspect (prebasis scale_u).
  float flat_hf, flat_su;
  // view_distance is a power of two: 1 / view_distance is exact and t * inv_vd
  // equals t / view_distance for every t.
  int vd_pow2;
  float inv_vd;
  // The world's diagonal is well inside view_distance: an in-world voxel's
  // entry distance t (at most the diagonal, plus rounding) never reaches it.
  int vd_far;
  // [B, 14] camera basis (src/kernels.cu:503-561; NULL: none).  Step items
  // write it in the prebasis modes; render-basis items read it.
  float* basis;
  // Layers y < ys of each world are staged in LDS (ys = Y: all).  With ys < Y
  // a step item first checks that its env's layers y >= ys are all air and
  // that the step cannot reach them (eyes and the raycast stay below ys);
  // rays that climb to ys then end as sky, which is the reference's result.
  // An env that fails the check is left untouched and listed in
  // fb[slot * (B + 2)]: [count, finished workgroups, env...] for a
  // full-height pass (k_env_fb), which empties the list when it ends.
  int ys;
  uint32_t* fb;
  // Seeded terrain (the k_*_terrain kernels only): column heights
  // ground_height + d, d in [-terrain_amp, terrain_amp], from terrain_seed.
  int terrain;                 // host side: 1 launches the k_*_terrain kernels
  int collision;               // host side: 1 launches the k_*_col step kernels
  uint32_t terrain_seed;
  int terrain_amp;
  // Ray-step counters of the counting step kernel (k_env<..., CNT = true>, a
  // measurement build only): [0] DDA voxel lookups of every pixel (a pixel
  // whose eye sits inside a solid counts its one lookup), [1] rays marched,
  // [2] pixels.  Unused by every other kernel.
  unsigned long long* ray_steps;
};

// Camera frame of one env, published by wave 0 through LDS.
struct Frame {
  float ex, ey, ez;            // eye
  float fx, fy, fz;            // forward
  float rx, rz;                // right (right_y == 0)
  float ux, uy, uz;            // up
  int fill_block;              // >= 0: eye inside this solid block, every ray hits it at t=0
  // Extent of the solid voxels at heights >= floor(eye y): x in [hx0, hx1],
  // z in [hz0, hz1], y <= hy1 (empty: hx0 > hx1, hy1 = INT_MIN).  Valid when
  // `lift` is set; used to end upward rays early (see trace()).
  int hx0, hx1, hz0, hz1, hy1;
  int lift;
  // inside: the eye voxel (ivx, ivy, ivz) lies in the world and every world
  // dimension is at most kMaxInsideDim; then every ray starts in the world
  // and trace() takes its in-world loop.  d*p / d*n: distance from the eye
  // to the voxel's upper / lower face per axis, (float)(iv + 1) - e and
  // (float)iv - e.
  int inside;
  int ivx, ivy, ivz;
  float dxp, dxn, dyp, dyn, dzp, dzn;
  // Flat-colour render: >= 0 when the eye voxel is in the world and solid
  // (every ray stops there on its first lookup); image-plane scales of a
  // render from a stored camera basis.
  int flat_fill;
  float su, sv;
  // Starting packed position of the in-world DDA per direction octant (see
  // OctPos), written by lanes 0-7 of each of the env's waves before the render.
  uint32_t pos0[8];
  // The env left this item untouched for the full-height pass (Params::fb).
  int deferred;
};

// Largest world dimension of the in-world DDA loop (10-bit packed position).
constexpr int kMaxInsideDim = 512;

// Frame fields are uniform over the waves of one env: read them into scalar
// registers.
__device__ __forceinline__ Frame uniform_frame(const Frame* src) {
  Frame f;
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(&f);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(Frame) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
  return f;
}

// ---------------------------------------------------------------------------
// World regeneration, closed form of generate_world_kernel + place_tree_kernel
// + place_house_kernel applied in that order (src/kernels.cu:1793-1946,
// minecu/__init__.py:258-268).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int8_t layer_block(int y, int gh) {
  return (y == 0) ? kBedrock : (y < gh - 1) ? kDirt : (y == gh - 1) ? kGrass : kAir;
}

// Value the house writes at offset (dy, dx, dz) of its 5x5x4 box (dy, dx, dz),
// or kNoWrite where it writes nothing (interior, door).
constexpr int8_t kNoWrite = -2;
__device__ __forceinline__ int8_t house_block(int dy, int dx, int dz) {
  if (dy == 0) return kCobble;                         // floor
  if (dy == 4) return kBrick;                          // roof
  const bool xe = (dx == 0 || dx == 4), ze = (dz == 0 || dz == 3);
  if (xe && ze) return kOakLog;                        // corner pillars (written last)
  if (dz == 0) return (dx == 2 && dy <= 2) ? kNoWrite : kPlanks;   // front wall, door
  if (dz == 3 || xe) return kPlanks;                   // back / side walls
  return kNoWrite;                                     // interior
}

__device__ __forceinline__ int8_t fresh_voxel(const Params& p, int y, int x, int z) {
  int8_t v = layer_block(y, p.ground_height);
  if (x == p.tree_x && z == p.tree_z) {
    const int dy = y - p.ground_height;
    if (dy >= 0 && dy < 3) v = kOakLog;
    else if (dy == 3) v = kLeaves;
  }
  const int hx = x - p.house_x, hy = y - p.ground_height, hz = z - p.house_z;
  if (hx >= 0 && hx < 5 && hz >= 0 && hz < 4 && hy >= 0 && hy <= 4) {
    const int8_t h = house_block(hy, hx, hz);
    if (h != kNoWrite) v = h;
  }
  return v;
}

// ---------------------------------------------------------------------------
// Seeded terrain (MineEnv(terrain="noise")).  No counterpart in the reference,
// whose world generator takes a seed and ignores it (src/kernels.cu:1800-1828).
// Env b's column (x, z) is solid up to height ground_height + d:
//   corners every 4 columns carry c = mix(seed, b, x / 4, z / 4) mod (2A + 1);
//   d = round(bilinear mean of the 4 corners around (x, z)) - A, in exact
//   integers: s = sum of corner * (4 - fx or fx) * (4 - fz or fz) over 16,
//   d = (s + 8) / 16 - A, with fx = x % 4, fz = z % 4.
// mix is MurmurHash3's 32-bit finaliser (fmix32, Austin Appleby, public
// domain) chained over the inputs.  The tree and the house stand on the
// surface of their own (x, z) corner column, with the flat world's shapes;
// house cells it does not write (interior, door) keep the terrain.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ int terrain_corner(const Params& p, int b, int cx, int cz) {
  const uint32_t h = fmix32(p.terrain_seed ^ fmix32((uint32_t)b * 0x9e3779b9u ^
                                                    fmix32((uint32_t)cx * 0x632be5abu +
                                                           (uint32_t)cz)));
  return (int)(h % (uint32_t)(2 * p.terrain_amp + 1));
}

// Solid height of env b's column (x, z) (ground_height in the flat world).
__device__ __forceinline__ int terrain_height(const Params& p, int b, int x, int z) {
  const int cx = x >> 2, cz = z >> 2, fx = x & 3, fz = z & 3;
  const int s = (terrain_corner(p, b, cx, cz) * (4 - fx) + terrain_corner(p, b, cx + 1, cz) * fx) *
                    (4 - fz) +
                (terrain_corner(p, b, cx, cz + 1) * (4 - fx) +
                 terrain_corner(p, b, cx + 1, cz + 1) * fx) * fz;
  return p.ground_height - p.terrain_amp + ((s + 8) >> 4);
}

__device__ __forceinline__ int8_t terrain_voxel(const Params& p, int b, int y, int x, int z,
                                                int tree_h, int house_h) {
  int8_t v = layer_block(y, terrain_height(p, b, x, z));
  if (x == p.tree_x && z == p.tree_z) {
    const int dy = y - tree_h;
    if (dy >= 0 && dy < 3) v = kOakLog;
    else if (dy == 3) v = kLeaves;
  }
  const int hx = x - p.house_x, hy = y - house_h, hz = z - p.house_z;
  if (hx >= 0 && hx < 5 && hz >= 0 && hz < 4 && hy >= 0 && hy <= 4) {
    const int8_t h = house_block(hy, hx, hz);
    if (h != kNoWrite) v = h;
  }
  return v;
}

__device__ __forceinline__ uint32_t pack4(int8_t a, int8_t b, int8_t c, int8_t d) {
  return (uint32_t)(uint8_t)a | ((uint32_t)(uint8_t)b << 8) | ((uint32_t)(uint8_t)c << 16) |
         ((uint32_t)(uint8_t)d << 24);
}

// 16 consecutive voxels starting at linear index i0 (i0 % 16 == 0), packed
// little-endian into a uint4.  Every index below is a compile-time constant
// after unrolling, so the chunk stays in registers.
__device__ __forceinline__ v4u fresh_chunk(const Params& p, int i0) {
  const int XZ = p.X * p.Z;
  uint32_t w[4] = {0u, 0u, 0u, 0u};
  if ((p.Z & 15) == 0) {
    // The chunk lies in one (y, x) row: uniform layer, then tree/house fix-ups.
    const int y = i0 / XZ, rem = i0 - y * XZ, x = rem / p.Z, z0 = rem - x * p.Z;
    const int8_t base = layer_block(y, p.ground_height);
    const int dy = y - p.ground_height;
    const bool tree_row = (x == p.tree_x) && dy >= 0 && dy <= 3;
    const int8_t tree_v = (dy < 3) ? kOakLog : kLeaves;
    const int hx = x - p.house_x;
    const bool house_row = hx >= 0 && hx < 5 && dy >= 0 && dy <= 4;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int z = z0 + k;
      int8_t v = base;
      if (tree_row && z == p.tree_z) v = tree_v;
      const int hz = z - p.house_z;
      if (house_row && hz >= 0 && hz < 4) {
        const int8_t h = house_block(dy, hx, hz);
        if (h != kNoWrite) v = h;
      }
      w[k >> 2] |= (uint32_t)(uint8_t)v << (8 * (k & 3));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int i = i0 + k;
      const int y = i / XZ, rem = i - y * XZ, x = rem / p.Z, z = rem - x * p.Z;
      w[k >> 2] |= (uint32_t)(uint8_t)fresh_voxel(p, y, x, z) << (8 * (k & 3));
    }
  }
  return v4u{w[0], w[1], w[2], w[3]};
}

// Write env b's fresh world to HBM (and to LDS when lds != nullptr).  TER:
// the seeded terrain world.
template <bool TER = false>
__device__ __forceinline__ void store_fresh_world(const Params& p, int b, int8_t* gworld,
                                                  int8_t* lds, int nvox, int tid, int nt) {
  if (TER) {
    const int XZ = p.X * p.Z;
    const int tree_h = terrain_height(p, b, p.tree_x, p.tree_z);
    const int house_h = terrain_height(p, b, p.house_x, p.house_z);
    if ((nvox & 3) == 0) {
      // 4 voxels per 32-bit store.
      typedef MC_GLOBAL uint32_t gu32;
      gu32* g = (gu32*)gworld;
      uint32_t* l = reinterpret_cast<uint32_t*>(lds);
      for (int c = tid; c < (nvox >> 2); c += nt) {
        uint32_t w = 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int i = (c << 2) + k;
          const int y = i / XZ, rem = i - y * XZ, x = rem / p.Z, z = rem - x * p.Z;
          w |= (uint32_t)(uint8_t)terrain_voxel(p, b, y, x, z, tree_h, house_h) << (8 * k);
        }
        g[c] = w;
        if (lds) l[c] = w;
      }
    } else {
      for (int i = tid; i < nvox; i += nt) {
        const int y = i / XZ, rem = i - y * XZ, x = rem / p.Z, z = rem - x * p.Z;
        const int8_t v = terrain_voxel(p, b, y, x, z, tree_h, house_h);
        gworld[i] = v;
        if (lds) lds[i] = v;
      }
    }
    return;
  }
  if ((p.Z & 15) == 0) {
    // 16-voxel chunks along z: chunk c is (y, x, zq) with c = (y X + x) zc + zq.
    // A row without tree or house voxels is its layer's block 16 times; the
    // chunk coordinates advance by the thread stride with carries.
    guint4* g = (guint4*)gworld;
    v4u* l = reinterpret_cast<v4u*>(lds);
    const int zc = p.Z >> 4, X = p.X, n16 = nvox >> 4;
    int zq = tid % zc, yx = tid / zc;
    int x = yx % X, y = yx / X;
    const int dzq = nt % zc, dyx = nt / zc, dx = dyx % X, dy = dyx / X;
    for (int c = tid; c < n16; c += nt) {
      const int hy = y - p.ground_height, hx = x - p.house_x;
      const bool special = ((x == p.tree_x) & (hy >= 0) & (hy <= 3)) |
                           ((hx >= 0) & (hx < 5) & (hy >= 0) & (hy <= 4));
      v4u v;
      if (special) {
        v = fresh_chunk(p, c << 4);
      } else {
        const uint32_t w = (uint32_t)(uint8_t)layer_block(y, p.ground_height) * 0x01010101u;
        v = v4u{w, w, w, w};
      }
      g[c] = v;
      if (lds) l[c] = v;
      zq += dzq;
      x += dx;
      y += dy;
      if (zq >= zc) { zq -= zc; ++x; }
      if (x >= X) { x -= X; ++y; }
    }
  } else if ((nvox & 15) == 0) {
    guint4* g = (guint4*)gworld;
    v4u* l = reinterpret_cast<v4u*>(lds);
    for (int c = tid; c < (nvox >> 4); c += nt) {
      const v4u v = fresh_chunk(p, c << 4);
      g[c] = v;
      if (lds) l[c] = v;
    }
  } else {
    const int XZ = p.X * p.Z;
    for (int i = tid; i < nvox; i += nt) {
      const int y = i / XZ, rem = i - y * XZ, x = rem / p.Z, z = rem - x * p.Z;
      const int8_t v = fresh_voxel(p, y, x, z);
      gworld[i] = v;
      if (lds) lds[i] = v;
    }
  }
}

// HBM -> LDS copy of one env's world.  Aligned worlds go by LDS-direct loads
// (global_load_lds_dwordx4): one wave instruction moves 64 consecutive
// 16-byte chunks from HBM into LDS, lane i's chunk at the instruction's LDS
// base + 16 i, with no vector register holding the data.  tid and nt count
// whole waves (tid's wave starts at tid - lane).
__device__ __forceinline__ void stage_world(int8_t* lds, const int8_t* gworld, int nvox, int tid,
                                            int nt) {
  typedef __attribute__((address_space(3))) void lds_void;
  if (((nvox | (int)(uintptr_t)gworld | (int)(uintptr_t)lds) & 15) == 0) {
    const int n16 = nvox >> 4;
    const int lane = tid & (kWave - 1);
    for (int c0 = __builtin_amdgcn_readfirstlane(tid - lane); c0 < n16; c0 += nt) {
      if (c0 + lane < n16)
        __builtin_amdgcn_global_load_lds((const void*)(gworld + (c0 + lane) * 16),
                                         (lds_void*)(lds + c0 * 16), 16, 0, 0);
    }
  } else {
    for (int i = tid; i < nvox; i += nt) lds[i] = gworld[i];
  }
}

// Whether any of the n voxels at g (this thread's share) is not air.  Four
// 16-byte loads per lane are in flight before the first test (out-of-range
// slots reload the last chunk, in bounds, and are not tested), so at 32^3
// the scan adds no round trip to the staging loads'.
__device__ __forceinline__ bool any_not_air(const int8_t* g, int n, int tid, int nt) {
  bool found = false;
  if ((((uintptr_t)g | (uintptr_t)n) & 15) == 0) {
    const guint4* g4 = (const guint4*)g;
    const int n16 = n >> 4;
    for (int c0 = tid; c0 < n16; c0 += 4 * nt) {
      v4u w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) w[k] = g4[min(c0 + k * nt, n16 - 1)];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        found |= (c0 + k * nt < n16) & ((w[k].x & w[k].y & w[k].z & w[k].w) != 0xffffffffu);
    }
  } else {
    for (int i = tid; i < n; i += nt) found |= g[i] != kAir;
  }
  return found;
}

// ---------------------------------------------------------------------------
// Camera frame (src/kernels.cu:326-345 eye + basis), plus the inside-solid
// shortcut: with the eye strictly inside the world box, every ray passes the
// AABB test with t_entry < 0, and with max_steps >= 1 and view_distance > 0 the
// first DDA iteration (t = 0, face -1) hits the eye voxel when it is solid; the
// shaded colour is then col * 1 * (1 - 0) + sky * 0 == col for every pixel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void finish_frame(const Params& p, Frame& f, const int8_t* world) {
  f.fill_block = -1;
  f.hx0 = f.hz0 = INT_MAX;
  f.hx1 = f.hz1 = f.hy1 = INT_MIN;
  f.lift = 0;
  if ((p.max_steps >= 1) & (p.view_distance > 0.0f) & (f.ex > 0.0f) & (f.ex < (float)p.X) &
      (f.ey > 0.0f) & (f.ey < (float)p.Y) & (f.ez > 0.0f) & (f.ez < (float)p.Z)) {
    const int vx = (int)floorf(f.ex), vy = (int)floorf(f.ey), vz = (int)floorf(f.ez);
    const int8_t blk = world[vy * p.X * p.Z + vx * p.Z + vz];
    if (blk >= 0 && blk < 13) f.fill_block = blk;
  }
  // floor(e) lies in [0, n) exactly when 0 <= e < n (false for NaN).
  f.inside = (f.ex >= 0.0f) & (f.ex < (float)p.X) & (f.ey >= 0.0f) & (f.ey < (float)p.Y) &
             (f.ez >= 0.0f) & (f.ez < (float)p.Z) & (p.X <= kMaxInsideDim) &
             (p.Y <= kMaxInsideDim) & (p.Z <= kMaxInsideDim);
  f.ivx = f.inside ? (int)floorf(f.ex) : 0;
  f.ivy = f.inside ? (int)floorf(f.ey) : 0;
  f.ivz = f.inside ? (int)floorf(f.ez) : 0;
  f.dxp = (float)(f.ivx + 1) - f.ex;
  f.dxn = (float)f.ivx - f.ex;
  f.dyp = (float)(f.ivy + 1) - f.ey;
  f.dyn = (float)f.ivy - f.ey;
  f.dzp = (float)(f.ivz + 1) - f.ez;
  f.dzn = (float)f.ivz - f.ez;
  // The flat render looks the eye voxel up first, with no slab test and no
  // view-distance limit (src/kernels.cu:1077-1083).
  f.flat_fill = -1;
  if (f.inside && p.max_steps >= 1) {
    const int8_t blk = world[f.ivy * p.X * p.Z + f.ivx * p.Z + f.ivz];
    if (blk >= 0) f.flat_fill = blk;
  }
  // The solid-extent scan reads whole 16-byte rows of z; an eye inside a
  // solid needs no ray at all.
  f.lift = (f.fill_block < 0) && (f.flat_fill < 0) && ((p.Z & 15) == 0);
  f.deferred = 0;
}

__device__ __forceinline__ Frame make_frame(const Params& p, float x, float y, float z,
                                            float sin_yaw, float cos_yaw, float sin_pitch,
                                            float cos_pitch, const int8_t* world) {
  Frame f;
  f.ex = x;
  f.ey = y + 1.6f;
  f.ez = z;
  f.fx = sin_yaw * cos_pitch;
  f.fy = sin_pitch;
  f.fz = cos_yaw * cos_pitch;
  f.rx = cos_yaw;
  f.rz = -sin_yaw;
  f.ux = -sin_yaw * sin_pitch;
  f.uy = cos_pitch;
  f.uz = -cos_yaw * sin_pitch;
  f.su = p.flat_su;
  f.sv = p.flat_hf;
  finish_frame(p, f, world);
  return f;
}

// Frame of env b read back from its stored camera basis (the prebasis render
// input, src/kernels.cu:593-621): eye, forward, right (x, z), up, scale_u,
// scale_v.
__device__ __forceinline__ Frame frame_from_basis(const Params& p, int b, const int8_t* world) {
  const float* bs = p.basis + (int64_t)b * 14;
  Frame f;
  f.ex = bs[0]; f.ey = bs[1]; f.ez = bs[2];
  f.fx = bs[3]; f.fy = bs[4]; f.fz = bs[5];
  f.rx = bs[6]; f.rz = bs[7];
  f.ux = bs[8]; f.uy = bs[9]; f.uz = bs[10];
  f.su = bs[11]; f.sv = bs[12];
  finish_frame(p, f, world);
  return f;
}

// Camera basis row of env b (src/kernels.cu:539-560); element 13 is padding
// and is not written.
__device__ __forceinline__ void store_basis(const Params& p, int b, const Frame& f) {
  float* bs = p.basis + (int64_t)b * 14;
  bs[0] = f.ex; bs[1] = f.ey; bs[2] = f.ez;
  bs[3] = f.fx; bs[4] = f.fy; bs[5] = f.fz;
  bs[6] = f.rx; bs[7] = f.rz;
  bs[8] = f.ux; bs[9] = f.uy; bs[10] = f.uz;
  bs[11] = p.flat_su; bs[12] = p.flat_hf;
}

// Wave-wide minima of xmin and zmin and maxima of xmax, zmax and ymax, by DPP
// lane moves (no LDS round trip): pairs and quads by quad permutes, half rows
// and rows by mirrors (every lane has a source), then rows 0-1 into row 1 and
// rows 0-3 into row 3 by row broadcasts; the results are read from lane 63.
// The broadcasts leave rows 0 and 2 undefined, which lane 63 never reads.  The
// five reductions advance together, so no move reads a register written by
// the instruction just before it.  Every lane must be active.
__device__ __forceinline__ void wave_extent(int& xmin, int& xmax, int& zmin, int& zmax,
                                            int& ymax) {
#define MC_EXT_STEP(CTRL, ROWS)                                                   \
  xmin = min(xmin, __builtin_amdgcn_mov_dpp(xmin, CTRL, ROWS, 0xf, false));       \
  zmin = min(zmin, __builtin_amdgcn_mov_dpp(zmin, CTRL, ROWS, 0xf, false));       \
  xmax = max(xmax, __builtin_amdgcn_mov_dpp(xmax, CTRL, ROWS, 0xf, false));       \
  zmax = max(zmax, __builtin_amdgcn_mov_dpp(zmax, CTRL, ROWS, 0xf, false));       \
  ymax = max(ymax, __builtin_amdgcn_mov_dpp(ymax, CTRL, ROWS, 0xf, false));
  MC_EXT_STEP(0xb1, 0xf)   // quad_perm 1,0,3,2
  MC_EXT_STEP(0x4e, 0xf)   // quad_perm 2,3,0,1
  MC_EXT_STEP(0x141, 0xf)  // row_half_mirror
  MC_EXT_STEP(0x140, 0xf)  // row_mirror
  MC_EXT_STEP(0x142, 0xa)  // row_bcast:15 into rows 1 and 3
  MC_EXT_STEP(0x143, 0xc)  // row_bcast:31 into rows 2 and 3
#undef MC_EXT_STEP
  xmin = __builtin_amdgcn_readlane(xmin, 63);
  zmin = __builtin_amdgcn_readlane(zmin, 63);
  xmax = __builtin_amdgcn_readlane(xmax, 63);
  zmax = __builtin_amdgcn_readlane(zmax, 63);
  ymax = __builtin_amdgcn_readlane(ymax, 63);
}

// Extent of the solid voxels at heights >= y0 (16-byte LDS chunks; Z % 16 ==
// 0), reduced over the env's threads into f (LDS) with LDS atomics.  Chunk c
// = (y X + x) zc + zq; a lane's first chunk and its stride are split into
// (y, x, zq) by shifts when X and zc are powers of two, else by division.
__device__ __forceinline__ void solid_extent(const Params& p, const int8_t* world, Frame* f,
                                             int y0, int tid, int nt) {
  const int X = p.X, Y = p.ys, zc = p.Z >> 4;
  int xmin = INT_MAX, xmax = INT_MIN, zmin = INT_MAX, zmax = INT_MIN, ymax = INT_MIN;
  if (y0 < 0) y0 = 0;
  if (y0 < Y) {
    const v4u* w4 = reinterpret_cast<const v4u*>(world);
    const int n16 = Y * X * zc;
    const int c_begin = y0 * X * zc + tid;
    int zq, x, y, dzq, dx, dy;
    if (((X & (X - 1)) | (zc & (zc - 1))) == 0) {
      const int lz = __builtin_ctz(zc), lx = __builtin_ctz(X);
      zq = c_begin & (zc - 1);
      x = (c_begin >> lz) & (X - 1);
      y = c_begin >> (lz + lx);
      dzq = nt & (zc - 1);
      dx = (nt >> lz) & (X - 1);
      dy = nt >> (lz + lx);
    } else {
      zq = c_begin % zc;
      const int yx = c_begin / zc;
      x = yx % X;
      y = yx / X;
      dzq = nt % zc;
      const int dyx = nt / zc;
      dx = dyx % X;
      dy = dyx / X;
    }
    for (int c = c_begin; c < n16; c += nt) {
      const v4u w = w4[c];
      const uint32_t m0 = ~w.x & 0x80808080u, m1 = ~w.y & 0x80808080u;
      const uint32_t m2 = ~w.z & 0x80808080u, m3 = ~w.w & 0x80808080u;
      if ((m0 | m1 | m2 | m3) != 0u) {
        const uint64_t lo = (uint64_t)m0 | ((uint64_t)m1 << 32);
        const uint64_t hi = (uint64_t)m2 | ((uint64_t)m3 << 32);
        const int fb = lo ? ((__ffsll((unsigned long long)lo) - 1) >> 3)
                          : 8 + ((__ffsll((unsigned long long)hi) - 1) >> 3);
        const int lb = hi ? 15 - (__clzll((long long)hi) >> 3) : 7 - (__clzll((long long)lo) >> 3);
        xmin = min(xmin, x);
        xmax = max(xmax, x);
        zmin = min(zmin, zq * 16 + fb);
        zmax = max(zmax, zq * 16 + lb);
        ymax = max(ymax, y);
      }
      zq += dzq;
      x += dx;
      y += dy;
      if (zq >= zc) { zq -= zc; ++x; }
      if (x >= X) { x -= X; ++y; }
    }
  }
  wave_extent(xmin, xmax, zmin, zmax, ymax);
  if ((tid & (kWave - 1)) == 0) {
    atomicMin(&f->hx0, xmin);
    atomicMax(&f->hx1, xmax);
    atomicMin(&f->hz0, zmin);
    atomicMax(&f->hz1, zmax);
    atomicMax(&f->hy1, ymax);
  }
}

// ---------------------------------------------------------------------------
// Correctly rounded sqrt, division and reciprocal for camera rays.
//
// The compiler's IEEE sequences (-fhip-fp32-correctly-rounded-divide-sqrt)
// wrap a core -- v_sqrt plus a one-ulp residual test; v_rcp, one Newton step
// and two FMA remainder corrections -- in operand scaling and special-case
// fix-ups (v_div_scale / v_div_fmas / v_div_fixup, v_cmp_class).  For normal
// operands whose quotient stays normal and whose exponents differ by less
// than 96 the scaling and fix-ups are identities, so the core alone returns
// the same bits.  ray_dir() checks that range per ray (len^2 in
// [2^-60, 2^60], every component 0 or at least 2^-60 in magnitude) and falls
// back to the full IEEE operations otherwise.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sqrt_core(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __int_as_float(__float_as_int(s) - 1);
  const float sp = __int_as_float(__float_as_int(s) + 1);
  const float rm = fmaf(-sm, s, x);
  const float rp = fmaf(-sp, s, x);
  float r = (rm <= 0.0f) ? sm : s;
  r = (rp > 0.0f) ? sp : r;
  return r;
}
// Reciprocal estimate of d refined by one Newton step (shared by every
// quotient with denominator d).
__device__ __forceinline__ float recip_core(float d) {
  const float r = __builtin_amdgcn_rcpf(d);
  const float e = fmaf(-d, r, 1.0f);
  return fmaf(e, r, r);
}
// Correctly rounded 1 / d: the Newton-refined estimate above, then one
// remainder correction (Markstein's reciprocal step, exact remainder by FMA).
// The same bits as 1.0f / d for every |d| in [2^-90, 2^90] (exhaustive GPU
// test).
__device__ __forceinline__ float rcp_cr(float d) {
  const float r = recip_core(d);
  const float e = fmaf(-d, r, 1.0f);
  return fmaf(e, r, r);
}
__device__ __forceinline__ float div_core(float n, float d, float r) {
  float q = n * r;
  float rem = fmaf(-d, q, n);
  q = fmaf(rem, r, q);
  rem = fmaf(-d, q, n);
  return fmaf(rem, r, q);
}

// Unit ray direction and per-axis reciprocals (big where a component is 0).
// RSQRT: the flat kernels' normalisation, a product with 1 / sqrt(len^2);
// otherwise the shaded kernels' division by sqrt(len^2).  A zero component
// may come out as -0 instead of +0 on the fast path; its only uses are
// comparisons with 0 and the reciprocal's zero test, where the two are equal.
template <bool RSQRT>
__device__ __forceinline__ void ray_dir(float& x, float& y, float& z, float& ix, float& iy,
                                        float& iz, float big) {
  const float len2 = x * x + y * y + z * z;
  // |c| is 0 or at least 2^-60 (0x21800000): its magnitude bits less one, as
  // an unsigned, are then at least 0x217fffff.  Bitwise ANDs, not a chain of
  // branches; a NaN or infinite component fails the len^2 range instead.
  auto ok = [](float c) { return (__float_as_uint(c) & 0x7fffffffu) - 1u >= 0x217fffffu; };
  const bool fast = (len2 >= 0x1p-60f) & (len2 <= 0x1p60f) & ok(x) & ok(y) & ok(z);
  if (fast) {
    const float len = sqrt_core(len2);
    const float rl = recip_core(len);
    if (RSQRT) {
      const float k = div_core(1.0f, len, rl);
      x *= k;
      y *= k;
      z *= k;
    } else {
      x = div_core(x, len, rl);
      y = div_core(y, len, rl);
      z = div_core(z, len, rl);
    }
    ix = (x != 0.0f) ? rcp_cr(x) : big;
    iy = (y != 0.0f) ? rcp_cr(y) : big;
    iz = (z != 0.0f) ? rcp_cr(z) : big;
  } else {
    const float len = sqrtf(len2);
    if (RSQRT) {
      const float k = 1.0f / len;
      x *= k;
      y *= k;
      z *= k;
    } else {
      x /= len;
      y /= len;
      z /= len;
    }
    ix = (x != 0.0f) ? 1.0f / x : big;
    iy = (y != 0.0f) ? 1.0f / y : big;
    iz = (z != 0.0f) ? 1.0f / z : big;
  }
}

// Starting packed position of the in-world DDA for each of the 8 direction
// octants (bit 0: ray_x >= 0, bit 1: ray_y >= 0, bit 2: ray_z >= 0).  Field
// per axis: the eye voxel counted in the direction of travel, biased by
// 512 - limit so that bit 9 of the field sets on reaching the limit; the limit
// is the world extent, or with the upward-ray exit (ray_y >= 0) the extent of
// the solids at or above eye height.  Uniform per env: 8 words in the env's
// LDS frame, read per ray by octant (kept out of scalar registers, which the
// camera frame and loop masks already fill).
struct OctPos {
  const uint32_t* v;  // Frame::pos0 in LDS
  __device__ __forceinline__ uint32_t pick(bool bx, bool by, bool bz) const {
    return v[(bx ? 1 : 0) | (by ? 2 : 0) | (bz ? 4 : 0)];
  }
  // The same word read again from LDS (not kept in a register meanwhile);
  // through an LDS-space pointer, so it stays a ds_read.
  __device__ __forceinline__ uint32_t reread(bool bx, bool by, bool bz) const {
    typedef __attribute__((address_space(3))) const volatile uint32_t lds_u32;
    return ((lds_u32*)v)[(bx ? 1 : 0) | (by ? 2 : 0) | (bz ? 4 : 0)];
  }
};

// The in-world DDA loop (the iterations of trace()'s in-world branch) for the
// lanes of `run`; the other active lanes keep their values.  Per iteration and
// lane: t = min(t_max); the axis is z when t_max_z == t, else y when
// t_max_y == t, else x (the reference's tie order, kernels.cu:424-441);
// t_max of that axis grows by |inv|, the LDS address va by that axis's byte
// step, the packed position pos by pinc (1, 1 << 10 or 1 << 20: also the
// axis, returned); then the voxel at va is read.  A lane stops when it leaves
// the world (a guard bit of pos sets), when the voxel is not air (signed
// value >= 0) or, without VDF, when t >= vd; every lane stops after
// max_steps iterations.  blk returns the last voxel read (unsigned byte).
//
// Written as one block so the loop control is a uniform counter and an exec
// update per iteration (the compiler's divergent-loop form spends about twice
// the scalar instructions), and the t_max adds run under per-axis exec masks
// (three adds in place of three adds and three selects: the step is bound by
// vector instruction issue); the arithmetic is the same IEEE binary32 adds and
// compares, so results are unchanged bit for bit.  Every operand of a
// vector instruction is a VGPR or an inline constant except at most one
// SGPR (a lane mask, the guard word or vd); a lane mask written by a vector
// compare is read by vector instructions no earlier than three instructions
// later.  exec is restored before the block ends.
template <bool VDF>
__device__ __forceinline__ void dda_loop(float& tx, float& ty, float& tz, float& t, float ix,
                                         float iy, float iz, uint32_t& va, uint32_t dx,
                                         uint32_t dy, uint32_t dz, uint32_t& pos, uint32_t& pinc,
                                         uint32_t& blk, uint64_t run, int max_steps, float vd) {
  uint64_t save, sa, mz, my, mc, mv;
  uint32_t d, w, step;
  const uint32_t k10 = 1u << 10, k20 = 1u << 20, kg = (1u << 9) | (1u << 19) | (1u << 29);
#define MC_DDA_ASM(VDCMP, VDAND)                                                   \
  asm volatile(                                                                    \
      "s_mov_b64 %[save], exec\n\t"                                               \
      "s_and_b64 exec, exec, %[run]\n\t"                                          \
      "s_mov_b32 %[step], 0\n\t"                                                  \
      "s_cbranch_execz .Ldda_end%=\n"                                              \
      ".Ldda_loop%=:\n\t"                                                         \
      "v_min3_f32 %[t], %[tx], %[ty], %[tz]\n\t"                                  \
      "s_add_u32 %[step], %[step], 1\n\t"                                         \
      "v_cmp_eq_f32_e64 %[mz], %[tz], %[t]\n\t"                                   \
      "v_cmp_eq_f32_e64 %[my], %[ty], %[t]\n\t" VDCMP                             \
      "s_mov_b64 %[sa], exec\n\t"                                                 \
      "s_andn2_b64 %[my], %[my], %[mz]\n\t"                                       \
      "s_or_b64 %[mc], %[my], %[mz]\n\t"                                          \
      "v_cndmask_b32_e64 %[d], %[dx], %[dy], %[my]\n\t"                           \
      "v_cndmask_b32_e64 %[d], %[d], %[dz], %[mz]\n\t"                            \
      "v_add_u32_e32 %[va], %[va], %[d]\n\t"                                      \
      "ds_read_u8 %[blk], %[va]\n\t"                                              \
      "v_cndmask_b32_e64 %[pi], 1, %[k10], %[my]\n\t"                             \
      "v_cndmask_b32_e64 %[pi], %[pi], %[k20], %[mz]\n\t"                         \
      "v_add_u32_e32 %[pos], %[pos], %[pi]\n\t"                                   \
      "s_andn2_b64 exec, %[sa], %[mc]\n\t"                                        \
      "v_add_f32_e64 %[tx], %[tx], |%[ix]|\n\t"                                   \
      "s_mov_b64 exec, %[my]\n\t"                                                 \
      "v_add_f32_e64 %[ty], %[ty], |%[iy]|\n\t"                                   \
      "s_mov_b64 exec, %[mz]\n\t"                                                 \
      "v_add_f32_e64 %[tz], %[tz], |%[iz]|\n\t"                                   \
      "s_mov_b64 exec, %[sa]\n\t"                                                 \
      "v_and_b32_e32 %[w], %[kg], %[pos]\n\t"                                     \
      "s_waitcnt lgkmcnt(0)\n\t"                                                  \
      "v_sub_u32_e32 %[w], %[blk], %[w]\n\t"                                      \
      "v_cmp_lt_i32_e32 vcc, 0x7f, %[w]\n\t"                                      \
      "s_and_b64 exec, exec, vcc\n\t" VDAND                                      \
      "s_cmp_ge_i32 %[step], %[ms]\n\t"                                           \
      "s_cbranch_scc1 .Ldda_end%=\n\t"                                            \
      "s_cbranch_execnz .Ldda_loop%=\n"                                            \
      ".Ldda_end%=:\n\t"                                                          \
      "s_mov_b64 exec, %[save]"                                                    \
      : [tx] "+v"(tx), [ty] "+v"(ty), [tz] "+v"(tz), [t] "+v"(t), [va] "+v"(va),     \
        [pos] "+v"(pos), [pi] "+v"(pinc), [blk] "+v"(blk), [d] "=&v"(d),             \
        [w] "=&v"(w), [save] "=&s"(save), [sa] "=&s"(sa), [mz] "=&s"(mz),           \
        [my] "=&s"(my), [mc] "=&s"(mc), [mv] "=&s"(mv), [step] "=&s"(step)          \
      : [ix] "v"(ix), [iy] "v"(iy), [iz] "v"(iz), [dx] "v"(dx), [dy] "v"(dy),        \
        [dz] "v"(dz), [k10] "v"(k10), [k20] "v"(k20), [kg] "s"(kg), [run] "s"(run),  \
        [ms] "s"(max_steps), [vd] "s"(vd)                                           \
      : "vcc", "scc")
  if (VDF)
    MC_DDA_ASM("", "");
  else
    MC_DDA_ASM("v_cmp_gt_f32_e64 %[mv], %[vd], %[t]\n\t", "s_and_b64 exec, exec, %[mv]\n\t");
#undef MC_DDA_ASM
}

template <bool LIFT>
__device__ __forceinline__ uint32_t oct_pos(const Params& p, const Frame& f, int k) {
  const int X = p.X, Y = p.Y, Z = p.Z;
  const bool bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
  int lim_x = X, lim_y = by ? p.ys : Y, lim_z = Z;
  if (LIFT && f.lift && by) {
    lim_y = max(0, min(p.ys, f.hy1 + 1));
    lim_x = max(0, bx ? min(X, f.hx1 + 1) : min(X, X - f.hx0));
    lim_z = max(0, bz ? min(Z, f.hz1 + 1) : min(Z, Z - f.hz0));
  }
  const int tx = bx ? f.ivx : X - 1 - f.ivx;
  const int ty = by ? f.ivy : Y - 1 - f.ivy;
  const int tz = bz ? f.ivz : Z - 1 - f.ivz;
  return (uint32_t)(tx + 512 - lim_x) | ((uint32_t)(ty + 512 - lim_y) << 10) |
         ((uint32_t)(tz + 512 - lim_z) << 20);
}

// ---------------------------------------------------------------------------
// One ray: AABB cull + DDA + face shade + fog (src/kernels.cu:51-79, 354-469).
// Writes the shaded colour in [0,1] to rgb.
//
// With the eye voxel in the world (f.inside) the slab test always passes
// (every slab interval contains t = 0), and once the DDA leaves the world
// box it never returns: each voxel coordinate moves one way only.  The
// reference's remaining iterations then see only out-of-bounds voxels and
// end without a hit, so leaving the world ends the ray as sky.  That loop
// tracks the position as three 10-bit fields of one word, each counting
// voxels in the direction of travel and biased so that its top bit sets on
// the step that leaves the world: one AND tests all three bounds.
//
// LIFT: for a ray with ray_y >= 0 the DDA's voxel y never decreases, so once
// the current voxel lies beyond every solid voxel at heights >= floor(eye y)
// (above hy1, or past [hx0, hx1] / [hz0, hz1] in the direction of travel),
// every voxel it can still visit is air or outside the world: the reference's
// result is sky.  Those bounds tighten the per-axis limits of the packed
// position, so the same test ends the ray there.
// ---------------------------------------------------------------------------
// CNT: *nlook receives the ray's DDA voxel lookups, the start voxel included
// (0 when the slab test culls the ray).  GLB: the world is the env's slab in
// global memory (worlds too large for LDS): the in-world walk never reads
// outside it (there is no margin) and runs as plain code.
template <bool LIFT, bool VDF, bool CNT = false, bool GLB = false>
__device__ __forceinline__ void trace(const Params& p, const Frame& f, const OctPos& op, float u,
                                      float v, const int8_t* __restrict__ world,
                                      const float* __restrict__ pal, float rgb[3],
                                      int* nlook = nullptr) {
  float ray_x = f.fx + u * f.rx + v * f.ux;
  float ray_y = f.fy + v * f.uy;
  float ray_z = f.fz + u * f.rz + v * f.uz;
  // Normalise by sqrt(len^2); one reciprocal per axis serves both the slab
  // test and the DDA: the reference computes the same expression twice
  // (kernels.cu:57-59, 395-397).
  float inv_x, inv_y, inv_z;
  ray_dir<false>(ray_x, ray_y, ray_z, inv_x, inv_y, inv_z, 1e30f);
  const float vd = p.view_distance;
  const int max_steps = p.max_steps;
  const int X = p.X, Y = p.Y, Z = p.Z, XZ = X * Z;
  const int step_x = (ray_x >= 0) ? 1 : -1;
  const int step_y = (ray_y >= 0) ? 1 : -1;
  const int step_z = (ray_z >= 0) ? 1 : -1;
  const float td_x = fabsf(inv_x), td_y = fabsf(inv_y), td_z = fabsf(inv_z);
  const int dix = (ray_x >= 0) ? Z : -Z, diy = (ray_y >= 0) ? XZ : -XZ, diz = step_z;
  int hit_block = -1;
  int hit_face = -1;
  float t = 0.0f;
  if (CNT) *nlook = 0;
  if (f.inside) {
    // dxp = dxn + 1 exactly: both differences are exact (Sterbenz), as
    // floor(e) <= e < floor(e) + 1.
    float tm_x = (f.dxn + ((ray_x >= 0) ? 1.0f : 0.0f)) * inv_x;
    float tm_y = (f.dyn + ((ray_y >= 0) ? 1.0f : 0.0f)) * inv_y;
    float tm_z = (f.dzn + ((ray_z >= 0) ? 1.0f : 0.0f)) * inv_z;
    constexpr uint32_t kGuard = (1u << 9) | (1u << 19) | (1u << 29);
    uint32_t pos = op.pick(step_x > 0, step_y > 0, step_z > 0);
    uint32_t pinc = 0u;  // the last step's increment of pos: its axis
    int blk;
    if (GLB) {
      // The reference's loop order, the lookup guarded: a position outside
      // the world reads the env's first voxel and discards it.
      const int8_t* vp = world + (f.ivy * XZ + f.ivx * Z + f.ivz);
      bool out = (pos & kGuard) != 0u;
      blk = out ? -1 : (int)*(out ? world : vp);
      bool lim = (max_steps <= 0) | (t >= vd);
      int step = 0;
      while (!(lim | out | (blk >= 0))) {
        t = fminf(fminf(tm_x, tm_y), tm_z);
        const bool cz = tm_z == t;
        const bool cy = !cz & (tm_y == t);
        const bool cx = !cz & !cy;
        tm_x = cx ? tm_x + td_x : tm_x;
        tm_y = cy ? tm_y + td_y : tm_y;
        tm_z = cz ? tm_z + td_z : tm_z;
        vp += cx ? dix : (cy ? diy : diz);
        pinc = cx ? 1u : (cy ? (1u << 10) : (1u << 20));
        pos += pinc;
        out = (pos & kGuard) != 0u;
        blk = out ? -1 : (int)*(out ? world : vp);
        ++step;
        lim = (step >= max_steps) | (!VDF & (t >= vd));
      }
    } else {
      // LDS address of the current voxel, as a 32-bit LDS offset.
      typedef __attribute__((address_space(3))) const uint8_t lds_u8;
      uint32_t va = (uint32_t)(uintptr_t)(lds_u8*)(world + (f.ivy * XZ + f.ivx * Z + f.ivz));
      uint32_t blk_u = (uint32_t)*(lds_u8*)(uintptr_t)va;
      // The reference's iteration is (limit check, lookup, advance); here the
      // first lookup precedes the loop, and each iteration advances, looks up
      // and tests (dda_loop); a lane runs while it is inside the world on air
      // and below the limits.
      const bool run = ((pos & kGuard) == 0u) & ((int8_t)blk_u < 0) & (max_steps > 0) & (t < vd);
      dda_loop<VDF>(tm_x, tm_y, tm_z, t, inv_x, inv_y, inv_z, va, (uint32_t)dix, (uint32_t)diy,
                    (uint32_t)diz, pos, pinc, blk_u, __builtin_amdgcn_ballot_w64(run), max_steps,
                    vd);
      blk = (int)(int8_t)blk_u;
    }
    // The walk's results are read back from vector registers (pos, blk, t,
    // pinc; the step count is the sum of pos's fields' advances, the start
    // word read again from LDS): per-lane booleans kept past the loop would
    // be merged into scalar lane masks on every iteration.
    const uint32_t moved = pos - op.reread(step_x > 0, step_y > 0, step_z > 0);
    const int steps = (int)((moved & 0x3ffu) + ((moved >> 10) & 0x3ffu) + (moved >> 20));
    if (CNT) *nlook = steps + 1;
    const bool lim_end = (steps >= max_steps) | (((steps == 0) | !VDF) & (t >= vd));
    asm volatile("" : "+v"(blk));  // a fresh test, not the loop's lane mask
    if (!lim_end & ((pos & kGuard) == 0u) & (blk >= 0)) {
      hit_block = blk;
      // Face from the last step's axis (x 0, y 1, z 2; none -1), by
      // independent selects rather than a nested chain.
      int face = (pinc == 1u) ? 0 : -1;
      face = (pinc == (1u << 10)) ? 1 : face;
      face = (pinc == (1u << 20)) ? 2 : face;
      hit_face = face;
    }
  } else {
    const float t1 = (0.0f - f.ex) * inv_x, t2 = ((float)X - f.ex) * inv_x;
    const float t3 = (0.0f - f.ey) * inv_y, t4 = ((float)Y - f.ey) * inv_y;
    const float t5 = (0.0f - f.ez) * inv_z, t6 = ((float)Z - f.ez) * inv_z;
    const float t_entry = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
    const float t_exit = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
    if (!(t_exit < 0 || t_entry > t_exit) && !(t_entry > vd)) {
      int vx = (int)floorf(f.ex), vy = (int)floorf(f.ey), vz = (int)floorf(f.ez);
      float tm_x = (ray_x >= 0) ? ((float)(vx + 1) - f.ex) * inv_x : ((float)vx - f.ex) * inv_x;
      float tm_y = (ray_y >= 0) ? ((float)(vy + 1) - f.ey) * inv_y : ((float)vy - f.ey) * inv_y;
      float tm_z = (ray_z >= 0) ? ((float)(vz + 1) - f.ez) * inv_z : ((float)vz - f.ez) * inv_z;
      // Linear voxel index kept in step with (vx, vy, vz); out-of-bounds
      // positions read voxel 0 (always valid) and discard it.
      int idx = vy * XZ + vx * Z + vz;
      bool inb = (unsigned)vx < (unsigned)X && (unsigned)vy < (unsigned)Y && (unsigned)vz < (unsigned)Z;
      int blk = inb ? (int)world[inb ? idx : 0] : -1;
      int face = -1;
      bool go = !(0 >= max_steps || t >= vd);
      int step = 0;
      while (go) {
        // blk >= 0 implies in bounds (out-of-bounds blk is -1).
        if ((blk >= 0) | ((!inb) & (t > t_exit))) break;
        const bool cx = tm_x < tm_y && tm_x < tm_z;
        const bool cy = !cx && (tm_y < tm_z);
        const bool cz = !cx && !cy;
        t = cx ? tm_x : (cy ? tm_y : tm_z);
        tm_x = cx ? tm_x + td_x : tm_x;
        tm_y = cy ? tm_y + td_y : tm_y;
        tm_z = cz ? tm_z + td_z : tm_z;
        vx = cx ? vx + step_x : vx;
        vy = cy ? vy + step_y : vy;
        vz = cz ? vz + step_z : vz;
        idx += cx ? dix : (cy ? diy : diz);
        face = cx ? 0 : (cy ? 1 : 2);
        inb = (unsigned)vx < (unsigned)X && (unsigned)vy < (unsigned)Y && (unsigned)vz < (unsigned)Z;
        const int raw = (int)world[inb ? idx : 0];
        blk = inb ? raw : -1;
        ++step;
        go = !(step >= max_steps || t >= vd);
      }
      if (go && inb && blk >= 0) {
        hit_block = blk;
        hit_face = face;
      }
      if (CNT) *nlook = step + 1;
    }
  }
  if ((unsigned)hit_block < 13u) {
    // Face shade by independent selects (no short-circuit branches).
    float shade = (hit_face == 0) ? 0.8f : 1.0f;
    shade = (hit_face == 2) ? 0.9f : shade;
    shade = ((hit_face == 1) & (step_y < 0)) ? 0.6f : shade;
    // t / vd, or t * (1 / vd) when vd is a power of two (the same value).
    float fog = fminf(p.vd_pow2 ? t * p.inv_vd : t / vd, 1.0f);
    fog = fog * fog;
    const float keep = 1.0f - fog;
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = pal[hit_block * 3 + c] * shade * keep + pal[kSkyBase + c] * fog;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) rgb[c] = pal[kSkyBase + c];
  }
}

// ---------------------------------------------------------------------------
// One ray of the flat-colour variants (src/kernels.cu:1040-1110, 625-686):
// no slab test, no view-distance limit, no shading; the DDA starts at the eye
// voxel and looks up at most max_steps voxels, skipping out-of-bounds ones.
// The reference normalises with rsqrtf; here the factor is 1 / sqrtf(len^2),
// both correctly rounded.  Returns the block hit, or -1.
//
// T = _Float16: the half-precision variant (src/kernels.cu:1127-1262): the
// same fp32 ray, a 65504 stand-in for 1/0, and t_max / t_delta rounded to
// binary16 (round to nearest even) then stepped and compared in binary16.
//
// Eye voxel in the world: leaving the world ends the ray (as in trace()).
// Otherwise a ray outside the world that moves away from it on some axis
// can never come back, which ends it too.
// ---------------------------------------------------------------------------
template <bool LIFT, typename T, bool CNT = false, bool GLB = false>
__device__ __forceinline__ int trace_flat(const Params& p, const Frame& f, const OctPos& op,
                                          float u, float v, const int8_t* __restrict__ world,
                                          int* nlook = nullptr) {
  constexpr float kInf = (sizeof(T) == 2) ? 65504.0f : 1e30f;
  float ray_x = f.fx + u * f.rx + v * f.ux;
  float ray_y = f.fy + v * f.uy;
  float ray_z = f.fz + u * f.rz + v * f.uz;
  float inv_x, inv_y, inv_z;
  ray_dir<true>(ray_x, ray_y, ray_z, inv_x, inv_y, inv_z, kInf);
  const int max_steps = p.max_steps;
  const int X = p.X, Y = p.Y, Z = p.Z, XZ = X * Z;
  const int step_x = (ray_x >= 0) ? 1 : -1;
  const int step_y = (ray_y >= 0) ? 1 : -1;
  const int step_z = (ray_z >= 0) ? 1 : -1;
  const T td_x = (T)fabsf(inv_x), td_y = (T)fabsf(inv_y), td_z = (T)fabsf(inv_z);
  const int dix = (ray_x >= 0) ? Z : -Z, diy = (ray_y >= 0) ? XZ : -XZ, diz = step_z;
  if (f.inside) {
    T tm_x = (T)((f.dxn + ((ray_x >= 0) ? 1.0f : 0.0f)) * inv_x);
    T tm_y = (T)((f.dyn + ((ray_y >= 0) ? 1.0f : 0.0f)) * inv_y);
    T tm_z = (T)((f.dzn + ((ray_z >= 0) ? 1.0f : 0.0f)) * inv_z);
    constexpr uint32_t kGuard = (1u << 9) | (1u << 19) | (1u << 29);
    uint32_t pos = op.pick(step_x > 0, step_y > 0, step_z > 0);
    // LDS address of the current voxel.
    const int8_t* vp = world + (f.ivy * XZ + f.ivx * Z + f.ivz);
    bool out = (pos & kGuard) != 0u;
    int blk = GLB ? (out ? -1 : (int)*(out ? world : vp)) : (int)*vp;
    int step = 0;
    bool lim = max_steps <= 0;
    while (!(lim | out | (blk >= 0))) {
      const bool cx = (tm_x < tm_y) & (tm_x < tm_z);
      const bool cy = !cx & (tm_y < tm_z);
      const bool cz = !cx & !cy;
      tm_x = cx ? tm_x + td_x : tm_x;
      tm_y = cy ? tm_y + td_y : tm_y;
      tm_z = cz ? tm_z + td_z : tm_z;
      vp += cx ? dix : (cy ? diy : diz);
      pos += cx ? 1u : (cy ? (1u << 10) : (1u << 20));
      out = (pos & kGuard) != 0u;
      // One step past the edge at most: a margin byte in LDS; a guarded read
      // of global memory.
      blk = GLB ? (out ? -1 : (int)*(out ? world : vp)) : (int)*vp;
      ++step;
      lim = step >= max_steps;
    }
    // Results from vector registers (see trace()): the step count is the sum
    // of pos's fields' advances.
    const uint32_t moved = pos - op.reread(step_x > 0, step_y > 0, step_z > 0);
    const int steps = (int)((moved & 0x3ffu) + ((moved >> 10) & 0x3ffu) + (moved >> 20));
    if (CNT) *nlook = steps + 1;
    asm volatile("" : "+v"(blk));  // a fresh test, not the loop's lane mask
    return ((steps < max_steps) & ((pos & kGuard) == 0u) & (blk >= 0)) ? blk : -1;
  }
  int vx = (int)floorf(f.ex), vy = (int)floorf(f.ey), vz = (int)floorf(f.ez);
  T tm_x = (T)((ray_x >= 0) ? ((float)(vx + 1) - f.ex) * inv_x : ((float)vx - f.ex) * inv_x);
  T tm_y = (T)((ray_y >= 0) ? ((float)(vy + 1) - f.ey) * inv_y : ((float)vy - f.ey) * inv_y);
  T tm_z = (T)((ray_z >= 0) ? ((float)(vz + 1) - f.ez) * inv_z : ((float)vz - f.ez) * inv_z);
  if (CNT) *nlook = 0;
  for (int step = 0; step < max_steps; ++step) {
    if (CNT) *nlook = step + 1;
    const bool inb = (unsigned)vx < (unsigned)X && (unsigned)vy < (unsigned)Y &&
                     (unsigned)vz < (unsigned)Z;
    if (inb) {
      const int blk = (int)world[vy * XZ + vx * Z + vz];
      if (blk >= 0) return blk;
    }
    const bool away = (vx < 0 && step_x < 0) || (vx >= X && step_x > 0) ||
                      (vy < 0 && step_y < 0) || (vy >= Y && step_y > 0) ||
                      (vz < 0 && step_z < 0) || (vz >= Z && step_z > 0);
    if (away) break;
    if (tm_x < tm_y && tm_x < tm_z) {
      tm_x += td_x;
      vx += step_x;
    } else if (tm_y < tm_z) {
      tm_y += td_y;
      vy += step_y;
    } else {
      tm_z += td_z;
      vz += step_z;
    }
  }
  return -1;
}

__device__ __forceinline__ uint32_t to_u8(float c) { return (uint32_t)(c * 255.0f) & 0xffu; }

// ---------------------------------------------------------------------------
// Render every pixel of env b into obs (layout [B,H,W,3]).
// OBS: 0 = uint8, 1 = float32.
// ---------------------------------------------------------------------------
// Image-plane coordinates of column px / row py, per variant:
//   kUvShaded   (a * half_fov) * aspect, b * half_fov     (src/kernels.cu:351-352)
//   kUvMinimal  (a * flat_hf) * aspect,  b * flat_hf      (src/kernels.cu:1040-1041)
//   kUvBasis    a * scale_u,             b * scale_v      (src/kernels.cu:624-625)
// with a = 2 (px + 0.5) / W - 1 and b = 1 - 2 (py + 0.5) / H.  kUvRaw keeps
// a and b, for renders whose scales come per env from a stored basis.
enum UvMode { kUvShaded = 0, kUvMinimal = 1, kUvBasis = 2, kUvRaw = 3 };
// x / n for a positive int n: a product with 2^-k when n = 2^k (the same
// value: scaling by a power of two is exact here), else the IEEE division.
__device__ __forceinline__ float div_by(float x, int n) {
  if ((n & (n - 1)) == 0) return x * __int_as_float((127 - (31 - __clz(n))) << 23);
  return x / (float)n;
}
__device__ __forceinline__ float u_of(const Params& p, int m, int px) {
  const float a = div_by(2.0f * (px + 0.5f), p.W) - 1.0f;
  if (m == kUvShaded) return a * p.half_fov * p.aspect;
  if (m == kUvMinimal) return a * p.flat_hf * p.aspect;
  if (m == kUvBasis) return a * p.flat_su;
  return a;
}
__device__ __forceinline__ float v_of(const Params& p, int m, int py) {
  const float c = 1.0f - div_by(2.0f * (py + 0.5f), p.H);
  if (m == kUvShaded) return c * p.half_fov;
  if (m == kUvMinimal || m == kUvBasis) return c * p.flat_hf;
  return c;
}

// Per-workgroup tables of u (per column) and v (per row), filled before the
// first barrier of an item when the image fits them.
constexpr int kMaxUV = 128;
__device__ __forceinline__ bool uv_tabled(const Params& p) { return p.W <= kMaxUV && p.H <= kMaxUV; }
__device__ __forceinline__ void fill_uv(const Params& p, int m, float* su, float* sv, int tid,
                                        int nt) {
  if (!uv_tabled(p)) return;
  if (su)
    for (int i = tid; i < p.W; i += nt) su[i] = u_of(p, m, i);
  for (int i = tid; i < p.H; i += nt) sv[i] = v_of(p, m, i);
}

// Constant-colour fill of env b's image with packed 0x00BBGGRR: three 16-byte
// phases of the repeating RGB pattern (chunk c starts at byte 16c, i.e. at
// channel phase c mod 3).  Needs 3 * W * H % 16 == 0.
__device__ __forceinline__ void fill_rgb8(uint8_t* out, int npix, uint32_t rgb, int tid, int nt) {
  const uint32_t r = rgb & 0xffu, g = (rgb >> 8) & 0xffu, bl = (rgb >> 16) & 0xffu;
  const uint32_t w0 = r | (g << 8) | (bl << 16) | (r << 24);
  const uint32_t w1 = g | (bl << 8) | (r << 16) | (g << 24);
  const uint32_t w2 = bl | (r << 8) | (g << 16) | (bl << 24);
  guint4* o4 = (guint4*)out;
  const int d3 = nt % 3;
  int ph = tid % 3;
  for (int c = tid; c < (npix * 3) >> 4; c += nt, ph = (ph + d3 >= 3) ? ph + d3 - 3 : ph + d3) {
    const uint32_t a = (ph == 0) ? w0 : (ph == 1) ? w1 : w2;
    const uint32_t bq = (ph == 0) ? w1 : (ph == 1) ? w2 : w0;
    const uint32_t cq = (ph == 0) ? w2 : (ph == 1) ? w0 : w1;
    o4[c] = v4u{a, bq, cq, a};
  }
}

// Constant-colour fill of env b's planar uint8 image [3, H, W] (packed
// 0x00BBGGRR): each plane is one byte repeated.  Needs H * W % 16 == 0.
__device__ __forceinline__ void fill_planes8(uint8_t* out, int npix, uint32_t rgb, int tid, int nt) {
  guint4* o4 = (guint4*)out;
  const int n16 = npix >> 4;
  for (int c = tid; c < 3 * n16; c += nt) {
    const int ch = (c >= n16) + (c >= 2 * n16);
    const uint32_t w = ((rgb >> (8 * ch)) & 0xffu) * 0x01010101u;
    o4[c] = v4u{w, w, w, w};
  }
}

// Constant-colour fill of env b's float32 image (c0, c1, c2) with 16-byte
// stores: interleaved, chunk c holds floats 4c .. 4c+3, whose channels start
// at phase 4c mod 3 = c mod 3; planar, each plane is one value repeated.
// Needs npix % 4 == 0 and a 16-byte aligned image.
__device__ __forceinline__ void fill_f32(float* out, int npix, float c0, float c1, float c2,
                                         bool planar, int tid, int nt) {
  gfloat4* o4 = (gfloat4*)out;
  const int n16 = (npix * 3) >> 2;
  if (planar) {
    const int per = npix >> 2;
    for (int c = tid; c < n16; c += nt) {
      const float v = (c < per) ? c0 : (c < 2 * per) ? c1 : c2;
      o4[c] = v4f{v, v, v, v};
    }
  } else {
    for (int c = tid; c < n16; c += nt) {
      const int ph = c % 3;
      const float a = (ph == 0) ? c0 : (ph == 1) ? c1 : c2;
      const float bq = (ph == 0) ? c1 : (ph == 1) ? c2 : c0;
      const float d = (ph == 0) ? c2 : (ph == 1) ? c0 : c1;
      o4[c] = v4f{a, bq, d, a};
    }
  }
}

// Pixels are traced in 8x8 tiles, one pixel per lane of a wave per tile, and
// a wave's loop over a tile ends when its longest ray ends: neighbouring
// pixels have rays of similar length, so square tiles waste fewer lane-steps
// than rows (measured on falling-agent frames at 16^3 / 32x32: 76% of the
// lane-steps useful with 8x8 tiles, 65% with 4-pixel quads spread over 8 rows).
// Tile t of an env goes to its wave t mod waves-per-env.

// Store packed 0x00BBGGRR at pixel p of a uint8 RGB image: a short and a byte,
// the short at the even byte offset.
__device__ __forceinline__ void store_rgb8(uint8_t* out, int p, uint32_t c) {
  typedef MC_GLOBAL uint16_t gu16;
  typedef MC_GLOBAL uint8_t gu8;
  const bool even = (p & 1) == 0;
  const uint32_t o = (uint32_t)p * 3u;
  *(gu16*)(out + (even ? o : o + 1u)) = (uint16_t)(even ? c : c >> 8);
  *(gu8*)(out + (even ? o + 2u : o)) = (uint8_t)(even ? c >> 16 : c);
}

// Tile loop over env b's image: fn(px, py) for every pixel, one per lane.
// The wave index is uniform: the tile row and column advance in scalar
// registers.
template <typename PixelFn>
__device__ __forceinline__ void for_tiles(const Params& p, int tid, int nt, PixelFn fn) {
  const int lane = tid & (kWave - 1), nwv = nt >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tx = lane & 7, ty = lane >> 3;
  const int tcols = (p.W + 7) >> 3, ntiles = tcols * ((p.H + 7) >> 3);
  const int dr = nwv / tcols, dc = nwv - dr * tcols;
  int tr = wv / tcols, tc = wv - tr * tcols;
  for (int t = wv; t < ntiles; t += nwv) {
    const int py = tr * 8 + ty, px = tc * 8 + tx;
    if ((py < p.H) & (px < p.W)) fn(px, py);
    tr += dr;
    tc += dc;
    if (tc >= tcols) {
      tc -= tcols;
      ++tr;
    }
  }
}

// uint8 image of env b from a per-pixel colour function returning packed
// 0x00BBGGRR.
// Interleaved images whose width is a multiple of 8 store each 8x8 tile as
// 48 dwords (8 rows of 24 bytes): the tile's packed pixels go through the
// wave's 64-dword LDS buffer `tile`, and lane 6r + d (r < 8, d < 6) builds
// dword d of row r from the two pixels it spans with one byte permute.  Rows
// below the image are not stored.  Other images store per pixel.
template <typename ColorFn>
__device__ __forceinline__ void render_rgb8(const Params& p, int b, int uvm, int tid, int nt,
                                            const float* su, const float* sv, ColorFn color,
                                            bool planar = false, uint32_t* tile = nullptr) {
  uint8_t* out = reinterpret_cast<uint8_t*>(p.obs) + (int64_t)b * p.W * p.H * 3;
  const bool tab = uv_tabled(p);
  // No u table (su null): u from its formula.
  const bool tab_u = tab && su != nullptr;
  if (!planar && tile && ((p.W & 7) == 0) && (((uintptr_t)out & 3u) == 0)) {
    const int lane = tid & (kWave - 1), nwv = nt >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tx = lane & 7, ty = lane >> 3;
    const int tcols = p.W >> 3, trows = (p.H + 7) >> 3, ntiles = tcols * trows;
    // Store role of this lane: row r, dword d, the tile pixels a and a + 1
    // it takes its bytes from, and the permute selector (bytes 0-3 of pixel
    // a, 4-7 of pixel a + 1, each 0x00BBGGRR).
    const int r = lane / 6, d = lane - 6 * r, dm = d % 3;
    const int a = r * 8 + (d / 3) * 4 + dm;
    const uint32_t sel = (dm == 0) ? 0x04020100u : (dm == 1) ? 0x05040201u : 0x06050402u;
    const uint32_t row_off = (uint32_t)(r * p.W * 3 + 4 * d);
    const int dr = nwv / tcols, dc = nwv - dr * tcols;
    int tr = wv / tcols, tc = wv - tr * tcols;
    typedef MC_GLOBAL uint32_t gu32;
    for (int t = wv; t < ntiles; t += nwv) {
      const int py = tr * 8 + ty, px = tc * 8 + tx;
      uint32_t c = 0u;
      if (py < p.H) c = color(tab_u ? su[px] : u_of(p, uvm, px), tab ? sv[py] : v_of(p, uvm, py));
      tile[lane] = c;
      __builtin_amdgcn_wave_barrier();
      if ((lane < 48) & (tr * 8 + r < p.H)) {
        const uint32_t w = __builtin_amdgcn_perm(tile[a + 1], tile[a], sel);
        *(gu32*)(out + (uint32_t)((tr * 8 * p.W + tc * 8) * 3) + row_off) = w;
      }
      __builtin_amdgcn_wave_barrier();
      tr += dr;
      tc += dc;
      if (tc >= tcols) {
        tc -= tcols;
        ++tr;
      }
    }
    return;
  }
  for_tiles(p, tid, nt, [&](int px, int py) {
    const uint32_t c = color(tab_u ? su[px] : u_of(p, uvm, px), tab ? sv[py] : v_of(p, uvm, py));
    if (planar) {
      typedef MC_GLOBAL uint8_t gu8;
      const int npix = p.W * p.H;
      gu8* o = (gu8*)(out + py * p.W + px);
      o[0] = (uint8_t)c;
      o[npix] = (uint8_t)(c >> 8);
      o[2 * npix] = (uint8_t)(c >> 16);
    } else {
      store_rgb8(out, py * p.W + px, c);
    }
  });
}

// Flat-colour uint8 render of env b (src/kernels.cu:985-1119 minimal,
// 568-700 prebasis, 835-982 smem: same image as prebasis, 1127-1262 with a
// binary16 DDA when T = _Float16).  flat: the packed colour table, 13 blocks
// then the sky; ids >= 13 show the sky.
// CNT: acc[0] += DDA lookups, acc[1] += rays marched, acc[2] += pixels.
template <bool LIFT, typename T, bool CNT = false, bool GLB = false>
__device__ __forceinline__ void render_env_flat(const Params& p, int b, const int8_t* world,
                                                const uint32_t* flat, const Frame& f,
                                                const OctPos& op, int uvm, int tid, int nt,
                                                const float* su, const float* sv,
                                                uint32_t* acc = nullptr, uint32_t* tile = nullptr) {
  const int npix = p.W * p.H;
  uint8_t* out = reinterpret_cast<uint8_t*>(p.obs) + (int64_t)b * npix * 3;
  const bool per_env_scale = (uvm == kUvRaw);
  auto color = [&](float u, float v) -> uint32_t {
    if (per_env_scale) { u = u * f.su; v = v * f.sv; }
    int h;
    if (f.flat_fill >= 0) {
      h = f.flat_fill;
      if (CNT) acc[0] += 1;
    } else {
      int n = 0;
      h = trace_flat<LIFT, T, CNT, GLB>(p, f, op, u, v, world, &n);
      if (CNT) acc[0] += n, acc[1] += 1;
    }
    if (CNT) acc[2] += 1;
    return flat[((unsigned)h < (unsigned)kFlatSky) ? h : kFlatSky];
  };
  if (f.flat_fill >= 0 && ((npix * 3) & 15) == 0) {
    fill_rgb8(out, npix, flat[f.flat_fill < kFlatSky ? f.flat_fill : kFlatSky], tid, nt);
    if (CNT && tid == 0) acc[0] += npix, acc[2] += npix;
    return;
  }
  render_rgb8(p, b, uvm, tid, nt, su, sv, color, false, tile);
}

// ---------------------------------------------------------------------------
// The shaded render of an env whose eye voxel lies in the world (trace()'s
// in-world branch for every pixel), with the per-column ray terms and the v
// table from LDS and the tile stores of render_rgb8 (in the float32 and
// planar kernels, per-pixel stores): the same arithmetic as trace(), laid out
// for instruction count.  The env's camera frame, face distances and address
// steps are held in vector registers for the whole image (as scalars they
// outnumber the scalar file and are reloaded per pixel); the ray set-up tests
// its fast-path range once, on the smallest component; the DDA block decides
// the hit itself (dda_in), so no per-lane state outlives it.
// ---------------------------------------------------------------------------

// Unit direction and reciprocals of a shaded ray (ray_dir<false>): the fast
// path takes rays whose components all have magnitude at least 2^-30 (then
// len^2 >= 3 * 2^-60 and no component or quotient is 0) and len^2 <= 2^60,
// a subset of ray_dir's; a NaN or infinite component fails the len^2 test.
// yy: y * y, computed once per image row by the caller.
__device__ __forceinline__ void ray_dir_in(float& x, float& y, float& z, float yy, float& ix,
                                           float& iy, float& iz) {
  const float len2 = x * x + yy + z * z;
  const float m = fminf(fminf(fabsf(x), fabsf(y)), fabsf(z));
  if ((m >= 0x1p-30f) & (len2 <= 0x1p60f)) {
    const float len = sqrt_core(len2);
    const float rl = recip_core(len);
    x = div_core(x, len, rl);
    y = div_core(y, len, rl);
    z = div_core(z, len, rl);
    ix = rcp_cr(x);
    iy = rcp_cr(y);
    iz = rcp_cr(z);
  } else {
    ray_dir<false>(x, y, z, ix, iy, iz, 1e30f);
  }
}

// dda_loop with the run test and the hit decision in the same block, for an
// eye voxel of air, max_steps > 0 and view_distance > 0.  Entry: the lanes
// whose starting packed position is in the world.  Exit:
// the lanes that ran max_steps iterations are those active at the start of
// the last iteration when the counter ended the loop (all lanes when
// max_steps <= 0, as trace()'s steps >= max_steps).  Returns the voxel byte
// of a hit (in the world, not limited by max_steps, and without VDF t < vd:
// trace()'s !lim_end & in-world test) or 0xffffffff; the caller's test
// `< 13` then also applies trace()'s valid-block test (air bytes are >= 128).
// t, pinc and the voxel byte are written by the first iteration; a lane that
// runs none has its start position outside the world, so its result is -1 and
// they are not read.
template <bool VDF>
__device__ __forceinline__ uint32_t dda_in(float& tx, float& ty, float& tz, float& t, float ix,
                                           float iy, float iz, uint32_t va, uint32_t dx,
                                           uint32_t dy, uint32_t dz, uint32_t& pos,
                                           uint32_t& pinc, int max_steps, float vd) {
  uint64_t save, sa, mz, my, mv;
  uint32_t d, w, step, hb, blk;
  const uint32_t k10 = 1u << 10, k20 = 1u << 20;  // the guard bits 9, 19, 29 are a literal
#define MC_DDA_IN_ASM(VDCMP, VDAND, VDEND, VDHIT)                                 \
  asm volatile(                                                                    \
      "s_mov_b64 %[save], exec\n\t"                                               \
      "s_mov_b64 %[sa], exec\n\t"                                                 \
      "s_mov_b32 %[step], 0\n\t"                                                  \
      "v_and_b32_e32 %[w], 0x20080200, %[pos]\n\t"                                     \
      "v_cmp_eq_u32_e32 vcc, 0, %[w]\n\t"                                         \
      "s_and_b64 exec, exec, vcc\n\t"                                             \
      "s_cbranch_execz .Ldin_end%=\n"                                              \
      ".Ldin_loop%=:\n\t"                                                         \
      "v_min3_f32 %[t], %[tx], %[ty], %[tz]\n\t"                                  \
      "s_add_u32 %[step], %[step], 1\n\t"                                         \
      "v_cmp_eq_f32_e64 %[mz], %[tz], %[t]\n\t"                                   \
      "v_cmp_eq_f32_e64 %[my], %[ty], %[t]\n\t" VDCMP                             \
      "s_mov_b64 %[sa], exec\n\t"                                                 \
      "s_andn2_b64 %[my], %[my], %[mz]\n\t"                                       \
      "v_cndmask_b32_e64 %[d], %[dx], %[dy], %[my]\n\t"                           \
      "v_cndmask_b32_e64 %[d], %[d], %[dz], %[mz]\n\t"                            \
      "v_add_u32_e32 %[va], %[va], %[d]\n\t"                                      \
      "ds_read_u8 %[blk], %[va]\n\t"                                              \
      "v_cndmask_b32_e64 %[pi], 1, %[k10], %[my]\n\t"                             \
      "v_cndmask_b32_e64 %[pi], %[pi], %[k20], %[mz]\n\t"                         \
      "v_add_u32_e32 %[pos], %[pos], %[pi]\n\t"                                   \
      "s_andn2_b64 exec, %[sa], %[my]\n\t"                                        \
      "s_andn2_b64 exec, exec, %[mz]\n\t"                                         \
      "v_add_f32_e64 %[tx], %[tx], |%[ix]|\n\t"                                   \
      "s_mov_b64 exec, %[my]\n\t"                                                 \
      "v_add_f32_e64 %[ty], %[ty], |%[iy]|\n\t"                                   \
      "s_mov_b64 exec, %[mz]\n\t"                                                 \
      "v_add_f32_e64 %[tz], %[tz], |%[iz]|\n\t"                                   \
      "s_mov_b64 exec, %[sa]\n\t"                                                 \
      "v_and_b32_e32 %[w], 0x20080200, %[pos]\n\t"                                     \
      "s_waitcnt lgkmcnt(0)\n\t"                                                  \
      "v_sub_u32_e32 %[w], %[blk], %[w]\n\t"                                      \
      "v_cmp_lt_i32_e32 vcc, 0x7f, %[w]\n\t"                                      \
      "s_and_b64 exec, exec, vcc\n\t" VDAND                                      \
      "s_cmp_ge_i32 %[step], %[ms]\n\t"                                           \
      "s_cbranch_scc1 .Ldin_end%=\n\t"                                            \
      "s_cbranch_execnz .Ldin_loop%=\n"                                            \
      ".Ldin_end%=:\n\t"                                                          \
      "s_cmp_ge_i32 %[step], %[ms]\n\t"                                           \
      "s_cselect_b64 %[sa], %[sa], 0\n\t"                                         \
      "s_mov_b64 exec, %[save]\n\t"                                               \
      "v_and_b32_e32 %[w], 0x20080200, %[pos]\n\t"                                     \
      "v_cmp_eq_u32_e64 %[mz], 0, %[w]\n\t" VDEND                                 \
      "s_andn2_b64 %[mz], %[mz], %[sa]\n\t" VDHIT                                 \
      "s_nop 1\n\t"                                                               \
      "v_cndmask_b32_e64 %[hb], -1, %[blk], %[mz]"                                \
      : [tx] "+v"(tx), [ty] "+v"(ty), [tz] "+v"(tz), [t] "=&v"(t), [va] "+v"(va),     \
        [pos] "+v"(pos), [pi] "=&v"(pinc), [blk] "=&v"(blk), [d] "=&v"(d),           \
        [w] "=&v"(w), [hb] "=&v"(hb), [save] "=&s"(save), [sa] "=&s"(sa),            \
        [mz] "=&s"(mz), [my] "=&s"(my), [mv] "=&s"(mv),                              \
        [step] "=&s"(step)                                                          \
      : [ix] "v"(ix), [iy] "v"(iy), [iz] "v"(iz), [dx] "v"(dx), [dy] "v"(dy),        \
        [dz] "v"(dz), [k10] "v"(k10), [k20] "v"(k20),                                \
        [ms] "s"(max_steps), [vd] "s"(vd)                                           \
      : "vcc", "scc")
  if (VDF)
    MC_DDA_IN_ASM("", "", "", "");
  else
    MC_DDA_IN_ASM("v_cmp_gt_f32_e64 %[mv], %[vd], %[t]\n\t", "s_and_b64 exec, exec, %[mv]\n\t",
                  "v_cmp_gt_f32_e64 %[mv], %[vd], %[t]\n\t", "s_and_b64 %[mz], %[mz], %[mv]\n\t");
#undef MC_DDA_IN_ASM
  return hb;
}

// OBS 0: uint8 (interleaved: render_rgb8's tile stores, width a multiple of 8
// and a dword-aligned image); OBS 1: float32; planar: [3, H, W] per pixel.
template <int OBS, bool LIFT, bool VDF>
// colA: the env's column terms of its camera rays, (fx + u rx, fz + u rz)
// per column (run_item); the row terms (v ux, v uz, fy + v uy and the
// square of the latter) are formed once per tile row, so a ray costs two
// additions (the reference's sums, same operands and order).
__device__ __forceinline__ void render_in(const Params& p, void* out_v, const int8_t* world,
                                          const float* pal, const Frame& f, const OctPos& op,
                                          int tid, int nt, const float2* colA, const float* sv,
                                          bool planar, uint32_t* tile) {
  typedef __attribute__((address_space(3))) const uint8_t lds_u8;
  typedef __attribute__((address_space(3))) const uint32_t lds_u32;
  const int X = p.X, Z = p.Z, XZ = X * Z;
  const float fy = f.fy, ux = f.ux, uy = f.uy, uz = f.uz;
  float dxn = f.dxn, dxp = f.dxp, dyn = f.dyn, dyp = f.dyp, dzn = f.dzn, dzp = f.dzp;
  uint32_t zp = (uint32_t)Z, zn = (uint32_t)-Z, xzp = (uint32_t)XZ, xzn = (uint32_t)-XZ;
  const uint32_t va0 = (uint32_t)(uintptr_t)(lds_u8*)(world + (f.ivy * XZ + f.ivx * Z + f.ivz));
  const uint32_t ob = (uint32_t)(uintptr_t)(lds_u32*)op.v;
  const uint32_t blk0 = __builtin_amdgcn_readfirstlane((uint32_t)*(lds_u8*)(uintptr_t)va0);
  // Scalars used inside the pixel loop, as values of their own (not parts of
  // the parameter block, which the register allocator reloads as a whole).
  float vd = p.view_distance;
  int max_steps = p.max_steps, W = p.W, H = p.H;
  asm volatile("" : "+s"(vd), "+s"(max_steps), "+s"(W), "+s"(H));
  const uint32_t go =
      __builtin_amdgcn_readfirstlane((((int8_t)blk0 < 0) & (max_steps > 0) & (vd > 0.0f)) ? 1u : 0u);
  const float sk0 = pal[kSkyBase], sk1 = pal[kSkyBase + 1], sk2 = pal[kSkyBase + 2];
  const uint32_t sky = to_u8(sk0) | (to_u8(sk1) << 8) | (to_u8(sk2) << 16);
  const int pow2 = __builtin_amdgcn_readfirstlane(p.vd_pow2);
  float ivd = p.inv_vd;
  // Uniform values moved to vector registers (see above).
  asm volatile("" : "+v"(dxn), "+v"(dxp), "+v"(dyn), "+v"(dyp), "+v"(dzn), "+v"(dzp));
  asm volatile("" : "+v"(zp), "+v"(zn), "+v"(xzp), "+v"(xzn), "+v"(ivd));
  // Row terms of the current tile row (row_terms).
  float rbx = 0.0f, rbz = 0.0f, rcy = 0.0f, ryy = 0.0f;
  auto row_terms = [&](int py) {
    const float v = sv[py];
    rbx = v * ux;
    rbz = v * uz;
    rcy = fy + v * uy;
    ryy = rcy * rcy;
  };
  // The shaded colour of the ray through column px of the current row, as
  // trace() computes it.  Returns whether the ray hits a block (else the
  // colour is the sky's).
  auto shade_ray = [&](int px, float& c0, float& c1, float& c2) -> bool {
    const float2 a = colA[px];
    float ray_x = a.x + rbx;
    float ray_y = rcy;
    float ray_z = a.y + rbz;
    float inv_x, inv_y, inv_z;
    ray_dir_in(ray_x, ray_y, ray_z, ryy, inv_x, inv_y, inv_z);
    const bool bx = ray_x >= 0, by = ray_y >= 0, bz = ray_z >= 0;
    // (d*n + 1) = d*p exactly (trace()).
    float tm_x = (bx ? dxp : dxn) * inv_x;
    float tm_y = (by ? dyp : dyn) * inv_y;
    float tm_z = (bz ? dzp : dzn) * inv_z;
    uint32_t pos = *(lds_u32*)(uintptr_t)(ob + (bx ? 4u : 0u) + (by ? 8u : 0u) + (bz ? 16u : 0u));
    uint32_t pinc = 0u;
    float t = 0.0f;
    uint32_t hb;
    if (go) {
      hb = dda_in<VDF>(tm_x, tm_y, tm_z, t, inv_x, inv_y, inv_z, va0, bx ? zp : zn,
                       by ? xzp : xzn, bz ? 1u : 0xffffffffu, pos, pinc, max_steps, vd);
    } else {
      // No iteration (trace(): steps 0, so lim_end is max_steps <= 0 or t = 0 >= vd):
      // the eye voxel, if solid, is hit at t = 0 with no face.
      hb = (((pos & ((1u << 9) | (1u << 19) | (1u << 29))) == 0u) & (max_steps > 0) &
            !(0.0f >= vd))
               ? blk0
               : 0xffffffffu;
    }
    c0 = sk0;
    c1 = sk1;
    c2 = sk2;
    const bool hit = hb < 13u;
    if (hit) {
      float shade = (pinc == 1u) ? 0.8f : 1.0f;
      shade = (pinc == (1u << 20)) ? 0.9f : shade;
      shade = ((pinc == (1u << 10)) & !by) ? 0.6f : shade;
      float fog = fminf(pow2 ? t * ivd : t / vd, 1.0f);
      fog = fog * fog;
      const float keep = 1.0f - fog;
      // hb * 3 floats as a 24-bit product (a full 32-bit multiply issues at a
      // quarter of the rate).
      uint32_t hoff;
      asm("v_mul_u32_u24_e32 %0, 12, %1" : "=v"(hoff) : "v"(hb));
      const float* hc = (const float*)((const char*)pal + hoff);
      c0 = hc[0] * shade * keep + sk0 * fog;
      c1 = hc[1] * shade * keep + sk1 * fog;
      c2 = hc[2] * shade * keep + sk2 * fog;
    }
    return hit;
  };
  if (OBS == 0 && !planar) {
    // render_rgb8's tile loop.
    uint8_t* out = reinterpret_cast<uint8_t*>(out_v);
    const int lane = tid & (kWave - 1), nwv = nt >> 6;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tx = lane & 7, ty = lane >> 3;
    const int tcols = W >> 3, trows = (H + 7) >> 3, ntiles = tcols * trows;
    const int r = lane / 6, d = lane - 6 * r, dm = d % 3;
    const int a = r * 8 + (d / 3) * 4 + dm;
    const uint32_t sel = (dm == 0) ? 0x04020100u : (dm == 1) ? 0x05040201u : 0x06050402u;
    const uint32_t row_off = (uint32_t)(r * W * 3 + 4 * d);
    const int dr = nwv / tcols, dc = nwv - dr * tcols;
    int tr = wv / tcols, tc = wv - tr * tcols;
    int row_tr = -1;
    typedef MC_GLOBAL uint32_t gu32;
    for (int tt = wv; tt < ntiles; tt += nwv) {
      const int py = tr * 8 + ty, px = tc * 8 + tx;
      if (tr != row_tr) {
        row_tr = tr;
        row_terms(min(py, H - 1));
      }
      uint32_t c = 0u;
      if (py < H) {
        float c0, c1, c2;
        c = shade_ray(px, c0, c1, c2) ? to_u8(c0) | (to_u8(c1) << 8) | (to_u8(c2) << 16) : sky;
      }
      tile[lane] = c;
      __builtin_amdgcn_wave_barrier();
      if ((lane < 48) & (tr * 8 + r < H)) {
        const uint32_t w = __builtin_amdgcn_perm(tile[a + 1], tile[a], sel);
        *(gu32*)(out + (uint32_t)((tr * 8 * W + tc * 8) * 3) + row_off) = w;
      }
      __builtin_amdgcn_wave_barrier();
      tr += dr;
      tc += dc;
      if (tc >= tcols) {
        tc -= tcols;
        ++tr;
      }
    }
    return;
  }
  const int npix = W * H;
  // for_tiles' walk over the tiles, with the row terms of each tile row.
  const int lane = tid & (kWave - 1), nwv = nt >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tx = lane & 7, ty = lane >> 3;
  const int tcols = (W + 7) >> 3, ntiles = tcols * ((H + 7) >> 3);
  const int dr = nwv / tcols, dc = nwv - dr * tcols;
  int tr = wv / tcols, tc = wv - tr * tcols;
  int row_tr = -1;
  for (int tt = wv; tt < ntiles; tt += nwv) {
    const int py = tr * 8 + ty, px = tc * 8 + tx;
    if (tr != row_tr) {
      row_tr = tr;
      row_terms(min(py, H - 1));
    }
    tr += dr;
    tc += dc;
    if (tc >= tcols) {
      tc -= tcols;
      ++tr;
    }
    if (!((py < H) & (px < W))) continue;
    float c0, c1, c2;
    shade_ray(px, c0, c1, c2);
    if (OBS == 0) {
      typedef MC_GLOBAL uint8_t gu8;
      gu8* o = (gu8*)(reinterpret_cast<uint8_t*>(out_v) + py * W + px);
      o[0] = (uint8_t)to_u8(c0);
      o[npix] = (uint8_t)to_u8(c1);
      o[2 * npix] = (uint8_t)to_u8(c2);
    } else if (planar) {
      typedef MC_GLOBAL float gf32;
      gf32* o = (gf32*)(reinterpret_cast<float*>(out_v) + (int64_t)py * W + px);
      o[0] = c0;
      o[npix] = c1;
      o[2 * npix] = c2;
    } else {
      typedef MC_GLOBAL float gf32;
      gf32* o = (gf32*)(reinterpret_cast<float*>(out_v) + ((int64_t)py * W + px) * 3);
      o[0] = c0;
      o[1] = c1;
      o[2] = c2;
    }
  }
}

// planar: the image is written [3, H, W] (the MC_OBS_*_CHW modes).  Callers
// pass a compile-time constant, so each kernel holds one store form only.
// IN32: float32 images of an in-world eye through render_in (the kernels
// built for the float32 modes only: in a kernel that also holds the uint8
// path, its registers would cost the uint8 render more than it gains).
template <int OBS, bool LIFT, bool VDF, bool CNT = false, bool GLB = false, bool IN32 = false>
__device__ __forceinline__ void render_env(const Params& p, int b, const int8_t* world,
                                           const float* pal, const Frame& f, const OctPos& op,
                                           int tid, int nt, const float* su, const float* sv,
                                           bool planar, uint32_t* acc = nullptr,
                                           uint32_t* tile = nullptr,
                                           const float2* colA = nullptr) {
  const int W = p.W, npix = W * p.H;
  const bool tab = uv_tabled(p);
  const bool tab_u = tab && su != nullptr;  // no u table: u from its formula
  auto ucol = [&](int px) { return tab_u ? su[px] : u_of(p, kUvShaded, px); };
  auto vrow = [&](int py) { return tab ? sv[py] : v_of(p, kUvShaded, py); };
  if (OBS == 0) {
    uint8_t* out = reinterpret_cast<uint8_t*>(p.obs) + (int64_t)b * npix * 3;
    if (f.fill_block >= 0 && ((npix * 3) & 15) == 0 && !planar) {
      fill_rgb8(out, npix,
                to_u8(pal[f.fill_block * 3]) | (to_u8(pal[f.fill_block * 3 + 1]) << 8) |
                    (to_u8(pal[f.fill_block * 3 + 2]) << 16),
                tid, nt);
      if (CNT && tid == 0) acc[0] += npix, acc[2] += npix;
      return;
    }
    if (!CNT && !GLB && colA && f.inside && f.fill_block < 0 && tab &&
        (planar || (tile && (W & 7) == 0 && (((uintptr_t)out & 3u) == 0)))) {
      render_in<0, LIFT, VDF>(p, out, world, pal, f, op, tid, nt, colA, sv, planar, tile);
      return;
    }
    // Packed 0x00BBGGRR per pixel.
    auto color = [&](float u, float v) -> uint32_t {
      float rgb[3];
      if (f.fill_block >= 0) {
        for (int c = 0; c < 3; ++c) rgb[c] = pal[f.fill_block * 3 + c];
        if (CNT) acc[0] += 1;
      } else {
        int n = 0;
        trace<LIFT, VDF, CNT, GLB>(p, f, op, u, v, world, pal, rgb, &n);
        if (CNT) acc[0] += n, acc[1] += 1;
      }
      if (CNT) acc[2] += 1;
      return to_u8(rgb[0]) | (to_u8(rgb[1]) << 8) | (to_u8(rgb[2]) << 16);
    };
    render_rgb8(p, b, kUvShaded, tid, nt, su, sv, color, planar, tile);
  } else {
    float* out = reinterpret_cast<float*>(p.obs) + (int64_t)b * npix * 3;
    if (IN32 && !CNT && !GLB && colA && f.inside && f.fill_block < 0 && tab) {
      render_in<1, LIFT, VDF>(p, out, world, pal, f, op, tid, nt, colA, sv, planar, tile);
      return;
    }
    typedef MC_GLOBAL float gf32;
    for_tiles(p, tid, nt, [&](int px, int py) {
      float rgb[3];
      if (f.fill_block >= 0) {
        for (int c = 0; c < 3; ++c) rgb[c] = pal[f.fill_block * 3 + c];
        if (CNT) acc[0] += 1;
      } else {
        int n = 0;
        trace<LIFT, VDF, CNT, GLB>(p, f, op, ucol(px), vrow(py), world, pal, rgb, &n);
        if (CNT) acc[0] += n, acc[1] += 1;
      }
      if (CNT) acc[2] += 1;
      if (planar) {
        gf32* o = (gf32*)(out + (int64_t)py * W + px);
        o[0] = rgb[0];
        o[npix] = rgb[1];
        o[2 * npix] = rgb[2];
      } else {
        gf32* o = (gf32*)(out + ((int64_t)py * W + px) * 3);
        o[0] = rgb[0];
        o[1] = rgb[1];
        o[2] = rgb[2];
      }
    });
  }
}

// Respawn height of env b: spawn_y on the flat world; on the seeded terrain
// (TER) spawn_y raised or lowered by the height of env b's spawn column
// (floor(spawn_x), floor(spawn_z)) relative to ground_height, so the agent
// appears the same distance above its own surface (no counterpart in the
// reference, whose world is flat).
template <bool TER>
__device__ __forceinline__ float spawn_of(const Params& p, int b) {
  if (!TER) return p.spawn_y;
  const int h = terrain_height(p, b, (int)floorf(p.spawn_x), (int)floorf(p.spawn_z));
  return p.spawn_y + (float)(h - p.ground_height);
}

// ---------------------------------------------------------------------------
// Per-env step logic, executed by the 64 lanes of wave 0 (values uniform
// across lanes except in the raycast).  Order: episode check
// (minecu/__init__.py:363-366), decode (kernels.cu:1477-1503), physics
// (1536-1583), raycast break/place (1609-1679), masked reset (1775-1785),
// camera (2330-2334), step_count update (minecu/__init__.py:435-436).
// ---------------------------------------------------------------------------
// Inputs of the per-env logic, loaded before the world staging completes so
// both HBM round trips overlap.
// Env b's inputs as loaded (load_env_in), decoded by env_update.
struct EnvIn {
  int32_t sc;                  // step_count[b] (done_mode 0)
  uint32_t dr;                 // do_reset[b]
  uint32_t a_lo, a_hi;         // the 8 action bytes: an int64 id, or the 8 int8 buttons
  float look_yaw, look_pitch;  // look[b] (buttons)
  float yaw, pitch, px, py, pz, vy;
  uint32_t og;
};

// sin and cos of two angles for the price of one evaluation: a on lane 0, b
// on lane 1, each result read back from its lane.  a and b must be the same
// on every lane, and lanes 0 and 1 active (a whole wave in uniform flow).
__device__ __forceinline__ void sincos_pair(float a, float b, float& sa, float& ca, float& sb,
                                            float& cb) {
  float s, c;
  mc_sincosf((threadIdx.x & (kWave - 1)) == 1 ? b : a, &s, &c);
  sa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 0));
  ca = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), 0));
  sb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s), 1));
  cb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c), 1));
}

// Every load of env b's inputs, unconditional and independent (issued
// together, one memory round trip; uniform choices pick a valid address whose
// value is then ignored).  The 8 bytes at actions + 8 b are the int64 id or
// the 8 buttons; a base address off 4-byte alignment reads them by byte.
__device__ __forceinline__ EnvIn load_env_in(const Params& p, int b) {
  EnvIn in;
  const bool m0 = p.done_mode == 0;
  in.sc = *(m0 ? p.step_count + b : reinterpret_cast<const int32_t*>(p.yaws + b));
  in.dr = p.do_reset[b];
  const uint8_t* src = reinterpret_cast<const uint8_t*>(p.actions) + (int64_t)b * 8;
  if (((uintptr_t)p.actions & 3u) == 0) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(src);
    in.a_lo = w[0];
    in.a_hi = w[1];
  } else {
    in.a_lo = in.a_hi = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      in.a_lo |= (uint32_t)src[k] << (8 * k);
      in.a_hi |= (uint32_t)src[4 + k] << (8 * k);
    }
  }
  const bool lk = p.action_kind != 1 && p.look;
  const float* lp = lk ? p.look + (int64_t)b * 2 : p.positions + (int64_t)b * 3;
  in.look_yaw = lp[0];
  in.look_pitch = lp[1];
  in.yaw = p.yaws[b];
  in.pitch = p.pitches[b];
  in.px = p.positions[b * 3 + 0];
  in.py = p.positions[b * 3 + 1];
  in.pz = p.positions[b * 3 + 2];
  in.vy = p.velocities[b * 3 + 1];
  in.og = p.on_ground[b];
  return in;
}

// spawn_y: env b's respawn height (spawn_of()).
template <bool COL = false>
__device__ __forceinline__ Frame env_update(const Params& p, int slot, int b, int8_t* world,
                                            int8_t* gworld, bool any_done, const EnvIn& in,
                                            bool upper_solid, float spawn_y) {
  const int lane = threadIdx.x & (kWave - 1);
  // Uniform parameters read once, as values the register allocator cannot
  // rematerialise: it would otherwise reload them from the kernel arguments
  // (a scalar-cache round trip) at each use.
  int episode_length = p.episode_length, action_kind = p.action_kind;
  float look_delta = p.look_delta;
  asm volatile("" : "+s"(episode_length), "+s"(action_kind), "+s"(look_delta));
  const int sc = (p.done_mode == 0) ? in.sc : 0;
  const bool done = (p.done_mode == 0) ? ((episode_length >= 0) && (sc >= episode_length))
                                       : ((in.dr & 0xffu) != 0u);
  // Buttons and look: from the discrete id (docs/environment.md:3-17; 1..6
  // press one button, 7..10 turn, other ids do nothing), by selects rather
  // than a branch chain, or as given.
  uint32_t bt_lo = in.a_lo, bt_hi = in.a_hi;
  float look_yaw = in.look_yaw, look_pitch = in.look_pitch;
  if (action_kind == 1) {
    const int64_t a64 = (int64_t)(((uint64_t)in.a_hi << 32) | in.a_lo);
    const int a = (a64 >= 1 && a64 <= 10) ? (int)a64 : 0;
    bt_lo = (a >= 1 && a <= 4) ? (1u << ((8 * (a - 1)) & 31)) : 0u;
    bt_hi = (a == 5 || a == 6) ? (1u << ((8 * (a - 5)) & 31)) : 0u;
    look_yaw = (a == 7) ? look_delta : (a == 8) ? -look_delta : 0.0f;
    look_pitch = (a == 9) ? look_delta : (a == 10) ? -look_delta : 0.0f;
  } else if (!p.look) {
    look_yaw = look_pitch = 0.0f;
  }

  // Decode (src/kernels.cu:1477-1503).
  int8_t bt[8];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    bt[k] = (int8_t)(bt_lo >> (8 * k));
    bt[4 + k] = (int8_t)(bt_hi >> (8 * k));
  }
  const float fwd = (float)bt[0] - (float)bt[1];
  const float str = (float)bt[3] - (float)bt[2];
  const bool jump = bt[4] != 0, brk = bt[5] != 0, place = bt[6] != 0;
  const float speed_mult = (bt[0] && bt[7]) ? p.sprint_mult : 1.0f;

  // Physics.
  float yaw = in.yaw + look_yaw;
  const float new_pitch = in.pitch + look_pitch;
  float pitch = fminf(fmaxf(new_pitch, -1.56f), 1.56f);
  float sin_yaw, cos_yaw, sin_pitch, cos_pitch;
  sincos_pair(yaw, pitch, sin_yaw, cos_yaw, sin_pitch, cos_pitch);
  const float speed = p.walk_speed * speed_mult;
  const float move_x = sin_yaw * fwd * speed * p.dt + cos_yaw * str * p.walk_speed * p.dt;
  const float move_z = cos_yaw * fwd * speed * p.dt - sin_yaw * str * p.walk_speed * p.dt;
  float vel_y = in.vy;
  bool og = (in.og & 0xffu) != 0u;
  if (jump && og) vel_y = p.jump_vel;
  vel_y += p.gravity * p.dt;
  float nx = in.px + move_x;
  float ny = in.py + vel_y * p.dt;
  float nz = in.pz + move_z;
  if (COL) {
    // Voxel collision (MineEnv(collision=True); no counterpart in the
    // reference, whose ground is the plane y = Y/4).  Solid = a block id >= 0
    // inside the world (layers above the staged ones count as air: an env
    // that could touch them is deferred to the full-height pass).
    //  1. x, z clamp to [1, dim - 1] first;
    //  2. the horizontal move is undone when the target column is solid at the
    //     old feet layer floor(py) or the one above;
    //  3. rising into a solid block at floor(ny + 1.8) stops: ny = py, vy = 0;
    //  4. the ground is the top of the highest solid voxel of the column at
    //     or below floor(py) (0 when there is none): ny <= ground lands there.
    nx = fminf(fmaxf(nx, 1.0f), (float)p.X - 1.0f);
    nz = fminf(fmaxf(nz, 1.0f), (float)p.Z - 1.0f);
    const int ylim = p.ys < p.Y ? p.ys : p.Y;
    auto solid = [&](int x, int y, int z) -> bool {
      if (!(((unsigned)x < (unsigned)p.X) & ((unsigned)y < (unsigned)ylim) &
            ((unsigned)z < (unsigned)p.Z)))
        return false;
      return world[(y * p.X + x) * p.Z + z] >= 0;
    };
    const int fy = (int)floorf(in.py);
    int cx = (int)floorf(nx), cz = (int)floorf(nz);
    if (solid(cx, fy, cz) || solid(cx, fy + 1, cz)) {
      nx = in.px;
      nz = in.pz;
      cx = (int)floorf(nx);
      cz = (int)floorf(nz);
    }
    if (vel_y > 0.0f && solid(cx, (int)floorf(ny + 1.8f), cz)) {
      ny = in.py;
      vel_y = 0.0f;
    }
    int g = 0;
    for (int y = min(fy, ylim - 1); y >= 0; --y)
      if (solid(cx, y, cz)) {
        g = y + 1;
        break;
      }
    if (ny <= (float)g) {
      ny = (float)g;
      vel_y = 0.0f;
      og = true;
    } else {
      og = false;
    }
  } else {
  const float ground_y = (float)(p.Y / 4);
  if (ny <= ground_y) {
    ny = ground_y;
    vel_y = 0.0f;
    og = true;
  } else {
    og = false;
  }
  nx = fminf(fmaxf(nx, 1.0f), (float)p.X - 1.0f);
  nz = fminf(fmaxf(nz, 1.0f), (float)p.Z - 1.0f);
  }

  // Staged height (Params::ys): the env runs here only when its layers
  // y >= ys are air, the raycast's samples (at most 3.975 from the eye) stay
  // below ys, and the rendered eye (the respawn point when done) lies in the
  // world below ys.  Otherwise nothing is written and the env is listed for
  // the full-height pass.
  if (p.ys < p.Y) {
    const float rx = done ? p.spawn_x : nx, rz = done ? p.spawn_z : nz;
    const float ry = (done ? spawn_y : ny) + 1.6f, ys = (float)p.ys;
    const bool ok = !upper_solid && (ny + 1.6f) + 4.5f < ys && rx >= 0.0f && rx < (float)p.X &&
                    ry >= 0.0f && ry < ys && rz >= 0.0f && rz < (float)p.Z;
    if (!ok) {
      if (lane == 0) {
        uint32_t* list = p.fb + (size_t)slot * (p.B + 2);
        const uint32_t k = atomicAdd(list, 1u);
        if (k < (uint32_t)p.B) list[2 + k] = (uint32_t)b;
      }
      Frame f;
      f.deferred = 1;
      return f;
    }
  }

  // Raycast: sample k of 32 on lane k, first solid in-bounds sample by ballot.
  float reward = 0.0f;
  if (brk || place) {
    const float dx = sin_yaw * cos_pitch, dy = sin_pitch, dz = cos_yaw * cos_pitch;
    const float ex = nx, ey = ny + 1.6f, ez = nz;
    bool inb = false, solid = false;
    int idx = 0;
    int8_t blk = kAir;
    if (lane < 32) {
      const float t = 0.1f + (float)lane * (4.0f / 32.0f);
      const int vx = (int)floorf(ex + dx * t);
      const int vy = (int)floorf(ey + dy * t);
      const int vz = (int)floorf(ez + dz * t);
      inb = ((unsigned)vx < (unsigned)p.X) & ((unsigned)vy < (unsigned)p.Y) &
            ((unsigned)vz < (unsigned)p.Z);
      if (inb) {
        idx = vy * p.X * p.Z + vx * p.Z + vz;
        blk = world[idx];
        solid = blk >= 0;
      }
    }
    const uint64_t hit = __ballot(solid);
    const uint64_t inb_mask = __ballot(inb);
    if (hit) {
      const int k = __ffsll((unsigned long long)hit) - 1;
      const int8_t hit_blk = (int8_t)__builtin_amdgcn_readlane((int)blk, k);  // k is uniform
      if (brk) {
        if (p.target_block < 0 || (int)hit_blk == p.target_block) reward = p.reward_value;
        if (lane == k) {
          world[idx] = kAir;
          if (!any_done) gworld[idx] = kAir;
        }
      } else {
        const uint64_t before = inb_mask & ((1ull << k) - 1ull);
        if (before) {
          const int pl = 63 - __clzll((long long)before);
          if (lane == pl) {
            world[idx] = kDirt;
            if (!any_done) gworld[idx] = kDirt;
          }
        }
      }
    }
  }

  // Masked respawn (after physics and raycast, as launch_step_uint8 orders it).
  if (done) {
    nx = p.spawn_x; ny = spawn_y; nz = p.spawn_z;
    vel_y = 0.0f;
    yaw = 0.0f;
    pitch = 0.0f;
    og = false;
  }
  if (lane == 0) {
    p.positions[b * 3 + 0] = nx;
    p.positions[b * 3 + 1] = ny;
    p.positions[b * 3 + 2] = nz;
    p.velocities[b * 3 + 1] = vel_y;
    if (done) {
      p.velocities[b * 3 + 0] = 0.0f;
      p.velocities[b * 3 + 2] = 0.0f;
    }
    p.yaws[b] = yaw;
    p.pitches[b] = pitch;
    p.on_ground[b] = og ? 1 : 0;
    p.rewards[b] = reward;
    if (p.done_mode == 0) {
      const int32_t nsc = done ? 1 : (int32_t)((uint32_t)sc + 1u);
      p.step_count[b] = nsc;
      p.do_reset[b] = done ? 1 : 0;
      if (p.dones) p.dones[b] = done ? 1 : 0;
      // Idempotent 0 -> 1 store: racing workgroups all write the same value,
      // and only a later launch reads it.  A plain store merges in each XCD's
      // L2 and is written back when the launch ends; an agent-scope atomic
      // store writes through, and 32K envs setting the same word in one step
      // serialised for about 0.7 ms.
      if (p.ring && episode_length >= 0 && nsc >= episode_length)
        p.ring[(slot + 1) % 3] = 1u;
    }
    if (p.forward_in) {
      p.forward_in[b] = fwd;
      p.strafe_in[b] = str;
      p.delta_yaw_in[b] = look_yaw;
      p.delta_pitch_in[b] = look_pitch;
      p.jump_in[b] = jump ? 1 : 0;
      p.do_break[b] = brk ? 1 : 0;
      p.do_place[b] = place ? 1 : 0;
      p.speed_mult_in[b] = speed_mult;
    }
    p.cameras[b * 5 + 0] = nx;
    p.cameras[b * 5 + 1] = ny;
    p.cameras[b * 5 + 2] = nz;
    p.cameras[b * 5 + 3] = yaw;
    p.cameras[b * 5 + 4] = pitch;
  }
  // Camera basis from the post-reset orientation (yaw = pitch = 0 after a
  // respawn, where sin/cos are exactly 0 and 1).
  if (done) {
    sin_yaw = 0.0f; cos_yaw = 1.0f; sin_pitch = 0.0f; cos_pitch = 1.0f;
  }
  return make_frame(p, nx, ny, nz, sin_yaw, cos_yaw, sin_pitch, cos_pitch, world);
}

// Both tables are read by every thread at clamped indices before either is
// stored: one memory round trip.
// flat: null in the kernels whose items keep the flat colours in their LDS
// lead (run_item).
__device__ __forceinline__ void load_palette(float* pal, uint32_t* flat, int tid) {
  const float pv = kPaletteTab[min(tid, kPaletteLen - 1)];
  const uint32_t fv = kFlatTab[min(tid, kFlatLen - 1)];
  if (tid < kPaletteLen) pal[tid] = pv;
  if (flat && tid < kFlatLen) flat[tid] = fv;
}

// Wave-0 logic of MineEnv.reset for env b (minecu/__init__.py:270-302):
// respawn when masked (reset_kernel, kernels.cu:1775-1785), step_count = 0,
// report "done on the next step" into the episode ring, refresh the camera.
__device__ __forceinline__ Frame reset_update(const Params& p, int slot, int b,
                                              const int8_t* world, float spawn_y) {
  const int lane = threadIdx.x & (kWave - 1);
  const bool r = p.mask ? (p.mask[b] != 0) : true;
  float x = p.positions[b * 3 + 0], y = p.positions[b * 3 + 1], z = p.positions[b * 3 + 2];
  float yaw = p.yaws[b], pitch = p.pitches[b];
  if (r) {
    x = p.spawn_x; y = spawn_y; z = p.spawn_z;
    yaw = 0.0f; pitch = 0.0f;
  }
  if (lane == 0) {
    p.do_reset[b] = r ? 1 : 0;
    if (r) {
      p.positions[b * 3 + 0] = x;
      p.positions[b * 3 + 1] = y;
      p.positions[b * 3 + 2] = z;
      p.velocities[b * 3 + 0] = 0.0f;
      p.velocities[b * 3 + 1] = 0.0f;
      p.velocities[b * 3 + 2] = 0.0f;
      p.yaws[b] = 0.0f;
      p.pitches[b] = 0.0f;
      p.on_ground[b] = 0;
      p.step_count[b] = 0;
    }
    const int32_t sc = r ? 0 : p.step_count[b];
    // Same idempotent flag store as the step's (read only by later launches).
    if (p.episode_length >= 0 && sc >= p.episode_length) p.ring[slot] = 1u;
    p.cameras[b * 5 + 0] = x;
    p.cameras[b * 5 + 1] = y;
    p.cameras[b * 5 + 2] = z;
    p.cameras[b * 5 + 3] = yaw;
    p.cameras[b * 5 + 4] = pitch;
  }
  float sy, cy, sp, cp;
  sincos_pair(yaw, pitch, sy, cy, sp, cp);
  return make_frame(p, x, y, z, sy, cy, sp, cp, world);
}

// Wave-0 camera of env b: from cameras [B,5], or from the agent state, which
// also refreshes cameras (MineEnv._render, minecu/__init__.py:446-466).
__device__ __forceinline__ Frame camera_update(const Params& p, int b, const int8_t* world,
                                               bool from_state) {
  const int lane = threadIdx.x & (kWave - 1);
  float x, y, z, yaw, pitch;
  if (from_state) {
    x = p.positions[b * 3 + 0]; y = p.positions[b * 3 + 1]; z = p.positions[b * 3 + 2];
    yaw = p.yaws[b]; pitch = p.pitches[b];
    if (lane == 0) {
      p.cameras[b * 5 + 0] = x;
      p.cameras[b * 5 + 1] = y;
      p.cameras[b * 5 + 2] = z;
      p.cameras[b * 5 + 3] = yaw;
      p.cameras[b * 5 + 4] = pitch;
    }
  } else {
    x = p.cameras[b * 5 + 0]; y = p.cameras[b * 5 + 1]; z = p.cameras[b * 5 + 2];
    yaw = p.cameras[b * 5 + 3]; pitch = p.cameras[b * 5 + 4];
  }
  float sy, cy, sp, cp;
  sincos_pair(yaw, pitch, sy, cy, sp, cp);
  return make_frame(p, x, y, z, sy, cy, sp, cp, world);
}

// ---------------------------------------------------------------------------
// Work items.  A workgroup of NT threads processes one item: epw = NT/(64 wpe)
// consecutive envs b0.. of one group, with wpe waves (64 wpe threads) per env.
// The envs of an item share the workgroup barriers; each has its own LDS world
// slot (world_stride bytes apart) and camera frame.
// ---------------------------------------------------------------------------
// Work-item kinds: a step, the reset render, a render from the [B, 5]
// cameras, a render from the agent state, a render from the [B, 14] basis.
enum ItemKind {
  kItemStep = 0,
  kItemResetRender = 1,
  kItemRenderCameras = 2,
  kItemRenderState = 3,
  kItemRenderBasis = 4
};
// Observation modes: shaded uint8, shaded float32, and the flat-colour uint8
// variants with the minimal kernel's and the prebasis kernel's image plane.
// Steps render the variant; the reset and state renders of the flat modes
// are shaded uint8, as the reference's _render() is (minecu/__init__.py:446-466).
enum ObsMode {
  kObsU8 = 0, kObsF32 = 1, kObsU8Minimal = 2, kObsU8Prebasis = 3, kObsU8Fp16 = 4,
  kObsU8Chw = 5, kObsF32Chw = 6  // the shaded images in planar [3, H, W] layout
};

__host__ __device__ __forceinline__ int world_stride_bytes(int X, int Y, int Z) {
  return (X * Y * Z + 15) & ~15;
}
// Item LDS: [lead][world 0] ... [world epw-1][trail].  The lead and the
// trail are at least one y-layer each, so the in-world DDA may read the voxel
// one step past its world's edge (which it then discards) without an address
// select; the tables that live there are written before the render and only
// ever read by such a walk as bytes it discards.
__host__ __device__ __forceinline__ int world_margin_bytes(int X, int Z) {
  return (X * Z + 15) & ~15;
}
__host__ __device__ __forceinline__ int round16(int n) { return (n + 15) & ~15; }
__host__ __device__ __forceinline__ bool uv_fits(int W, int H) {
  return W <= kMaxUV && H <= kMaxUV;
}
// The lead holds, when the image fits the u and v tables (fill_uv), either
// - for the shaded images: the v table (H floats), then per env of the item
//   the column terms of its camera rays (render_in's colA: W float pairs),
//   with u taken from its formula where a render needs it; or
// - for the other images: the u table (W floats), the v table (H floats),
//   then the flat colours (kFlatLen words, the flat images' colour table).
__host__ __device__ __forceinline__ int lead_bytes(int epw, int X, int Z, int W, int H) {
  const int uv = uv_fits(W, H) ? round16(4 * (W + H)) : 0;
  const int flat = uv + round16(4 * kFlatLen);
  const int col = uv_fits(W, H) ? round16(4 * H) + epw * W * 8 : 0;
  const int m = world_margin_bytes(X, Z);
  const int t = flat > col ? flat : col;
  return m > t ? m : t;
}
// The trail holds the per-wave tile buffers of the uint8 render (64 words
// per wave of the nt-thread workgroup).
__host__ __device__ __forceinline__ int trail_bytes(int nt, int X, int Z) {
  const int m = world_margin_bytes(X, Z), t = (nt / kWave) * kWave * 4;
  return m > t ? m : t;
}
__host__ __device__ __forceinline__ int item_lds_bytes(int nt, int epw, int X, int Y, int Z,
                                                       int W, int H) {
  return lead_bytes(epw, X, Z, W, H) + epw * world_stride_bytes(X, Y, Z) + trail_bytes(nt, X, Z);
}

// upper: one word per wave of the workgroup (staged-height check).
// CNT: count ray steps into p.ray_steps (the measurement build of the step).
// GLB: the env's world stays in global memory (worlds that do not fit LDS):
// one env per workgroup, nothing staged, and the world rewritten on a done
// step only after the whole workgroup has rendered.
// F32: a kernel for the float32 modes only (MC_OBS_F32, and with CHW
// MC_OBS_F32_CHW), whose render is render_env's IN32 form.
template <int NT, int KIND, bool LIFT, bool CHW = false, bool TER = false, bool COL = false,
          bool CNT = false, bool GLB = false, bool F32 = false>
__device__ __forceinline__ void run_item(const Params& p, int obs_mode, int b0, int wpe,
                                         int8_t* lds, const float* pal, const uint32_t* flat,
                                         Frame* frames, float* su, float* sv, int* upper) {
  // A 64-thread workgroup is one wave: one env per item, one wave per env
  // (a compile-time constant there, so every per-env stride folds).
  if (NT == kWave) wpe = 1;
  // The per-env phases before the render (loads, env logic, extent scan) are
  // serial latency on the item's path; their waves issue ahead of rendering
  // waves (wave priority, lowered again before the render).
  __builtin_amdgcn_s_setprio(3);
  const int tid = threadIdx.x, wave = tid / kWave;
  const int epw = (NT / kWave) / wpe;
  const int e = wave / wpe;                                 // env slot within the item
  const int et = tid - e * wpe * kWave;                     // thread index within the env
  const int ent = wpe * kWave;                              // threads per env
  const int b = b0 + e;
  const bool active = (e < epw) && (b < p.B);
  const int nvox = p.X * p.Y * p.Z;
  // Staged layers (Params::ys; always all of them outside step items).
  const bool staged = !GLB && (KIND == kItemStep) && p.ys < p.Y;
  const int nstage = staged ? p.X * p.ys * p.Z : nvox;
  int8_t* gworld = p.voxels + (int64_t)b * nvox;
  const int lead = lead_bytes(epw, p.X, p.Z, p.W, p.H);
  const int wstride = world_stride_bytes(p.X, staged ? p.ys : p.Y, p.Z);
  int8_t* world = GLB ? gworld : lds + lead + e * wstride;
  // Flat images for steps and camera renders of the flat modes; shaded
  // uint8 for their reset and state renders.
  const bool flat_kind = (KIND == kItemStep || KIND == kItemRenderCameras);
  const int uvm = (KIND == kItemRenderBasis) ? kUvRaw
                  : (flat_kind && (obs_mode == kObsU8Minimal || obs_mode == kObsU8Fp16))
                      ? kUvMinimal
                  : (flat_kind && obs_mode == kObsU8Prebasis) ? kUvBasis
                                                              : kUvShaded;
  const bool flat_img = (uvm != kUvShaded);
  // Shaded renders with u and v tables take their camera rays' column terms
  // from LDS (render_in's colA, filled below).
  const bool coltab = !GLB && uvm == kUvShaded && uv_tabled(p);
  // Tables in the lead (lead_bytes) and the per-wave tile buffers in the
  // trail (trail_bytes); a world in global memory keeps them in static LDS.
  uint32_t* tile;
  if constexpr (GLB) {
    __shared__ uint32_t s_tile[NT / kWave][kWave];
    tile = s_tile[threadIdx.x >> 6];
  } else {
    tile = reinterpret_cast<uint32_t*>(lds + lead + epw * wstride) + (tid >> 6) * kWave;
    if (coltab) {
      su = nullptr;
      sv = reinterpret_cast<float*>(lds);
    } else {
      su = reinterpret_cast<float*>(lds);
      sv = su + p.W;
    }
    if (flat_img) {
      uint32_t* fl = reinterpret_cast<uint32_t*>(lds + (uv_tabled(p) ? round16(4 * (p.W + p.H)) : 0));
      const uint32_t fv = kFlatTab[min(tid, kFlatLen - 1)];
      if (tid < kFlatLen) fl[tid] = fv;
      flat = fl;
    }
  }
  const int slot = p.slot_dev ? (int)*p.slot_dev : p.slot;
  // The flag word is read here and tested after the world is staged, so its
  // load overlaps the env's other loads.
  uint32_t ring_word = 0u;
  if (KIND == kItemStep && p.done_mode == 0 && p.ring) {
    ring_word = p.ring[slot];
    if (b == 0 && et == 0) p.ring[(slot + 2) % 3] = 0u;
  }
  const bool regen = (KIND == kItemResetRender) && (p.ring[3] != 0u);
  // Every lane loads (env 0's inputs when its env is past the batch): no
  // divergent region whose end would wait for the loads.
  EnvIn in;
  if (KIND == kItemStep) in = load_env_in(p, active ? b : 0);
  fill_uv(p, uvm, su, sv, tid, NT);
  if (active) {
    if (regen) store_fresh_world<TER>(p, b, gworld, GLB ? nullptr : world, nvox, et, ent);
    else if (!GLB) stage_world(world, gworld, nstage, et, ent);
  }
  if (staged) {
    const bool found = active && any_not_air(gworld + nstage, nvox - nstage, et, ent);
    const bool w = __ballot(found) != 0ull;
    if ((tid & (kWave - 1)) == 0) upper[wave] = w ? 1 : 0;
  }
  __syncthreads();
  const bool any_done = ring_word != 0u;
  if (active && et < kWave) {
    bool upper_solid = false;
    if (staged)
      for (int k = 0; k < wpe; ++k) upper_solid |= upper[e * wpe + k] != 0;
    Frame f;
    if (KIND == kItemStep)
      f = env_update<COL>(p, slot, b, world, gworld, any_done, in, upper_solid, spawn_of<TER>(p, b));
    else if (KIND == kItemResetRender) f = reset_update(p, slot, b, world, spawn_of<TER>(p, b));
    else if (KIND == kItemRenderBasis) f = frame_from_basis(p, b, world);
    else f = camera_update(p, b, world, KIND == kItemRenderState);
    if (et == 0) {
      // The solid-extent scan reads 16-byte chunks: a world in global memory
      // must be 16-byte aligned for it.
      if (GLB && ((uintptr_t)world & 15u)) f.lift = 0;
      frames[e] = f;
      if (KIND == kItemStep && p.basis && !f.deferred) store_basis(p, b, f);
    }
  }
  __syncthreads();
  const bool go = active && !frames[e].deferred;
  // Column terms of the env's camera rays for render_in (shaded renders
  // with u and v tables), published by the barrier below.
  float2* colA = coltab ? reinterpret_cast<float2*>(lds + round16(4 * p.H)) + e * p.W : nullptr;
  if (coltab && go) {
    const float fx = frames[e].fx, fz = frames[e].fz, rx = frames[e].rx, rz = frames[e].rz;
    for (int px = et; px < p.W; px += ent) {
      const float u = u_of(p, kUvShaded, px);
      colA[px] = make_float2(fx + u * rx, fz + u * rz);
    }
  }
  if (LIFT) {
    if (go && frames[e].lift)
      solid_extent(p, world, &frames[e], (int)floorf(frames[e].ey), et, ent);
    __syncthreads();
  } else if (coltab) {
    __syncthreads();
  }
  // An eye inside a solid block sees that block in every pixel: a constant
  // uint8 image, written without the camera frame or the octant words.
  const int npix = p.W * p.H;
  int fill = -1;
  // CHW kernels: only the planar uint8 image has a block fill (whole planes).
  // Float32 images fill with float stores (16-byte aligned images only).  The
  // interleaved uint8 step kernel never sees a float32 mode (those run the F32
  // and CHW kernels), so it compiles without that path.
  constexpr bool kMayF32 = F32 || CHW || TER || COL || GLB || KIND != kItemStep;
  const bool f32_img =
      kMayF32 && !flat_img && (obs_mode == kObsF32 || obs_mode == kObsF32Chw);
  const bool can_fill = f32_img ? ((npix & 3) == 0) & (((uintptr_t)p.obs & 15u) == 0)
                                : (((npix * 3) & 15) == 0) & (!CHW | ((npix & 15) == 0));
  if (go && can_fill)
    fill = __builtin_amdgcn_readfirstlane(flat_img ? frames[e].flat_fill : frames[e].fill_block);
  uint32_t acc[3] = {0u, 0u, 0u};
  __builtin_amdgcn_s_setprio(0);
  if (fill >= 0) {
    if (CNT && et == 0) acc[0] += npix, acc[2] += npix;
    if (f32_img) {
      // trace()'s colour for the eye's block at t = 0, face -1: col * 1 * 1 + sky * 0.
      float* out = reinterpret_cast<float*>(p.obs) + (int64_t)b * npix * 3;
      fill_f32(out, npix, pal[fill * 3], pal[fill * 3 + 1], pal[fill * 3 + 2], CHW, et, ent);
    } else {
      const uint32_t rgb = flat_img ? flat[fill < kFlatSky ? fill : kFlatSky]
                                    : to_u8(pal[fill * 3]) | (to_u8(pal[fill * 3 + 1]) << 8) |
                                          (to_u8(pal[fill * 3 + 2]) << 16);
      uint8_t* out = reinterpret_cast<uint8_t*>(p.obs) + (int64_t)b * npix * 3;
      if (CHW) fill_planes8(out, npix, rgb, et, ent);
      else fill_rgb8(out, npix, rgb, et, ent);
    }
    if (!GLB && KIND == kItemStep && any_done)
      store_fresh_world<TER>(p, b, gworld, nullptr, nvox, et, ent);
  } else if (go) {
    // Every wave of the env writes the same 8 octant words before it reads
    // them; a wave's LDS operations complete in order.
    if ((et & (kWave - 1)) < 8)
      frames[e].pos0[et & (kWave - 1)] = oct_pos<LIFT>(p, frames[e], et & (kWave - 1));
    __builtin_amdgcn_wave_barrier();
    const Frame f = uniform_frame(&frames[e]);
    const OctPos op{frames[e].pos0};
    const bool f32 = obs_mode == (CHW ? kObsF32Chw : kObsF32);
    if (F32 && p.vd_far)
      render_env<1, LIFT, true, CNT, GLB, true>(p, b, world, pal, f, op, et, ent, su, sv, CHW, acc,
                                                tile, colA);
    else if (F32)
      render_env<1, LIFT, false, CNT, GLB, true>(p, b, world, pal, f, op, et, ent, su, sv, CHW,
                                                 acc, tile, colA);
    else if (uvm != kUvShaded && obs_mode == kObsU8Fp16)
      render_env_flat<LIFT, _Float16, CNT, GLB>(p, b, world, flat, f, op, uvm, et, ent, su, sv,
                                                acc, tile);
    else if (uvm != kUvShaded)
      render_env_flat<LIFT, float, CNT, GLB>(p, b, world, flat, f, op, uvm, et, ent, su, sv, acc,
                                             tile);
    else if (f32 && p.vd_far)
      render_env<1, LIFT, true, CNT, GLB>(p, b, world, pal, f, op, et, ent, su, sv, CHW, acc);
    else if (f32)
      render_env<1, LIFT, false, CNT, GLB>(p, b, world, pal, f, op, et, ent, su, sv, CHW, acc);
    else if (p.vd_far)
      render_env<0, LIFT, true, CNT, GLB>(p, b, world, pal, f, op, et, ent, su, sv, CHW, acc,
                                          tile, colA);
    else
      render_env<0, LIFT, false, CNT, GLB>(p, b, world, pal, f, op, et, ent, su, sv, CHW, acc,
                                           tile, colA);
    if (!GLB && KIND == kItemStep && any_done)
      store_fresh_world<TER>(p, b, gworld, nullptr, nvox, et, ent);
  }
  if (GLB && KIND == kItemStep && any_done) {
    // Every wave of the env has rendered from the world in global memory:
    // only now may it be rewritten (one env per workgroup, so the barrier
    // covers the env's waves; any_done is the same for every workgroup).
    __syncthreads();
    if (active && !frames[e].deferred) store_fresh_world<TER>(p, b, gworld, nullptr, nvox, et, ent);
  }
  if (CNT && p.ray_steps) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      uint32_t v = acc[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if ((tid & (kWave - 1)) == 0 && v) atomicAdd(&p.ray_steps[k], (unsigned long long)v);
    }
  }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// One group, one item per workgroup: grid = ceil(B / epw).  CHW: the planar
// image modes, in kernels of their own.  CNT: the measurement build of the
// step, counting ray steps into p.ray_steps.
// Register caps: gfx950 runs 7 waves per SIMD only when a wave holds at most
// 96 scalar registers in total (measured: 94 gives 7, 102 gives 6, whatever
// the compiler's occupancy estimate says), so the step kernels are held to 96
// (the excess spills to vector-register lanes, as before); the 256-thread
// step kernel is also held to 72 vector registers (7 waves per SIMD, 7
// workgroups per CU at 32^3), with no scratch.
template <int NT, int KIND, bool LIFT, bool CHW = false, bool CNT = false, bool F32 = false>
__global__ __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu((NT == 256 && KIND == kItemStep) ? 7 : 1)))
__attribute__((amdgpu_num_sgpr(96)))
void k_env(Params p, int obs_mode, int wpe) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_upper[NT / kWave];
  load_palette(s_pal, nullptr, threadIdx.x);
  const int epw = (NT / kWave) / wpe;
  run_item<NT, KIND, LIFT, CHW, false, false, CNT, false, F32>(
      p, obs_mode, blockIdx.x * epw, wpe, s_lds, s_pal, nullptr, s_frames, nullptr, nullptr,
      s_upper);
}

// Full-height pass after a staged-height step launch (Params::ys): the envs
// that launch deferred, listed in fb[slot * (B + 2)], one env per workgroup
// (all NT / 64 waves on it).  Launched with p.ys = Y; the list is complete
// because the step launch precedes this one on the stream.  The last
// workgroup to finish empties the list for the next launch on this slot.
template <int NT, bool LIFT, bool CHW = false>
__global__ __launch_bounds__(NT) void k_env_fb(Params p, int obs_mode) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_upper[NT / kWave];
  const int slot = p.slot_dev ? (int)*p.slot_dev : p.slot;
  uint32_t* list = p.fb + (size_t)slot * (p.B + 2);
  const int n = (int)min(list[0], (uint32_t)p.B);
  if (n == 0) return;  // nothing deferred: the list is already empty
  load_palette(s_pal, nullptr, threadIdx.x);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    __syncthreads();  // the previous env is done with LDS
    run_item<NT, kItemStep, LIFT, CHW>(p, obs_mode, (int)list[2 + i], NT / kWave, s_lds, s_pal,
                                       nullptr, s_frames, nullptr, nullptr, s_upper);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&list[1], 1u) == gridDim.x - 1) {
      list[0] = 0u;
      list[1] = 0u;
      __threadfence();
    }
  }
}

// k_env and k_env_fb for the seeded terrain world (interleaved images only):
// kernels of their own, so the flat-world kernels keep their code.
template <int NT, int KIND, bool LIFT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_num_sgpr(96)))
void k_env_terrain(Params p, int obs_mode, int wpe) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_upper[NT / kWave];
  load_palette(s_pal, nullptr, threadIdx.x);
  const int epw = (NT / kWave) / wpe;
  run_item<NT, KIND, LIFT, false, true>(p, obs_mode, blockIdx.x * epw, wpe, s_lds, s_pal, nullptr,
                                        s_frames, nullptr, nullptr, s_upper);
}

template <int NT, bool LIFT>
__global__ __launch_bounds__(NT) void k_env_fb_terrain(Params p, int obs_mode) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_upper[NT / kWave];
  const int slot = p.slot_dev ? (int)*p.slot_dev : p.slot;
  uint32_t* list = p.fb + (size_t)slot * (p.B + 2);
  const int n = (int)min(list[0], (uint32_t)p.B);
  if (n == 0) return;
  load_palette(s_pal, nullptr, threadIdx.x);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    __syncthreads();
    run_item<NT, kItemStep, LIFT, false, true>(p, obs_mode, (int)list[2 + i], NT / kWave, s_lds,
                                               s_pal, nullptr, s_frames, nullptr, nullptr, s_upper);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&list[1], 1u) == gridDim.x - 1) {
      list[0] = 0u;
      list[1] = 0u;
      __threadfence();
    }
  }
}

// Step kernels with voxel collision (flat or seeded terrain world, TER).
template <int NT, bool LIFT, bool TER>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_num_sgpr(96)))
void k_env_col(Params p, int obs_mode, int wpe) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_upper[NT / kWave];
  load_palette(s_pal, nullptr, threadIdx.x);
  const int epw = (NT / kWave) / wpe;
  run_item<NT, kItemStep, LIFT, false, TER, true>(p, obs_mode, blockIdx.x * epw, wpe, s_lds,
                                                  s_pal, nullptr, s_frames, nullptr, nullptr, s_upper);
}

template <int NT, bool LIFT, bool TER>
__global__ __launch_bounds__(NT) void k_env_fb_col(Params p, int obs_mode) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_upper[NT / kWave];
  const int slot = p.slot_dev ? (int)*p.slot_dev : p.slot;
  uint32_t* list = p.fb + (size_t)slot * (p.B + 2);
  const int n = (int)min(list[0], (uint32_t)p.B);
  if (n == 0) return;
  load_palette(s_pal, nullptr, threadIdx.x);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    __syncthreads();
    run_item<NT, kItemStep, LIFT, false, TER, true>(p, obs_mode, (int)list[2 + i], NT / kWave,
                                                    s_lds, s_pal, nullptr, s_frames, nullptr, nullptr,
                                                    s_upper);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&list[1], 1u) == gridDim.x - 1) {
      list[0] = 0u;
      list[1] = 0u;
      __threadfence();
    }
  }
}

// Worlds too large for one workgroup's LDS: the env's world stays in global
// memory, one env per 256-thread workgroup (four waves), the upward-ray exit
// built in.  LIST: the envs a staged-height step deferred (fb list, as
// k_env_fb), else env blockIdx.x.
template <int KIND, bool CHW, bool TER, bool COL, bool LIST>
__global__ __launch_bounds__(256) void k_env_glb(Params p, int obs_mode) {
  __shared__ float s_pal[kPaletteLen];
  __shared__ uint32_t s_flat[kFlatLen];
  __shared__ Frame s_frames[4];
  __shared__ float s_u[kMaxUV], s_v[kMaxUV];
  __shared__ int s_upper[4];
  load_palette(s_pal, s_flat, threadIdx.x);
  if (!LIST) {
    run_item<256, KIND, true, CHW, TER, COL, false, true>(p, obs_mode, blockIdx.x, 4, nullptr,
                                                          s_pal, s_flat, s_frames, s_u, s_v,
                                                          s_upper);
    return;
  }
  const int slot = p.slot_dev ? (int)*p.slot_dev : p.slot;
  uint32_t* list = p.fb + (size_t)slot * (p.B + 2);
  const int n = (int)min(list[0], (uint32_t)p.B);
  if (n == 0) return;  // nothing deferred: the list is already empty
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    __syncthreads();  // the previous env is done with the shared frame
    run_item<256, kItemStep, true, CHW, TER, COL, false, true>(p, obs_mode, (int)list[2 + i], 4,
                                                               nullptr, s_pal, s_flat, s_frames,
                                                               s_u, s_v, s_upper);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&list[1], 1u) == gridDim.x - 1) {
      list[0] = 0u;
      list[1] = 0u;
      __threadfence();
    }
  }
}

// Several env groups (uneven batch sizes, world sizes, resolutions) in one
// persistent launch.  Items are numbered group by group (groups ordered by the
// host, most expensive first); each workgroup takes the next item from an
// atomic counter until none is left, and the last workgroup to finish resets
// the counters for the next launch.
constexpr int kMaxGroups = 64;

struct MultiEntry {
  const Params* desc;      // static per-group parameters (device memory)
  const void* actions;
  const float* look;
  uint8_t* dones;
  int action_kind;
  int slot;
  int item_begin;          // first global item of this group
  int wpe;
  int obs_mode;
  int pad;
};

struct MultiArgs {
  int n_groups;
  int total_items;
  uint32_t* counters;          // [2] device words, zero between launches
  const MultiEntry* entries;   // [n_groups] device memory
};

static_assert(sizeof(float) * (kPaletteLen + 2 * kMaxUV) + sizeof(uint32_t) * kFlatLen +
                      sizeof(Frame) * (512 / kWave) + sizeof(int) * (kMaxGroups + 1 + 512 / kWave) +
                      sizeof(uint32_t) * 512 + 64 <=
                  kStaticLdsReserve,
              "static LDS of the env kernels exceeds its reserve");

// PERSIST: a grid sized to the resident capacity whose workgroups loop over
// items taken from an atomic counter; otherwise one workgroup per item (the
// loop is compiled out; the hardware dispatcher then balances the load, and
// the kernel needs fewer registers: 72 VGPRs against 111 on gfx950).
// CHW / TER / COL: a launch over groups of one mode family (planar images,
// seeded terrain, voxel collision), as the single-group kernels separate them.
// Register caps as for k_env: 96 scalar registers; the one-item form also 72
// vector registers (7 waves per SIMD).
template <int NT, bool LIFT, bool PERSIST, bool CHW = false, bool TER = false, bool COL = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_num_sgpr(96)))
__attribute__((amdgpu_waves_per_eu(PERSIST ? 1 : 7)))
void k_multi(MultiArgs a) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_item;
  __shared__ int s_upper[NT / kWave];
  // The entry table and the descriptors are written by the host before the
  // launch and only read here: constant address space, so uniform reads are
  // scalar loads.
  typedef __attribute__((address_space(4))) const MultiEntry CEntry;
  typedef __attribute__((address_space(4))) const uint32_t CWord;
  CEntry* entries = (CEntry*)a.entries;
  load_palette(s_pal, nullptr, threadIdx.x);
  // First item of group `lane` (past the last group: never <= an item).
  const int lane = threadIdx.x & (kWave - 1);
  const int my_begin = (lane < a.n_groups) ? entries[lane].item_begin : 0x7fffffff;
  int item = blockIdx.x;
  for (;;) {
    if (PERSIST) {
      __syncthreads();  // the previous item is done with LDS and s_item
      if (threadIdx.x == 0) s_item = (int)atomicAdd(&a.counters[0], 1u);
    }
    __syncthreads();
    if (PERSIST) item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= a.total_items) break;
    // Items are numbered group by group: the item's group is the number of
    // groups that begin at or before it, less one.
    const int g = __builtin_amdgcn_readfirstlane(__popcll(__ballot(my_begin <= item)) - 1);
    CEntry& me = entries[g];
    // The descriptor is uniform over the workgroup: scalar loads into scalar
    // registers (the vector file is needed by the ray loops).
    Params p;
    {
      CWord* src = (CWord*)me.desc;
      uint32_t* dst = reinterpret_cast<uint32_t*>(&p);
#pragma unroll
      for (int i = 0; i < (int)(sizeof(Params) / 4); ++i) dst[i] = src[i];
    }
    p.actions = me.actions;
    p.look = me.look;
    p.dones = me.dones;
    p.action_kind = me.action_kind;
    p.slot = me.slot;
    const int wpe = me.wpe;
    const int epw = (NT / kWave) / wpe;
    run_item<NT, kItemStep, LIFT, CHW, TER, COL>(p, me.obs_mode, (item - me.item_begin) * epw, wpe,
                                                 s_lds, s_pal, nullptr, s_frames, nullptr, nullptr, s_upper);
    if (!PERSIST) break;
  }
  if (PERSIST && threadIdx.x == 0) {
    __threadfence();
    const uint32_t prev = atomicAdd(&a.counters[1], 1u);
    if (prev == gridDim.x - 1) {
      a.counters[0] = 0u;
      a.counters[1] = 0u;
      __threadfence();
    }
  }
}

// Full-height pass after a batched launch whose groups stage fewer layers
// (Params::ys, fb in their descriptors): group by group, the envs each group
// deferred, one env per workgroup with its whole world in LDS (the launch's
// LDS fits the largest).  The last workgroup done with a group's list empties
// it for the next launch on that slot.
template <int NT, bool LIFT, bool CHW = false, bool TER = false, bool COL = false>
__global__ __launch_bounds__(NT) void k_multi_fb(MultiArgs a) {
  extern __shared__ __align__(16) int8_t s_lds[];
  __shared__ float s_pal[kPaletteLen];
  __shared__ Frame s_frames[NT / kWave];
  __shared__ int s_upper[NT / kWave];
  typedef __attribute__((address_space(4))) const MultiEntry CEntry;
  typedef __attribute__((address_space(4))) const uint32_t CWord;
  CEntry* entries = (CEntry*)a.entries;
  bool palette = false;
  for (int g = 0; g < a.n_groups; ++g) {
    CEntry& me = entries[g];
    Params p;
    {
      CWord* src = (CWord*)me.desc;
      uint32_t* dst = reinterpret_cast<uint32_t*>(&p);
#pragma unroll
      for (int i = 0; i < (int)(sizeof(Params) / 4); ++i) dst[i] = src[i];
    }
    if (!p.fb || p.ys >= p.Y) continue;
    uint32_t* list = p.fb + (size_t)me.slot * (p.B + 2);
    const int n = (int)min(list[0], (uint32_t)p.B);
    if (n == 0) continue;
    if (!palette) {
      load_palette(s_pal, nullptr, threadIdx.x);
      palette = true;
    }
    p.actions = me.actions;
    p.look = me.look;
    p.dones = me.dones;
    p.action_kind = me.action_kind;
    p.slot = me.slot;
    p.ys = p.Y;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
      __syncthreads();  // the previous env is done with LDS
      run_item<NT, kItemStep, LIFT, CHW, TER, COL>(p, me.obs_mode, (int)list[2 + i], NT / kWave,
                                                   s_lds, s_pal, nullptr, s_frames, nullptr, nullptr,
                                                   s_upper);
    }
    if (threadIdx.x == 0) {
      __threadfence();
      if (atomicAdd(&list[1], 1u) == gridDim.x - 1) {
        list[0] = 0u;
        list[1] = 0u;
        __threadfence();
      }
    }
  }
}

}  // namespace mc
