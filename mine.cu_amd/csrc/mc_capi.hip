// mc_capi.hip - C ABI of libminecu_hip.so (declared in include/minecu_hip.h):
// argument validation, launch geometry and the per-group env handle.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "mc_launch.h"
#include "mc_util_kernels.h"
#include "minecu_hip.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MC_E_HIP, "%s: %s", what, hipGetErrorString(e));
  return MC_OK;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// World bytes one workgroup may stage in LDS on the current device: the
// device's per-workgroup LDS maximum less the kernels' static reserve, capped
// by the compiled budget.  Queried once per device; without a device the
// compiled budget applies (validation only, nothing is launched then).
int lds_budget() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return mc::kMaxLdsWorld;
  }
  int v = cache[dev].load(std::memory_order_relaxed);
  if (v > 0) return v;
  int mx = 0;
  if (hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
      mx <= mc::kStaticLdsReserve) {
    (void)hipGetLastError();
    return mc::kMaxLdsWorld;
  }
  v = std::min(mc::kMaxLdsWorld, mx - mc::kStaticLdsReserve);
  cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

int check_dims(int B, int X, int Y, int Z) {
  if (B < 1) return fail(MC_E_INVALID, "batch_size must be >= 1 (got %d)", B);
  if (X < 1 || Y < 1 || Z < 1)
    return fail(MC_E_INVALID, "world dims must be >= 1 (got %d x %d x %d)", X, Y, Z);
  // Worlds that do not fit one workgroup's LDS run from global memory; the
  // in-world walk packs each coordinate into 10 bits.
  if (X > mc::kMaxInsideDim || Y > mc::kMaxInsideDim || Z > mc::kMaxInsideDim)
    return fail(MC_E_UNSUPPORTED, "world %dx%dx%d: each dimension must be at most %d", X, Y, Z,
                mc::kMaxInsideDim);
  return MC_OK;
}

int check_image(int B, int W, int H) {
  if (W < 1 || H < 1) return fail(MC_E_INVALID, "render size must be >= 1 (got %dx%d)", W, H);
  if ((int64_t)W * H * 3 > (int64_t)1 << 30)
    return fail(MC_E_INVALID, "render size %dx%d too large", W, H);
  (void)B;
  return MC_OK;
}

int check_collision(const mc_config* c) {
  if (c->collision != 0 && c->collision != 1)
    return fail(MC_E_INVALID, "collision must be 0 or 1 (got %d)", c->collision);
  if (c->collision && (c->obs_mode == MC_OBS_U8_CHW || c->obs_mode == MC_OBS_F32_CHW))
    return fail(MC_E_UNSUPPORTED, "collision renders the interleaved obs modes only");
  return MC_OK;
}

int check_terrain(const mc_config* c) {
  if (c->terrain == MC_TERRAIN_FLAT) return check_collision(c);
  if (c->terrain != MC_TERRAIN_NOISE)
    return fail(MC_E_INVALID, "terrain must be one of MC_TERRAIN_* (got %d)", c->terrain);
  if (c->terrain_amplitude < 0 || c->terrain_amplitude > 8)
    return fail(MC_E_INVALID, "terrain_amplitude must be in [0, 8] (got %d)", c->terrain_amplitude);
  if (c->ground_height - c->terrain_amplitude < 2)
    return fail(MC_E_INVALID, "ground_height - terrain_amplitude must be >= 2 (got %d - %d)",
                c->ground_height, c->terrain_amplitude);
  if (c->obs_mode == MC_OBS_U8_CHW || c->obs_mode == MC_OBS_F32_CHW)
    return fail(MC_E_UNSUPPORTED, "MC_TERRAIN_NOISE renders the interleaved obs modes only");
  return check_collision(c);
}


int validate_config(const mc_config* c) {
  if (!c) return fail(MC_E_INVALID, "config is NULL");
  int r = check_dims(c->batch_size, c->world_x, c->world_y, c->world_z);
  if (r) return r;
  r = check_image(c->batch_size, c->width, c->height);
  if (r) return r;
  if (c->obs_mode < MC_OBS_U8 || c->obs_mode > MC_OBS_F32_CHW)
    return fail(MC_E_INVALID, "obs_mode must be one of MC_OBS_* (got %d)", c->obs_mode);
  return check_terrain(c);
}

int validate_state(const mc_state* s) {
  if (!s) return fail(MC_E_INVALID, "state is NULL");
  if (!s->voxels || !s->positions || !s->velocities || !s->yaws || !s->pitches ||
      !s->on_ground || !s->step_count || !s->do_reset || !s->cameras || !s->obs || !s->rewards)
    return fail(MC_E_INVALID, "every state pointer must be non-NULL");
  const int ctrl = !!s->forward_in + !!s->strafe_in + !!s->delta_yaw_in + !!s->delta_pitch_in +
                   !!s->jump_in + !!s->do_break + !!s->do_place + !!s->speed_mult;
  if (ctrl != 0 && ctrl != 8)
    return fail(MC_E_INVALID, "the decode buffers must be all NULL or all non-NULL");
  return MC_OK;
}

mc::Params make_params(const mc_config* c, const mc_state* s) {
  mc::Params p;
  memset(&p, 0, sizeof p);
  p.B = c->batch_size;
  p.X = c->world_x;
  p.Y = c->world_y;
  p.Z = c->world_z;
  p.W = c->width;
  p.H = c->height;
  p.max_steps = c->max_steps;
  p.view_distance = c->view_distance;
  p.half_fov = mc_half_fov_tan(c->fov_degrees);
  p.aspect = (float)c->width / (float)c->height;
  p.flat_hf = mc_half_fov_tan_flat(c->fov_degrees);
  p.flat_su = p.flat_hf * p.aspect;
  {
    int e = 0;
    const float m = frexpf(c->view_distance, &e);
    p.vd_pow2 = (c->view_distance > 0.0f && m == 0.5f && e - 1 >= -126 && e - 1 <= 126) ? 1 : 0;
    p.inv_vd = p.vd_pow2 ? ldexpf(1.0f, 1 - e) : 0.0f;
  }
  {
    const double d2 = (double)c->world_x * c->world_x + (double)c->world_y * c->world_y +
                      (double)c->world_z * c->world_z;
    const double vd = c->view_distance;
    p.vd_far = (vd > 0.0 && d2 * 1.001 < vd * vd) ? 1 : 0;
  }
  p.dt = c->dt;
  p.gravity = c->gravity;
  p.walk_speed = c->walk_speed;
  p.sprint_mult = c->sprint_mult;
  p.jump_vel = c->jump_vel;
  // The reference narrows target_block to int8 (bindings.cpp:531).
  p.target_block = (int)(int8_t)c->target_block;
  p.reward_value = c->reward_value;
  p.spawn_x = c->spawn_x;
  p.spawn_y = c->spawn_y;
  p.spawn_z = c->spawn_z;
  p.ground_height = c->ground_height;
  p.terrain = c->terrain == MC_TERRAIN_NOISE ? 1 : 0;
  p.terrain_seed = c->seed;
  p.terrain_amp = p.terrain ? c->terrain_amplitude : 0;
  p.collision = c->collision ? 1 : 0;
  p.tree_x = c->tree_x;
  p.tree_z = c->tree_z;
  p.house_x = c->house_x;
  p.house_z = c->house_z;
  p.episode_length = c->episode_length;
  p.look_delta = c->look_delta;
  if (s) {
    p.voxels = s->voxels;
    p.positions = s->positions;
    p.velocities = s->velocities;
    p.yaws = s->yaws;
    p.pitches = s->pitches;
    p.on_ground = s->on_ground;
    p.step_count = s->step_count;
    p.do_reset = s->do_reset;
    p.cameras = s->cameras;
    p.obs = s->obs;
    p.rewards = s->rewards;
    p.basis = s->basis;
    p.forward_in = s->forward_in;
    p.strafe_in = s->strafe_in;
    p.delta_yaw_in = s->delta_yaw_in;
    p.delta_pitch_in = s->delta_pitch_in;
    p.jump_in = s->jump_in;
    p.do_break = s->do_break;
    p.do_place = s->do_place;
    p.speed_mult_in = s->speed_mult;
  }
  p.ys = c->world_y;
  p.fb = nullptr;
  return p;
}

// Launch shape of the per-env work: NT threads per workgroup, wpe waves per
// env, so one workgroup holds epw = NT / (64 wpe) envs.
// lift: build the upward-ray early exit into the kernel (pays off when rays
// climb through tall air columns; costs a scan and registers otherwise).
struct Shape {
  int nt;
  int wpe;
  int lift;
};

bool valid_shape(Shape s) {
  return (s.nt == 64 || s.nt == 256 || s.nt == 512) &&
         (s.wpe == 1 || s.wpe == 2 || s.wpe == 4 || s.wpe == 8) && s.wpe * 64 <= s.nt &&
         (s.lift == 0 || s.lift == 1);
}

// Default before any tuning: one wave per env up to 1024 pixels, four above,
// raised until the workgroups that fit one CU's LDS by their world hold at
// least 16 waves (one env per workgroup); the upward-ray exit on.  Measured on
// MI355X, env-steps/s at (64,1) / (256,4) / (512,8): 16^3 32x32 144M / 109M /
// 69M; 32^3 32x32 17.5M / 38.1M / 35.2M; 48^3 32x32 3.0M / 8.1M / 11.1M.  The
// upward-ray exit: at 16^3 / 32x32 falling-phase steps 20% faster, landed
// steps 7% slower, 7% overall; at 32^3 / 64x64 1.8x.
Shape default_shape(int W, int H, int X, int Y, int Z) {
  int wpe = ((int64_t)W * H <= 1024) ? 1 : 4;
  const int per_cu =
      std::max(1, mc::kLdsPerWorkgroup / mc::item_lds_bytes(mc::kWave, 1, X, Y, Z, W, H));
  while (wpe < 8 && per_cu * wpe < 16) wpe = (wpe == 1) ? 4 : 8;
  return Shape{wpe * mc::kWave, wpe, 1};
}

Shape default_shape(const mc_config& c) {
  return default_shape(c.width, c.height, c.world_x, c.world_y, c.world_z);
}

int epw_of(Shape s) { return (s.nt / mc::kWave) / s.wpe; }

// Raise wpe until the envs of one workgroup fit the LDS budget.
Shape fit_lds(Shape s, int X, int Y, int Z, int W, int H) {
  while (mc::item_lds_bytes(s.nt, epw_of(s), X, Y, Z, W, H) > lds_budget() && s.wpe * 64 < s.nt)
    s.wpe *= 2;
  return s;
}

bool planar_mode(int obs_mode) {
  return obs_mode == MC_OBS_U8_CHW || obs_mode == MC_OBS_F32_CHW;
}

// Staged layers below Y (Params::ys) apply to step items with a deferred-env
// list only; every other launch stages whole worlds.
// count: the ray-step counting build of the step (flat world, interleaved
// obs, p.ray_steps set).
int launch_env(int kind, const mc::Params& q, int obs_mode, Shape s, hipStream_t st,
               bool count = false) {
  mc::Params p = q;
  if (kind != mc::kItemStep || !p.fb || p.ys < 1 || p.ys > p.Y) {
    p.ys = p.Y;
    p.fb = nullptr;
  }
  // Planar images come from MineEnv's own step, reset and render only.
  if (planar_mode(obs_mode) && kind != mc::kItemStep && kind != mc::kItemResetRender &&
      kind != mc::kItemRenderState)
    return fail(MC_E_UNSUPPORTED, "the MC_OBS_*_CHW modes render from MineEnv state only");
  s = fit_lds(s, p.X, p.ys, p.Z, p.W, p.H);
  if (mc::item_lds_bytes(s.nt, epw_of(s), p.X, p.ys, p.Z, p.W, p.H) > lds_budget()) {
    // Too large for LDS even at one env per workgroup: the world stays in
    // global memory (whole height; no counting build there).
    if (count) return fail(MC_E_UNSUPPORTED, "ray-step counting needs an LDS-resident world");
    p.ys = p.Y;
    p.fb = nullptr;
    const hipError_t e =
        mcl::kernels(256, true).glb(kind, p, obs_mode, false, dim3(p.B), st);
    if (e != hipSuccess) return fail(MC_E_HIP, "global-world launch: %s", hipGetErrorString(e));
    return MC_OK;
  }
  const size_t lds = (size_t)mc::item_lds_bytes(s.nt, epw_of(s), p.X, p.ys, p.Z, p.W, p.H);
  const mcl::KernelSet& ks = mcl::kernels(s.nt, s.lift != 0);
  const hipError_t e = count ? ks.env_count(p, obs_mode, s.wpe, lds, st)
                             : ks.env(kind, p, obs_mode, s.wpe, lds, st);
  if (e != hipSuccess) return fail(MC_E_HIP, "env kernel launch: %s", hipGetErrorString(e));
  return MC_OK;
}

// Full-height pass over the envs a staged-height step deferred (p.fb, the
// same slot): one env per workgroup, whole world in LDS.
int launch_deferred(const mc::Params& q, int obs_mode, int lift, hipStream_t st) {
  mc::Params p = q;
  p.ys = p.Y;
  // 512 threads for worlds above 80 KiB (the workgroup's tile buffers sit in
  // its LDS trail, so the size follows the width).
  const int nt = mc::item_lds_bytes(256, 1, p.X, p.Y, p.Z, p.W, p.H) > 80 * 1024 ? 512 : 256;
  const int full = mc::item_lds_bytes(nt, 1, p.X, p.Y, p.Z, p.W, p.H);
  if (full > lds_budget()) {
    // The whole world does not fit LDS: the deferred envs run from global
    // memory.
    const hipError_t e =
        mcl::kernels(256, true).glb(mc::kItemStep, p, obs_mode, true, dim3(std::min(p.B, 256)), st);
    if (e != hipSuccess) return fail(MC_E_HIP, "deferred-env launch: %s", hipGetErrorString(e));
    return MC_OK;
  }
  const hipError_t e = mcl::kernels(nt, lift != 0).fb(p, obs_mode, dim3(std::min(p.B, 64)),
                                                      (size_t)full, st);
  if (e != hipSuccess) return fail(MC_E_HIP, "deferred-env launch: %s", hipGetErrorString(e));
  return MC_OK;
}

// Layers a MineEnv step stages in LDS: the fresh world's solids end at
// ground_height + 4 (the house roof; + terrain_amplitude on the seeded
// terrain); the agent's eye stays below
// max(spawn_y, Y/4 + jump apex) + 1.6 (the physics has no block collision:
// only gravity, jumps and the ground clamp at Y/4), and the raycast reaches
// 4 above the eye.  Envs outside these bounds are deferred to the
// full-height pass, so this only sizes the fast path.
int staged_layers(const mc_config& c) {
  const int Y = c.world_y;
  if (c.world_x > mc::kMaxInsideDim || Y > mc::kMaxInsideDim || c.world_z > mc::kMaxInsideDim)
    return Y;
  // On the seeded terrain an env respawns up to terrain_amplitude above spawn_y
  // (its spawn column's surface, spawn_of() in mc_device.h).
  const int amp = c.terrain == MC_TERRAIN_NOISE ? c.terrain_amplitude : 0;
  const float spawn_top = c.spawn_y + (float)amp;
  if (!(c.gravity < 0.0f) || !(c.jump_vel >= 0.0f) || !(spawn_top < (float)Y)) return Y;
  const float apex = c.jump_vel * c.jump_vel / (-2.0f * c.gravity);
  const float eye = std::max(spawn_top, (float)(Y / 4) + apex) + 1.6f;
  const int top = c.ground_height + amp;
  // With collision the agent can stand on any solid: up to the house roof.
  const float eye_c = c.collision ? std::max(eye, (float)(top + 5) + apex + 1.6f) : eye;
  const float need = std::max((float)(top + 5), std::ceil(eye_c + 4.5f) + 1.0f);
  if (!(need < (float)Y)) return Y;
  return std::max(1, (int)need);
}

}  // namespace

struct mc_env {
  mc_config cfg;
  mc_state st;
  mc::Params base;
  uint32_t* ring;     // device words: episode flag ring [0..2], reset flag [3],
                      // device-resident slot [4] (graph mode)
  mc::Params* desc;   // device copy of `base` (with ring), read by batched launches
  int slot;           // ring slot the next step launch reads (host-tracked mode)
  int graph_mode;     // 1: the slot lives in ring[4] and each step advances it
  Shape shape;        // launch shape for this group (default or tuned)
  int shape_default;  // shape not set by the caller: steps size it for ys
  int ys;             // layers a step stages in LDS (world_y: all)
  uint32_t* fb;       // [3][B + 2] deferred-env lists (NULL when ys == world_y)
  unsigned long long* ray_counter;  // non-NULL: steps run the ray-step counting build
};

extern "C" {

const char* mc_last_error(void) { return g_err; }
int mc_version(void) { return 1; }

int mc_generate_world(int8_t* voxels, int B, int X, int Y, int Z, int ground_height,
                      unsigned int seed, void* stream) {
  (void)seed;  // accepted and unused, as in the reference (src/kernels.cu:1800)
  if (!voxels) return fail(MC_E_INVALID, "voxels is NULL");
  if (B < 1 || X < 1 || Y < 1 || Z < 1) return fail(MC_E_INVALID, "bad dims");
  const int64_t total = (int64_t)B * X * Y * Z;
  const int64_t wxyz = (int64_t)X * Y * Z;
  const int64_t items = (wxyz % 16 == 0) ? total / 16 : total;
  const int64_t blocks = (items + 255) / 256;
  hipLaunchKernelGGL(mc::k_generate_world, dim3((unsigned)blocks), dim3(256), 0, S(stream),
                     voxels, total, X * Z, Y, ground_height);
  return check_launch("generate_world");
}

int mc_place_tree(int8_t* voxels, int B, int X, int Y, int Z, int tree_x, int tree_z,
                  int ground_height, void* stream) {
  if (!voxels) return fail(MC_E_INVALID, "voxels is NULL");
  if (B < 1 || X < 1 || Y < 1 || Z < 1) return fail(MC_E_INVALID, "bad dims");
  hipLaunchKernelGGL(mc::k_place_tree, dim3((B + 255) / 256), dim3(256), 0, S(stream), voxels,
                     B, X, Y, Z, tree_x, tree_z, ground_height);
  return check_launch("place_tree");
}

int mc_place_house(int8_t* voxels, int B, int X, int Y, int Z, int house_x, int house_z,
                   int ground_height, void* stream) {
  if (!voxels) return fail(MC_E_INVALID, "voxels is NULL");
  if (B < 1 || X < 1 || Y < 1 || Z < 1) return fail(MC_E_INVALID, "bad dims");
  hipLaunchKernelGGL(mc::k_place_house, dim3((B + 255) / 256), dim3(256), 0, S(stream), voxels,
                     B, X, Y, Z, house_x, house_z, ground_height);
  return check_launch("place_house");
}

int mc_regen_worlds(const mc_config* cfg, int8_t* voxels, const uint8_t* mask, void* stream) {
  if (!cfg || !voxels) return fail(MC_E_INVALID, "config/voxels is NULL");
  int r = check_dims(cfg->batch_size, cfg->world_x, cfg->world_y, cfg->world_z);
  if (r) return r;
  r = check_terrain(cfg);
  if (r) return r;
  mc::Params p = make_params(cfg, nullptr);
  p.voxels = voxels;
  p.mask = mask;
  if (p.terrain) hipLaunchKernelGGL(mc::k_regen_terrain, dim3(p.B), dim3(256), 0, S(stream), p);
  else hipLaunchKernelGGL(mc::k_regen, dim3(p.B), dim3(256), 0, S(stream), p);
  return check_launch("regen_worlds");
}

int mc_reset(float* positions, float* velocities, float* yaws, float* pitches, uint8_t* on_ground,
             const uint8_t* do_reset, int B, float spawn_x, float spawn_y, float spawn_z,
             void* stream) {
  if (!positions || !velocities || !yaws || !pitches || !on_ground || !do_reset)
    return fail(MC_E_INVALID, "reset: NULL pointer");
  if (B < 1) return fail(MC_E_INVALID, "batch_size must be >= 1");
  hipLaunchKernelGGL(mc::k_reset, dim3((B + 255) / 256), dim3(256), 0, S(stream), positions,
                     velocities, yaws, pitches, on_ground, do_reset, B, spawn_x, spawn_y,
                     spawn_z);
  return check_launch("reset");
}

int mc_render(const int8_t* voxels, const float* cameras, void* obs, int obs_mode, int B, int X,
              int Y, int Z, int W, int H, int max_steps, float view_distance, float fov_degrees,
              void* stream) {
  if (!voxels || !cameras || !obs) return fail(MC_E_INVALID, "render: NULL pointer");
  int r = check_dims(B, X, Y, Z);
  if (r) return r;
  r = check_image(B, W, H);
  if (r) return r;
  if (obs_mode < MC_OBS_U8 || obs_mode > MC_OBS_U8_FP16)
    return fail(MC_E_INVALID, "bad obs_mode %d", obs_mode);
  mc_config c;
  memset(&c, 0, sizeof c);
  c.batch_size = B;
  c.world_x = X; c.world_y = Y; c.world_z = Z;
  c.width = W; c.height = H;
  c.max_steps = max_steps;
  c.view_distance = view_distance;
  c.fov_degrees = fov_degrees;
  mc::Params p = make_params(&c, nullptr);
  p.voxels = const_cast<int8_t*>(voxels);
  p.cameras = const_cast<float*>(cameras);
  p.obs = obs;
  return launch_env(mc::kItemRenderCameras, p, obs_mode, default_shape(W, H, X, Y, Z), S(stream));
}

int mc_render_state(const mc_config* cfg, const mc_state* st, void* stream) {
  int r = validate_config(cfg);
  if (r) return r;
  r = validate_state(st);
  if (r) return r;
  mc::Params p = make_params(cfg, st);
  return launch_env(mc::kItemRenderState, p, cfg->obs_mode, default_shape(*cfg), S(stream));
}

int mc_step(const mc_config* cfg, const mc_state* st, const void* actions, int action_kind,
            const float* look, uint8_t* dones_out, void* workspace, void* stream) {
  int r = validate_config(cfg);
  if (r) return r;
  r = validate_state(st);
  if (r) return r;
  if (!actions) return fail(MC_E_INVALID, "actions is NULL");
  if (action_kind != MC_ACT_BUTTONS && action_kind != MC_ACT_DISCRETE)
    return fail(MC_E_INVALID, "bad action_kind %d", action_kind);
  if (action_kind == MC_ACT_BUTTONS && !look) return fail(MC_E_INVALID, "look is NULL");
  if (!workspace) return fail(MC_E_INVALID, "workspace is NULL");
  mc::Params p = make_params(cfg, st);
  p.actions = actions;
  p.action_kind = action_kind;
  p.look = look;
  p.dones = dones_out;
  p.ring = reinterpret_cast<uint32_t*>(workspace);
  p.slot = 0;
  hipLaunchKernelGGL(mc::k_ring_setup, dim3(1), dim3(1024), 0, S(stream), p, 0);
  r = check_launch("episode pre-pass");
  if (r) return r;
  return launch_env(mc::kItemStep, p, cfg->obs_mode, default_shape(*cfg), S(stream));
}

namespace {
// Scratch of mc_step_static: the episode flag ring (the MC_WORKSPACE_BYTES
// that mc_step takes from its caller) followed by one deferred-env list,
// [count, finished workgroups, env...], of at most MC_STEP_STATIC_MAX_BATCH
// envs.  One copy per GPU (a file-scope device array), about 1 MiB, held until
// the process ends.  In host code `scratch` only names the symbol: its device
// address comes from hipGetSymbolAddress, once per device.
constexpr int kStaticRingWords = MC_WORKSPACE_BYTES / 4;
constexpr size_t kStaticScratchWords = kStaticRingWords + (size_t)MC_STEP_STATIC_MAX_BATCH + 2;
static __device__ __align__(256) uint32_t scratch[kStaticScratchWords];

constexpr int kMaxDevices = 64;
std::atomic<uint32_t*> g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;

// Device address of the current device's scratch (looked up on its first use).
int static_scratch(uint32_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(MC_E_HIP, "mc_step_static: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(MC_E_UNSUPPORTED, "mc_step_static: device %d beyond %d", dev, kMaxDevices);
  uint32_t* ptr = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (!ptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ptr = g_scratch_ptr[dev].load(std::memory_order_relaxed);
    if (!ptr) {
      void* addr = nullptr;
      e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !addr)
        return fail(MC_E_HIP, "mc_step_static: scratch address: %s", hipGetErrorString(e));
      ptr = static_cast<uint32_t*>(addr);
      g_scratch_ptr[dev].store(ptr, std::memory_order_release);
    }
  }
  *out = ptr;
  return MC_OK;
}
}  // namespace

int mc_step_static(const mc_config* cfg, const mc_state* st, const void* actions, int action_kind,
                   const float* look, uint8_t* dones_out, void* stream) {
  if (cfg && cfg->batch_size > MC_STEP_STATIC_MAX_BATCH)
    return fail(MC_E_INVALID, "mc_step_static: batch_size %d exceeds its scratch limit of %d envs",
                cfg->batch_size, MC_STEP_STATIC_MAX_BATCH);
  int r = validate_config(cfg);
  if (r) return r;
  r = validate_state(st);
  if (r) return r;
  if (!actions) return fail(MC_E_INVALID, "actions is NULL");
  if (action_kind != MC_ACT_BUTTONS && action_kind != MC_ACT_DISCRETE)
    return fail(MC_E_INVALID, "bad action_kind %d", action_kind);
  if (action_kind == MC_ACT_BUTTONS && !look) return fail(MC_E_INVALID, "look is NULL");
  uint32_t* scr = nullptr;
  r = static_scratch(&scr);
  if (r) return r;
  mc::Params p = make_params(cfg, st);
  p.actions = actions;
  p.action_kind = action_kind;
  p.look = look;
  p.dones = dones_out;
  p.ring = scr;
  p.slot = 0;
  // Staged height, as MineEnv's steps run it: envs the staged launch cannot
  // run are listed in the scratch's deferred-env list for the full-height pass.
  const mc_config& c = *cfg;
  const int ys = staged_layers(c);
  Shape s = default_shape(c);
  if (ys < c.world_y) {
    p.ys = ys;
    p.fb = scr + kStaticRingWords;
    const Shape d = default_shape(c.width, c.height, c.world_x, ys, c.world_z);
    s = Shape{d.nt, d.wpe, s.lift};
  }
  // The ring words and the list header start each call at zero (the set-up
  // kernel writes them; the array keeps the last call's contents).
  hipLaunchKernelGGL(mc::k_ring_setup, dim3(1), dim3(1024), 0, S(stream), p, 0);
  r = check_launch("episode pre-pass");
  if (r) return r;
  r = launch_env(mc::kItemStep, p, c.obs_mode, s, S(stream));
  if (!r && p.fb) r = launch_deferred(p, c.obs_mode, s.lift, S(stream));
  return r;
}

namespace {
int reset_with_ring(const mc_config* cfg, const mc_state* st, const uint8_t* mask,
                    uint32_t* ring, int slot, const uint32_t* slot_dev, Shape shape,
                    void* stream) {
  mc::Params p = make_params(cfg, st);
  p.mask = mask;
  p.ring = ring;
  p.slot = slot;
  p.slot_dev = slot_dev;
  hipLaunchKernelGGL(mc::k_ring_setup, dim3(1), dim3(1024), 0, S(stream), p, 1);
  const int r = check_launch("reset pre-pass");
  if (r) return r;
  return launch_env(mc::kItemResetRender, p, cfg->obs_mode, shape, S(stream));
}
}  // namespace

int mc_env_reset_state(const mc_config* cfg, const mc_state* st, const uint8_t* mask,
                       void* workspace, void* stream) {
  int r = validate_config(cfg);
  if (r) return r;
  r = validate_state(st);
  if (r) return r;
  if (!workspace) return fail(MC_E_INVALID, "workspace is NULL");
  return reset_with_ring(cfg, st, mask, reinterpret_cast<uint32_t*>(workspace), 0, nullptr,
                         default_shape(*cfg), stream);
}

int mc_env_create(const mc_config* cfg, const mc_state* st, mc_env** out) {
  if (!out) return fail(MC_E_INVALID, "out is NULL");
  *out = nullptr;
  int r = validate_config(cfg);
  if (r) return r;
  r = validate_state(st);
  if (r) return r;
  mc_env* e = new (std::nothrow) mc_env;
  if (!e) return fail(MC_E_INVALID, "out of host memory");
  e->cfg = *cfg;
  e->st = *st;
  e->base = make_params(cfg, st);
  e->ring = nullptr;
  e->desc = nullptr;
  e->slot = 0;
  e->graph_mode = 0;
  e->shape = default_shape(*cfg);
  e->shape_default = 1;
  e->ys = staged_layers(*cfg);
  e->fb = nullptr;
  e->ray_counter = nullptr;
  hipError_t he = hipMalloc(reinterpret_cast<void**>(&e->ring), MC_WORKSPACE_BYTES);
  if (he == hipSuccess) he = hipMemset(e->ring, 0, MC_WORKSPACE_BYTES);
  if (he == hipSuccess && e->ys < cfg->world_y) {
    const size_t fb_bytes = sizeof(uint32_t) * 3 * ((size_t)cfg->batch_size + 2);
    he = hipMalloc(reinterpret_cast<void**>(&e->fb), fb_bytes);
    if (he == hipSuccess) he = hipMemset(e->fb, 0, fb_bytes);
  }
  if (he == hipSuccess) {
    e->base.ring = e->ring;
    he = hipMalloc(reinterpret_cast<void**>(&e->desc), sizeof(mc::Params));
  }
  if (he == hipSuccess) {
    // Batched steps read this copy: it carries the staged height and the
    // deferred-env lists (the host-side base stages whole worlds).
    mc::Params d = e->base;
    if (e->fb) {
      d.ys = e->ys;
      d.fb = e->fb;
    }
    he = hipMemcpy(e->desc, &d, sizeof(mc::Params), hipMemcpyHostToDevice);
  }
  if (he != hipSuccess) {
    if (e->ring) (void)hipFree(e->ring);
    if (e->desc) (void)hipFree(e->desc);
    if (e->fb) (void)hipFree(e->fb);
    delete e;
    return fail(MC_E_HIP, "workspace: %s", hipGetErrorString(he));
  }
  *out = e;
  return MC_OK;
}

int mc_env_set_shape(mc_env* env, int threads, int waves_per_env, int lift) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  const Shape d = default_shape(env->cfg);
  env->shape_default = (threads == 0 && waves_per_env == 0);
  if (env->shape_default) {
    env->shape = Shape{d.nt, d.wpe, lift < 0 ? d.lift : lift};
  } else {
    env->shape = Shape{threads, waves_per_env, lift < 0 ? env->shape.lift : lift};
  }
  if (!valid_shape(env->shape)) {
    env->shape = d;
    env->shape_default = 1;
    return fail(MC_E_INVALID, "shape (%d threads, %d waves/env, lift %d) not supported",
                threads, waves_per_env, lift);
  }
  return MC_OK;
}

int mc_env_get_shape(const mc_env* env, int* threads, int* waves_per_env, int* lift) {
  if (!env || !threads || !waves_per_env) return fail(MC_E_INVALID, "NULL argument");
  const Shape s = fit_lds(env->shape, env->cfg.world_x, env->cfg.world_y, env->cfg.world_z,
                          env->cfg.width, env->cfg.height);
  *threads = s.nt;
  *waves_per_env = s.wpe;
  if (lift) *lift = s.lift;
  return MC_OK;
}

int mc_env_get_config(const mc_env* env, mc_config* out) {
  if (!env || !out) return fail(MC_E_INVALID, "NULL argument");
  *out = env->cfg;
  return MC_OK;
}

int mc_env_sync(mc_env* env, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  mc::Params p = env->base;
  p.ring = env->ring;
  p.slot = env->slot;
  if (env->graph_mode) p.slot_dev = env->ring + 4;
  hipLaunchKernelGGL(mc::k_ring_setup, dim3(1), dim3(1024), 0, S(stream), p, 0);
  return check_launch("episode sync");
}

int mc_env_destroy(mc_env* env) {
  if (!env) return MC_OK;
  if (env->ring) (void)hipFree(env->ring);
  if (env->desc) (void)hipFree(env->desc);
  if (env->fb) (void)hipFree(env->fb);
  delete env;
  return MC_OK;
}

int mc_env_step(mc_env* env, const void* actions, int action_kind, const float* look,
                uint8_t* dones_out, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  if (!actions) return fail(MC_E_INVALID, "actions is NULL");
  if (action_kind != MC_ACT_BUTTONS && action_kind != MC_ACT_DISCRETE)
    return fail(MC_E_INVALID, "bad action_kind %d", action_kind);
  if (action_kind == MC_ACT_BUTTONS && !look) return fail(MC_E_INVALID, "look is NULL");
  mc::Params p = env->base;
  p.actions = actions;
  p.action_kind = action_kind;
  p.look = look;
  p.dones = dones_out;
  p.ring = env->ring;
  p.slot = env->slot;
  if (env->graph_mode) p.slot_dev = env->ring + 4;
  // Staged height: the step launch defers the envs it cannot run to a
  // full-height pass over the same slot, launched right behind it.
  Shape s = env->shape;
  if (env->fb) {
    p.ys = env->ys;
    p.fb = env->fb;
    if (env->shape_default) {
      const mc_config& c = env->cfg;
      const Shape d = default_shape(c.width, c.height, c.world_x, env->ys, c.world_z);
      s = Shape{d.nt, d.wpe, s.lift};
    }
  }
  p.ray_steps = env->ray_counter;
  int r = launch_env(mc::kItemStep, p, env->cfg.obs_mode, s, S(stream), env->ray_counter != nullptr);
  if (!r && env->fb) r = launch_deferred(p, env->cfg.obs_mode, s.lift, S(stream));
  if (r) return r;
  if (env->graph_mode) {
    hipLaunchKernelGGL(mc::k_slot_advance, dim3(1), dim3(1), 0, S(stream), env->ring + 4);
    return check_launch("slot advance");
  }
  env->slot = (env->slot + 1) % 3;
  return MC_OK;
}

int mc_env_set_ray_counter(mc_env* env, unsigned long long* counters) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  if (counters && (env->cfg.terrain != MC_TERRAIN_FLAT || env->cfg.collision ||
                   planar_mode(env->cfg.obs_mode)))
    return fail(MC_E_UNSUPPORTED, "ray-step counting covers the flat-world interleaved modes");
  env->ray_counter = counters;
  return MC_OK;
}

int mc_env_set_graph_mode(mc_env* env, int on, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  if ((on != 0) == (env->graph_mode != 0)) return MC_OK;
  if (on) {
    hipLaunchKernelGGL(mc::k_slot_set, dim3(1), dim3(1), 0, S(stream), env->ring + 4,
                       (uint32_t)env->slot);
    const int r = check_launch("slot set");
    if (r) return r;
    env->graph_mode = 1;
    return MC_OK;
  }
  uint32_t v = 0;
  hipError_t he = hipMemcpyAsync(&v, env->ring + 4, sizeof v, hipMemcpyDeviceToHost, S(stream));
  if (he == hipSuccess) he = hipStreamSynchronize(S(stream));
  if (he != hipSuccess) return fail(MC_E_HIP, "slot read-back: %s", hipGetErrorString(he));
  env->slot = (int)(v % 3u);
  env->graph_mode = 0;
  return MC_OK;
}

int mc_env_reset(mc_env* env, const uint8_t* mask, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  return reset_with_ring(&env->cfg, &env->st, mask, env->ring, env->slot,
                         env->graph_mode ? env->ring + 4 : nullptr, env->shape, stream);
}

int mc_env_render(mc_env* env, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  mc::Params p = env->base;
  return launch_env(mc::kItemRenderState, p, env->cfg.obs_mode, env->shape, S(stream));
}

namespace {
// Variant-specific arguments of the reference-signature steps: world extents,
// the prebasis basis output, and the prebasis_full episode bookkeeping
// (step_count non-NULL: do_reset is computed from step_count >= episode_length
// and step_count advanced in the kernel, episode_length <= 0 meaning none).
struct RefStep {
  int X, Y, Z;
  float* basis;
  int32_t* step_count;
  int episode_length;
};

int launch_step_ref(int obs_mode, const RefStep& v, int8_t* voxels, float* positions,
                    float* velocities,
                    float* yaws, float* pitches, uint8_t* on_ground, const int8_t* buttons,
                    const float* look, const uint8_t* do_reset, float* forward_in,
                    float* strafe_in, float* delta_yaw_in, float* delta_pitch_in,
                    uint8_t* jump_in, uint8_t* do_break, uint8_t* do_place, float* speed_mult,
                    float* cameras, void* obs_buffer, float* rewards, int B, int ws, int W,
                    int H, int max_steps, float view_distance, float fov_degrees, float dt,
                    float gravity, float walk_speed, float sprint_mult, float jump_vel,
                    int8_t target_block, float reward_value, float spawn_x, float spawn_y,
                    float spawn_z, void* stream) {
  if (!voxels || !positions || !velocities || !yaws || !pitches || !on_ground || !buttons ||
      !look || !do_reset || !cameras || !obs_buffer || !rewards)
    return fail(MC_E_INVALID, "step: NULL pointer");
  const bool any_ctrl = forward_in || strafe_in || delta_yaw_in || delta_pitch_in || jump_in ||
                        do_break || do_place || speed_mult;
  const bool all_ctrl = forward_in && strafe_in && delta_yaw_in && delta_pitch_in && jump_in &&
                        do_break && do_place && speed_mult;
  if (any_ctrl && !all_ctrl)
    return fail(MC_E_INVALID, "step: decode buffers must be all NULL or all non-NULL");
  (void)ws;
  int r = check_dims(B, v.X, v.Y, v.Z);
  if (r) return r;
  r = check_image(B, W, H);
  if (r) return r;
  mc_config c;
  memset(&c, 0, sizeof c);
  c.batch_size = B;
  c.world_x = v.X;
  c.world_y = v.Y;
  c.world_z = v.Z;
  c.width = W;
  c.height = H;
  c.max_steps = max_steps;
  c.view_distance = view_distance;
  c.fov_degrees = fov_degrees;
  c.dt = dt;
  c.gravity = gravity;
  c.walk_speed = walk_speed;
  c.sprint_mult = sprint_mult;
  c.jump_vel = jump_vel;
  c.target_block = target_block;
  c.reward_value = reward_value;
  c.spawn_x = spawn_x;
  c.spawn_y = spawn_y;
  c.spawn_z = spawn_z;
  c.episode_length = -1;
  c.obs_mode = obs_mode;
  mc::Params p = make_params(&c, nullptr);
  p.voxels = voxels;
  p.positions = positions;
  p.velocities = velocities;
  p.yaws = yaws;
  p.pitches = pitches;
  p.on_ground = on_ground;
  p.do_reset = const_cast<uint8_t*>(do_reset);
  p.cameras = cameras;
  p.obs = obs_buffer;
  p.rewards = rewards;
  p.actions = buttons;
  p.action_kind = MC_ACT_BUTTONS;
  p.look = look;
  p.ring = nullptr;
  p.basis = v.basis;
  if (v.step_count) {
    p.done_mode = 0;
    p.step_count = v.step_count;
    p.episode_length = v.episode_length > 0 ? v.episode_length : -1;
  } else {
    p.done_mode = 1;
  }
  if (all_ctrl) {
    p.forward_in = forward_in;
    p.strafe_in = strafe_in;
    p.delta_yaw_in = delta_yaw_in;
    p.delta_pitch_in = delta_pitch_in;
    p.jump_in = jump_in;
    p.do_break = do_break;
    p.do_place = do_place;
    p.speed_mult_in = speed_mult;
  }
  return launch_env(mc::kItemStep, p, obs_mode, default_shape(W, H, v.X, v.Y, v.Z), S(stream));
}
}  // namespace

int mc_launch_step_uint8(int8_t* voxels, float* positions, float* velocities, float* yaws,
                         float* pitches, uint8_t* on_ground, const int8_t* buttons,
                         const float* look, const uint8_t* do_reset, float* forward_in,
                         float* strafe_in, float* delta_yaw_in, float* delta_pitch_in,
                         uint8_t* jump_in, uint8_t* do_break, uint8_t* do_place,
                         float* speed_mult, float* cameras, uint8_t* obs_buffer, float* rewards,
                         int batch_size, int world_size, int width, int height, int max_steps,
                         float view_distance, float fov_degrees, float dt, float gravity,
                         float walk_speed, float sprint_mult, float jump_vel, int8_t target_block,
                         float reward_value, float spawn_x, float spawn_y, float spawn_z,
                         void* stream) {
  return launch_step_ref(MC_OBS_U8, RefStep{world_size, world_size, world_size, nullptr, nullptr, 0},
                         voxels, positions, velocities, yaws, pitches, on_ground,
                         buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
                         delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras,
                         obs_buffer, rewards, batch_size, world_size, width, height, max_steps,
                         view_distance, fov_degrees, dt, gravity, walk_speed, sprint_mult,
                         jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z, stream);
}

int mc_launch_step(int8_t* voxels, float* positions, float* velocities, float* yaws,
                   float* pitches, uint8_t* on_ground, const int8_t* buttons, const float* look,
                   const uint8_t* do_reset, float* forward_in, float* strafe_in,
                   float* delta_yaw_in, float* delta_pitch_in, uint8_t* jump_in,
                   uint8_t* do_break, uint8_t* do_place, float* speed_mult, float* cameras,
                   float* obs_buffer, float* rewards, int batch_size, int world_size, int width,
                   int height, int max_steps, float view_distance, float fov_degrees, float dt,
                   float gravity, float walk_speed, float sprint_mult, float jump_vel,
                   int8_t target_block, float reward_value, float spawn_x, float spawn_y,
                   float spawn_z, void* stream) {
  return launch_step_ref(MC_OBS_F32, RefStep{world_size, world_size, world_size, nullptr, nullptr, 0},
                         voxels, positions, velocities, yaws, pitches, on_ground,
                         buttons, look, do_reset, forward_in, strafe_in, delta_yaw_in,
                         delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras,
                         obs_buffer, rewards, batch_size, world_size, width, height, max_steps,
                         view_distance, fov_degrees, dt, gravity, walk_speed, sprint_mult,
                         jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z, stream);
}

int mc_launch_render_uint8(const int8_t* voxels, const float* cameras, uint8_t* output,
                           int batch_size, int world_size, int width, int height, int max_steps,
                           float view_distance, float fov_degrees, void* stream) {
  return mc_render(voxels, cameras, output, MC_OBS_U8, batch_size, world_size, world_size,
                   world_size, width, height, max_steps, view_distance, fov_degrees, stream);
}

int mc_launch_render(const int8_t* voxels, const float* cameras, float* output, int batch_size,
                     int world_size, int width, int height, int max_steps, float view_distance,
                     float fov_degrees, void* stream) {
  return mc_render(voxels, cameras, output, MC_OBS_F32, batch_size, world_size, world_size,
                   world_size, width, height, max_steps, view_distance, fov_degrees, stream);
}

#define MC_REF_STEP_ARGS                                                                     \
  voxels, positions, velocities, yaws, pitches, on_ground, buttons, look, do_reset, forward_in, \
      strafe_in, delta_yaw_in, delta_pitch_in, jump_in, do_break, do_place, speed_mult, cameras, \
      obs_buffer, rewards, batch_size
#define MC_REF_STEP_TAIL                                                                      \
  width, height, max_steps, view_distance, fov_degrees, dt, gravity, walk_speed, sprint_mult,    \
      jump_vel, target_block, reward_value, spawn_x, spawn_y, spawn_z, stream

int mc_launch_step_uint8_minimal(int8_t* voxels, float* positions, float* velocities,
                                 float* yaws, float* pitches, uint8_t* on_ground,
                                 const int8_t* buttons, const float* look,
                                 const uint8_t* do_reset, float* forward_in, float* strafe_in,
                                 float* delta_yaw_in, float* delta_pitch_in, uint8_t* jump_in,
                                 uint8_t* do_break, uint8_t* do_place, float* speed_mult,
                                 float* cameras, uint8_t* obs_buffer, float* rewards,
                                 int batch_size, int world_size, int width, int height,
                                 int max_steps, float view_distance, float fov_degrees, float dt,
                                 float gravity, float walk_speed, float sprint_mult,
                                 float jump_vel, int8_t target_block, float reward_value,
                                 float spawn_x, float spawn_y, float spawn_z, void* stream) {
  return launch_step_ref(MC_OBS_U8_MINIMAL,
                         RefStep{world_size, world_size, world_size, nullptr, nullptr, 0},
                         MC_REF_STEP_ARGS, world_size, MC_REF_STEP_TAIL);
}

int mc_launch_step_uint8_fp16(int8_t* voxels, float* positions, float* velocities, float* yaws,
                              float* pitches, uint8_t* on_ground, const int8_t* buttons,
                              const float* look, const uint8_t* do_reset, float* forward_in,
                              float* strafe_in, float* delta_yaw_in, float* delta_pitch_in,
                              uint8_t* jump_in, uint8_t* do_break, uint8_t* do_place,
                              float* speed_mult, float* cameras, uint8_t* obs_buffer,
                              float* rewards, int batch_size, int world_size, int width,
                              int height, int max_steps, float view_distance, float fov_degrees,
                              float dt, float gravity, float walk_speed, float sprint_mult,
                              float jump_vel, int8_t target_block, float reward_value,
                              float spawn_x, float spawn_y, float spawn_z, void* stream) {
  return launch_step_ref(MC_OBS_U8_FP16,
                         RefStep{world_size, world_size, world_size, nullptr, nullptr, 0},
                         MC_REF_STEP_ARGS, world_size, MC_REF_STEP_TAIL);
}

int mc_launch_step_uint8_prebasis(int8_t* voxels, float* positions, float* velocities,
                                  float* yaws, float* pitches, uint8_t* on_ground,
                                  const int8_t* buttons, const float* look,
                                  const uint8_t* do_reset, float* forward_in, float* strafe_in,
                                  float* delta_yaw_in, float* delta_pitch_in, uint8_t* jump_in,
                                  uint8_t* do_break, uint8_t* do_place, float* speed_mult,
                                  float* cameras, float* basis, uint8_t* obs_buffer,
                                  float* rewards, int batch_size, int world_x, int world_y,
                                  int world_z, int width, int height, int max_steps,
                                  float view_distance, float fov_degrees, float dt, float gravity,
                                  float walk_speed, float sprint_mult, float jump_vel,
                                  int8_t target_block, float reward_value, float spawn_x,
                                  float spawn_y, float spawn_z, void* stream) {
  if (!basis) return fail(MC_E_INVALID, "step: NULL basis");
  return launch_step_ref(MC_OBS_U8_PREBASIS,
                         RefStep{world_x, world_y, world_z, basis, nullptr, 0},
                         MC_REF_STEP_ARGS, world_y, MC_REF_STEP_TAIL);
}

int mc_launch_step_uint8_prebasis_full(
    int8_t* voxels, float* positions, float* velocities, float* yaws, float* pitches,
    uint8_t* on_ground, const int8_t* buttons, const float* look, uint8_t* do_reset,
    int32_t* step_count, float* forward_in, float* strafe_in, float* delta_yaw_in,
    float* delta_pitch_in, uint8_t* jump_in, uint8_t* do_break, uint8_t* do_place,
    float* speed_mult, float* cameras, float* basis, uint8_t* obs_buffer, float* rewards,
    int batch_size, int world_x, int world_y, int world_z, int width, int height, int max_steps,
    float view_distance, float fov_degrees, float dt, float gravity, float walk_speed,
    float sprint_mult, float jump_vel, int8_t target_block, float reward_value, float spawn_x,
    float spawn_y, float spawn_z, int episode_length, void* stream) {
  if (!basis || !step_count) return fail(MC_E_INVALID, "step: NULL basis or step_count");
  return launch_step_ref(MC_OBS_U8_PREBASIS,
                         RefStep{world_x, world_y, world_z, basis, step_count, episode_length},
                         MC_REF_STEP_ARGS, world_y, MC_REF_STEP_TAIL);
}

int mc_launch_step_uint8_smem(int8_t* voxels, float* positions, float* velocities, float* yaws,
                              float* pitches, uint8_t* on_ground, const int8_t* buttons,
                              const float* look, const uint8_t* do_reset, float* forward_in,
                              float* strafe_in, float* delta_yaw_in, float* delta_pitch_in,
                              uint8_t* jump_in, uint8_t* do_break, uint8_t* do_place,
                              float* speed_mult, float* cameras, float* basis,
                              uint8_t* obs_buffer, float* rewards, int batch_size,
                              int world_size, int width, int height, int max_steps,
                              float view_distance, float fov_degrees, float dt, float gravity,
                              float walk_speed, float sprint_mult, float jump_vel,
                              int8_t target_block, float reward_value, float spawn_x,
                              float spawn_y, float spawn_z, void* stream) {
  if (!basis) return fail(MC_E_INVALID, "step: NULL basis");
  return launch_step_ref(MC_OBS_U8_PREBASIS,
                         RefStep{world_size, world_size, world_size, basis, nullptr, 0},
                         MC_REF_STEP_ARGS, world_size, MC_REF_STEP_TAIL);
}
#undef MC_REF_STEP_ARGS
#undef MC_REF_STEP_TAIL

int mc_launch_precompute_camera_basis(const float* cameras, float* basis, int batch_size,
                                      int width, int height, float fov_degrees, void* stream) {
  if (!cameras || !basis) return fail(MC_E_INVALID, "camera basis: NULL pointer");
  if (batch_size < 1) return fail(MC_E_INVALID, "batch_size must be >= 1 (got %d)", batch_size);
  int r = check_image(batch_size, width, height);
  if (r) return r;
  mc_config c;
  memset(&c, 0, sizeof c);
  c.batch_size = batch_size;
  c.width = width;
  c.height = height;
  c.fov_degrees = fov_degrees;
  mc::Params p = make_params(&c, nullptr);
  p.cameras = const_cast<float*>(cameras);
  p.basis = basis;
  hipLaunchKernelGGL(mc::k_camera_basis, dim3((batch_size + 255) / 256), dim3(256), 0,
                     S(stream), p);
  return check_launch("camera basis launch");
}

int mc_launch_render_uint8_minimal(const int8_t* voxels, const float* cameras, uint8_t* output,
                                   int batch_size, int world_size, int width, int height,
                                   int max_steps, float view_distance, float fov_degrees,
                                   void* stream) {
  return mc_render(voxels, cameras, output, MC_OBS_U8_MINIMAL, batch_size, world_size,
                   world_size, world_size, width, height, max_steps, view_distance, fov_degrees,
                   stream);
}

int mc_launch_render_uint8_fp16(const int8_t* voxels, const float* cameras, uint8_t* output,
                                int batch_size, int world_size, int width, int height,
                                int max_steps, float view_distance, float fov_degrees,
                                void* stream) {
  return mc_render(voxels, cameras, output, MC_OBS_U8_FP16, batch_size, world_size, world_size,
                   world_size, width, height, max_steps, view_distance, fov_degrees, stream);
}

int mc_launch_render_uint8_prebasis(const int8_t* voxels, const float* basis, uint8_t* output,
                                    int batch_size, int world_size, int width, int height,
                                    int max_steps, void* stream) {
  if (!voxels || !basis || !output) return fail(MC_E_INVALID, "render: NULL pointer");
  int r = check_dims(batch_size, world_size, world_size, world_size);
  if (r) return r;
  r = check_image(batch_size, width, height);
  if (r) return r;
  mc_config c;
  memset(&c, 0, sizeof c);
  c.batch_size = batch_size;
  c.world_x = c.world_y = c.world_z = world_size;
  c.width = width;
  c.height = height;
  c.max_steps = max_steps;
  mc::Params p = make_params(&c, nullptr);
  p.voxels = const_cast<int8_t*>(voxels);
  p.basis = const_cast<float*>(basis);
  p.obs = output;
  return launch_env(mc::kItemRenderBasis, p, MC_OBS_U8_PREBASIS,
                    default_shape(width, height, world_size, world_size, world_size), S(stream));
}

int mc_launch_render_uint8_smem(const int8_t* voxels, const float* basis, uint8_t* output,
                                int batch_size, int world_size, int width, int height,
                                int max_steps, void* stream) {
  return mc_launch_render_uint8_prebasis(voxels, basis, output, batch_size, world_size, width,
                                         height, max_steps, stream);
}

// ---------------------------------------------------------------------------
// Batched groups: one persistent launch over several env groups.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
constexpr int kEntryRing = 8;   // descriptor tables in flight

// X, Y, Z, W, H, obs_mode, world (terrain << 1 | collision)
using BucketKey = std::tuple<int, int, int, int, int, int, int>;

BucketKey bucket_of(const mc_config& c) {
  return BucketKey{c.world_x, c.world_y, c.world_z, c.width, c.height, c.obs_mode,
                   (c.terrain << 1) | (c.collision ? 1 : 0)};
}

struct Occupancy {
  int nt, lift, family, lds, blocks;
};

// Mode family of a group (mc_launch.h): each family's groups run kernels of
// their own, launched back to back in one batched step.
int family_of(const mc_config& c) {
  if (c.obs_mode == MC_OBS_U8_CHW || c.obs_mode == MC_OBS_F32_CHW) return mcl::kFamilyPlanar;
  if (c.collision)
    return c.terrain == MC_TERRAIN_NOISE ? mcl::kFamilyCollisionTerrain : mcl::kFamilyCollision;
  if (c.terrain == MC_TERRAIN_NOISE) return mcl::kFamilyTerrain;
  return mcl::kFamilyFlat;
}
}  // namespace

struct mc_batch {
  // Work counters, one pair per descriptor slot: launches in flight at the
  // same time (on different streams) never share a counter.
  uint32_t* counters = nullptr;               // device [kEntryRing][2]
  mc::MultiEntry* dev_entries = nullptr;      // [kEntryRing][kMaxGroups]
  mc::MultiEntry* host_entries = nullptr;     // pinned, same shape
  hipEvent_t ev[kEntryRing] = {};
  bool ev_live[kEntryRing] = {};
  int next = 0;
  int persistent = 1;
  int cus = 0;
  std::vector<Occupancy> occ;
  std::map<BucketKey, Shape> tuned;
};

namespace {
int occupancy_blocks(mc_batch* b, int nt, bool lift, int family, int lds) {
  for (const auto& o : b->occ)
    if (o.nt == nt && o.lift == (int)lift && o.family == family && o.lds == lds) return o.blocks;
  const int nb = mcl::kernels(nt, lift).multi_blocks(family, (size_t)lds);
  b->occ.push_back({nt, (int)lift, family, lds, nb});
  return nb;
}

int launch_multi(mc_batch* b, int nt, bool lift, int family, const mc::MultiArgs& a, int lds,
                 hipStream_t st) {
  const int grid =
      b->persistent
          ? std::max(1, std::min(a.total_items, occupancy_blocks(b, nt, lift, family, lds) * b->cus))
          : std::max(1, a.total_items);
  const hipError_t e =
      mcl::kernels(nt, lift).multi(a, family, b->persistent != 0, grid, (size_t)lds, st);
  if (e != hipSuccess) return fail(MC_E_HIP, "batched step launch: %s", hipGetErrorString(e));
  return MC_OK;
}
}  // namespace

extern "C" {

int mc_batch_create(mc_batch** out) {
  if (!out) return fail(MC_E_INVALID, "out is NULL");
  *out = nullptr;
  mc_batch* b = new (std::nothrow) mc_batch;
  if (!b) return fail(MC_E_INVALID, "out of host memory");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t tbl = sizeof(mc::MultiEntry) * mc::kMaxGroups * kEntryRing;
  const size_t ctr = 2 * sizeof(uint32_t) * kEntryRing;
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->counters), ctr);
  if (e == hipSuccess) e = hipMemset(b->counters, 0, ctr);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->dev_entries), tbl);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&b->host_entries), tbl, 0);
  for (int i = 0; i < kEntryRing && e == hipSuccess; ++i)
    e = hipEventCreateWithFlags(&b->ev[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    mc_batch_destroy(b);
    return fail(MC_E_HIP, "batch create: %s", hipGetErrorString(e));
  }
  *out = b;
  return MC_OK;
}

int mc_batch_destroy(mc_batch* b) {
  if (!b) return MC_OK;
  for (int i = 0; i < kEntryRing; ++i)
    if (b->ev[i]) {
      if (b->ev_live[i]) (void)hipEventSynchronize(b->ev[i]);
      (void)hipEventDestroy(b->ev[i]);
    }
  if (b->counters) (void)hipFree(b->counters);
  if (b->dev_entries) (void)hipFree(b->dev_entries);
  if (b->host_entries) (void)hipHostFree(b->host_entries);
  delete b;
  return MC_OK;
}

int mc_batch_set_persistent(mc_batch* b, int persistent) {
  if (!b) return fail(MC_E_INVALID, "batch is NULL");
  b->persistent = persistent ? 1 : 0;
  return MC_OK;
}

int mc_batch_step(mc_batch* b, mc_env* const* envs, const void* const* actions,
                  const int* action_kinds, const float* const* looks, uint8_t* const* dones,
                  int n, void* stream) {
  if (!b || !envs || !actions || !action_kinds) return fail(MC_E_INVALID, "NULL argument");
  if (n < 1 || n > mc::kMaxGroups)
    return fail(MC_E_INVALID, "group count %d outside [1, %d]", n, mc::kMaxGroups);
  for (int i = 0; i < n; ++i) {
    if (!envs[i] || !actions[i]) return fail(MC_E_INVALID, "group %d: NULL env/actions", i);
    const int k = action_kinds[i];
    if (k != MC_ACT_BUTTONS && k != MC_ACT_DISCRETE)
      return fail(MC_E_INVALID, "group %d: bad action kind %d", i, k);
    if (k == MC_ACT_BUTTONS && (!looks || !looks[i]))
      return fail(MC_E_INVALID, "group %d: look is NULL", i);
    for (int j = 0; j < i; ++j)
      if (envs[j] == envs[i]) return fail(MC_E_INVALID, "group %d listed twice", i);
    if (envs[i]->graph_mode)
      return fail(MC_E_UNSUPPORTED, "group %d is in graph mode (its slot is device-resident)", i);
  }
  // One launch per mode family present (flat-world interleaved groups, planar
  // images, seeded terrain, voxel collision), each with its own width: 512
  // threads if one of its groups wants 8 waves per env, else 256.
  struct G {
    int idx, wpe, items;
    double cost;
  };
  struct Fam {
    std::vector<G> gs;
    int nt = 256, lds = 0, lds_full = 0;
    bool lift = false;
  };
  Fam fam[mcl::kFamilies];
  for (int i = 0; i < n; ++i) {
    Fam& f = fam[family_of(envs[i]->cfg)];
    if (envs[i]->shape.wpe == 8) f.nt = 512;
    f.lift = f.lift || envs[i]->shape.lift;
    f.gs.push_back(G{i, 0, 0, 0.0});
  }
  for (Fam& f : fam) {
    for (G& g : f.gs) {
      const int i = g.idx;
      const mc_config& c = envs[i]->cfg;
      const int ys = envs[i]->fb ? envs[i]->ys : c.world_y;  // layers the descriptor stages
      Shape s{f.nt, std::min(envs[i]->shape.wpe, f.nt / mc::kWave), f.lift ? 1 : 0};
      s = fit_lds(s, c.world_x, ys, c.world_z, c.width, c.height);
      const int stride = mc::world_stride_bytes(c.world_x, c.world_y, c.world_z);
      const int item_lds =
          mc::item_lds_bytes(s.nt, epw_of(s), c.world_x, ys, c.world_z, c.width, c.height);
      if (item_lds > lds_budget())
        return fail(MC_E_UNSUPPORTED, "group %d: world does not fit the LDS budget", i);
      f.lds = std::max(f.lds, item_lds);
      if (envs[i]->fb) {
        const int full =
            mc::item_lds_bytes(512, 1, c.world_x, c.world_y, c.world_z, c.width, c.height);
        if (full > lds_budget())
          return fail(MC_E_UNSUPPORTED, "group %d: world does not fit the LDS budget", i);
        f.lds_full = std::max(f.lds_full, full);
      }
      const int epw = epw_of(s);
      const double per_env = (double)c.width * c.height * (1.0 + c.world_y / 16.0) + stride;
      g = G{i, s.wpe, (c.batch_size + epw - 1) / epw, per_env * epw};
    }
    // Most expensive items first, so the tail of the persistent loop is short.
    std::stable_sort(f.gs.begin(), f.gs.end(),
                     [](const G& x, const G& y) { return x.cost > y.cost; });
  }
  const int slot = b->next;
  b->next = (b->next + 1) % kEntryRing;
  if (b->ev_live[slot]) {
    const hipError_t e = hipEventSynchronize(b->ev[slot]);
    if (e != hipSuccess) return fail(MC_E_HIP, "batch table wait: %s", hipGetErrorString(e));
  }
  // The families' entries lie one after another in this slot's table.
  mc::MultiEntry* he = b->host_entries + (size_t)slot * mc::kMaxGroups;
  mc::MultiEntry* de = b->dev_entries + (size_t)slot * mc::kMaxGroups;
  int at = 0;
  int first[mcl::kFamilies], total[mcl::kFamilies];
  for (int fi = 0; fi < mcl::kFamilies; ++fi) {
    first[fi] = at;
    int items = 0;
    for (const G& g : fam[fi].gs) {
      const int i = g.idx;
      mc::MultiEntry& m = he[at++];
      m.desc = envs[i]->desc;
      m.actions = actions[i];
      m.look = looks ? looks[i] : nullptr;
      m.dones = dones ? dones[i] : nullptr;
      m.action_kind = action_kinds[i];
      m.slot = envs[i]->slot;
      m.item_begin = items;
      m.wpe = g.wpe;
      m.obs_mode = envs[i]->cfg.obs_mode;
      m.pad = 0;
      items += g.items;
    }
    total[fi] = items;
  }
  hipError_t e = hipMemcpyAsync(de, he, sizeof(mc::MultiEntry) * n, hipMemcpyHostToDevice,
                                S(stream));
  if (e != hipSuccess) return fail(MC_E_HIP, "batch table upload: %s", hipGetErrorString(e));
  for (int fi = 0; fi < mcl::kFamilies; ++fi) {
    const Fam& f = fam[fi];
    if (f.gs.empty()) continue;
    mc::MultiArgs a;
    a.n_groups = (int)f.gs.size();
    a.total_items = total[fi];
    // Launches of one slot run one after another on `stream`; each leaves the
    // counters at zero for the next.
    a.counters = b->counters + 2 * slot;
    a.entries = de + first[fi];
    const int r = launch_multi(b, f.nt, f.lift, fi, a, f.lds, S(stream));
    if (r) return r;
    if (f.lds_full > 0) {
      // Full-height pass over the envs the family's staged groups deferred.
      const hipError_t ef = mcl::kernels(f.lds_full > 80 * 1024 ? 512 : 256, f.lift)
                                .multi_fb(a, fi, dim3(64), (size_t)f.lds_full, S(stream));
      if (ef != hipSuccess)
        return fail(MC_E_HIP, "batched deferred-env launch: %s", hipGetErrorString(ef));
    }
  }
  e = hipEventRecord(b->ev[slot], S(stream));
  if (e != hipSuccess) return fail(MC_E_HIP, "batch event: %s", hipGetErrorString(e));
  b->ev_live[slot] = true;
  for (int i = 0; i < n; ++i) envs[i]->slot = (envs[i]->slot + 1) % 3;
  return MC_OK;
}

}  // extern "C"

namespace {
// Scratch copy of one env group for timing candidates without touching the
// caller's state.
struct Scratch {
  mc_state st{};
  std::vector<void*> allocs;
  ~Scratch() {
    for (void* p : allocs) (void)hipFree(p);
  }
  bool alloc(void** p, size_t n) {
    if (hipMalloc(p, n) != hipSuccess) return false;
    allocs.push_back(*p);
    return hipMemset(*p, 0, n) == hipSuccess;
  }
};

// Deterministic discrete actions (uniform ids 0..10) for tuning runs.
std::vector<int64_t> tuning_actions(int steps, int B) {
  std::vector<int64_t> a((size_t)steps * B);
  uint64_t x = 0x9e3779b97f4a7c15ull;
  for (auto& v : a) {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    v = (int64_t)(x % 11u);
  }
  return a;
}
}  // namespace

extern "C" {

int mc_batch_autotune(mc_batch* b, mc_env* const* envs, int n, int steps, void* stream) {
  if (!b || !envs || n < 1) return fail(MC_E_INVALID, "NULL argument");
  if (n > mc::kMaxGroups) return fail(MC_E_INVALID, "at most %d groups", mc::kMaxGroups);
  if (steps < 1) steps = 12;
  // Buckets of this mix; every group of a bucket gets the same shape.
  std::vector<BucketKey> keys;
  std::vector<int> bucket(n);
  std::vector<double> bucket_cost;
  int64_t total = 0;
  for (int i = 0; i < n; ++i) {
    if (!envs[i]) return fail(MC_E_INVALID, "env %d is NULL", i);
    for (int j = 0; j < i; ++j)
      if (envs[j] == envs[i]) return fail(MC_E_INVALID, "group %d listed twice", i);
    if (envs[i]->graph_mode)
      return fail(MC_E_UNSUPPORTED, "group %d is in graph mode (its slot is device-resident)", i);
    const mc_config& c = envs[i]->cfg;
    const BucketKey k = bucket_of(c);
    const int j = (int)(std::find(keys.begin(), keys.end(), k) - keys.begin());
    if (j == (int)keys.size()) {
      keys.push_back(k);
      bucket_cost.push_back(0.0);
    }
    bucket[i] = j;
    bucket_cost[j] += (double)c.batch_size *
                      ((double)c.width * c.height * c.world_y + 2.0 * c.world_x * c.world_y * c.world_z);
    total += c.batch_size;
  }
  // A scratch mirror of the mix (same configs, batches scaled to at most
  // kTuneEnvs envs in all, no episode limit so the cost is stationary) is
  // stepped through mc_batch_step itself: shapes are judged in the launch
  // they will run in, where one group's shape sets the workgroup size and the
  // LDS of every workgroup.  The caller's groups are never touched.
  constexpr int64_t kTuneEnvs = 1 << 17;
  const double scale = std::min(1.0, (double)kTuneEnvs / (double)total);
  struct Mirror {
    Scratch sc;
    mc_env* env = nullptr;
    ~Mirror() {
      if (env) mc_env_destroy(env);
    }
  };
  std::vector<std::unique_ptr<Mirror>> mir;
  int maxB = 0;
  for (int i = 0; i < n; ++i) {
    mc_config c = envs[i]->cfg;
    c.batch_size = std::max(1, (int)std::lround(c.batch_size * scale));
    c.episode_length = -1;
    const int B = c.batch_size;
    const size_t nvox = (size_t)c.world_x * c.world_y * c.world_z;
    const size_t obs_b = (size_t)c.width * c.height * 3 *
                         ((c.obs_mode == MC_OBS_F32 || c.obs_mode == MC_OBS_F32_CHW) ? 4 : 1);
    mir.emplace_back(new (std::nothrow) Mirror);
    if (!mir.back()) return fail(MC_E_INVALID, "out of host memory");
    Scratch& sc = mir.back()->sc;
    mc_state& st = sc.st;
    const bool ok =
        sc.alloc((void**)&st.voxels, B * nvox) && sc.alloc((void**)&st.positions, B * 12) &&
        sc.alloc((void**)&st.velocities, B * 12) && sc.alloc((void**)&st.yaws, B * 4) &&
        sc.alloc((void**)&st.pitches, B * 4) && sc.alloc((void**)&st.on_ground, B) &&
        sc.alloc((void**)&st.step_count, B * 4) && sc.alloc((void**)&st.do_reset, B) &&
        sc.alloc((void**)&st.cameras, B * 20) && sc.alloc(&st.obs, B * obs_b) &&
        sc.alloc((void**)&st.rewards, B * 4) && sc.alloc((void**)&st.basis, (size_t)B * 14 * 4);
    if (!ok) return fail(MC_E_HIP, "autotune: scratch allocation failed");
    int r = mc_env_create(&c, &st, &mir.back()->env);
    if (!r) r = mc_regen_worlds(&c, st.voxels, nullptr, stream);
    if (!r) r = mc_env_reset(mir.back()->env, nullptr, stream);
    if (r) return r;
    maxB = std::max(maxB, B);
  }
  constexpr int kActionSteps = 64;
  Scratch abuf;
  void* acts = nullptr;
  if (!abuf.alloc(&acts, (size_t)kActionSteps * maxB * 8))
    return fail(MC_E_HIP, "autotune: scratch allocation failed");
  const std::vector<int64_t> ha = tuning_actions(kActionSteps, maxB);
  if (hipMemcpy(acts, ha.data(), ha.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
    return fail(MC_E_HIP, "autotune: action upload failed");
  const hipStream_t st = S(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail(MC_E_HIP, "autotune: event create failed");
  }
  std::vector<mc_env*> es(n);
  std::vector<const void*> as(n);
  std::vector<int> kinds(n, MC_ACT_DISCRETE);
  for (int i = 0; i < n; ++i) es[i] = mir[i]->env;
  int tick = 0;
  // One step of the mirror under configuration c (per-bucket shapes and the
  // launch form); with timing, the launch is bracketed by events and waited.
  struct Config {
    std::vector<Shape> shape;
    int persistent;
  };
  auto step_as = [&](const Config& c, float* ms) -> int {
    for (int i = 0; i < n; ++i) mir[i]->env->shape = c.shape[bucket[i]];
    b->persistent = c.persistent;
    for (int i = 0; i < n; ++i) as[i] = (const int64_t*)acts + (size_t)(tick % kActionSteps) * maxB;
    ++tick;
    if (ms && hipEventRecord(e0, st) != hipSuccess) return fail(MC_E_HIP, "autotune: event");
    const int e = mc_batch_step(b, es.data(), as.data(), kinds.data(), nullptr, nullptr, n, stream);
    if (e || !ms) return e;
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess)
      return fail(MC_E_HIP, "autotune: timing failed");
    return MC_OK;
  };
  // Mean launch time of x and y over `steps` launches each, alternating
  // launch by launch: the mirror's cost drifts as its agents move, and the
  // alternation gives both the same drift.
  auto duel = [&](const Config& x, const Config& y, float* tx, float* ty) -> int {
    int e = step_as(x, nullptr);
    if (!e) e = step_as(y, nullptr);
    *tx = *ty = 0.0f;
    for (int k = 0; k < 2 * steps && !e; ++k) {
      float ms = 0.0f;
      e = step_as((k & 1) ? y : x, &ms);
      *((k & 1) ? ty : tx) += ms;
    }
    return e;
  };
  const int caller_form = b->persistent;
  Config cur{std::vector<Shape>(keys.size()), caller_form};
  for (size_t j = 0; j < keys.size(); ++j) {
    const int X = std::get<0>(keys[j]), Y = std::get<1>(keys[j]), Z = std::get<2>(keys[j]);
    cur.shape[j] = default_shape(std::get<3>(keys[j]), std::get<4>(keys[j]), X, Y, Z);
  }
  const Config defaults = cur;
  // Warm-up: agents land and the mirror reaches its steady state.
  int r = MC_OK;
  for (int t = 0; t < 20 && !r; ++t) r = step_as(cur, nullptr);
  // Launch form first (it decides how the buckets share the device).
  if (!r) {
    Config other = cur;
    other.persistent = !cur.persistent;
    float tc = 0.0f, to = 0.0f;
    r = duel(cur, other, &tc, &to);
    if (!r && to < tc) cur = other;
  }
  // Coordinate descent, most expensive bucket first: a bucket takes a
  // candidate shape that beats the current configuration by more than 1.5%.
  std::vector<int> order(keys.size());
  for (size_t j = 0; j < order.size(); ++j) order[j] = (int)j;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return bucket_cost[x] > bucket_cost[y]; });
  const Shape cands[] = {{64, 1, 1},  {256, 1, 1}, {256, 2, 1}, {256, 4, 1}, {512, 8, 1},
                         {64, 1, 0},  {256, 1, 0}, {256, 2, 0}, {256, 4, 0}, {512, 8, 0}};
  for (int j : order) {
    const int X = std::get<0>(keys[j]), Y = std::get<1>(keys[j]), Z = std::get<2>(keys[j]);
    for (const Shape& cand : cands) {
      if (r) break;
      const Shape& now = cur.shape[j];
      if (cand.nt == now.nt && cand.wpe == now.wpe && cand.lift == now.lift) continue;
      const int W = std::get<3>(keys[j]), H = std::get<4>(keys[j]);
      if (fit_lds(cand, X, Y, Z, W, H).wpe != cand.wpe) continue;  // does not fit as asked
      Config next = cur;
      next.shape[j] = cand;
      float tc = 0.0f, tn = 0.0f;
      r = duel(cur, next, &tc, &tn);
      if (!r && tn < tc * 0.985f) cur = next;
    }
  }
  // The result must beat the default configuration side by side.
  if (!r) {
    float tc = 0.0f, td = 0.0f;
    r = duel(cur, defaults, &tc, &td);
    if (!r && !(tc < td * 0.99f)) cur = defaults;
  }
  if (!r && hipStreamSynchronize(st) != hipSuccess) r = fail(MC_E_HIP, "autotune: sync failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (r) {
    b->persistent = caller_form;
    return r;
  }
  b->persistent = cur.persistent;
  for (size_t j = 0; j < keys.size(); ++j) b->tuned[keys[j]] = cur.shape[j];
  for (int i = 0; i < n; ++i) envs[i]->shape = cur.shape[bucket[i]];
  return MC_OK;
}

int mc_batch_tuned_shape(const mc_batch* b, const mc_config* cfg, int* threads,
                         int* waves_per_env, int* lift) {
  if (!b || !cfg || !threads || !waves_per_env) return fail(MC_E_INVALID, "NULL argument");
  auto it = b->tuned.find(bucket_of(*cfg));
  if (it == b->tuned.end()) return fail(MC_E_INVALID, "shape not tuned");
  *threads = it->second.nt;
  *waves_per_env = it->second.wpe;
  if (lift) *lift = it->second.lift;
  return MC_OK;
}

int mc_debug_sincosf(const float* x, float* s, float* c, int n, void* stream) {
  if (!x || !s || !c || n < 0) return fail(MC_E_INVALID, "debug_sincosf: bad args");
  if (n == 0) return MC_OK;
  hipLaunchKernelGGL(mc::k_debug_sincos, dim3((n + 255) / 256), dim3(256), 0, S(stream), x, s, c,
                     n);
  return check_launch("debug_sincosf");
}

int mc_debug_fast_math(int mode, uint32_t begin, uint32_t count, uint32_t* out, void* stream) {
  if (!out || mode < 0 || mode > 2) return fail(MC_E_INVALID, "debug_fast_math: bad args");
  if (count == 0) return MC_OK;
  hipLaunchKernelGGL(mc::k_debug_fast_math, dim3(4096), dim3(256), 0, S(stream), mode, begin,
                     count, out);
  return check_launch("debug_fast_math");
}

int mc_debug_f16(const float* a, const float* b, uint16_t* ha, uint16_t* sum, uint8_t* lt, int n,
                 void* stream) {
  if (!a || !b || !ha || !sum || !lt || n < 0) return fail(MC_E_INVALID, "debug_f16: bad args");
  if (n == 0) return MC_OK;
  hipLaunchKernelGGL(mc::k_debug_f16, dim3((n + 255) / 256), dim3(256), 0, S(stream), a, b, ha,
                     sum, lt, n);
  return check_launch("debug_f16");
}

}  // extern "C"
