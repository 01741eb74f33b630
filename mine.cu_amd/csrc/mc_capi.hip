// mc_capi.hip - C ABI of libminecu_hip.so (declared in include/minecu_hip.h):
// argument validation, launch geometry and the per-group env handle.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>

#include "mc_device.h"
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

// Workgroup size per env: one wave while an observation has at most 1024
// pixels, four waves above that.
constexpr int kSmallImage = 1024;

int world_lds_bytes(int X, int Y, int Z) { return ((X * Y * Z) + 15) & ~15; }

int check_dims(int B, int X, int Y, int Z) {
  if (B < 1) return fail(MC_E_INVALID, "batch_size must be >= 1 (got %d)", B);
  if (X < 1 || Y < 1 || Z < 1)
    return fail(MC_E_INVALID, "world dims must be >= 1 (got %d x %d x %d)", X, Y, Z);
  if ((int64_t)X * Y * Z > mc::kMaxLdsWorld)
    return fail(MC_E_UNSUPPORTED, "world %dx%dx%d exceeds the %d-byte LDS-resident limit", X, Y,
                Z, mc::kMaxLdsWorld);
  return MC_OK;
}

int check_image(int B, int W, int H) {
  if (W < 1 || H < 1) return fail(MC_E_INVALID, "render size must be >= 1 (got %dx%d)", W, H);
  if ((int64_t)W * H * 3 > (int64_t)1 << 30)
    return fail(MC_E_INVALID, "render size %dx%d too large", W, H);
  (void)B;
  return MC_OK;
}

int validate_config(const mc_config* c) {
  if (!c) return fail(MC_E_INVALID, "config is NULL");
  int r = check_dims(c->batch_size, c->world_x, c->world_y, c->world_z);
  if (r) return r;
  r = check_image(c->batch_size, c->width, c->height);
  if (r) return r;
  if (c->obs_mode != MC_OBS_U8 && c->obs_mode != MC_OBS_F32)
    return fail(MC_E_INVALID, "obs_mode must be MC_OBS_U8 or MC_OBS_F32 (got %d)", c->obs_mode);
  return MC_OK;
}

int validate_state(const mc_state* s) {
  if (!s) return fail(MC_E_INVALID, "state is NULL");
  if (!s->voxels || !s->positions || !s->velocities || !s->yaws || !s->pitches ||
      !s->on_ground || !s->step_count || !s->do_reset || !s->cameras || !s->obs || !s->rewards)
    return fail(MC_E_INVALID, "every state pointer must be non-NULL");
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
  }
  return p;
}

// Launch one of the per-env kernels with the right template arguments.
enum class Kind { Step, ResetRender, Render };

template <int OBS, int NT>
int launch_env_kernel(Kind kind, const mc::Params& p, int from_state, hipStream_t st) {
  const size_t lds = (size_t)world_lds_bytes(p.X, p.Y, p.Z);
  switch (kind) {
    case Kind::Step:
      hipLaunchKernelGGL((mc::k_step<OBS, NT>), dim3(p.B), dim3(NT), lds, st, p);
      break;
    case Kind::ResetRender:
      hipLaunchKernelGGL((mc::k_reset_render<OBS, NT>), dim3(p.B), dim3(NT), lds, st, p);
      break;
    case Kind::Render:
      hipLaunchKernelGGL((mc::k_render<OBS, NT>), dim3(p.B), dim3(NT), lds, st, p, from_state);
      break;
  }
  return check_launch("env kernel launch");
}

int launch_env(Kind kind, const mc::Params& p, int obs_mode, int from_state, hipStream_t st) {
  const bool small = (int64_t)p.W * p.H <= kSmallImage;
  if (obs_mode == MC_OBS_U8)
    return small ? launch_env_kernel<0, 64>(kind, p, from_state, st)
                 : launch_env_kernel<0, 256>(kind, p, from_state, st);
  return small ? launch_env_kernel<1, 64>(kind, p, from_state, st)
               : launch_env_kernel<1, 256>(kind, p, from_state, st);
}

}  // namespace

struct mc_env {
  mc_config cfg;
  mc_state st;
  mc::Params base;
  uint32_t* ring;   // device words: episode flag ring [0..2] + reset flag [3]
  int slot;         // ring slot the next step launch reads
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
  const int r = check_dims(cfg->batch_size, cfg->world_x, cfg->world_y, cfg->world_z);
  if (r) return r;
  mc::Params p = make_params(cfg, nullptr);
  p.voxels = voxels;
  p.mask = mask;
  hipLaunchKernelGGL((mc::k_regen<256>), dim3(p.B), dim3(256), 0, S(stream), p);
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
  if (obs_mode != MC_OBS_U8 && obs_mode != MC_OBS_F32)
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
  return launch_env(Kind::Render, p, obs_mode, 0, S(stream));
}

int mc_render_state(const mc_config* cfg, const mc_state* st, void* stream) {
  int r = validate_config(cfg);
  if (r) return r;
  r = validate_state(st);
  if (r) return r;
  mc::Params p = make_params(cfg, st);
  return launch_env(Kind::Render, p, cfg->obs_mode, 1, S(stream));
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
  return launch_env(Kind::Step, p, cfg->obs_mode, 0, S(stream));
}

namespace {
int reset_with_ring(const mc_config* cfg, const mc_state* st, const uint8_t* mask,
                    uint32_t* ring, int slot, void* stream) {
  mc::Params p = make_params(cfg, st);
  p.mask = mask;
  p.ring = ring;
  p.slot = slot;
  hipLaunchKernelGGL(mc::k_ring_setup, dim3(1), dim3(1024), 0, S(stream), p, 1);
  const int r = check_launch("reset pre-pass");
  if (r) return r;
  return launch_env(Kind::ResetRender, p, cfg->obs_mode, 0, S(stream));
}
}  // namespace

int mc_env_reset_state(const mc_config* cfg, const mc_state* st, const uint8_t* mask,
                       void* workspace, void* stream) {
  int r = validate_config(cfg);
  if (r) return r;
  r = validate_state(st);
  if (r) return r;
  if (!workspace) return fail(MC_E_INVALID, "workspace is NULL");
  return reset_with_ring(cfg, st, mask, reinterpret_cast<uint32_t*>(workspace), 0, stream);
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
  e->slot = 0;
  hipError_t he = hipMalloc(reinterpret_cast<void**>(&e->ring), MC_WORKSPACE_BYTES);
  if (he == hipSuccess) he = hipMemset(e->ring, 0, MC_WORKSPACE_BYTES);
  if (he != hipSuccess) {
    if (e->ring) (void)hipFree(e->ring);
    delete e;
    return fail(MC_E_HIP, "workspace: %s", hipGetErrorString(he));
  }
  *out = e;
  return MC_OK;
}

int mc_env_sync(mc_env* env, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  mc::Params p = env->base;
  p.ring = env->ring;
  p.slot = env->slot;
  hipLaunchKernelGGL(mc::k_ring_setup, dim3(1), dim3(1024), 0, S(stream), p, 0);
  return check_launch("episode sync");
}

int mc_env_destroy(mc_env* env) {
  if (!env) return MC_OK;
  if (env->ring) (void)hipFree(env->ring);
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
  const int r = launch_env(Kind::Step, p, env->cfg.obs_mode, 0, S(stream));
  if (r == MC_OK) env->slot = (env->slot + 1) % 3;
  return r;
}

int mc_env_reset(mc_env* env, const uint8_t* mask, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  return reset_with_ring(&env->cfg, &env->st, mask, env->ring, env->slot, stream);
}

int mc_env_render(mc_env* env, void* stream) {
  if (!env) return fail(MC_E_INVALID, "env is NULL");
  mc::Params p = env->base;
  return launch_env(Kind::Render, p, env->cfg.obs_mode, 1, S(stream));
}

namespace {
int launch_step_ref(int obs_mode, int8_t* voxels, float* positions, float* velocities,
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
  int r = check_dims(B, ws, ws, ws);
  if (r) return r;
  r = check_image(B, W, H);
  if (r) return r;
  mc_config c;
  memset(&c, 0, sizeof c);
  c.batch_size = B;
  c.world_x = c.world_y = c.world_z = ws;
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
  p.done_mode = 1;
  p.ring = nullptr;
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
  return launch_env(Kind::Step, p, obs_mode, 0, S(stream));
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
  return launch_step_ref(MC_OBS_U8, voxels, positions, velocities, yaws, pitches, on_ground,
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
  return launch_step_ref(MC_OBS_F32, voxels, positions, velocities, yaws, pitches, on_ground,
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

int mc_debug_sincosf(const float* x, float* s, float* c, int n, void* stream) {
  if (!x || !s || !c || n < 0) return fail(MC_E_INVALID, "debug_sincosf: bad args");
  if (n == 0) return MC_OK;
  hipLaunchKernelGGL(mc::k_debug_sincos, dim3((n + 255) / 256), dim3(256), 0, S(stream), x, s, c,
                     n);
  return check_launch("debug_sincosf");
}

}  // extern "C"
