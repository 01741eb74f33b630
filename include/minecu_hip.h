/* minecu_hip.h - C ABI of libminecu_hip.so, the gfx950 (MI355X) batched
 * voxel-environment step.
 *
 * The reference puts a pybind/libtorch module (src/bindings.cpp:985-1269) over
 * `extern "C" void launch_*(raw pointers, ints, floats)` launchers
 * (src/kernels.cu:2076-3245).  This header is the equivalent boundary: plain
 * pointers, sizes and an explicit stream; no torch types.  Every function
 * returns 0 on success or a negative MC_E* code; mc_last_error() gives the
 * message of the calling thread's last failure.  All launches are
 * asynchronous on `stream` (a hipStream_t; NULL = the null stream), and
 * nothing here synchronises the device.
 *
 * Layouts (identical to the reference's tensors, minecu/__init__.py:214-256):
 *   voxels      int8   [B, Y, X, Z]   block ids, AIR = -1
 *   positions   f32    [B, 3]         x, y, z (feet)
 *   velocities  f32    [B, 3]         only [:,1] is integrated
 *   yaws, pitches f32  [B]
 *   on_ground   bool   [B]  (1 byte)
 *   step_count  int32  [B]
 *   do_reset    bool   [B]
 *   cameras     f32    [B, 5]         x, y, z, yaw, pitch
 *   obs         uint8 or f32 [B, H, W, 3]  ([B, 3, H, W] in the MC_OBS_*_CHW modes)
 *   rewards     f32    [B]
 *   buttons     int8   [B, 8]   look f32 [B, 2]   (multi-hot + radians)
 *   actions     int64  [B]      discrete ids 0..10 (docs/environment.md:3-17)
 */
#ifndef MINECU_HIP_H
#define MINECU_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MC_OK = 0,
  MC_E_INVALID = -1,   /* bad argument (shape, null pointer, range) */
  MC_E_HIP = -2,       /* a HIP runtime call failed */
  MC_E_UNSUPPORTED = -3
};

/* Observation encodings (the reference's `obs_dtype`, minecu/__init__.py:167). */
enum {
  MC_OBS_U8 = 0,          /* shaded + fogged, truncated to uint8  (render_kernel_uint8, kernels.cu:300-470) */
  MC_OBS_F32 = 1,         /* shaded + fogged float32 in [0,1]     (render_kernel, kernels.cu:87-293) */
  /* Flat block colours, no fog, no view-distance limit (kernels.cu:479-495).  Steps render
   * the flat image; the reset render and mc_render_state stay shaded uint8, as MineEnv's
   * _render() does for every uint8 variant (minecu/__init__.py:446-466). */
  MC_OBS_U8_MINIMAL = 2,  /* obs_dtype "uint8_minimal"  (render_kernel_uint8_minimal, kernels.cu:985-1119) */
  MC_OBS_U8_PREBASIS = 3, /* "uint8_prebasis" / "_full" / "uint8_smem": image plane from the [B,14]
                             camera basis (render_kernel_uint8_prebasis, kernels.cu:568-700; the
                             smem kernel, kernels.cu:835-982, renders the same image); steps
                             write st->basis when it is non-NULL */
  MC_OBS_U8_FP16 = 4,     /* "uint8_fp16": the minimal image plane with a binary16 DDA
                             (render_kernel_uint8_fp16, kernels.cu:1127-1262) */
  /* No counterpart in the reference: the MC_OBS_U8 / MC_OBS_F32 images, same values, in planar
   * [B, 3, H, W] layout -- the layout a convolutional policy reads, which the reference's
   * training example produces with a permute of every observation (examples/train_wood.py:53). */
  MC_OBS_U8_CHW = 5,
  MC_OBS_F32_CHW = 6
};

/* World generators (mc_config.terrain).  No counterpart in the reference, whose generator ignores
 * its seed (kernels.cu:1800-1828): MC_TERRAIN_NOISE gives each env a seeded height map, column
 * heights ground_height + d with d in [-terrain_amplitude, terrain_amplitude] (integer value noise
 * on a 4-column lattice), with the tree and the house standing on it. */
enum { MC_TERRAIN_FLAT = 0, MC_TERRAIN_NOISE = 1 };

/* Action encodings accepted by the step. */
enum {
  MC_ACT_BUTTONS = 0,  /* int8 [B,8] multi-hot + f32 [B,2] look (minecu/__init__.py:304-331) */
  MC_ACT_DISCRETE = 1  /* int64 [B] ids in [0,11); others act as no-op                        */
};

/* Static configuration of one env group: the MineEnv constructor arguments
 * (minecu/__init__.py:141-165) plus derived spawn/tree/house coordinates
 * (minecu/__init__.py:196-206). */
typedef struct mc_config {
  int batch_size;
  int world_x, world_y, world_z;
  int width, height;
  int max_steps;
  float view_distance;
  float fov_degrees;
  float dt, gravity, walk_speed, sprint_mult, jump_vel;
  int target_block;        /* < 0: every block rewards */
  float reward_value;
  float spawn_x, spawn_y, spawn_z;
  int ground_height;
  unsigned int seed;       /* unused by the flat world, as in the reference (kernels.cu:1800);
                              the MC_TERRAIN_NOISE height maps are drawn from it */
  int tree_x, tree_z, house_x, house_z;
  int episode_length;      /* < 0: no episode limit (the reference's None) */
  float look_delta;        /* radians per discrete look action */
  int obs_mode;            /* MC_OBS_* */
  int terrain;             /* MC_TERRAIN_* (0: the reference's flat world) */
  int terrain_amplitude;   /* MC_TERRAIN_NOISE: 0..8, and ground_height - amplitude >= 2 */
  int collision;           /* 1: voxel collision in place of the reference's ground plane at
                              world_y / 4 (kernels.cu:1566); no counterpart in the reference */
} mc_config;

/* Device pointers of one env group's state (caller-owned, mutated in place). */
typedef struct mc_state {
  int8_t* voxels;
  float* positions;
  float* velocities;
  float* yaws;
  float* pitches;
  uint8_t* on_ground;
  int32_t* step_count;
  uint8_t* do_reset;
  float* cameras;
  void* obs;
  float* rewards;
  float* basis;            /* f32 [B,14] camera basis (kernels.cu:503-510), or NULL */
  /* The decoded controls of the last step, MineEnv's public decode buffers
   * (minecu/__init__.py:243-251, written by decode_actions_kernel,
   * kernels.cu:1459-1504): f32 [B] forward_in, strafe_in, delta_yaw_in,
   * delta_pitch_in, speed_mult; bool [B] jump_in, do_break, do_place.  All
   * NULL (not written) or all non-NULL. */
  float* forward_in;
  float* strafe_in;
  float* delta_yaw_in;
  float* delta_pitch_in;
  uint8_t* jump_in;
  uint8_t* do_break;
  uint8_t* do_place;
  float* speed_mult;
} mc_state;

/* Minimum workspace (device bytes) for the stateless mc_step / mc_env_reset_state. */
#define MC_WORKSPACE_BYTES 256

const char* mc_last_error(void);
int mc_version(void);

/* ---- world generation -------------------------------------------------- */

/* Replaces launch_generate_world (src/kernels.cu:2281-2291; binding
 * bindings.cpp:357-371): flat layers by y, BEDROCK/DIRT/GRASS/AIR. */
int mc_generate_world(int8_t* voxels, int batch_size, int world_x, int world_y, int world_z,
                      int ground_height, unsigned int seed, void* stream);

/* Replaces launch_place_tree (src/kernels.cu:2293-2302; bindings.cpp:373-388). */
int mc_place_tree(int8_t* voxels, int batch_size, int world_x, int world_y, int world_z,
                  int tree_x, int tree_z, int ground_height, void* stream);

/* Replaces launch_place_house (src/kernels.cu:2304-2313; bindings.cpp:390-406). */
int mc_place_house(int8_t* voxels, int batch_size, int world_x, int world_y, int world_z,
                   int house_x, int house_z, int ground_height, void* stream);

/* One pass writing generate_world + place_tree + place_house for the envs whose
 * mask byte is set (mask NULL = all).  Replaces the three-launch regeneration
 * MineEnv performs (minecu/__init__.py:258-268, 296-300, 439-442). */
int mc_regen_worlds(const mc_config* cfg, int8_t* voxels, const uint8_t* mask, void* stream);

/* ---- agent ------------------------------------------------------------- */

/* Replaces launch_reset (src/kernels.cu:2268-2279; bindings.cpp:333-355). */
int mc_reset(float* positions, float* velocities, float* yaws, float* pitches,
             uint8_t* on_ground, const uint8_t* do_reset, int batch_size, float spawn_x,
             float spawn_y, float spawn_z, void* stream);

/* ---- render ------------------------------------------------------------ */

/* Replaces launch_render_uint8 (src/kernels.cu:2094-2108; bindings.cpp:244-271)
 * and launch_render (kernels.cu:2078-2092; bindings.cpp:215-242), selected by
 * obs_mode.  cameras f32 [B,5]; world extents are X, Y, Z per axis. */
int mc_render(const int8_t* voxels, const float* cameras, void* obs, int obs_mode,
              int batch_size, int world_x, int world_y, int world_z, int width, int height,
              int max_steps, float view_distance, float fov_degrees, void* stream);

/* ---- step -------------------------------------------------------------- */

/* One full environment step for every env of the group, fused into one
 * gfx950 kernel after a one-workgroup episode pre-pass.  Replaces
 * launch_step_uint8 (src/kernels.cu:2594-2652) / launch_step
 * (kernels.cu:2497-2591) together with the Python bookkeeping around them
 * (minecu/__init__.py:362-444: episode check, step_count update, regeneration
 * of every world when any env is done).  `actions` is int8 [B,8] buttons with
 * `look` f32 [B,2] (MC_ACT_BUTTONS) or int64 [B] ids (MC_ACT_DISCRETE, look
 * ignored).  dones_out (bool [B], may be NULL) receives the done mask, as does
 * st->do_reset.  workspace: >= MC_WORKSPACE_BYTES of device memory, private to
 * this call until it completes on `stream`. */
int mc_step(const mc_config* cfg, const mc_state* st, const void* actions, int action_kind,
            const float* look, uint8_t* dones_out, void* workspace, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory or pass a workspace: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  mc_step without the workspace argument, with
 * the same results bit for bit: the episode flag ring and a deferred-env list
 * (worlds taller than the layers a step stages in LDS run their full-height pass
 * from it, as MineEnv's steps do) live in a file-scope device array of about
 * 1 MiB per GPU, held until the process ends.  batch_size above
 * MC_STEP_STATIC_MAX_BATCH is refused (MC_E_INVALID).  No call allocates; the
 * first call on a device looks the array's address up (hipGetSymbolAddress), so
 * make it outside a graph capture.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one device (different streams, threads or
 * processes' graphs) race and must not. */
#define MC_STEP_STATIC_MAX_BATCH 262142
int mc_step_static(const mc_config* cfg, const mc_state* st, const void* actions, int action_kind,
                   const float* look, uint8_t* dones_out, void* stream);

/* MineEnv.reset(mask) (minecu/__init__.py:270-302) in two launches: masked
 * respawn, step_count = 0, regeneration of every world when mask is NULL or
 * any mask byte is set, camera update and render. */
int mc_env_reset_state(const mc_config* cfg, const mc_state* st, const uint8_t* mask,
                       void* workspace, void* stream);

/* Camera update + render of the current state (MineEnv._render,
 * minecu/__init__.py:446-466). */
int mc_render_state(const mc_config* cfg, const mc_state* st, void* stream);

/* ---- reference-signature launchers ------------------------------------
 * Same argument lists as the reference's extern "C" launchers (cubic
 * world_size = voxels.size(1), bindings.cpp:231,546), plus a stream, so the
 * reference's pybind layer could bind them unchanged.  The decode buffers
 * (forward_in .. speed_mult) may all be NULL, in which case they are not
 * written; when non-NULL they receive what decode_actions_kernel writes. */

/* Replaces launch_step_uint8 (src/kernels.cu:2594-2652; binding
 * bindings.cpp:492-575): decode, physics, raycast break/place, reset masked by
 * do_reset, camera update, shaded uint8 render -- one fused kernel. */
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
                         void* stream);

/* Replaces launch_step (src/kernels.cu:2497-2591; bindings.cpp:408-490), the
 * float32-observation step.  Unlike the reference it does not synchronise and
 * keeps no process-global graph (kernels.cu:2338-2382), so fresh action
 * pointers are always honoured. */
int mc_launch_step(int8_t* voxels, float* positions, float* velocities, float* yaws,
                   float* pitches, uint8_t* on_ground, const int8_t* buttons, const float* look,
                   const uint8_t* do_reset, float* forward_in, float* strafe_in,
                   float* delta_yaw_in, float* delta_pitch_in, uint8_t* jump_in,
                   uint8_t* do_break, uint8_t* do_place, float* speed_mult, float* cameras,
                   float* obs_buffer, float* rewards, int batch_size, int world_size, int width,
                   int height, int max_steps, float view_distance, float fov_degrees, float dt,
                   float gravity, float walk_speed, float sprint_mult, float jump_vel,
                   int8_t target_block, float reward_value, float spawn_x, float spawn_y,
                   float spawn_z, void* stream);

/* Replaces launch_render_uint8 (src/kernels.cu:2094-2108). */
int mc_launch_render_uint8(const int8_t* voxels, const float* cameras, uint8_t* output,
                           int batch_size, int world_size, int width, int height, int max_steps,
                           float view_distance, float fov_degrees, void* stream);

/* Replaces launch_render (src/kernels.cu:2078-2092). */
int mc_launch_render(const int8_t* voxels, const float* cameras, float* output, int batch_size,
                     int world_size, int width, int height, int max_steps, float view_distance,
                     float fov_degrees, void* stream);

/* Replaces launch_step_uint8_minimal (src/kernels.cu:2716-2774; binding
 * bindings.cpp:632-682): the fused step with the flat minimal render. */
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
                                 float spawn_x, float spawn_y, float spawn_z, void* stream);

/* Replaces launch_step_uint8_fp16 (src/kernels.cu:2776-2835; binding
 * bindings.cpp:684-734): the fused step with the binary16-DDA flat render. */
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
                              float spawn_x, float spawn_y, float spawn_z, void* stream);

/* Replaces launch_step_uint8_prebasis (src/kernels.cu:2993-3076; binding
 * bindings.cpp:736-792): the fused step writing basis f32 [B,14] and the flat
 * prebasis render; world extents per axis. */
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
                                  float spawn_y, float spawn_z, void* stream);

/* Replaces launch_step_uint8_prebasis_full (src/kernels.cu:3081-3168; binding
 * bindings.cpp:794-875): as the prebasis step, plus the episode check
 * (do_reset = step_count >= episode_length; episode_length <= 0: no limit) and
 * the step_count update in the kernel; world regeneration stays with the
 * caller, as in the reference. */
int mc_launch_step_uint8_prebasis_full(
    int8_t* voxels, float* positions, float* velocities, float* yaws, float* pitches,
    uint8_t* on_ground, const int8_t* buttons, const float* look, uint8_t* do_reset,
    int32_t* step_count, float* forward_in, float* strafe_in, float* delta_yaw_in,
    float* delta_pitch_in, uint8_t* jump_in, uint8_t* do_break, uint8_t* do_place,
    float* speed_mult, float* cameras, float* basis, uint8_t* obs_buffer, float* rewards,
    int batch_size, int world_x, int world_y, int world_z, int width, int height, int max_steps,
    float view_distance, float fov_degrees, float dt, float gravity, float walk_speed,
    float sprint_mult, float jump_vel, int8_t target_block, float reward_value, float spawn_x,
    float spawn_y, float spawn_z, int episode_length, void* stream);

/* Replaces launch_step_uint8_smem with the argument list its binding declares
 * (bindings.cpp:195-210, 930-981: cubic world_size); the image equals the
 * prebasis step's. */
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
                              float spawn_y, float spawn_z, void* stream);

/* Replaces launch_precompute_camera_basis (src/kernels.cu:2160-2170). */
int mc_launch_precompute_camera_basis(const float* cameras, float* basis, int batch_size,
                                      int width, int height, float fov_degrees, void* stream);

/* Replaces launch_render_uint8_minimal (src/kernels.cu:2128-2142). */
int mc_launch_render_uint8_minimal(const int8_t* voxels, const float* cameras, uint8_t* output,
                                   int batch_size, int world_size, int width, int height,
                                   int max_steps, float view_distance, float fov_degrees,
                                   void* stream);

/* Replaces launch_render_uint8_fp16 (src/kernels.cu:2144-2158). */
int mc_launch_render_uint8_fp16(const int8_t* voxels, const float* cameras, uint8_t* output,
                                int batch_size, int world_size, int width, int height,
                                int max_steps, float view_distance, float fov_degrees,
                                void* stream);

/* Replaces launch_render_uint8_prebasis (src/kernels.cu:2172-2189; binding
 * bindings.cpp:877-899) and launch_render_uint8_smem (kernels.cu:2191-2211):
 * flat render from a stored basis f32 [B,14]. */
int mc_launch_render_uint8_prebasis(const int8_t* voxels, const float* basis, uint8_t* output,
                                    int batch_size, int world_size, int width, int height,
                                    int max_steps, void* stream);
int mc_launch_render_uint8_smem(const int8_t* voxels, const float* basis, uint8_t* output,
                                int batch_size, int world_size, int width, int height,
                                int max_steps, void* stream);

/* ---- env handle (no reference counterpart: per-group state instead of the
 * reference's process-global StepGraphState, kernels.cu:2338-2382) ------- */

typedef struct mc_env mc_env;

/* Validates cfg/st once, allocates the group's workspace (a 4-word episode
 * flag ring: each step launch reads the "some env is done now" word left by
 * the previous launch, so mc_env_step is a single kernel launch). */
int mc_env_create(const mc_config* cfg, const mc_state* st, mc_env** out);
/* Recomputes the episode flag from step_count (after the caller rewrites
 * state, e.g. a checkpoint restore, or at start-up). */
int mc_env_sync(mc_env* env, void* stream);
int mc_env_destroy(mc_env* env);
int mc_env_step(mc_env* env, const void* actions, int action_kind, const float* look,
                uint8_t* dones_out, void* stream);
/* Graph mode (no reference counterpart; the reference captures its 6-launch
 * sequence in a process-global CUDA graph, kernels.cu:2338-2382): with on = 1
 * the group's episode ring slot moves to device memory and each mc_env_step
 * launches the step plus a one-thread slot advance, so a stream capture of
 * mc_env_step (e.g. torch.cuda.graph) replays correctly any number of times.
 * on = 0 reads the slot back (synchronises `stream`).  Groups in graph mode
 * cannot join mc_batch_step. */
int mc_env_set_graph_mode(mc_env* env, int on, void* stream);
int mc_env_reset(mc_env* env, const uint8_t* mask, void* stream);
int mc_env_render(mc_env* env, void* stream);

/* Measurement hook (no reference counterpart): with counters non-NULL (device,
 * 3 x u64, accumulated, never cleared here) the group's steps run a build of the
 * step kernel that adds [0] the DDA voxel lookups of every pixel (the start voxel
 * included; a pixel whose eye sits inside a solid counts its one lookup), [1] the
 * rays marched and [2] the pixels rendered.  Flat world, interleaved obs modes;
 * envs run by the full-height pass are not counted.  NULL restores the
 * production kernel.  Results are the same either way. */
int mc_env_set_ray_counter(mc_env* env, unsigned long long* counters);

/* Launch shape of a group's step: `threads` per workgroup (64, 256 or 512),
 * `waves_per_env` (1, 2, 4 or 8) 64-lane waves per env, so a workgroup holds
 * threads / (64 waves_per_env) envs, and `lift` (0/1; -1 keeps the current
 * value): build in the exact early exit of upward rays past every solid voxel
 * above eye height.  (0, 0, lift) restores the default size; the autotuner
 * below sets all three.  Results do not depend on the shape.  get returns the
 * shape after the LDS fit (lift may be NULL). */
int mc_env_set_shape(mc_env* env, int threads, int waves_per_env, int lift);
int mc_env_get_shape(const mc_env* env, int* threads, int* waves_per_env, int* lift);
/* Copy of the group's configuration (for schedulers that weigh groups by
 * their per-step work). */
int mc_env_get_config(const mc_env* env, mc_config* out);

/* ---- batched groups (no reference counterpart: the reference runs one
 * fixed-shape MineEnv per process-global graph) --------------------------- */

typedef struct mc_batch mc_batch;

int mc_batch_create(mc_batch** out);
int mc_batch_destroy(mc_batch* b);
/* Launch form of mc_batch_step: 1 (default) persistent workgroups that take
 * work items from an atomic counter; 0: one workgroup per work item.
 * mc_batch_autotune times both on the given mix and keeps the faster. */
int mc_batch_set_persistent(mc_batch* b, int persistent);

/* One step of n (<= 64) env groups -- any mix of batch sizes, world sizes,
 * resolutions and obs modes -- in ONE kernel launch.  Per group: actions,
 * action kind, look (buttons only; `looks` may be NULL when every group is
 * discrete) and dones (NULL entries or a NULL array: not written).  Each
 * group's results equal an mc_env_step of that group alone, bit for bit. */
int mc_batch_step(mc_batch* b, mc_env* const* envs, const void* const* actions,
                  const int* action_kinds, const float* const* looks, uint8_t* const* dones,
                  int n, void* stream);

/* Autotuner for a mix of groups: steps a scratch mirror of the mix (same
 * configs, never the caller's state) through mc_batch_step and, bucket by
 * bucket (distinct world, image and obs mode; most work first), keeps the
 * launch shape that shortens the whole mixed launch; the result must also beat
 * the default shapes side by side.  Then picks the launch form (persistent or
 * not, mc_batch_set_persistent).  Timings are best of two runs of `steps`
 * steps.  Sets the shape of every listed env and records each bucket's shape
 * in the batch (mc_batch_tuned_shape).  Synchronises `stream`. */
int mc_batch_autotune(mc_batch* b, mc_env* const* envs, int n, int steps, void* stream);
int mc_batch_tuned_shape(const mc_batch* b, const mc_config* cfg, int* threads,
                         int* waves_per_env, int* lift);

/* ---- serving (no reference counterpart) ----------------------------------
 * A continuous-batching server: a loop thread owns n_streams HIP streams and
 * merges the pending step requests of every admitted group (at most one in
 * flight per group; up to 64 groups per launch) into one mc_batch_step
 * launch per free stream, so requests arriving while a launch runs join the
 * next one and overlap it on the device.  Groups are mc_env handles whose
 * state stays resident; admission is FIFO under an HBM budget (`bytes` per
 * group as the caller accounts it).  Latency = completion seen by the loop
 * minus submit time.  Action/look/dones pointers must stay valid until the
 * request completes; a group's outputs are those of its last completed step.
 * mc_server_autotune must not run while requests are pending. */

typedef struct mc_server mc_server;

typedef struct mc_server_record {
  int64_t req_id;
  int64_t group;
  int64_t t_submit_ns;   /* steady clock */
  int64_t t_launch_ns;
  int64_t t_done_ns;     /* -1: the launch failed */
} mc_server_record;

enum { MC_GROUP_WAITING = 0, MC_GROUP_ADMITTED = 1, MC_GROUP_REMOVED = 2 };

int mc_server_create(int64_t hbm_budget_bytes, int n_streams, mc_server** out);
int mc_server_destroy(mc_server* s);
const char* mc_server_last_error(mc_server* s);
int mc_server_add_group(mc_server* s, mc_env* env, int64_t bytes, int batch_size, int* gid);
int mc_server_group_state(mc_server* s, int gid, int* state);
int mc_server_remove_group(mc_server* s, int gid);
int mc_server_submit(mc_server* s, int gid, const void* actions, int action_kind,
                     const float* look, uint8_t* dones, int64_t* req_id);
/* Closed loop: after each completion the group's next of `count` steps is
 * submitted at once (actions at actions_base + k * stride_bytes). */
int mc_server_set_replay(mc_server* s, int gid, const void* actions_base, int64_t stride_bytes,
                         int64_t count, int action_kind, const float* look_base,
                         int64_t look_stride_bytes, uint8_t* dones);
int mc_server_start_replay(mc_server* s, int gid);
/* Open loop at planned times: request k of group gids[k] (actions[k], of
 * action_kind, no look) arrives t_offset_ns[k] after this call; offsets must
 * be non-decreasing.  The loop thread queues each one at its arrival time and
 * records that time as its submit time, so a late release counts against its
 * latency.  Adds to any schedule still running. */
int mc_server_schedule(mc_server* s, const int* gids, const void* const* actions,
                       int action_kind, const int64_t* t_offset_ns, int64_t n);
/* Blocks until no request is pending, in flight, scheduled or left to replay. */
int mc_server_drain(mc_server* s);
int mc_server_take_records(mc_server* s, mc_server_record* out, int max, int* n);
int mc_server_stats(mc_server* s, int64_t* launches, int64_t* env_steps, int64_t* requests,
                    int64_t* used_bytes);
int mc_server_autotune(mc_server* s, int steps, void* stream);

/* ---- test hooks -------------------------------------------------------- */

/* Device evaluation of the shared sin/cos (mc_math.h) on n inputs. */
int mc_debug_sincosf(const float* x, float* s, float* c, int n, void* stream);
/* Scan of the ray-setup arithmetic (the cores of correctly rounded 1/x, sqrt
 * and n/d used on the ray fast path) against the full IEEE operations on the
 * device.  mode 0: reciprocals of float bit patterns [begin, begin+count) with
 * |x| in [2^-90, 2^90]; mode 1: square roots of patterns in [2^-60, 2^60];
 * mode 2: count hashed quotients.  out (device, 3 x u32, zeroed by the caller):
 * += inputs checked, += mismatches, one mismatching input. */
int mc_debug_fast_math(int mode, uint32_t begin, uint32_t count, uint32_t* out, void* stream);
/* Device binary16 arithmetic of the half-precision DDA on n float pairs:
 * ha = half(a), sum = half(a) + half(b) (binary16 bits), lt = half(a) < half(b). */
int mc_debug_f16(const float* a, const float* b, uint16_t* ha, uint16_t* sum, uint8_t* lt, int n,
                 void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MINECU_HIP_H */
