// mc_kernels.hip - launch wrappers of the step and render kernels for one
// workgroup width MC_NT (64, 256 or 512) and one setting MC_LIFT (0/1) of
// the upward-ray exit.  The Makefile compiles this file once per pair; each
// object exports the KernelSet kKernels<NT>_<LIFT> (mc_launch.h).
#include "mc_launch.h"

#ifndef MC_NT
#error "compile with -DMC_NT=64|256|512 -DMC_LIFT=0|1"
#endif

#define MC_CAT3(a, b, c) a##b##_##c
#define MC_SET_NAME(nt, lift) MC_CAT3(kKernels, nt, lift)

namespace mcl {
namespace {

constexpr int NT = MC_NT;
constexpr bool LIFT = MC_LIFT != 0;

bool planar(int obs_mode) { return obs_mode == mc::kObsU8Chw || obs_mode == mc::kObsF32Chw; }

template <int KIND>
void launch_kind(const mc::Params& p, int obs_mode, int wpe, size_t lds, hipStream_t st) {
  const int epw = (NT / mc::kWave) / wpe;
  const dim3 grid((p.B + epw - 1) / epw);
  // The planar modes, the float32 modes, the seeded terrain and voxel
  // collision run kernels of their own (template constants), so the
  // flat-world interleaved uint8 kernels keep their code unchanged.
  if (KIND == mc::kItemStep && p.collision && p.terrain)
    hipLaunchKernelGGL((mc::k_env_col<NT, LIFT, true>), grid, dim3(NT), lds, st, p, obs_mode, wpe);
  else if (KIND == mc::kItemStep && p.collision)
    hipLaunchKernelGGL((mc::k_env_col<NT, LIFT, false>), grid, dim3(NT), lds, st, p, obs_mode, wpe);
  else if (p.terrain)
    hipLaunchKernelGGL((mc::k_env_terrain<NT, KIND, LIFT>), grid, dim3(NT), lds, st, p, obs_mode,
                       wpe);
  else if (KIND != mc::kItemRenderCameras && KIND != mc::kItemRenderBasis &&
           obs_mode == mc::kObsF32Chw)
    hipLaunchKernelGGL((mc::k_env<NT, KIND, LIFT, true, false, true>), grid, dim3(NT), lds, st, p,
                       obs_mode, wpe);
  else if (KIND != mc::kItemRenderCameras && KIND != mc::kItemRenderBasis && planar(obs_mode))
    hipLaunchKernelGGL((mc::k_env<NT, KIND, LIFT, true>), grid, dim3(NT), lds, st, p, obs_mode,
                       wpe);
  else if (KIND != mc::kItemRenderCameras && KIND != mc::kItemRenderBasis &&
           obs_mode == mc::kObsF32)
    hipLaunchKernelGGL((mc::k_env<NT, KIND, LIFT, false, false, true>), grid, dim3(NT), lds, st,
                       p, obs_mode, wpe);
  else
    hipLaunchKernelGGL((mc::k_env<NT, KIND, LIFT>), grid, dim3(NT), lds, st, p, obs_mode, wpe);
}

hipError_t env(int kind, const mc::Params& p, int obs_mode, int wpe, size_t lds, hipStream_t st) {
  switch (kind) {
    case mc::kItemStep: launch_kind<mc::kItemStep>(p, obs_mode, wpe, lds, st); break;
    case mc::kItemResetRender: launch_kind<mc::kItemResetRender>(p, obs_mode, wpe, lds, st); break;
    case mc::kItemRenderCameras:
      launch_kind<mc::kItemRenderCameras>(p, obs_mode, wpe, lds, st);
      break;
    case mc::kItemRenderBasis: launch_kind<mc::kItemRenderBasis>(p, obs_mode, wpe, lds, st); break;
    default: launch_kind<mc::kItemRenderState>(p, obs_mode, wpe, lds, st); break;
  }
  return hipGetLastError();
}

hipError_t env_count(const mc::Params& p, int obs_mode, int wpe, size_t lds, hipStream_t st) {
  const int epw = (NT / mc::kWave) / wpe;
  hipLaunchKernelGGL((mc::k_env<NT, mc::kItemStep, LIFT, false, true>), dim3((p.B + epw - 1) / epw),
                     dim3(NT), lds, st, p, obs_mode, wpe);
  return hipGetLastError();
}

#if MC_NT >= 256
hipError_t fb(const mc::Params& p, int obs_mode, dim3 grid, size_t lds, hipStream_t st) {
  if (p.collision && p.terrain)
    hipLaunchKernelGGL((mc::k_env_fb_col<NT, LIFT, true>), grid, dim3(NT), lds, st, p, obs_mode);
  else if (p.collision)
    hipLaunchKernelGGL((mc::k_env_fb_col<NT, LIFT, false>), grid, dim3(NT), lds, st, p, obs_mode);
  else if (p.terrain)
    hipLaunchKernelGGL((mc::k_env_fb_terrain<NT, LIFT>), grid, dim3(NT), lds, st, p, obs_mode);
  else if (planar(obs_mode))
    hipLaunchKernelGGL((mc::k_env_fb<NT, LIFT, true>), grid, dim3(NT), lds, st, p, obs_mode);
  else
    hipLaunchKernelGGL((mc::k_env_fb<NT, LIFT>), grid, dim3(NT), lds, st, p, obs_mode);
  return hipGetLastError();
}

template <bool PERSIST, bool CHW, bool TER, bool COL>
void multi_launch(const mc::MultiArgs& a, int grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((mc::k_multi<NT, LIFT, PERSIST, CHW, TER, COL>), dim3(grid), dim3(NT), lds, st,
                     a);
}

template <bool PERSIST>
void multi_family(const mc::MultiArgs& a, int family, int grid, size_t lds, hipStream_t st) {
  switch (family) {
    case kFamilyPlanar: multi_launch<PERSIST, true, false, false>(a, grid, lds, st); break;
    case kFamilyTerrain: multi_launch<PERSIST, false, true, false>(a, grid, lds, st); break;
    case kFamilyCollision: multi_launch<PERSIST, false, false, true>(a, grid, lds, st); break;
    case kFamilyCollisionTerrain: multi_launch<PERSIST, false, true, true>(a, grid, lds, st); break;
    default: multi_launch<PERSIST, false, false, false>(a, grid, lds, st); break;
  }
}

hipError_t multi(const mc::MultiArgs& a, int family, bool persist, int grid, size_t lds,
                 hipStream_t st) {
  if (persist) multi_family<true>(a, family, grid, lds, st);
  else multi_family<false>(a, family, grid, lds, st);
  return hipGetLastError();
}

hipError_t multi_fb(const mc::MultiArgs& a, int family, dim3 grid, size_t lds, hipStream_t st) {
  switch (family) {
    case kFamilyPlanar:
      hipLaunchKernelGGL((mc::k_multi_fb<NT, LIFT, true, false, false>), grid, dim3(NT), lds, st, a);
      break;
    case kFamilyTerrain:
      hipLaunchKernelGGL((mc::k_multi_fb<NT, LIFT, false, true, false>), grid, dim3(NT), lds, st, a);
      break;
    case kFamilyCollision:
      hipLaunchKernelGGL((mc::k_multi_fb<NT, LIFT, false, false, true>), grid, dim3(NT), lds, st, a);
      break;
    case kFamilyCollisionTerrain:
      hipLaunchKernelGGL((mc::k_multi_fb<NT, LIFT, false, true, true>), grid, dim3(NT), lds, st, a);
      break;
    default: hipLaunchKernelGGL((mc::k_multi_fb<NT, LIFT>), grid, dim3(NT), lds, st, a); break;
  }
  return hipGetLastError();
}

template <bool CHW, bool TER, bool COL>
int blocks_of(size_t lds) {
  int nb = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, mc::k_multi<NT, LIFT, true, CHW, TER, COL>, NT, lds);
  if (e != hipSuccess || nb < 1) {
    (void)hipGetLastError();
    nb = 1;
  }
  return nb;
}

int multi_blocks(int family, size_t lds) {
  switch (family) {
    case kFamilyPlanar: return blocks_of<true, false, false>(lds);
    case kFamilyTerrain: return blocks_of<false, true, false>(lds);
    case kFamilyCollision: return blocks_of<false, false, true>(lds);
    case kFamilyCollisionTerrain: return blocks_of<false, true, true>(lds);
    default: return blocks_of<false, false, false>(lds);
  }
}
#endif

#if MC_NT == 256 && MC_LIFT
template <int KIND>
void glb_kind(const mc::Params& p, int obs_mode, bool list, dim3 grid, hipStream_t st) {
  const dim3 blk(256);
  if (KIND == mc::kItemStep && p.collision) {
    if (p.terrain && list)
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, true, true, true>), grid, blk, 0, st, p, obs_mode);
    else if (p.terrain)
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, true, true, false>), grid, blk, 0, st, p, obs_mode);
    else if (list)
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, false, true, true>), grid, blk, 0, st, p, obs_mode);
    else
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, false, true, false>), grid, blk, 0, st, p, obs_mode);
  } else if (p.terrain) {
    if (KIND == mc::kItemStep && list)
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, true, false, true>), grid, blk, 0, st, p, obs_mode);
    else
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, true, false, false>), grid, blk, 0, st, p, obs_mode);
  } else if (KIND != mc::kItemRenderCameras && KIND != mc::kItemRenderBasis && planar(obs_mode)) {
    if (KIND == mc::kItemStep && list)
      hipLaunchKernelGGL((mc::k_env_glb<KIND, true, false, false, true>), grid, blk, 0, st, p, obs_mode);
    else
      hipLaunchKernelGGL((mc::k_env_glb<KIND, true, false, false, false>), grid, blk, 0, st, p, obs_mode);
  } else {
    if (KIND == mc::kItemStep && list)
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, false, false, true>), grid, blk, 0, st, p, obs_mode);
    else
      hipLaunchKernelGGL((mc::k_env_glb<KIND, false, false, false, false>), grid, blk, 0, st, p, obs_mode);
  }
}

hipError_t glb(int kind, const mc::Params& p, int obs_mode, bool list, dim3 grid, hipStream_t st) {
  switch (kind) {
    case mc::kItemStep: glb_kind<mc::kItemStep>(p, obs_mode, list, grid, st); break;
    case mc::kItemResetRender: glb_kind<mc::kItemResetRender>(p, obs_mode, false, grid, st); break;
    case mc::kItemRenderCameras: glb_kind<mc::kItemRenderCameras>(p, obs_mode, false, grid, st); break;
    case mc::kItemRenderBasis: glb_kind<mc::kItemRenderBasis>(p, obs_mode, false, grid, st); break;
    default: glb_kind<mc::kItemRenderState>(p, obs_mode, false, grid, st); break;
  }
  return hipGetLastError();
}
#endif

}  // namespace

const KernelSet& MC_SET_NAME(MC_NT, MC_LIFT)() {
#if MC_NT == 256 && MC_LIFT
  static const KernelSet set = {env, env_count, fb, multi, multi_fb, multi_blocks, glb};
#elif MC_NT >= 256
  static const KernelSet set = {env, env_count, fb, multi, multi_fb, multi_blocks, nullptr};
#else
  static const KernelSet set = {env, env_count, nullptr, nullptr, nullptr, nullptr, nullptr};
#endif
  return set;
}

}  // namespace mcl
