// mc_launch.h - internal interface between the host logic (mc_capi.hip) and
// the kernel launch wrappers of mc_kernels.hip.
//
// mc_kernels.hip is compiled once per workgroup width (64, 256, 512 threads)
// and setting of the upward-ray exit (lift 0/1); each object holds the step
// and render kernels of that width and exports one KernelSet.  The six
// objects compile in parallel, and every kernel instantiation lives in
// exactly one of them.
#pragma once
#include <hip/hip_runtime.h>

#include "mc_device.h"

namespace mcl {

// Mode families of batched groups: each runs kernels of its own.
enum { kFamilyFlat = 0, kFamilyPlanar = 1, kFamilyTerrain = 2, kFamilyCollision = 3,
       kFamilyCollisionTerrain = 4, kFamilies = 5 };

struct KernelSet {
  // One item per workgroup over a group: grid = ceil(B / epw) with epw = NT /
  // (64 wpe).  Dispatches on the item kind, the obs mode (interleaved or
  // planar) and the world (flat or seeded terrain, plane or voxel collision).
  hipError_t (*env)(int kind, const mc::Params& p, int obs_mode, int wpe, size_t lds,
                    hipStream_t st);
  // The step of a flat-world interleaved group with ray-step counting
  // (p.ray_steps; a measurement build of the step kernel).
  hipError_t (*env_count)(const mc::Params& p, int obs_mode, int wpe, size_t lds, hipStream_t st);
  // Full-height pass over the envs a staged-height step deferred (NULL for
  // the 64-thread set).
  hipError_t (*fb)(const mc::Params& p, int obs_mode, dim3 grid, size_t lds, hipStream_t st);
  // Several groups of one mode family (kFamily*) in one launch (NULL for the
  // 64-thread set).
  hipError_t (*multi)(const mc::MultiArgs& a, int family, bool persist, int grid, size_t lds,
                      hipStream_t st);
  hipError_t (*multi_fb)(const mc::MultiArgs& a, int family, dim3 grid, size_t lds,
                         hipStream_t st);
  // Resident workgroups per CU of the persistent multi kernel at `lds` bytes.
  int (*multi_blocks)(int family, size_t lds);
  // Worlds too large for one workgroup's LDS: the item with the world in
  // global memory, one env per 256-thread workgroup; list: the envs a staged
  // step deferred (p.fb).  Non-NULL in the 256-thread lift set only.
  hipError_t (*glb)(int kind, const mc::Params& p, int obs_mode, bool list, dim3 grid,
                    hipStream_t st);
};

const KernelSet& kKernels64_0();
const KernelSet& kKernels64_1();
const KernelSet& kKernels256_0();
const KernelSet& kKernels256_1();
const KernelSet& kKernels512_0();
const KernelSet& kKernels512_1();

inline const KernelSet& kernels(int nt, bool lift) {
  if (nt == 64) return lift ? kKernels64_1() : kKernels64_0();
  if (nt == 256) return lift ? kKernels256_1() : kKernels256_0();
  return lift ? kKernels512_1() : kKernels512_0();
}

}  // namespace mcl
