// Kernel launch entry used by the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include "steel_dev.h"

enum : uint32_t {
  STAGE_COLLIDE = 1u,   // shapes + broadphase (+ narrowphase and contact sort with STAGE_NARROW)
  STAGE_NARROW = 2u,
  STAGE_PREPARE = 4u,
  STAGE_SOLVE = 8u,
  STAGE_ALL = 15u,
};

// marks (optional): four events recorded before collide, prepare, solve and after solve.
hipError_t steel_launch_substep(const WorldDev* worlds, uint32_t nworlds, uint32_t block, hipStream_t stream, uint32_t stages,
                                hipEvent_t* marks = nullptr);
