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

// Dynamic LDS bytes k_solve needs for a world of NL links and ND dofs (its link mirror).
size_t steel_solve_lds_bytes(uint32_t NL, uint32_t ND);
// Upper bound of k_solve's dynamic LDS (worlds needing more are refused at batch creation).
static constexpr size_t STEEL_SOLVE_LDS_MAX = 64 * 1024;

// Workgroups of k_narrow per launch for a device of `cus` compute units (four 4-wave groups per CU).
inline uint32_t steel_narrow_groups(uint32_t cus) { return cus * 4; }

// solve_lds: dynamic LDS of k_solve, the maximum of steel_solve_lds_bytes over the batch's worlds.
// pair_offs: device scratch of nworlds + 1 words (per-world pair offsets for k_narrow).
// narrow_groups: k_narrow's grid (steel_narrow_groups).
// marks (optional): four events recorded before collide, prepare, solve and after solve.
hipError_t steel_launch_substep(const WorldDev* worlds, uint32_t nworlds, uint32_t block, size_t solve_lds, uint32_t* pair_offs,
                                uint32_t narrow_groups, hipStream_t stream, uint32_t stages, hipEvent_t* marks = nullptr);
