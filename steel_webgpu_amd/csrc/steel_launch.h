// Kernel launch entry used by the host runtime.
#pragma once
#include <hip/hip_runtime.h>
#include "steel_dev.h"
#include "steel_team.cuh"

enum : uint32_t {
  STAGE_COLLIDE = 1u,   // shapes + broadphase (+ narrowphase and contact sort with STAGE_NARROW)
  STAGE_NARROW = 2u,
  STAGE_PREPARE = 4u,
  STAGE_SOLVE = 8u,
  STAGE_ALL = 15u,
};

// k_solve's per-link LDS mirror: a link of n dofs takes n^2 + 14n + 21 floats.
__host__ __device__ inline uint32_t mirror_floats(uint32_t n) { return n * n + 14 * n + 21; }
// k_solve's dynamic LDS for a world: link records and tree (2 NL uint4), per-link contributions (8 NL floats),
// joint velocities (ND floats), with tw_lds the twists and impulses of its NB bodies (64 bytes each), with nest
// the 256 partial sums of the momentum restart test (1 KB), and the link mirror (mfloats floats).
__host__ __device__ inline size_t solve_lds_bytes(uint32_t NL, uint32_t ND, uint32_t mfloats, uint32_t NB, bool tw_lds, bool nest = false) {
  size_t b = (size_t)NL * 64 + (((size_t)ND * 4 + 15) & ~(size_t)15);
  if (tw_lds) b += (size_t)(NB > 0 ? NB : 1) * 64;
  if (nest) b += 1024;
  return b + (((size_t)mfloats * 4 + 15) & ~(size_t)15);
}
// Upper bound of k_solve's dynamic LDS without the twists (worlds needing more are refused at creation).
static constexpr size_t STEEL_SOLVE_LDS_MAX = 64 * 1024;

// Shapes whose sorted AABBs and candidate runs k_broad keeps in LDS (36 bytes each, 144 KB); a world with
// more shapes sweeps from global memory.
static constexpr uint32_t STEEL_BROAD_LDS_MAX = 4096;
// Words of k_prepare's dynamic LDS: per-body color masks and first-pending slots, NB * (NCOLW + 1) words
// (64 KB); a world needing more colors serially from global masks.
static constexpr uint32_t STEEL_PREP_LDS_MAX_WORDS = 16384;

// Dynamic LDS over the worlds of one launch.  k_solve: the largest world's bytes without and with the twists
// in LDS; twists go to LDS when that still lets four waves per SIMD stay resident (160 KB of LDS per CU
// shared by 16 waves: 16 / (block / 64) workgroups).  k_broad: the largest world's shape count, capped.
// k_prepare: the largest world's coloring words, capped.
// A world whose k_solve LDS (without the twists) exceeds STEEL_SOLVE_LDS_MAX keeps its link records and mirror in
// global memory (WorldDev::mscr, solve_lds_bytes of it); a launch holding one runs the solve's global-mirror
// instantiation (mgl), in which every world solves through the generic-pointer passes.
struct LaunchLds {
  size_t notw = 0, tw = 0;
  uint32_t ns = 0, cw = 0;
  bool mgl = false;
  bool cell = false;  // some world has a cell broadphase bound (W_CELL)
  void add(uint32_t NL, uint32_t ND, uint32_t mfloats, uint32_t NB, uint32_t NS, uint32_t NCOLW, bool nest, bool has_cell = false) {
    ns = ns > NS ? ns : NS;
    cell = cell || has_cell;
    uint32_t c = (NB > 0 ? NB : 1) * (NCOLW + 1);
    cw = cw > c ? cw : c;
    if (solve_lds_bytes(NL, ND, mfloats, NB, false, nest) > STEEL_SOLVE_LDS_MAX) { mgl = true; return; }
    notw = notw > solve_lds_bytes(NL, ND, mfloats, NB, false, nest) ? notw : solve_lds_bytes(NL, ND, mfloats, NB, false, nest);
    tw = tw > solve_lds_bytes(NL, ND, mfloats, NB, true, nest) ? tw : solve_lds_bytes(NL, ND, mfloats, NB, true, nest);
  }
  static size_t budget(uint32_t block) { return (size_t)160 * 1024 / (16 / (block / 64)) - 1024; }
  bool tw_lds(uint32_t block) const { return !mgl && tw <= budget(block); }
  size_t bytes(uint32_t block) const { return tw_lds(block) ? tw : notw; }
  uint32_t broad_cap() const { return ns < STEEL_BROAD_LDS_MAX ? ns : STEEL_BROAD_LDS_MAX; }
  size_t broad_bytes() const { return (size_t)(broad_cap() > 0 ? broad_cap() : 1) * 36 + 16; }
  uint32_t prep_words() const { return cw < STEEL_PREP_LDS_MAX_WORDS ? cw : STEEL_PREP_LDS_MAX_WORDS; }
};

// Workgroups of k_narrow per launch for a device of `cus` compute units: eight 4-wave groups per CU, the autotuner's
// choice on the 4096-world quadruped batch (more waves in flight hide the hull loads).
inline uint32_t steel_narrow_groups(uint32_t cus) { return cus * 8; }

// sl: dynamic LDS over the launch's worlds.
// pair_offs: device scratch of nworlds + 1 words (per-world pair offsets for k_narrow).
// narrow_groups: k_narrow's grid (steel_narrow_groups).
// marks (optional): four events recorded before collide, prepare, solve and after solve.
hipError_t steel_launch_substep(const WorldDev* worlds, uint32_t nworlds, uint32_t block, const LaunchLds& sl, uint32_t* pair_offs,
                                uint32_t narrow_groups, hipStream_t stream, uint32_t stages, hipEvent_t* marks = nullptr);

// One large world over the whole device (cooperative launches of the stage kernels, GRID_BS threads per workgroup).
// steel_grid_groups: workgroups per launch that can all be resident on a device of `cus` compute units.
// g: grid team scratch of grid_scratch_words(groups) words, zeroed once.
hipError_t steel_grid_groups(uint32_t cus, uint32_t* groups);
hipError_t steel_launch_substep_grid(const WorldDev* world, uint32_t groups, uint32_t* g, uint32_t* pair_offs, uint32_t narrow_groups,
                                     hipStream_t stream, uint32_t stages, hipEvent_t* marks = nullptr);
