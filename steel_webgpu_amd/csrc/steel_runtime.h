// Internal declarations shared by the host runtime files (batch: steel_api.cpp, serving: steel_server.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <map>
#include <utility>
#include "steel_dev.h"
#include "steel_launch.h"
#include "../../include/steel_amd.h"

// Sets the thread's last-error message and returns code.
int fail(int code, const char* fmt, ...);
#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(STEEL_ERROR, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

enum Kind { K_HDR, K_STATF, K_STATU, K_POSES, K_VEL, K_ART, K_AKEY, K_AVAL, K_CTR, K_STATS, K_PAIRS, K_CRAW, K_CONTACTS, K_ROWS, K_SCHED,
            K_SKEY, K_SVAL, K_TW, K_IMP, K_AQD, K_POSES0, K_VEL0, K_ART0, K_WARM, K_MIRROR, K_COUNT };

struct Sizes { size_t b[K_COUNT]; };

inline uint32_t world_mirror_floats(const uint32_t* h, const float* statF);
// statF (when given) sizes k_solve's global link mirror of a world too large for its LDS (K_MIRROR).
inline int world_sizes(const uint32_t* h, Sizes& s, const float* statF = nullptr) {
  uint32_t NB = h[W_NB], NS = h[W_NS], ND = h[W_ND], MAXC = h[W_MAXC], MAXP = h[W_MAXP], AH = h[W_AH], ROWS = h[W_ROWS];
  if (MAXC < NS || MAXC == 0) return fail(STEEL_BAD_ARGUMENT, "world capacity MAXC=%u is below its shape count %u", MAXC, NS);
  if (h[W_NDA] > 64 || h[W_MAXLP] > 64) return fail(STEEL_BAD_ARGUMENT, "articulations of more than 64 dofs or links are not supported");
  if (h[W_MAXPATH] > 64 || h[W_MAXPATH] * h[W_JM] > 192) return fail(STEEL_BAD_ARGUMENT, "articulation root paths deeper than 64 links are not supported");
  if (h[W_UH] > MAXC || h[W_MAXU] > MAXC) return fail(STEEL_BAD_ARGUMENT, "unit capacities exceed the contact capacity");
  SchedLayout L = sched_layout(MAXC, h[W_MAXU], h[W_UH], h[W_NCOLW] * 32, NB, h[W_NCOLW], h[W_NL], h[W_WH], h[W_MAXDENSE]);
  s.b[K_HDR] = STEEL_WORLD_HEADER_WORDS * 4;
  s.b[K_STATF] = (size_t)h[W_NSTATF] * 16;
  s.b[K_STATU] = (size_t)std::max(h[W_NSTATU], 1u) * 4;
  s.b[K_POSES] = (size_t)h[W_NPOSES] * 16;
  s.b[K_VEL] = (size_t)std::max(NB, 1u) * 32;
  s.b[K_ART] = (size_t)h[W_NART] * 4;
  s.b[K_AKEY] = (size_t)2 * AH * 4;
  s.b[K_AVAL] = (size_t)2 * AH * STEEL_AVN * 4;
  s.b[K_CTR] = 32;
  s.b[K_STATS] = 64;
  s.b[K_PAIRS] = (size_t)2 * MAXP * 4;
  s.b[K_CRAW] = (size_t)6 * MAXC * 16;
  s.b[K_CONTACTS] = (size_t)6 * MAXC * 16;
  s.b[K_ROWS] = (size_t)MAXC * ROWS * 16;
  s.b[K_SCHED] = (size_t)L.TOTAL * 4;
  s.b[K_SKEY] = (size_t)2 * MAXC * 8;
  s.b[K_SVAL] = (size_t)2 * MAXC * 4;
  s.b[K_TW] = (size_t)std::max(NB, 1u) * 32;
  s.b[K_IMP] = (size_t)std::max(NB, 1u) * 32;
  s.b[K_AQD] = (size_t)std::max(ND, 1u) * 4;
  s.b[K_POSES0] = s.b[K_POSES];
  s.b[K_VEL0] = s.b[K_VEL];
  s.b[K_ART0] = s.b[K_ART];
  s.b[K_WARM] = std::max((size_t)h[W_WH] * 10 * 4, (size_t)4);
  s.b[K_MIRROR] = 0;
  if (statF) {
    size_t mb = solve_lds_bytes(h[W_NL], h[W_ND], world_mirror_floats(h, statF), NB, false, h[W_NEST] != 0);
    if (mb > STEEL_SOLVE_LDS_MAX) s.b[K_MIRROR] = mb;
  }
  return STEEL_OK;
}


inline int buf_kind(uint32_t buf) {
  static const int map[STEEL_BUF_COUNT] = {K_POSES, K_VEL, K_ART, K_AKEY, K_AVAL, K_PAIRS, K_CONTACTS, K_ROWS, K_SCHED, K_CTR, K_STATS,
                                           K_WARM};
  return buf < STEEL_BUF_COUNT ? map[buf] : -1;
}

// Pointers of one world's buffers, given the base address of each kind's slice (mirror: the world has a K_MIRROR slice).
inline WorldDev world_dev_at(char* const* base, bool mirror = false) {
  WorldDev x;
  x.hdr = (const uint32_t*)base[K_HDR];
  x.statF = (const float4*)base[K_STATF];
  x.statU = (const uint32_t*)base[K_STATU];
  x.poses = (float4*)base[K_POSES];
  x.vel = (float4*)base[K_VEL];
  x.art = (float*)base[K_ART];
  x.akey = (uint32_t*)base[K_AKEY];
  x.aval = (uint32_t*)base[K_AVAL];
  x.ctr = (uint32_t*)base[K_CTR];
  x.stats = (uint32_t*)base[K_STATS];
  x.pairs = (uint32_t*)base[K_PAIRS];
  x.craw = (float4*)base[K_CRAW];
  x.contacts = (float4*)base[K_CONTACTS];
  x.rows = (float4*)base[K_ROWS];
  x.sched = (uint32_t*)base[K_SCHED];
  x.skey = (uint64_t*)base[K_SKEY];
  x.sval = (uint32_t*)base[K_SVAL];
  x.tw = (float4*)base[K_TW];
  x.imp = (float4*)base[K_IMP];
  x.aqd = (float*)base[K_AQD];
  x.warm = (uint32_t*)base[K_WARM];
  x.mscr = mirror ? (float4*)base[K_MIRROR] : nullptr;
  return x;
}

// Floats of k_solve's link mirror for a world: the sum over its links of mirror_floats(ndof) (statF link
// records: ndof in the .w of the second float4).
inline uint32_t world_mirror_floats(const uint32_t* h, const float* statF) {
  uint32_t t = 0;
  for (uint32_t l = 0; l < h[W_NL]; l++) t += mirror_floats((uint32_t)statF[((size_t)h[W_SF_LINK] + l * 6 + 1) * 4 + 3]);
  return t;
}
// Adds the world to sl; with global_mirror false, refuses worlds whose mirror does not fit k_solve's LDS.
inline int world_solve_lds(const uint32_t* h, const float* statF, LaunchLds& sl, bool global_mirror = true) {
  uint32_t mf = world_mirror_floats(h, statF);
  if (!global_mirror && solve_lds_bytes(h[W_NL], h[W_ND], mf, h[W_NB], false, h[W_NEST] != 0) > STEEL_SOLVE_LDS_MAX)
    return fail(STEEL_UNSUPPORTED, "a world of %u articulation links exceeds the solve kernel's LDS link mirror", h[W_NL]);
  sl.add(h[W_NL], h[W_ND], mf, h[W_NB], h[W_NS], h[W_NCOLW], h[W_NEST] != 0, h[W_CELL] != 0);
  return STEEL_OK;
}

// Kinds cleared to zero when a world starts or restarts.
static constexpr int K_ZEROED[] = {K_AKEY, K_AVAL, K_CTR, K_STATS, K_TW, K_IMP, K_AQD, K_ROWS, K_SCHED, K_WARM};

struct steel_device {
  int ordinal;
  hipDeviceProp_t prop;
  unsigned char* noalloc_scratch = nullptr;  // this device's address of steel_noalloc.hip's scratch array
  std::map<std::string, std::pair<uint32_t, uint32_t>> tune_cache;  // autotuner: signature -> (block, narrow per CU)
  bool tune_file_loaded = false;
};

// Takes the device address of the no-allocation entry point's scratch array on the current device.
int steel_noalloc_init(steel_device* dev);
