// Device-side world table shared by the host runtime and the kernels.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

// Per-world device pointers.  Static arrays and state are the packed-world arrays of
// include/steel_world.h; the rest is per-world scratch owned by the runtime.
struct WorldDev {
  const uint32_t* hdr;     // header words (STEEL_WORLD_HEADER_WORDS)
  const float4* statF;
  const uint32_t* statU;
  float4* poses;
  float4* vel;
  float* art;
  uint32_t* akey;          // friction-anchor keys [2*AH]
  uint32_t* aval;          // anchor values [2*AH*AVN]
  uint32_t* ctr;           // [0] pairs, [1] contacts, [2] substep count (parity), [3] control flags, [4-5] frame clock
  uint32_t* stats;         // [16], the reference's stats words
  uint32_t* pairs;         // [2*MAXP]
  float4* craw;            // [6*MAXC] narrowphase contacts, then
  float4* contacts;        // [6*MAXC] sorted contacts (5 vec4 + shape mu pair per record)
  float4* rows;            // [MAXC*ROWS]
  uint32_t* sched;         // schedule tables (sched_layout)
  uint64_t* skey;          // [2*MAXC] contact sort keys (ping-pong)
  uint32_t* sval;          // [2*MAXC]
  float4* tw;              // [2*NB] solver twists
  float4* imp;             // [2*NB] accumulated link impulses
  float* aqd;              // [ND] solver joint velocities
  uint32_t* warm;          // contact warm start: keys [2*WH] (two parities), then values [2*WH*4] (float bits)
  float4* mscr;            // k_solve's link records and mirror in global memory, for a world too large for its LDS (else null)
};

// Offsets of the schedule tables inside WorldDev::sched (uint32 words), from the world's capacities.
struct SchedLayout {
  uint32_t HKEY, UKEY, UTPW, HCNT, CSLOT, CLOCAL, USIZE, USTART, CSTART, META, CORDER, RK0, RK1, RK2, KEPT, OSTART, LEAD, SLOT, ROWOF,
      RUNIT, PCOL, PORD, UTC, UTP, RCON, UWOFF, WALK, HEADS, DSLOT, WCOL, WMASK, WCOUNT, LINKHIT, JSPLIT, SLOTT, NPAIRL, UF, UFF, WWIN,
      WSLOT, FPX, TOTAL;
};

__host__ __device__ inline SchedLayout sched_layout(uint32_t MAXC, uint32_t MAXU, uint32_t UH, uint32_t NCOL, uint32_t NB, uint32_t NCOLW,
                                                    uint32_t NL, uint32_t WH, uint32_t MAXDENSE) {
  SchedLayout s;
  uint32_t o = 0;
  s.HKEY = o; o += 2 * UH;                                 // run heads' body-pair keys (uint64, pair_key)
  s.UKEY = o; o += 2 * MAXU;                               // units' body-pair keys (uint64)
  s.UTPW = o; o += 2 * MAXU;                               // coloring: each unit's largest (color + 1, walk index) (uint64)
  s.HCNT = o; o += UH;
  s.CSLOT = o; o += MAXC;
  s.CLOCAL = o; o += MAXC;
  s.USIZE = o; o += MAXU;
  s.USTART = o; o += MAXU + 1;
  s.CSTART = o; o += NCOL + 2;
  s.META = o; o += 8;
  s.CORDER = o; o += MAXC;
  s.RK0 = o; o += MAXC;
  s.RK1 = o; o += MAXC;
  s.RK2 = o; o += MAXC;
  s.KEPT = o; o += MAXC;
  s.OSTART = o; o += MAXU;
  s.LEAD = o; o += MAXC;
  s.SLOT = o; o += MAXC * 26;
  s.ROWOF = o; o += MAXC;
  s.RUNIT = o; o += MAXC;
  s.PCOL = o; o += MAXC;
  s.PORD = o; o += MAXC;
  s.UTC = o; o += MAXU;
  s.UTP = o; o += MAXU;
  s.RCON = o; o += MAXC;
  s.UWOFF = o; o += UH;
  s.WALK = o; o += MAXC;
  s.HEADS = o; o += MAXC + 1;
  s.DSLOT = o; o += MAXC + 1;
  s.WCOL = o; o += MAXC;
  s.WMASK = o; o += (NB > 0 ? NB : 1) * NCOLW;
  s.WCOUNT = o; o += NCOL + 2;
  s.LINKHIT = o; o += NB > 0 ? NB : 1;
  s.JSPLIT = o; o += NL > 0 ? NL : 1;
  s.SLOTT = o; o += MAXC * 4;
  s.NPAIRL = o; o += 4;
  s.UF = o; o += (NB > 0 ? NB : 1) + (NL > 0 ? NL : 1);   // sleep islands: union-find parent per node (NB + NA <= NB + NL)
  s.UFF = o; o += (NB > 0 ? NB : 1) + (NL > 0 ? NL : 1);  // island flags
  o = (o + 1) & ~1u;
  s.WWIN = o; o += 2 * WH;                                 // warm table: winning row order per slot (uint64)
  s.WSLOT = o; o += WH ? MAXC + MAXDENSE : 0;              // warm table: slot of each row
  s.FPX = o; o += MAXC * 4;                                // prepare: each listed contact's reduction inputs (float4)
  s.TOTAL = o;
  return s;
}
