// steel_world_step_noalloc: a world step that allocates nothing, for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The caller holds the world's static arrays and state in device memory.  Everything else the step kernels use
// (broadphase pairs, raw and sorted contacts, rows, schedule tables, sort keys, solver twists, impulses and
// joint velocities, the one-world table and the narrowphase pair offsets) lives in `scratch`, a file-scope
// device array of STEEL_NOALLOC_SCRATCH_BYTES that every GPU holds once, for the life of the process.  Its
// device address differs per GPU: steel_noalloc_init (called by steel_device_open with the device current)
// takes it with hipGetSymbolAddress and keeps it in the steel_device; host code uses only that pointer.
//
// Every call on one GPU shares that GPU's array and overwrites it: calls that overlap on one GPU (two host
// threads, or two streams) race.  The tables the kernels expect to start at zero (rows, schedule, twists,
// impulses, joint velocities) are cleared by a fill kernel at the start of every call; nothing in the array
// carries from one call to the next, since the step rebuilds these tables every substep.
#include <hip/hip_runtime.h>
#include "steel_runtime.h"

static __device__ __align__(256) unsigned char scratch[STEEL_NOALLOC_SCRATCH_BYTES];

namespace {

// Scratch kinds the step uses beyond the caller's arrays, in layout order.
constexpr int K_SCRATCH[] = {K_PAIRS, K_CRAW, K_CONTACTS, K_ROWS, K_SCHED, K_SKEY, K_SVAL, K_TW, K_IMP, K_AQD};
// Of those, the ones a world starts with at zero (as a batch clears them at reset).
constexpr int K_SCRATCH_ZEROED[] = {K_ROWS, K_SCHED, K_TW, K_IMP, K_AQD};
constexpr size_t TABLE_BYTES = 256;   // one WorldDev record
constexpr size_t OFFS_BYTES = 256;    // k_narrow's pair offsets of one world (2 words)

struct Layout {
  size_t off[K_COUNT];
  size_t total;
};

Layout scratch_layout(const Sizes& sz) {
  Layout L{};
  size_t o = TABLE_BYTES + OFFS_BYTES;
  for (int k : K_SCRATCH) { L.off[k] = o; o += align256(sz.b[k]); }
  L.total = o;
  return L;
}

// Writes the one-world table and zeroes up to five ranges of the scratch array (uint4 stores; every range
// offset and length is a multiple of 16).
struct ZeroRanges {
  uint64_t off[5], len[5];
  uint32_t n;
};

__global__ void k_noalloc_setup(unsigned char* base, WorldDev w, ZeroRanges z) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *(WorldDev*)base = w;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t r = 0; r < z.n; r++) {
    uint4* p = (uint4*)(base + z.off[r]);
    const uint64_t n = z.len[r] / 16;
    for (uint64_t i = t0; i < n; i += stride) p[i] = make_uint4(0u, 0u, 0u, 0u);
  }
}

}  // namespace

int steel_noalloc_init(steel_device* dev) {
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return fail(STEEL_ERROR, "hipGetSymbolAddress(scratch) on device %d: %s", dev->ordinal, hipGetErrorString(e));
  dev->noalloc_scratch = (unsigned char*)p;
  return STEEL_OK;
}

extern "C" int steel_world_step_noalloc(steel_device* dev, const steel_world_desc* host, const steel_world_device_arrays* d,
                                        uint32_t frames, void* stream) {
  if (!dev || !host || !d || !host->header || !host->statF) return fail(STEEL_BAD_ARGUMENT, "null argument");
  if (!dev->noalloc_scratch) return fail(STEEL_ERROR, "device %d has no scratch address", dev->ordinal);
  if (!d->header || !d->statF || !d->statU || !d->poses || !d->vel || !d->art || !d->anchor_keys || !d->anchor_vals || !d->warm ||
      !d->counters || !d->stats)
    return fail(STEEL_BAD_ARGUMENT, "null device array");
  Sizes sz;
  int rc = world_sizes(host->header, sz);
  if (rc) return rc;
  Layout L = scratch_layout(sz);
  if (L.total > STEEL_NOALLOC_SCRATCH_BYTES)
    return fail(STEEL_BAD_ARGUMENT, "the world needs %zu bytes of step scratch; steel_world_step_noalloc holds %zu", L.total,
                (size_t)STEEL_NOALLOC_SCRATCH_BYTES);
  LaunchLds sl;
  rc = world_solve_lds(host->header, host->statF, sl, false);
  if (rc) return rc;
  if (frames == 0) return STEEL_OK;
  HIPCHK(hipSetDevice(dev->ordinal));
  unsigned char* base = dev->noalloc_scratch;
  WorldDev w;
  w.hdr = d->header;
  w.statF = (const float4*)d->statF;
  w.statU = d->statU;
  w.poses = (float4*)d->poses;
  w.vel = (float4*)d->vel;
  w.art = d->art;
  w.akey = d->anchor_keys;
  w.aval = d->anchor_vals;
  w.ctr = d->counters;
  w.stats = d->stats;
  w.warm = d->warm;
  w.pairs = (uint32_t*)(base + L.off[K_PAIRS]);
  w.craw = (float4*)(base + L.off[K_CRAW]);
  w.contacts = (float4*)(base + L.off[K_CONTACTS]);
  w.rows = (float4*)(base + L.off[K_ROWS]);
  w.sched = (uint32_t*)(base + L.off[K_SCHED]);
  w.skey = (uint64_t*)(base + L.off[K_SKEY]);
  w.sval = (uint32_t*)(base + L.off[K_SVAL]);
  w.tw = (float4*)(base + L.off[K_TW]);
  w.imp = (float4*)(base + L.off[K_IMP]);
  w.aqd = (float*)(base + L.off[K_AQD]);
  ZeroRanges z{};
  for (int k : K_SCRATCH_ZEROED) { z.off[z.n] = L.off[k]; z.len[z.n] = align256(sz.b[k]); z.n++; }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_noalloc_setup, dim3(256), dim3(256), 0, s, base, w, z);
  HIPCHK(hipGetLastError());
  const WorldDev* table = (const WorldDev*)base;
  uint32_t* offs = (uint32_t*)(base + TABLE_BYTES);
  const uint32_t groups = steel_narrow_groups((uint32_t)dev->prop.multiProcessorCount);
  const uint32_t S = host->header[W_SUBSTEPS];
  for (uint32_t f = 0; f < frames; f++)
    for (uint32_t k = 0; k < S; k++) HIPCHK(steel_launch_substep(table, 1, 64, sl, offs, groups, s, STAGE_ALL));
  return STEEL_OK;
}
