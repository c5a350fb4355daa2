// Serving runtime (C ABI in include/steel_amd.h, "Serving").
//
// One device pool of the caller's HBM budget holds every admitted world as a single block (its buffers as
// 256-byte aligned slices, the same per-world layout a batch uses).  A first-fit free list with coalescing
// hands out blocks; admission that does not fit returns STEEL_OUT_OF_MEMORY without allocating, so the
// caller keeps the request queued.  Worlds are stepped in lanes: each lane has its own stream and device
// table of world pointers, and one lane frame is S substeps x (k_collide, k_prepare, k_solve) over the worlds
// the lane holds when the frame is launched.  Worlds join and leave a lane between its frames; a released
// world's block returns to the pool only after the lane frame that last used it has finished.  Lanes run
// concurrently, so a lane of small worlds is not held back by a lane of large ones.
#include <map>
#include <unordered_map>
#include "steel_runtime.h"

namespace {

// First-fit allocator over [0, cap) with a coalescing free list (offset -> size).
struct PoolAlloc {
  size_t cap = 0, used = 0;
  std::map<size_t, size_t> free_;
  void init(size_t c) { cap = c; used = 0; free_.clear(); if (c) free_[0] = c; }
  bool alloc(size_t n, size_t& off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      off = it->first;
      size_t rest = it->second - n;
      free_.erase(it);
      if (rest) free_[off + n] = rest;
      used += n;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    used -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
    }
  }
  size_t largest_free() const {
    size_t m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
  }
};

struct ServerWorld {
  uint32_t lane = 0;
  size_t off = 0, bytes = 0;         // pool block
  size_t koff[K_COUNT], ksize[K_COUNT];
  WorldDev dev;
  uint32_t NL = 0, ND = 0, NB = 0, NS = 0, NCOLW = 0, mfloats = 0, NEST = 0;  // dynamic LDS inputs
};

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  bool in_flight = false;             // a frame was launched and not yet waited for by step/read
  std::vector<uint32_t> members;      // worlds the next frame steps
  bool dirty = true;                  // members changed since the device table was written
  uint32_t nframe = 0;                // worlds in the last launched frame
  WorldDev* d_table = nullptr;
  WorldDev* h_table = nullptr;        // pinned staging of the table
  uint32_t* pair_offs = nullptr;      // k_narrow's per-world pair offsets (cap + 1 words)
  uint32_t cap = 0;
  uint32_t substeps = 0;              // substeps per frame of the lane's worlds; set while the lane is empty
  uint32_t block = 64;
  LaunchLds sl;                       // dynamic LDS over the lane's worlds
  std::vector<uint32_t> pending_free; // released while a frame may still use them
};

}  // namespace

struct steel_server {
  steel_device* dev = nullptr;
  char* pool = nullptr;
  PoolAlloc alloc;
  hipStream_t upload = nullptr;
  std::vector<Lane> lanes;
  std::unordered_map<uint32_t, ServerWorld> worlds;
  uint32_t next_id = 1;
};

static void free_world(steel_server* s, uint32_t id) {
  auto it = s->worlds.find(id);
  if (it == s->worlds.end()) return;
  s->alloc.release(it->second.off, it->second.bytes);
  s->worlds.erase(it);
}

// Return the blocks of worlds released from lanes whose frames have finished (does not wait).
static void reclaim(steel_server* s) {
  for (Lane& O : s->lanes)
    if (!O.pending_free.empty() && (!O.in_flight || hipEventQuery(O.ev_end) == hipSuccess)) {
      for (uint32_t x : O.pending_free) free_world(s, x);
      O.pending_free.clear();
    }
}

// Wait for the lane's in-flight frame, then return the blocks of worlds released meanwhile.
static int lane_settle(steel_server* s, Lane& L) {
  if (L.in_flight) {
    HIPCHK(hipEventSynchronize(L.ev_end));
    L.in_flight = false;
  }
  for (uint32_t id : L.pending_free) free_world(s, id);
  L.pending_free.clear();
  return STEEL_OK;
}

extern "C" {

int steel_server_create(steel_device* dev, size_t budget_bytes, uint32_t nlanes, steel_server** out) {
  if (!dev || !out || nlanes == 0 || nlanes > 64 || budget_bytes < 4096) return fail(STEEL_BAD_ARGUMENT, "bad server arguments");
  HIPCHK(hipSetDevice(dev->ordinal));
  steel_server* s = new steel_server();
  s->dev = dev;
  size_t budget = budget_bytes & ~(size_t)255;
  if (hipMalloc(&s->pool, budget) != hipSuccess) {
    delete s;
    return fail(STEEL_OUT_OF_MEMORY, "pool of %zu bytes does not fit on the device", budget);
  }
  s->alloc.init(budget);
  s->lanes.resize(nlanes);
  bool ok = hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking) == hipSuccess;
  for (Lane& L : s->lanes) {
    ok = ok && hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreate(&L.ev_start) == hipSuccess && hipEventCreate(&L.ev_end) == hipSuccess;
  }
  if (!ok) {
    steel_server_destroy(s);
    return fail(STEEL_ERROR, "stream/event creation failed");
  }
  *out = s;
  return STEEL_OK;
}

int steel_server_destroy(steel_server* s) {
  if (!s) return STEEL_OK;
  (void)hipSetDevice(s->dev->ordinal);
  for (Lane& L : s->lanes) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (L.d_table) (void)hipFree(L.d_table);
    if (L.h_table) (void)hipHostFree(L.h_table);
    if (L.pair_offs) (void)hipFree(L.pair_offs);
    if (L.ev_start) (void)hipEventDestroy(L.ev_start);
    if (L.ev_end) (void)hipEventDestroy(L.ev_end);
    if (L.stream) (void)hipStreamDestroy(L.stream);
  }
  if (s->upload) { (void)hipStreamSynchronize(s->upload); (void)hipStreamDestroy(s->upload); }
  if (s->pool) (void)hipFree(s->pool);
  delete s;
  return STEEL_OK;
}

int steel_server_configure_lane(steel_server* s, uint32_t lane, uint32_t block) {
  if (!s || lane >= s->lanes.size()) return fail(STEEL_BAD_ARGUMENT, "bad lane %u", lane);
  if (block != 64 && block != 128 && block != 256 && block != 1024) return fail(STEEL_BAD_ARGUMENT, "block must be 64, 128, 256 or 1024");
  s->lanes[lane].block = block;
  return STEEL_OK;
}

int steel_server_world_bytes(const steel_world_desc* w, size_t* bytes) {
  if (!w || !w->header || !bytes) return fail(STEEL_BAD_ARGUMENT, "null argument");
  Sizes sz;
  int rc = world_sizes(w->header, sz);
  if (rc) return rc;
  size_t t = 0;
  for (int k = 0; k < K_COUNT; k++) t += align256(sz.b[k]);
  *bytes = t;
  return STEEL_OK;
}

int steel_server_admit(steel_server* s, const steel_world_desc* w, uint32_t lane, uint32_t* id) {
  if (!s || !w || !id || lane >= s->lanes.size()) return fail(STEEL_BAD_ARGUMENT, "bad admit arguments");
  if (!w->header || !w->statF || !w->statU || !w->poses0 || !w->vel0 || !w->art0) return fail(STEEL_BAD_ARGUMENT, "null world array");
  Lane& L = s->lanes[lane];
  uint32_t S = w->header[W_SUBSTEPS];
  if (L.members.empty()) L.substeps = 0;  // an empty lane takes the substep count of the next world it admits
  if (L.substeps && L.substeps != S) return fail(STEEL_BAD_ARGUMENT, "lane %u steps %u substeps per frame, the world %u", lane, L.substeps, S);
  Sizes sz;
  int rc = world_sizes(w->header, sz);
  if (rc) return rc;
  LaunchLds probe;
  rc = world_solve_lds(w->header, w->statF, probe, false);
  if (rc) return rc;
  ServerWorld W;
  W.lane = lane;
  size_t t = 0;
  for (int k = 0; k < K_COUNT; k++) { W.koff[k] = t; W.ksize[k] = sz.b[k]; t += align256(sz.b[k]); }
  W.bytes = t;
  reclaim(s);  // blocks of released worlds whose last frame has finished
  if (!s->alloc.alloc(W.bytes, W.off))
    return fail(STEEL_OUT_OF_MEMORY, "pool holds %zu of %zu bytes; the world needs %zu (largest free block %zu)", s->alloc.used,
                s->alloc.cap, W.bytes, s->alloc.largest_free());
  HIPCHK(hipSetDevice(s->dev->ordinal));
  char* base[K_COUNT];
  for (int k = 0; k < K_COUNT; k++) base[k] = s->pool + W.off + W.koff[k];
  struct { int k; const void* src; } srcs[] = {{K_HDR, w->header}, {K_STATF, w->statF}, {K_STATU, w->statU},
                                               {K_POSES, w->poses0}, {K_VEL, w->vel0}, {K_ART, w->art0},
                                               {K_POSES0, w->poses0}, {K_VEL0, w->vel0}, {K_ART0, w->art0}};
  hipError_t e = hipSuccess;
  for (auto& p : srcs)
    if (e == hipSuccess) e = hipMemcpyAsync(base[p.k], p.src, W.ksize[p.k], hipMemcpyHostToDevice, s->upload);
  for (int k : K_ZEROED) if (e == hipSuccess) e = hipMemsetAsync(base[k], 0, W.ksize[k], s->upload);
  if (e == hipSuccess) e = hipStreamSynchronize(s->upload);
  if (e != hipSuccess) {
    s->alloc.release(W.off, W.bytes);
    return fail(STEEL_ERROR, "world upload failed: %s", hipGetErrorString(e));
  }
  W.dev = world_dev_at(base);
  W.NL = w->header[W_NL]; W.ND = w->header[W_ND]; W.NB = w->header[W_NB]; W.NS = w->header[W_NS]; W.NCOLW = w->header[W_NCOLW];
  W.NEST = w->header[W_NEST];
  W.mfloats = world_mirror_floats(w->header, w->statF);
  uint32_t wid = s->next_id++;
  s->worlds[wid] = W;
  L.substeps = S;
  L.members.push_back(wid);
  L.dirty = true;
  *id = wid;
  return STEEL_OK;
}

int steel_server_release(steel_server* s, uint32_t id) {
  if (!s) return fail(STEEL_BAD_ARGUMENT, "null server");
  auto it = s->worlds.find(id);
  if (it == s->worlds.end()) return fail(STEEL_BAD_ARGUMENT, "unknown world %u", id);
  Lane& L = s->lanes[it->second.lane];
  auto m = std::find(L.members.begin(), L.members.end(), id);
  if (m == L.members.end()) return fail(STEEL_BAD_ARGUMENT, "world %u was already released", id);
  L.members.erase(m);
  L.dirty = true;
  if (L.in_flight) L.pending_free.push_back(id);
  else free_world(s, id);
  return STEEL_OK;
}

int steel_server_step_lane(steel_server* s, uint32_t lane) {
  if (!s || lane >= s->lanes.size()) return fail(STEEL_BAD_ARGUMENT, "bad lane %u", lane);
  Lane& L = s->lanes[lane];
  HIPCHK(hipSetDevice(s->dev->ordinal));
  int rc = lane_settle(s, L);
  if (rc) return rc;
  uint32_t n = (uint32_t)L.members.size();
  if (L.dirty) {
    if (n > L.cap) {
      uint32_t cap = std::max<uint32_t>(n, std::max<uint32_t>(256, L.cap * 2));
      if (L.d_table) HIPCHK(hipFree(L.d_table));
      if (L.h_table) HIPCHK(hipHostFree(L.h_table));
      if (L.pair_offs) HIPCHK(hipFree(L.pair_offs));
      L.d_table = nullptr; L.h_table = nullptr; L.pair_offs = nullptr; L.cap = 0;
      HIPCHK(hipMalloc(&L.d_table, sizeof(WorldDev) * cap));
      HIPCHK(hipHostMalloc(&L.h_table, sizeof(WorldDev) * cap));
      HIPCHK(hipMalloc(&L.pair_offs, sizeof(uint32_t) * (cap + 1)));
      L.cap = cap;
    }
    L.sl = LaunchLds();
    for (uint32_t i = 0; i < n; i++) {
      const ServerWorld& W = s->worlds[L.members[i]];
      L.h_table[i] = W.dev;
      L.sl.add(W.NL, W.ND, W.mfloats, W.NB, W.NS, W.NCOLW, W.NEST != 0);
    }
    if (n) HIPCHK(hipMemcpyAsync(L.d_table, L.h_table, sizeof(WorldDev) * n, hipMemcpyHostToDevice, L.stream));
    L.dirty = false;
  }
  L.nframe = n;
  if (n == 0) return STEEL_OK;
  HIPCHK(hipEventRecord(L.ev_start, L.stream));
  for (uint32_t k = 0; k < L.substeps; k++) HIPCHK(steel_launch_substep(L.d_table, n, L.block, L.sl, L.pair_offs,
                                                                       steel_narrow_groups((uint32_t)s->dev->prop.multiProcessorCount), L.stream, STAGE_ALL));
  HIPCHK(hipEventRecord(L.ev_end, L.stream));
  L.in_flight = true;
  return STEEL_OK;
}

int steel_server_lane_poll(steel_server* s, uint32_t lane, int* done, uint32_t* nworlds, float* gpu_ms) {
  if (!s || lane >= s->lanes.size() || !done) return fail(STEEL_BAD_ARGUMENT, "bad poll arguments");
  Lane& L = s->lanes[lane];
  if (nworlds) *nworlds = L.nframe;
  if (gpu_ms) *gpu_ms = 0.0f;
  if (!L.in_flight) { *done = 1; return STEEL_OK; }
  hipError_t e = hipEventQuery(L.ev_end);
  if (e == hipErrorNotReady) { *done = 0; return STEEL_OK; }
  if (e != hipSuccess) return fail(STEEL_ERROR, "lane %u: %s", lane, hipGetErrorString(e));
  *done = 1;
  if (gpu_ms) HIPCHK(hipEventElapsedTime(gpu_ms, L.ev_start, L.ev_end));
  return STEEL_OK;
}

int steel_server_read(steel_server* s, uint32_t id, uint32_t buf, size_t byte_off, size_t bytes, void* dst) {
  if (!s || (!dst && bytes)) return fail(STEEL_BAD_ARGUMENT, "null argument");
  auto it = s->worlds.find(id);
  int k = buf_kind(buf);
  if (it == s->worlds.end() || k < 0) return fail(STEEL_BAD_ARGUMENT, "unknown world %u or buffer %u", id, buf);
  const ServerWorld& W = it->second;
  if (byte_off > W.ksize[k] || bytes > W.ksize[k] - byte_off) return fail(STEEL_BAD_ARGUMENT, "read of %zu bytes at %zu past the buffer's %zu", bytes, byte_off, W.ksize[k]);
  HIPCHK(hipSetDevice(s->dev->ordinal));
  HIPCHK(hipStreamSynchronize(s->lanes[W.lane].stream));
  HIPCHK(hipMemcpy(dst, s->pool + W.off + W.koff[k] + byte_off, bytes, hipMemcpyDeviceToHost));
  return STEEL_OK;
}

int steel_server_write(steel_server* s, uint32_t id, uint32_t buf, size_t byte_off, size_t bytes, const void* src) {
  if (!s || (!src && bytes)) return fail(STEEL_BAD_ARGUMENT, "null argument");
  auto it = s->worlds.find(id);
  int k = buf_kind(buf);
  if (it == s->worlds.end() || k < 0) return fail(STEEL_BAD_ARGUMENT, "unknown world %u or buffer %u", id, buf);
  const ServerWorld& W = it->second;
  if (byte_off > W.ksize[k] || bytes > W.ksize[k] - byte_off) return fail(STEEL_BAD_ARGUMENT, "write of %zu bytes at %zu past the buffer's %zu", bytes, byte_off, W.ksize[k]);
  HIPCHK(hipSetDevice(s->dev->ordinal));
  HIPCHK(hipStreamSynchronize(s->lanes[W.lane].stream));
  HIPCHK(hipMemcpy(s->pool + W.off + W.koff[k] + byte_off, src, bytes, hipMemcpyHostToDevice));
  return STEEL_OK;
}

int steel_server_info(steel_server* s, size_t* used, size_t* capacity, uint32_t* nworlds) {
  if (!s) return fail(STEEL_BAD_ARGUMENT, "null server");
  reclaim(s);
  if (used) *used = s->alloc.used;
  if (capacity) *capacity = s->alloc.cap;
  if (nworlds) *nworlds = (uint32_t)s->worlds.size();
  return STEEL_OK;
}

}  // extern "C"
