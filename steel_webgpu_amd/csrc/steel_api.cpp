// Host runtime and C ABI of the Steel step library (declared in include/steel_amd.h): device and batch.
//
// A batch keeps every per-world buffer in one device arena per buffer kind (one hipMalloc each, 256-byte
// aligned slices), a device table of per-world pointers, a copy of the initial state for reset, and one
// HIP stream.  A frame is S substeps x (k_collide, k_prepare, k_solve); with graphs on, the frame is
// captured once and replayed, since nothing in it depends on the host (the substep parity lives in each
// world's device counter).
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include "steel_runtime.h"

thread_local std::string g_err;
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}


struct steel_batch {
  steel_device* dev;
  uint32_t n;
  uint32_t substeps;
  uint32_t block = 256;
  LaunchLds sl;                    // dynamic LDS over the batch's worlds
  uint32_t* pair_offs = nullptr;   // k_narrow's per-world pair offsets (n + 1 words)
  uint32_t narrow_groups = 0;
  bool use_graph = true;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  char* arena[K_COUNT] = {};
  std::vector<size_t> off[K_COUNT];
  std::vector<size_t> size[K_COUNT];
  size_t total[K_COUNT] = {};
  WorldDev* d_worlds = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint32_t> headers;  // n * STEEL_WORLD_HEADER_WORDS
  bool wide = false;               // one large world over the whole device (steel_batch_set_wide)
  uint32_t wide_groups = 0;        // workgroups of its cooperative launches
  uint32_t* grid_scratch = nullptr;
};

extern "C" {

const char* steel_last_error(void) { return g_err.c_str(); }

int steel_device_open(int ordinal, steel_device** out) {
  if (!out) return fail(STEEL_BAD_ARGUMENT, "null output");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(STEEL_UNSUPPORTED, "no HIP device is available");
  if (ordinal < 0 || ordinal >= count) return fail(STEEL_BAD_ARGUMENT, "device ordinal %d out of range (%d devices)", ordinal, count);
  steel_device* d = new steel_device();
  d->ordinal = ordinal;
  if (hipSetDevice(ordinal) != hipSuccess || hipGetDeviceProperties(&d->prop, ordinal) != hipSuccess) {
    delete d;
    return fail(STEEL_UNSUPPORTED, "cannot open HIP device %d", ordinal);
  }
  if (std::strncmp(d->prop.gcnArchName, "gfx950", 6) != 0) {
    std::string arch = d->prop.gcnArchName;
    delete d;
    return fail(STEEL_UNSUPPORTED, "device %d is %s; this library is built for gfx950", ordinal, arch.c_str());
  }
  int rc = steel_noalloc_init(d);
  if (rc) { delete d; return rc; }
  *out = d;
  return STEEL_OK;
}

int steel_device_close(steel_device* dev) {
  delete dev;
  return STEEL_OK;
}

int steel_device_info(steel_device* dev, char* name, size_t name_len, size_t* total_mem, uint32_t* num_cus) {
  if (!dev) return fail(STEEL_BAD_ARGUMENT, "null device");
  if (name && name_len) { std::snprintf(name, name_len, "%s (%s)", dev->prop.name, dev->prop.gcnArchName); }
  if (total_mem) *total_mem = dev->prop.totalGlobalMem;
  if (num_cus) *num_cus = (uint32_t)dev->prop.multiProcessorCount;
  return STEEL_OK;
}

int steel_estimate_bytes(uint32_t nworlds, const steel_world_desc* worlds, size_t* bytes) {
  if (!bytes || (nworlds && !worlds)) return fail(STEEL_BAD_ARGUMENT, "null argument");
  size_t tot = 0;
  for (uint32_t i = 0; i < nworlds; i++) {
    Sizes s;
    int rc = world_sizes(worlds[i].header, s, worlds[i].statF);
    if (rc) return rc;
    for (int k = 0; k < K_COUNT; k++) tot += align256(s.b[k]);
  }
  tot += align256((size_t)nworlds * sizeof(WorldDev));
  *bytes = tot;
  return STEEL_OK;
}

static void batch_free(steel_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (int k = 0; k < K_COUNT; k++) if (b->arena[k]) (void)hipFree(b->arena[k]);
  if (b->d_worlds) (void)hipFree(b->d_worlds);
  if (b->pair_offs) (void)hipFree(b->pair_offs);
  if (b->grid_scratch) (void)hipFree(b->grid_scratch);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

int steel_batch_reset(steel_batch* b, int32_t world) {
  if (!b) return fail(STEEL_BAD_ARGUMENT, "null batch");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  uint32_t lo = 0, hi = b->n;
  if (world >= 0) {
    if ((uint32_t)world >= b->n) return fail(STEEL_BAD_ARGUMENT, "world %d out of range", world);
    lo = (uint32_t)world; hi = lo + 1;
  }
  if (world < 0) {
    HIPCHK(hipMemcpyAsync(b->arena[K_POSES], b->arena[K_POSES0], b->total[K_POSES], hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->arena[K_VEL], b->arena[K_VEL0], b->total[K_VEL], hipMemcpyDeviceToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->arena[K_ART], b->arena[K_ART0], b->total[K_ART], hipMemcpyDeviceToDevice, b->stream));
    for (int k : K_ZEROED) HIPCHK(hipMemsetAsync(b->arena[k], 0, b->total[k], b->stream));
  } else {
    for (uint32_t i = lo; i < hi; i++) {
      HIPCHK(hipMemcpyAsync(b->arena[K_POSES] + b->off[K_POSES][i], b->arena[K_POSES0] + b->off[K_POSES0][i], b->size[K_POSES][i], hipMemcpyDeviceToDevice, b->stream));
      HIPCHK(hipMemcpyAsync(b->arena[K_VEL] + b->off[K_VEL][i], b->arena[K_VEL0] + b->off[K_VEL0][i], b->size[K_VEL][i], hipMemcpyDeviceToDevice, b->stream));
      HIPCHK(hipMemcpyAsync(b->arena[K_ART] + b->off[K_ART][i], b->arena[K_ART0] + b->off[K_ART0][i], b->size[K_ART][i], hipMemcpyDeviceToDevice, b->stream));
      for (int k : K_ZEROED) HIPCHK(hipMemsetAsync(b->arena[k] + b->off[k][i], 0, b->size[k][i], b->stream));
    }
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return STEEL_OK;
}

int steel_batch_create(steel_device* dev, uint32_t nworlds, const steel_world_desc* worlds, steel_batch** out) {
  if (!dev || !out || nworlds == 0 || !worlds) return fail(STEEL_BAD_ARGUMENT, "bad batch arguments");
  HIPCHK(hipSetDevice(dev->ordinal));
  steel_batch* b = new steel_batch();
  b->dev = dev;
  b->n = nworlds;
  b->substeps = worlds[0].header[W_SUBSTEPS];
  b->headers.resize((size_t)nworlds * STEEL_WORLD_HEADER_WORDS);
  for (int k = 0; k < K_COUNT; k++) { b->off[k].resize(nworlds); b->size[k].resize(nworlds); }
  for (uint32_t i = 0; i < nworlds; i++) {
    const steel_world_desc& w = worlds[i];
    if (!w.header || !w.statF || !w.statU || !w.poses0 || !w.vel0 || !w.art0) { delete b; return fail(STEEL_BAD_ARGUMENT, "world %u: null array", i); }
    if (w.header[W_SUBSTEPS] != b->substeps) { delete b; return fail(STEEL_BAD_ARGUMENT, "world %u: all worlds of a batch step the same substep count", i); }
    Sizes s;
    int rc = world_sizes(w.header, s, w.statF);
    if (rc) { delete b; return rc; }
    for (int k = 0; k < K_COUNT; k++) { b->off[k][i] = b->total[k]; b->size[k][i] = s.b[k]; b->total[k] += align256(s.b[k]); }
    std::memcpy(&b->headers[(size_t)i * STEEL_WORLD_HEADER_WORDS], w.header, STEEL_WORLD_HEADER_WORDS * 4);
    rc = world_solve_lds(w.header, w.statF, b->sl);
    if (rc) { delete b; return rc; }
  }
  for (int k = 0; k < K_COUNT; k++) {
    if (hipMalloc(&b->arena[k], std::max<size_t>(b->total[k], 256)) != hipSuccess) {
      batch_free(b);
      delete b;
      return fail(STEEL_OUT_OF_MEMORY, "device allocation of %zu bytes failed", b->total[k]);
    }
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess) {
    batch_free(b);
    delete b;
    return fail(STEEL_ERROR, "stream/event creation failed");
  }
  // Upload static arrays and initial state through one pinned staging buffer per kind.
  std::vector<char> host;
  auto upload = [&](int k, const void* (*get)(const steel_world_desc&), const steel_world_desc* ws) -> int {
    host.assign(b->total[k], 0);
    for (uint32_t i = 0; i < nworlds; i++) std::memcpy(host.data() + b->off[k][i], get(ws[i]), b->size[k][i]);
    HIPCHK(hipMemcpy(b->arena[k], host.data(), b->total[k], hipMemcpyHostToDevice));
    return STEEL_OK;
  };
  int rc = 0;
  rc |= upload(K_HDR, [](const steel_world_desc& w) -> const void* { return w.header; }, worlds);
  rc |= upload(K_STATF, [](const steel_world_desc& w) -> const void* { return w.statF; }, worlds);
  rc |= upload(K_STATU, [](const steel_world_desc& w) -> const void* { return w.statU; }, worlds);
  rc |= upload(K_POSES0, [](const steel_world_desc& w) -> const void* { return w.poses0; }, worlds);
  rc |= upload(K_VEL0, [](const steel_world_desc& w) -> const void* { return w.vel0; }, worlds);
  rc |= upload(K_ART0, [](const steel_world_desc& w) -> const void* { return w.art0; }, worlds);
  if (rc) { batch_free(b); delete b; return STEEL_ERROR; }
  std::vector<WorldDev> wd(nworlds);
  for (uint32_t i = 0; i < nworlds; i++) {
    char* base[K_COUNT];
    for (int k = 0; k < K_COUNT; k++) base[k] = b->arena[k] + b->off[k][i];
    wd[i] = world_dev_at(base, b->size[K_MIRROR][i] > 0);
  }
  b->narrow_groups = steel_narrow_groups((uint32_t)dev->prop.multiProcessorCount);
  if (hipMalloc(&b->pair_offs, sizeof(uint32_t) * (nworlds + 1)) != hipSuccess ||
      hipMalloc(&b->d_worlds, sizeof(WorldDev) * nworlds) != hipSuccess ||
      hipMemcpy(b->d_worlds, wd.data(), sizeof(WorldDev) * nworlds, hipMemcpyHostToDevice) != hipSuccess) {
    batch_free(b);
    delete b;
    return fail(STEEL_OUT_OF_MEMORY, "world table allocation failed");
  }
  rc = steel_batch_reset(b, -1);
  if (rc) { batch_free(b); delete b; return rc; }
  *out = b;
  return STEEL_OK;
}

int steel_batch_destroy(steel_batch* b) {
  if (!b) return STEEL_OK;
  (void)hipSetDevice(b->dev->ordinal);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  delete b;
  return STEEL_OK;
}

int steel_batch_configure(steel_batch* b, uint32_t block, int use_graph) {
  if (!b) return fail(STEEL_BAD_ARGUMENT, "null batch");
  if (block != 64 && block != 128 && block != 256 && block != 1024) return fail(STEEL_BAD_ARGUMENT, "block must be 64, 128, 256 or 1024");
  if (b->gexec && (block != b->block || (use_graph != 0) != b->use_graph)) {
    (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr;
    (void)hipGraphDestroy(b->graph); b->graph = nullptr;
  }
  b->block = block;
  b->use_graph = use_graph != 0;
  return STEEL_OK;
}

// One substep's launches: one workgroup per world and stage, or for a wide batch the device-wide launches.
static hipError_t launch_substep(steel_batch* b, uint32_t stages, hipEvent_t* marks = nullptr) {
  if (b->wide) return steel_launch_substep_grid(b->d_worlds, b->wide_groups, b->grid_scratch, b->pair_offs, b->narrow_groups, b->stream, stages, marks);
  return steel_launch_substep(b->d_worlds, b->n, b->block, b->sl, b->pair_offs, b->narrow_groups, b->stream, stages, marks);
}

// Device-wide stepping stays off until its launches run clean on the GPU.
static constexpr bool kWideEnabled = false;
static bool steel_wide_enabled() { return kWideEnabled; }

static int launch_frame(steel_batch* b) {
  for (uint32_t s = 0; s < b->substeps; s++) HIPCHK(launch_substep(b, STAGE_ALL));
  return STEEL_OK;
}

int steel_batch_set_wide(steel_batch* b, int enable) {
  if (!b) return fail(STEEL_BAD_ARGUMENT, "null batch");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  if (!enable) { b->wide = false; return STEEL_OK; }
  if (!steel_wide_enabled())
    return fail(STEEL_UNSUPPORTED, "device-wide stepping is switched off in this build: its cooperative launches faulted the GPU "
                                   "and the cause is not yet found");
  const uint32_t* h = b->headers.data();
  if (b->n != 1) return fail(STEEL_UNSUPPORTED, "device-wide stepping takes a batch of one world, not %u", b->n);
  if (b->size[K_MIRROR][0] == 0) return fail(STEEL_UNSUPPORTED, "the world fits one solve workgroup; step it as a batch");
  if (h[W_SLEEP] || h[W_WH] || h[W_NEST] || h[W_VELITERS])
    return fail(STEEL_UNSUPPORTED, "device-wide stepping does not support sleeping, contact warm start, Nesterov momentum or velocity iterations");
  if (!b->grid_scratch) {
    HIPCHK(steel_grid_groups((uint32_t)b->dev->prop.multiProcessorCount, &b->wide_groups));
    const size_t bytes = grid_scratch_words(b->wide_groups) * 4;
    HIPCHK(hipMalloc(&b->grid_scratch, bytes));
    HIPCHK(hipMemset(b->grid_scratch, 0, bytes));
  }
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr;
    (void)hipGraphDestroy(b->graph); b->graph = nullptr;
  }
  b->wide = true;
  return STEEL_OK;
}

int steel_batch_step(steel_batch* b, uint32_t frames) {
  if (!b) return fail(STEEL_BAD_ARGUMENT, "null batch");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  if (frames == 0) return STEEL_OK;
  const bool graph = b->use_graph && !b->wide;
  if (graph && !b->gexec) {
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_frame(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(STEEL_ERROR, "graph capture failed: %s", hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  }
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  for (uint32_t f = 0; f < frames; f++) {
    if (graph) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    else { int rc = launch_frame(b); if (rc) return rc; }
  }
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  b->timed = true;
  return STEEL_OK;
}

int steel_batch_substep_stages(steel_batch* b, uint32_t stages) {
  if (!b) return fail(STEEL_BAD_ARGUMENT, "null batch");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  HIPCHK(launch_substep(b, stages));
  return STEEL_OK;
}

int steel_batch_step_timed(steel_batch* b, uint32_t frames, float* frame_ms) {
  if (!b || (frames && !frame_ms)) return fail(STEEL_BAD_ARGUMENT, "null argument");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  std::vector<hipEvent_t> ev(frames + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  int rc = STEEL_OK;
  HIPCHK(hipEventRecord(ev[0], b->stream));
  for (uint32_t f = 0; f < frames && rc == STEEL_OK; f++) {
    rc = steel_batch_step(b, 1);
    if (rc == STEEL_OK && hipEventRecord(ev[f + 1], b->stream) != hipSuccess) rc = fail(STEEL_ERROR, "event record failed");
  }
  if (rc == STEEL_OK && hipEventSynchronize(ev[frames]) != hipSuccess) rc = fail(STEEL_ERROR, "event sync failed");
  for (uint32_t f = 0; f < frames && rc == STEEL_OK; f++) {
    if (hipEventElapsedTime(&frame_ms[f], ev[f], ev[f + 1]) != hipSuccess) rc = fail(STEEL_ERROR, "event timing failed");
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

int steel_batch_profile_kernels(steel_batch* b, uint32_t substeps, double kernel_ms[3]) {
  if (!b || !kernel_ms) return fail(STEEL_BAD_ARGUMENT, "null argument");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  std::vector<hipEvent_t> ev(4 * substeps);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  int rc = STEEL_OK;
  for (uint32_t s = 0; s < substeps && rc == STEEL_OK; s++)
    if (launch_substep(b, STAGE_ALL, &ev[4 * s]) != hipSuccess) rc = fail(STEEL_ERROR, "launch failed");
  if (rc == STEEL_OK && hipStreamSynchronize(b->stream) != hipSuccess) rc = fail(STEEL_ERROR, "sync failed");
  for (int k = 0; k < 3; k++) kernel_ms[k] = 0.0;
  for (uint32_t s = 0; s < substeps && rc == STEEL_OK; s++)
    for (int k = 0; k < 3; k++) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[4 * s + k], ev[4 * s + k + 1]) != hipSuccess) { rc = fail(STEEL_ERROR, "event timing failed"); break; }
      kernel_ms[k] += ms / substeps;
    }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}

// A wide batch's barriers record a wait they gave up on (grid_barrier); its results are then not the step's.
static int check_wide(steel_batch* b) {
  if (!b->grid_scratch) return STEEL_OK;
  uint32_t flag = 0;
  HIPCHK(hipMemcpy(&flag, b->grid_scratch + 2, 4, hipMemcpyDeviceToHost));
  if (flag) return fail(STEEL_ERROR, "a device-wide launch's workgroups did not all reach a barrier");
  return STEEL_OK;
}

int steel_batch_sync(steel_batch* b) {
  if (!b) return fail(STEEL_BAD_ARGUMENT, "null batch");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  HIPCHK(hipStreamSynchronize(b->stream));
  return check_wide(b);
}


int steel_batch_buffer_size(steel_batch* b, uint32_t world, uint32_t buf, size_t* bytes) {
  if (!b || !bytes) return fail(STEEL_BAD_ARGUMENT, "null argument");
  int k = buf_kind(buf);
  if (k < 0 || world >= b->n) return fail(STEEL_BAD_ARGUMENT, "bad buffer %u or world %u", buf, world);
  *bytes = b->size[k][world];
  return STEEL_OK;
}

int steel_batch_read(steel_batch* b, uint32_t world, uint32_t buf, size_t byte_off, size_t bytes, void* dst) {
  if (!b || (!dst && bytes)) return fail(STEEL_BAD_ARGUMENT, "null argument");
  int k = buf_kind(buf);
  if (k < 0 || world >= b->n) return fail(STEEL_BAD_ARGUMENT, "bad buffer %u or world %u", buf, world);
  if (byte_off > b->size[k][world] || bytes > b->size[k][world] - byte_off) return fail(STEEL_BAD_ARGUMENT, "read of %zu bytes at %zu past the buffer's %zu", bytes, byte_off, b->size[k][world]);
  HIPCHK(hipSetDevice(b->dev->ordinal));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (int rc = check_wide(b)) return rc;
  HIPCHK(hipMemcpy(dst, b->arena[k] + b->off[k][world] + byte_off, bytes, hipMemcpyDeviceToHost));
  return STEEL_OK;
}

int steel_batch_write(steel_batch* b, uint32_t world, uint32_t buf, size_t byte_off, size_t bytes, const void* src) {
  if (!b || (!src && bytes)) return fail(STEEL_BAD_ARGUMENT, "null argument");
  int k = buf_kind(buf);
  if (k < 0 || world >= b->n) return fail(STEEL_BAD_ARGUMENT, "bad buffer %u or world %u", buf, world);
  if (byte_off > b->size[k][world] || bytes > b->size[k][world] - byte_off) return fail(STEEL_BAD_ARGUMENT, "write of %zu bytes at %zu past the buffer's %zu", bytes, byte_off, b->size[k][world]);
  HIPCHK(hipSetDevice(b->dev->ordinal));
  HIPCHK(hipMemcpyAsync(b->arena[k] + b->off[k][world] + byte_off, src, bytes, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return STEEL_OK;
}

int steel_batch_stats(steel_batch* b, uint32_t world, uint32_t out[16], double* gpu_ms) {
  if (!b) return fail(STEEL_BAD_ARGUMENT, "null batch");
  if (out) { int rc = steel_batch_read(b, world, STEEL_BUF_STATS, 0, 64, out); if (rc) return rc; }
  if (gpu_ms) {
    *gpu_ms = 0.0;
    if (b->timed) {
      HIPCHK(hipEventSynchronize(b->ev1));
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      *gpu_ms = ms;
    }
  }
  return STEEL_OK;
}

// ---- Autotuner
// The batch's signature: what decides which schedule is fastest (the worlds' sizes and the substep count).
static std::string tune_signature(const steel_batch* b) {
  uint64_t nb = 0, ns = 0, nl = 0, mb = 0, ms = 0, ml = 0;
  for (uint32_t i = 0; i < b->n; i++) {
    const uint32_t* h = &b->headers[(size_t)i * STEEL_WORLD_HEADER_WORDS];
    nb += h[W_NB]; ns += h[W_NS]; nl += h[W_NL];
    mb = std::max<uint64_t>(mb, h[W_NB]); ms = std::max<uint64_t>(ms, h[W_NS]); ml = std::max<uint64_t>(ml, h[W_NL]);
  }
  char buf[256];
  std::snprintf(buf, sizeof buf, "n%u_b%llu_%llu_s%llu_%llu_l%llu_%llu_S%u", b->n, (unsigned long long)nb, (unsigned long long)mb,
                (unsigned long long)ns, (unsigned long long)ms, (unsigned long long)nl, (unsigned long long)ml, b->substeps);
  return buf;
}

static std::string tune_device_key(const steel_device* d) {
  char buf[160];
  std::snprintf(buf, sizeof buf, "%s_%dcu", d->prop.gcnArchName, d->prop.multiProcessorCount);
  for (char* c = buf; *c; c++) if (*c == ' ' || *c == ':') *c = '_';
  return buf;
}

// Persistent cache (STEEL_TUNE_CACHE): one line per entry, "<device key> <signature> <block> <narrow per CU>".
static std::mutex g_tune_file_mu;
static void tune_file_load(steel_device* d) {
  const char* path = std::getenv("STEEL_TUNE_CACHE");
  if (!path || !*path || d->tune_file_loaded) return;
  d->tune_file_loaded = true;
  std::lock_guard<std::mutex> lk(g_tune_file_mu);
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  const std::string dk = tune_device_key(d);
  char key[160], sig[256];
  unsigned blk, npc;
  while (std::fscanf(f, "%159s %255s %u %u", key, sig, &blk, &npc) == 4)
    if (dk == key && (blk == 64 || blk == 128 || blk == 256) && (npc == 2 || npc == 4 || npc == 8)) d->tune_cache[sig] = {blk, npc};
  std::fclose(f);
}
static void tune_file_append(const steel_device* d, const std::string& sig, uint32_t blk, uint32_t npc) {
  const char* path = std::getenv("STEEL_TUNE_CACHE");
  if (!path || !*path) return;
  std::lock_guard<std::mutex> lk(g_tune_file_mu);
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%s %s %u %u\n", tune_device_key(d).c_str(), sig.c_str(), blk, npc);
  std::fclose(f);
}

static void batch_set_schedule(steel_batch* b, uint32_t block, uint32_t narrow_groups) {
  if (b->gexec && (block != b->block || narrow_groups != b->narrow_groups)) {
    (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr;
    (void)hipGraphDestroy(b->graph); b->graph = nullptr;
  }
  b->block = block;
  b->narrow_groups = narrow_groups;
}

// The state a frame carries to the next (everything else is rebuilt every substep).
static constexpr int K_STATE[] = {K_POSES, K_VEL, K_ART, K_AKEY, K_AVAL, K_CTR, K_STATS, K_WARM};

int steel_batch_autotune(steel_batch* b, uint32_t frames, uint32_t flags, steel_tune_result* out) {
  if (!b || !out) return fail(STEEL_BAD_ARGUMENT, "null argument");
  if (frames == 0 || frames > 1000) return fail(STEEL_BAD_ARGUMENT, "trial frames must be 1..1000");
  HIPCHK(hipSetDevice(b->dev->ordinal));
  *out = steel_tune_result{};
  const uint32_t cus = (uint32_t)b->dev->prop.multiProcessorCount;
  const std::string sig = tune_signature(b);
  tune_file_load(b->dev);
  if (!(flags & STEEL_TUNE_REMEASURE)) {
    auto it = b->dev->tune_cache.find(sig);
    if (it != b->dev->tune_cache.end()) {
      batch_set_schedule(b, it->second.first, cus * it->second.second);
      *out = steel_tune_result{it->second.first, it->second.second, 1u, 0u, 0.0f};
      return STEEL_OK;
    }
  }
  // Snapshot of the state, restored after each trial and at the end.
  size_t tot = 0;
  for (int k : K_STATE) tot += b->total[k];
  char* snap = nullptr;
  HIPCHK(hipStreamSynchronize(b->stream));
  if (hipMalloc(&snap, tot) != hipSuccess) return fail(STEEL_OUT_OF_MEMORY, "autotune: no room for a %zu-byte state snapshot", tot);
  auto copy_state = [&](bool save) -> hipError_t {
    size_t o = 0;
    for (int k : K_STATE) {
      hipError_t e = save ? hipMemcpyAsync(snap + o, b->arena[k], b->total[k], hipMemcpyDeviceToDevice, b->stream)
                          : hipMemcpyAsync(b->arena[k], snap + o, b->total[k], hipMemcpyDeviceToDevice, b->stream);
      if (e != hipSuccess) return e;
      o += b->total[k];
    }
    return hipStreamSynchronize(b->stream);
  };
  const uint32_t block0 = b->block, groups0 = b->narrow_groups;
  int rc = STEEL_OK;
  if (copy_state(true) != hipSuccess) rc = fail(STEEL_ERROR, "autotune: state snapshot failed");
  uint32_t best_b = block0, best_g = groups0 / (cus ? cus : 1);
  float best = 0.0f;
  std::vector<float> fm(frames);
  static const uint32_t blocks[] = {64, 128, 256}, per_cu[] = {2, 4, 8};
  for (uint32_t bs : blocks)
    for (uint32_t g : per_cu) {
      if (rc) break;
      batch_set_schedule(b, bs, cus * g);
      // one untimed frame captures the graph and warms the caches; then the timed trial
      rc = steel_batch_step(b, 1);
      if (!rc) rc = steel_batch_step_timed(b, frames, fm.data());
      if (!rc && copy_state(false) != hipSuccess) rc = fail(STEEL_ERROR, "autotune: state restore failed");
      if (rc) break;
      float ms = 0.0f;
      for (float v : fm) ms += v;
      ms /= (float)frames;
      out->trials++;
      if (best == 0.0f || ms < best) { best = ms; best_b = bs; best_g = g; }
    }
  if (!rc && copy_state(false) != hipSuccess) rc = fail(STEEL_ERROR, "autotune: state restore failed");
  (void)hipFree(snap);
  if (rc) { batch_set_schedule(b, block0, groups0); return rc; }
  batch_set_schedule(b, best_b, cus * best_g);
  b->dev->tune_cache[sig] = {best_b, best_g};
  tune_file_append(b->dev, sig, best_b, best_g);
  out->block = best_b;
  out->narrow_per_cu = best_g;
  out->frame_ms = best;
  return STEEL_OK;
}

int steel_batch_device_bytes(steel_batch* b, size_t* bytes) {
  if (!b || !bytes) return fail(STEEL_BAD_ARGUMENT, "null argument");
  size_t t = 0;
  for (int k = 0; k < K_COUNT; k++) t += b->total[k];
  *bytes = t + sizeof(WorldDev) * b->n;
  return STEEL_OK;
}

}  // extern "C"
