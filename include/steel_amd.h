/* C ABI of the MI355X Steel step library (libsteel_amd.so).
 *
 * The drop-in boundary for the reference's step path: the JS `Sim` (packages/steel-webgpu/src/sim.js:63-835,
 * typed in src/sim.d.ts:53-92) keeps its API and scene packing and calls these functions where it used
 * WebGPU.  Plain pointers and sizes only; every call returns 0 (STEEL_OK) or a negative status, with the
 * message in steel_last_error().  The library owns all device memory: create copies the caller's arrays,
 * read copies into caller memory, destroy frees the world's memory but never the device.
 *
 * A "batch" is N independent worlds stepped together (one launch per kernel for all of them); a plain
 * per-world Sim is a batch of one.
 */
#ifndef STEEL_AMD_H
#define STEEL_AMD_H

#include <stddef.h>
#include <stdint.h>
#include "steel_world.h"

#ifdef __cplusplus
extern "C" {
#endif

#define STEEL_OK 0
#define STEEL_ERROR -1
#define STEEL_UNSUPPORTED -2   /* no usable GPU; the JS side maps it to UnsupportedError (src/device.js:1-9) */
#define STEEL_BAD_ARGUMENT -3
#define STEEL_OUT_OF_MEMORY -4

/* Buffer ids for steel_batch_read / steel_batch_write (names of sim.js:503-519). */
enum steel_buffer {
  STEEL_BUF_POSES = 0,     /* float4[W_NPOSES]: body COM pose, shape poses, AABBs, previous poses */
  STEEL_BUF_VEL = 1,       /* float4[2*NB] */
  STEEL_BUF_ART = 2,       /* float[W_NART] */
  STEEL_BUF_ANCHOR_KEYS = 3,  /* uint32[2*AH] */
  STEEL_BUF_ANCHOR_VALS = 4,  /* uint32[2*AH*STEEL_AVN] */
  STEEL_BUF_PAIRS = 5,     /* uint32[2*MAXP], the last substep's broadphase pairs (unordered) */
  STEEL_BUF_CONTACTS = 6,  /* float4[6*MAXC], the last substep's contacts in canonical order */
  STEEL_BUF_ROWS = 7,      /* float4[MAXC*ROWS] */
  STEEL_BUF_SCHED = 8,     /* uint32 schedule tables */
  STEEL_BUF_COUNTERS = 9,  /* uint32[8]: pairs, contacts, substep count, control flags (bit 0: the user set joint
                              targets, so a recorded target path no longer drives them), then the frame clock in
                              seconds as a double (words 4-5, the reference's Sim.phase), 2 reserved */
  STEEL_BUF_STATS = 10,    /* uint32[16]: the reference's stats words 0-8 (src/wgsl.js:3879-3890); 9-11 the last
                              substep's candidate pairs, rows and units; 12-15 sums since reset of pairs, kept
                              contacts, rows and the substep count */
  STEEL_BUF_WARM = 11,     /* uint32[10*W_WH] contact warm start (two parities): keys [2*W_WH], then per slot the
                              normal impulse and the tangent impulse as a world vector (float bits) [2*W_WH*4] */
  STEEL_BUF_COUNT = 12
};

/* Stage mask for steel_batch_substep_stages (stage-isolated parity runs). */
#define STEEL_STAGE_COLLIDE 1u
#define STEEL_STAGE_NARROW 2u
#define STEEL_STAGE_PREPARE 4u
#define STEEL_STAGE_SOLVE 8u
#define STEEL_STAGE_ALL 15u

typedef struct steel_device steel_device;
typedef struct steel_batch steel_batch;

/* One packed world (include/steel_world.h): the arrays sim.js:95-611 builds, plus the initial state. */
typedef struct steel_world_desc {
  const uint32_t* header;  /* STEEL_WORLD_HEADER_WORDS words */
  const float* statF;      /* 4 * header[W_NSTATF] floats */
  const uint32_t* statU;   /* header[W_NSTATU] words (at least 1) */
  const float* poses0;     /* 4 * header[W_NPOSES] floats */
  const float* vel0;       /* 8 * header[W_NB] floats */
  const float* art0;       /* header[W_NART] floats */
} steel_world_desc;

/* Thread-local message of the last failed call. */
const char* steel_last_error(void);

/* Device acquisition (src/device.js:1-27 initDevice). Returns STEEL_UNSUPPORTED without a GPU. */
int steel_device_open(int ordinal, steel_device** out);
int steel_device_close(steel_device* dev);
/* Name and memory of the device, for logs. */
int steel_device_info(steel_device* dev, char* name, size_t name_len, size_t* total_mem, uint32_t* num_cus);

/* Batch of worlds: new Sim(...) for each (sim.js:64-611); arrays are copied. */
int steel_batch_create(steel_device* dev, uint32_t nworlds, const steel_world_desc* worlds, steel_batch** out);
int steel_batch_destroy(steel_batch* b);
/* Sim.reset (sim.js:613-623): initial poses/vel/art, zero anchors, counters and stats. world < 0: all. */
int steel_batch_reset(steel_batch* b, int32_t world);
/* Sim.step(frames) (sim.js:692-701): frames x substeps x (collide, prepare, solve) on the batch stream,
 * asynchronous.  With graphs enabled the frame is a captured hipGraph replayed per frame. */
int steel_batch_step(steel_batch* b, uint32_t frames);
/* One substep with only the given stages (STEEL_STAGE_*), for stage-isolated parity checks. */
int steel_batch_substep_stages(steel_batch* b, uint32_t stages);
/* steel_batch_step one frame at a time with a HIP event after each; frame_ms[f] is frame f's GPU time. */
int steel_batch_step_timed(steel_batch* b, uint32_t frames, float* frame_ms);
/* Average GPU time per launch of k_collide, k_prepare and k_solve over `substeps` substeps launched one
 * kernel at a time with HIP events between them on the batch's stream (the state advances). */
int steel_batch_profile_kernels(steel_batch* b, uint32_t substeps, double kernel_ms[3]);
/* Wait for the batch's stream. */
int steel_batch_sync(steel_batch* b);
/* Sim.read / queue.writeBuffer on one world's buffer (sim.js:716-802); read synchronises the stream. */
int steel_batch_read(steel_batch* b, uint32_t world, uint32_t buf, size_t byte_off, size_t bytes, void* dst);
int steel_batch_write(steel_batch* b, uint32_t world, uint32_t buf, size_t byte_off, size_t bytes, const void* src);
/* Size in bytes of one world's buffer `buf` (worlds of a batch may differ in size). */
int steel_batch_buffer_size(steel_batch* b, uint32_t world, uint32_t buf, size_t* bytes);
/* Sim.stats (sim.js:822-834): the 16 stats words of one world and the GPU time of the last step. */
int steel_batch_stats(steel_batch* b, uint32_t world, uint32_t out[16], double* gpu_ms);
/* Launch options: workgroup size of prepare/solve (64, 128 or 256; 1024 for batches of large worlds, whose solve then
 * keeps every world's link mirror behind generic pointers) and hipGraph capture on/off. */
int steel_batch_configure(steel_batch* b, uint32_t block, int use_graph);
/* Device-wide stepping of one large world (Sim.step, sim.js:692-701, for a world past the reference's limits of
 * 256 articulations, 16-bit body ids and one prepare/solve workgroup, sim.js:159, :663): every stage but the
 * narrowphase runs as one cooperative launch over the whole device instead of one workgroup.  The batch must
 * hold one world whose solve keeps its link mirror in global memory, without sleeping, contact warm start,
 * Nesterov momentum or velocity iterations (STEEL_UNSUPPORTED otherwise); results are those of the one-workgroup
 * step, bit for bit.  enable 0 returns to one workgroup per stage.  Frames are not captured as graphs. */
int steel_batch_set_wide(steel_batch* b, int enable);

/* The launch schedule the autotuner picked for a batch. */
typedef struct steel_tune_result {
  uint32_t block;          /* prepare/solve workgroup size: 64, 128 or 256 */
  uint32_t narrow_per_cu;  /* k_narrow workgroups (4 waves each) per compute unit: 2, 4 or 8 */
  uint32_t from_cache;     /* 1: the device's cache held this batch's signature, nothing was measured */
  uint32_t trials;         /* schedules measured */
  float frame_ms;          /* the chosen schedule's mean frame time over the trial (0 when cached) */
} steel_tune_result;
#define STEEL_TUNE_REMEASURE 1u  /* measure even when the device's cache holds the batch's signature */
/* Autotuner (no counterpart in the reference, whose dispatch sizes are fixed in sim.js:648-690): times each
 * schedule (workgroup size x narrowphase grid) over `frames` frames from the batch's current state, restoring
 * the state after every trial, and keeps the fastest; the batch's state is unchanged on return (the step's
 * results do not depend on the schedule).  The choice is cached per device under the batch's signature
 * (world count, body/shape/link totals and maxima, substeps); with STEEL_TUNE_CACHE naming a file, the cache
 * is also read from and appended to that file, so later processes on the same device model skip the trials. */
int steel_batch_autotune(steel_batch* b, uint32_t frames, uint32_t flags, steel_tune_result* out);
/* Device bytes the batch holds (admission by HBM budget). */
int steel_batch_device_bytes(steel_batch* b, size_t* bytes);
/* Device bytes a batch of these worlds would hold, without creating it. */
int steel_estimate_bytes(uint32_t nworlds, const steel_world_desc* worlds, size_t* bytes);

/* ---- Serving (no counterpart in the reference, which steps one Sim per device; this is the request loop's
 * device side).  A server owns one device pool of `budget_bytes` (the HBM budget).  Worlds are admitted one
 * at a time into a lane (a size class with its own stream); a lane frame steps every world the lane holds at
 * launch by one Sim.step(1) frame, so worlds join and leave between frames while other lanes keep running.
 * Each admitted world is stepped exactly as it would be alone (same kernels, no cross-world reads). */
typedef struct steel_server steel_server;
int steel_server_create(steel_device* dev, size_t budget_bytes, uint32_t nlanes, steel_server** out);
int steel_server_destroy(steel_server* s);
/* Workgroup size (64, 128, 256 or 1024) of a lane's stage kernels. */
int steel_server_configure_lane(steel_server* s, uint32_t lane, uint32_t block);
/* Pool bytes one world takes. */
int steel_server_world_bytes(const steel_world_desc* w, size_t* bytes);
/* Copy the world into the pool and add it to `lane` from the lane's next frame on.  STEEL_OUT_OF_MEMORY, with
 * nothing allocated, when the pool cannot hold it now (the caller keeps it queued and retries). */
int steel_server_admit(steel_server* s, const steel_world_desc* w, uint32_t lane, uint32_t* id);
/* Remove a world from its lane; its pool block is reused once the lane's in-flight frame has finished. */
int steel_server_release(steel_server* s, uint32_t id);
/* Launch the lane's next frame over its current worlds (asynchronous; waits for the lane's previous frame). */
int steel_server_step_lane(steel_server* s, uint32_t lane);
/* done = 1 once the lane's last launched frame has finished; nworlds = worlds it stepped; gpu_ms = its GPU time. */
int steel_server_lane_poll(steel_server* s, uint32_t lane, int* done, uint32_t* nworlds, float* gpu_ms);
/* Sim.read of an admitted world (waits for its lane's in-flight frame). */
int steel_server_read(steel_server* s, uint32_t id, uint32_t buf, size_t byte_off, size_t bytes, void* dst);
/* Sim's queue.writeBuffer for an admitted world (setJointTargets, setBodyWrench, setBodyPose/Twist, the control
   flags word): waits for its lane's in-flight frame, then writes, so the next lane frame sees the input. */
int steel_server_write(steel_server* s, uint32_t id, uint32_t buf, size_t byte_off, size_t bytes, const void* src);
/* Pool bytes in use, pool capacity, admitted worlds. */
int steel_server_info(steel_server* s, size_t* used, size_t* capacity, uint32_t* nworlds);

/* ---- Step without allocation.  No counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * caller holds one world's static arrays and state in device memory (device pointers; each array has the size
 * steel_batch_buffer_size reports for that world, header/statF/statU the sizes of steel_world_desc).  The step
 * scratch is a fixed array of STEEL_NOALLOC_SCRATCH_BYTES per GPU inside the library, so a world whose scratch
 * exceeds it is refused with STEEL_BAD_ARGUMENT.  Calls on one GPU share that array: calls that overlap on one
 * GPU (other threads or streams) race.  The steps run on `stream` (a hipStream_t; null: the default stream),
 * asynchronously, with the kernels and results of steel_batch_step on a batch of this one world. */
#define STEEL_NOALLOC_SCRATCH_BYTES (16u << 20)
typedef struct steel_world_device_arrays {
  const uint32_t* header;  /* device copies of the static arrays of steel_world_desc */
  const float* statF;
  const uint32_t* statU;
  float* poses;            /* state, as the STEEL_BUF_* buffers of the same name */
  float* vel;
  float* art;
  uint32_t* anchor_keys;
  uint32_t* anchor_vals;
  uint32_t* warm;
  uint32_t* counters;      /* zero at reset, like stats, anchors and warm */
  uint32_t* stats;
} steel_world_device_arrays;
/* `host` gives the world's header and statF in host memory (read for sizing only; the other pointers of
 * `host` are not read). */
int steel_world_step_noalloc(steel_device* dev, const steel_world_desc* host, const steel_world_device_arrays* arrays,
                             uint32_t frames, void* stream);

#ifdef __cplusplus
}
#endif
#endif
