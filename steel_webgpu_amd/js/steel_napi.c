/* Node-API addon over the C ABI of libsteel_amd.so (include/steel_amd.h): the binding the reference's JS host
 * (packages/steel-webgpu/src/sim.js) loads in place of WebGPU when it runs under Node.  Handles are externals;
 * buffers cross as typed arrays (copied by the library); every failed call throws an Error whose `code` is
 * the status name (STEEL_UNSUPPORTED, ...) and whose message is steel_last_error().  steel_amd.js wraps it. */
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/steel_amd.h"

#define CALL(env, expr)                                              \
  do {                                                               \
    if ((expr) != napi_ok) {                                         \
      napi_throw_error((env), "STEEL_NAPI", "Node-API call failed"); \
      return NULL;                                                   \
    }                                                                \
  } while (0)

static const char* status_name(int rc) {
  switch (rc) {
    case STEEL_UNSUPPORTED: return "STEEL_UNSUPPORTED";
    case STEEL_BAD_ARGUMENT: return "STEEL_BAD_ARGUMENT";
    case STEEL_OUT_OF_MEMORY: return "STEEL_OUT_OF_MEMORY";
    default: return "STEEL_ERROR";
  }
}

/* Throws for a failed status; returns 1 when it threw. */
static int check(napi_env env, int rc) {
  if (rc == STEEL_OK) return 0;
  napi_throw_error(env, status_name(rc), steel_last_error());
  return 1;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok || argc < want) {
    napi_throw_type_error(env, "STEEL_BAD_ARGUMENT", "missing arguments");
    return 1;
  }
  return 0;
}

static int get_ptr(napi_env env, napi_value v, void** out) {
  if (napi_get_value_external(env, v, out) != napi_ok || !*out) {
    napi_throw_type_error(env, "STEEL_BAD_ARGUMENT", "expected a handle");
    return 1;
  }
  return 0;
}

static int get_u32(napi_env env, napi_value v, uint32_t* out) {
  if (napi_get_value_uint32(env, v, out) != napi_ok) {
    napi_throw_type_error(env, "STEEL_BAD_ARGUMENT", "expected an unsigned integer");
    return 1;
  }
  return 0;
}

/* Data pointer and byte length of a typed array or ArrayBuffer view. */
static int get_bytes(napi_env env, napi_value v, void** data, size_t* len) {
  bool is_ta = false;
  if (napi_is_typedarray(env, v, &is_ta) != napi_ok || !is_ta) {
    napi_throw_type_error(env, "STEEL_BAD_ARGUMENT", "expected a typed array");
    return 1;
  }
  napi_typedarray_type t;
  size_t n = 0, off = 0;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &t, &n, data, &ab, &off) != napi_ok) return 1;
  size_t el = 1;
  switch (t) {
    case napi_int16_array: case napi_uint16_array: el = 2; break;
    case napi_int32_array: case napi_uint32_array: case napi_float32_array: el = 4; break;
    case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: el = 8; break;
    default: el = 1;
  }
  *len = n * el;
  return 0;
}

static napi_value make_u32(napi_env env, uint32_t x) {
  napi_value v;
  napi_create_uint32(env, x, &v);
  return v;
}

static napi_value last_error(napi_env env, napi_callback_info info) {
  napi_value v;
  CALL(env, napi_create_string_utf8(env, steel_last_error(), NAPI_AUTO_LENGTH, &v));
  return v;
}

static napi_value device_open(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (get_args(env, info, 1, argv)) return NULL;
  int32_t ordinal = 0;
  CALL(env, napi_get_value_int32(env, argv[0], &ordinal));
  steel_device* dev = NULL;
  if (check(env, steel_device_open(ordinal, &dev))) return NULL;
  napi_value v;
  CALL(env, napi_create_external(env, dev, NULL, NULL, &v));
  return v;
}

static napi_value device_close(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  void* dev;
  if (get_args(env, info, 1, argv) || get_ptr(env, argv[0], &dev)) return NULL;
  if (check(env, steel_device_close((steel_device*)dev))) return NULL;
  return NULL;
}

static napi_value device_info(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  void* dev;
  if (get_args(env, info, 1, argv) || get_ptr(env, argv[0], &dev)) return NULL;
  char name[256];
  size_t mem = 0;
  uint32_t cus = 0;
  if (check(env, steel_device_info((steel_device*)dev, name, sizeof name, &mem, &cus))) return NULL;
  napi_value o, s, m;
  CALL(env, napi_create_object(env, &o));
  CALL(env, napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s));
  CALL(env, napi_create_double(env, (double)mem, &m));
  CALL(env, napi_set_named_property(env, o, "name", s));
  CALL(env, napi_set_named_property(env, o, "totalMem", m));
  CALL(env, napi_set_named_property(env, o, "cus", make_u32(env, cus)));
  return o;
}

/* batchCreate(device, [{header, statF, statU, poses0, vel0, art0}, ...]) -> handle */
static napi_value batch_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  void* dev;
  if (get_args(env, info, 2, argv) || get_ptr(env, argv[0], &dev)) return NULL;
  uint32_t n = 0;
  bool is_arr = false;
  CALL(env, napi_is_array(env, argv[1], &is_arr));
  if (!is_arr) { napi_throw_type_error(env, "STEEL_BAD_ARGUMENT", "expected an array of worlds"); return NULL; }
  CALL(env, napi_get_array_length(env, argv[1], &n));
  if (n == 0 || n > (1u << 20)) { napi_throw_range_error(env, "STEEL_BAD_ARGUMENT", "bad world count"); return NULL; }
  steel_world_desc* d = (steel_world_desc*)calloc(n, sizeof(steel_world_desc));
  if (!d) { napi_throw_error(env, "STEEL_OUT_OF_MEMORY", "host allocation failed"); return NULL; }
  static const char* names[6] = {"header", "statF", "statU", "poses0", "vel0", "art0"};
  for (uint32_t i = 0; i < n; i++) {
    napi_value w;
    if (napi_get_element(env, argv[1], i, &w) != napi_ok) { free(d); return NULL; }
    const void* p[6];
    for (int k = 0; k < 6; k++) {
      napi_value f;
      void* data = NULL;
      size_t len = 0;
      if (napi_get_named_property(env, w, names[k], &f) != napi_ok || get_bytes(env, f, &data, &len)) { free(d); return NULL; }
      if (k == 0 && len < STEEL_WORLD_HEADER_WORDS * 4) {
        free(d);
        napi_throw_range_error(env, "STEEL_BAD_ARGUMENT", "header shorter than STEEL_WORLD_HEADER_WORDS words");
        return NULL;
      }
      p[k] = data;
    }
    d[i].header = (const uint32_t*)p[0]; d[i].statF = (const float*)p[1]; d[i].statU = (const uint32_t*)p[2];
    d[i].poses0 = (const float*)p[3]; d[i].vel0 = (const float*)p[4]; d[i].art0 = (const float*)p[5];
  }
  steel_batch* b = NULL;
  int rc = steel_batch_create((steel_device*)dev, n, d, &b);
  free(d);
  if (check(env, rc)) return NULL;
  napi_value v;
  CALL(env, napi_create_external(env, b, NULL, NULL, &v));
  return v;
}

static napi_value batch_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  void* b;
  if (get_args(env, info, 1, argv) || get_ptr(env, argv[0], &b)) return NULL;
  check(env, steel_batch_destroy((steel_batch*)b));
  return NULL;
}

static napi_value batch_reset(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  void* b;
  int32_t world = -1;
  if (get_args(env, info, 2, argv) || get_ptr(env, argv[0], &b)) return NULL;
  CALL(env, napi_get_value_int32(env, argv[1], &world));
  check(env, steel_batch_reset((steel_batch*)b, world));
  return NULL;
}

static napi_value batch_step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  void* b;
  uint32_t frames;
  if (get_args(env, info, 2, argv) || get_ptr(env, argv[0], &b) || get_u32(env, argv[1], &frames)) return NULL;
  check(env, steel_batch_step((steel_batch*)b, frames));
  return NULL;
}

static napi_value batch_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  void* b;
  if (get_args(env, info, 1, argv) || get_ptr(env, argv[0], &b)) return NULL;
  check(env, steel_batch_sync((steel_batch*)b));
  return NULL;
}

/* batchRead(batch, world, buf, byteOffset, dst) fills dst (dst.byteLength bytes). */
static napi_value batch_read(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  void *b, *dst;
  uint32_t world, buf;
  double off;
  size_t len;
  if (get_args(env, info, 5, argv) || get_ptr(env, argv[0], &b) || get_u32(env, argv[1], &world) || get_u32(env, argv[2], &buf)) return NULL;
  CALL(env, napi_get_value_double(env, argv[3], &off));
  if (off < 0) { napi_throw_range_error(env, "STEEL_BAD_ARGUMENT", "negative offset"); return NULL; }
  if (get_bytes(env, argv[4], &dst, &len)) return NULL;
  check(env, steel_batch_read((steel_batch*)b, world, buf, (size_t)off, len, dst));
  return NULL;
}

/* batchWrite(batch, world, buf, byteOffset, src) */
static napi_value batch_write(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  void *b, *src;
  uint32_t world, buf;
  double off;
  size_t len;
  if (get_args(env, info, 5, argv) || get_ptr(env, argv[0], &b) || get_u32(env, argv[1], &world) || get_u32(env, argv[2], &buf)) return NULL;
  CALL(env, napi_get_value_double(env, argv[3], &off));
  if (off < 0) { napi_throw_range_error(env, "STEEL_BAD_ARGUMENT", "negative offset"); return NULL; }
  if (get_bytes(env, argv[4], &src, &len)) return NULL;
  check(env, steel_batch_write((steel_batch*)b, world, buf, (size_t)off, len, src));
  return NULL;
}

static napi_value batch_buffer_size(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  void* b;
  uint32_t world, buf;
  if (get_args(env, info, 3, argv) || get_ptr(env, argv[0], &b) || get_u32(env, argv[1], &world) || get_u32(env, argv[2], &buf)) return NULL;
  size_t bytes = 0;
  if (check(env, steel_batch_buffer_size((steel_batch*)b, world, buf, &bytes))) return NULL;
  napi_value v;
  CALL(env, napi_create_double(env, (double)bytes, &v));
  return v;
}

/* batchStats(batch, world, out: Uint32Array(16)) -> gpu milliseconds of the last step */
static napi_value batch_stats(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  void *b, *out;
  uint32_t world;
  size_t len;
  if (get_args(env, info, 3, argv) || get_ptr(env, argv[0], &b) || get_u32(env, argv[1], &world) || get_bytes(env, argv[2], &out, &len)) return NULL;
  if (len < 64) { napi_throw_range_error(env, "STEEL_BAD_ARGUMENT", "stats need 16 words"); return NULL; }
  double ms = 0;
  if (check(env, steel_batch_stats((steel_batch*)b, world, (uint32_t*)out, &ms))) return NULL;
  napi_value v;
  CALL(env, napi_create_double(env, ms, &v));
  return v;
}

static napi_value init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback fn; } fns[] = {
      {"lastError", last_error}, {"deviceOpen", device_open}, {"deviceClose", device_close}, {"deviceInfo", device_info},
      {"batchCreate", batch_create}, {"batchDestroy", batch_destroy}, {"batchReset", batch_reset}, {"batchStep", batch_step},
      {"batchSync", batch_sync}, {"batchRead", batch_read}, {"batchWrite", batch_write}, {"batchBufferSize", batch_buffer_size},
      {"batchStats", batch_stats},
  };
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
    napi_value f;
    if (napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f) != napi_ok) return NULL;
    if (napi_set_named_property(env, exports, fns[i].name, f) != napi_ok) return NULL;
  }
  napi_value c;
  napi_create_object(env, &c);
  napi_set_named_property(env, c, "OK", make_u32(env, 0));
  napi_set_named_property(env, c, "HEADER_WORDS", make_u32(env, STEEL_WORLD_HEADER_WORDS));
  napi_set_named_property(env, exports, "constants", c);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
