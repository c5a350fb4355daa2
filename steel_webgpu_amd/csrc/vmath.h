// Device vector/quaternion math for the Steel step kernels (gfx950).
// Operation order follows the reference's WGSL helpers (packages/steel-webgpu/src/wgsl.js:47-66,
// artwgsl.js:58-95) term by term; the kernels are compiled with -ffp-contract=off so every float op
// rounds once, as written.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HD __host__ __device__ __forceinline__

struct f3 { float x, y, z; };
struct f2 { float x, y; };
struct q4 { float x, y, z, w; };

HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
HD f2 mk2(float x, float y) { f2 r; r.x = x; r.y = y; return r; }
HD q4 mk4(float x, float y, float z, float w) { q4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
HD q4 mk4(f3 a, float w) { return mk4(a.x, a.y, a.z, w); }
HD f3 xyz(q4 a) { return mk3(a.x, a.y, a.z); }
HD f3 xyz(float4 a) { return mk3(a.x, a.y, a.z); }
HD q4 Q(float4 a) { return mk4(a.x, a.y, a.z, a.w); }
HD float4 F4(q4 a) { return make_float4(a.x, a.y, a.z, a.w); }
HD float4 F4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
HD f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
HD f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
HD f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
HD f3& operator+=(f3& a, f3 b) { a = a + b; return a; }
HD f3& operator-=(f3& a, f3 b) { a = a - b; return a; }
HD q4 operator+(q4 a, q4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
HD q4 operator*(q4 a, float s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }
HD q4 operator*(float s, q4 a) { return mk4(s * a.x, s * a.y, s * a.z, s * a.w); }
HD f2 operator+(f2 a, f2 b) { return mk2(a.x + b.x, a.y + b.y); }
HD f2 operator-(f2 a, f2 b) { return mk2(a.x - b.x, a.y - b.y); }
HD f2 operator*(f2 a, float s) { return mk2(a.x * s, a.y * s); }
HD f2 operator*(float s, f2 a) { return mk2(s * a.x, s * a.y); }

HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HD float dot(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
HD float dot4(q4 a, q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
HD f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
HD float len(f3 a) { return sqrtf(dot(a, a)); }
HD float len(f2 a) { return sqrtf(dot(a, a)); }
// normalize(v) = v * (1 / sqrt(dot(v, v)))
HD f3 nrm(f3 a) { float il = 1.0f / sqrtf(dot(a, a)); return a * il; }
HD f2 nrm(f2 a) { float il = 1.0f / sqrtf(dot(a, a)); return a * il; }
HD q4 nrm4(q4 a) { float il = 1.0f / sqrtf(dot4(a, a)); return a * il; }
HD float rsq(float x) { return 1.0f / sqrtf(x); }
HD f3 vabs(f3 a) { return mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
// min/max/clamp with the selection rule the oracle uses (b if b is strictly smaller/larger).
HD float mn(float a, float b) { return b < a ? b : a; }
HD float mx(float a, float b) { return a < b ? b : a; }
HD float clampf_(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
HD bool vne0(f3 a) { return a.x != 0.0f || a.y != 0.0f || a.z != 0.0f; }
HD float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
HD f3 with_comp(f3 v, int k, float x) { if (k == 0) v.x = x; else if (k == 1) v.y = x; else v.z = x; return v; }
HD void setc(f3& v, int k, float s) { if (k == 0) v.x = s; else if (k == 1) v.y = s; else v.z = s; }

HD f3 qrot(q4 q, f3 v) {
  f3 u = xyz(q); float s = q.w;
  return (2.0f * dot(u, v)) * u + (s * s - dot(u, u)) * v + (2.0f * s) * cross(u, v);
}
HD q4 qconj(q4 q) { return mk4(-q.x, -q.y, -q.z, q.w); }
HD f3 qinvrot(q4 q, f3 v) { return qrot(qconj(q), v); }
HD q4 qmul(q4 a, q4 b) {
  f3 v = a.w * xyz(b) + b.w * xyz(a) + cross(xyz(a), xyz(b));
  return mk4(v.x, v.y, v.z, a.w * b.w - dot(xyz(a), xyz(b)));
}

struct m3 { f3 c[3]; };  // column major
HD f3 mul(const m3& m, f3 v) { return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z; }
HD m3 tr(const m3& m) {
  m3 r;
  r.c[0] = mk3(m.c[0].x, m.c[1].x, m.c[2].x);
  r.c[1] = mk3(m.c[0].y, m.c[1].y, m.c[2].y);
  r.c[2] = mk3(m.c[0].z, m.c[1].z, m.c[2].z);
  return r;
}
HD m3 mul(const m3& a, const m3& b) { m3 r; for (int j = 0; j < 3; j++) r.c[j] = mul(a, b.c[j]); return r; }
HD m3 madd(const m3& a, const m3& b) { m3 r; for (int j = 0; j < 3; j++) r.c[j] = a.c[j] + b.c[j]; return r; }
HD m3 msub(const m3& a, const m3& b) { m3 r; for (int j = 0; j < 3; j++) r.c[j] = a.c[j] - b.c[j]; return r; }
HD m3 mscale(const m3& a, float s) { m3 r; for (int j = 0; j < 3; j++) r.c[j] = a.c[j] * s; return r; }
HD m3 mscale(float s, const m3& a) { m3 r; for (int j = 0; j < 3; j++) r.c[j] = s * a.c[j]; return r; }
HD m3 qmat(q4 q) { m3 m; m.c[0] = qrot(q, mk3(1, 0, 0)); m.c[1] = qrot(q, mk3(0, 1, 0)); m.c[2] = qrot(q, mk3(0, 0, 1)); return m; }
HD m3 skew3(f3 r) { m3 m; m.c[0] = mk3(0, r.z, -r.y); m.c[1] = mk3(-r.z, 0, r.x); m.c[2] = mk3(r.y, -r.x, 0); return m; }
HD m3 outer3(f3 x, f3 y) { m3 m; m.c[0] = x * y.x; m.c[1] = x * y.y; m.c[2] = x * y.z; return m; }
HD m3 m3zero() { m3 m; m.c[0] = m.c[1] = m.c[2] = mk3(0, 0, 0); return m; }

// Cephes (S. L. Moshier) single-precision sin/cos/atan, written out so every compiler rounds them the
// same: the kernels and the CPU oracle share the result bit for bit.
HD void sincos_reduce(float xx, float& x, int& j) {
  const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f, FOPI = 1.27323954473516f;
  j = (int)(FOPI * xx);
  float y = (float)j;
  if (j & 1) { j += 1; y += 1.0f; }
  x = ((xx - y * DP1) - y * DP2) - y * DP3;
}
HD float sin_poly(float x) { float z = x * x; return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x; }
HD float cos_poly(float x) { float z = x * x; return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f; }
HD float sin_c(float xx) {
  float sign = 1.0f;
  if (xx < 0.0f) { sign = -1.0f; xx = -xx; }
  float x; int j; sincos_reduce(xx, x, j);
  j &= 7;
  if (j > 3) { sign = -sign; j -= 4; }
  float y = (j == 1 || j == 2) ? cos_poly(x) : sin_poly(x);
  return sign < 0.0f ? -y : y;
}
HD float cos_c(float xx) {
  if (xx < 0.0f) xx = -xx;
  float x; int j; sincos_reduce(xx, x, j);
  j &= 7;
  float sign = 1.0f;
  if (j > 3) { j -= 4; sign = -sign; }
  if (j > 1) sign = -sign;
  float y = (j == 1 || j == 2) ? sin_poly(x) : cos_poly(x);
  return sign < 0.0f ? -y : y;
}
HD float atan_c(float xx) {
  float sign = 1.0f, x = xx;
  if (x < 0.0f) { sign = -1.0f; x = -x; }
  float y;
  if (x > 2.414213562373095f) { y = 1.5707963267948966f; x = -(1.0f / x); }
  else if (x > 0.4142135623730950f) { y = 0.7853981633974483f; x = (x - 1.0f) / (x + 1.0f); }
  else y = 0.0f;
  float z = x * x;
  y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
  return sign < 0.0f ? -y : y;
}
HD float atan2_c(float y, float x) {
  if (x == 0.0f) { if (y > 0.0f) return 1.5707963267948966f; if (y < 0.0f) return -1.5707963267948966f; return 0.0f; }
  float z = atan_c(y / x);
  if (x < 0.0f) z = (y < 0.0f) ? z - 3.14159265358979f : z + 3.14159265358979f;
  return z;
}
HD q4 qaxis(f3 axis, float angle) { float s = sin_c(0.5f * angle); return mk4(axis * s, cos_c(0.5f * angle)); }

HD uint32_t hash_u32(uint32_t x) {  // lowbias32 (wgsl.js:61-65)
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}
HD uint32_t fbits(float f) { return __float_as_uint(f); }
HD float ubits(uint32_t u) { return __uint_as_float(u); }
HD uint32_t fflip_u(uint32_t i) { return i ^ (((i >> 31) != 0 ? 0xffffffffu : 0u) | 0x80000000u); }
HD uint32_t tu(float f) { return f <= 0.0f ? 0u : (uint32_t)f; }
HD int32_t ti(float f) { return (int32_t)f; }
